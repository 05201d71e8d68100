"""CPU-side checks of the product: the C ABI library loads and exports every
declared entry point, the host topology (pattern, edges, slot maps) is
bit-identical to the oracle's restatement of the reference, and the Python
surface raises the reference's exception types.  No GPU compute is called."""

import ctypes
import os
import re

import numpy as np
import pytest

import chysx
from chysx import scenes
from oracle.oracle import OracleCloth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chysx_cloth.h")
LIB = os.path.join(ROOT, "ai-phyiscs_amd", "chysx", "libchysx_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chysx_[a-z0-9_]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.chysx_version.restype = ctypes.c_char_p
    assert lib.chysx_version().decode() == chysx.__version__ == "0.1.0"


def test_c_abi_error_codes_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.chysx_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    assert lib.chysx_cloth_create(ctypes.byref(h)) == 0
    # step(dt <= 0) with buffers set -> invalid_argument (cloth_simulator.cu:868-871)
    assert lib.chysx_cloth_set_external_buffers(h, ctypes.c_size_t(16), ctypes.c_size_t(16), 4,
                                                ctypes.c_size_t(0), ctypes.c_size_t(0)) == 0
    assert lib.chysx_cloth_step(h, ctypes.c_float(-1.0), ctypes.c_size_t(0)) == 1
    assert b"dt must be positive" in lib.chysx_last_error()
    # null handle -> invalid_argument
    assert lib.chysx_cloth_step(None, ctypes.c_float(0.1), ctypes.c_size_t(0)) == 1
    assert lib.chysx_cloth_destroy(h) == 0
    # a batch's statistics for all cloths: the array length must be the cloth count
    b = ctypes.c_void_p()
    assert lib.chysx_batch_create(ctypes.byref(b)) == 0
    n = ctypes.c_int(-1)
    assert lib.chysx_batch_num_cloths(b, ctypes.byref(n)) == 0 and n.value == 0
    iters = (ctypes.c_int * 1)()
    assert lib.chysx_batch_stats_all(b, 1, iters, None, None, None) == 1
    assert b"number of cloths" in lib.chysx_last_error()
    assert lib.chysx_batch_stats_all(b, 0, None, None, None, None) == 0
    assert lib.chysx_batch_destroy(b) == 0


def oracle_topology(scene, springs_active):
    oc = OracleCloth(64)
    oc.set_mesh(scene.tris, scene.n)
    p = scene.pos.astype(np.float64)
    oc.build_springs(p, 1.0)
    oc.build_stretch(p, 1.0)
    oc.set_pins(scene.pin_idx, scene.pin_targets, scene.pin_k)
    oc.set_params(springs_active=springs_active)
    oc.build_topology()
    ro, co = oc.pattern()
    e = oc.edges()
    ts, ss, ps = oc.slots(scene.tris.shape[0], e.shape[0] if springs_active else 0, scene.pin_idx.size)
    return ro, co, e, ts, ss, ps


@pytest.mark.parametrize("name", ["cfg1", "cfg2_64", "cfg3_33", "random"])
@pytest.mark.parametrize("springs", [False, True])
def test_topology_bit_identical_to_oracle(name, springs):
    if name == "cfg1":
        s = scenes.config1()
    elif name == "cfg2_64":
        s = scenes.config2(w=64, steps=1)
    elif name == "cfg3_33":
        s = scenes.config3(w=33, steps=1)
    else:
        rng = np.random.default_rng(3)
        s = scenes.config1()
        s.tris = rng.integers(0, s.n, size=(150, 3)).astype(np.int32)
    d = chysx.topology_query(s.n, s.tris, springs, s.pin_idx)
    ro, co, e, ts, ss, ps = oracle_topology(s, springs)
    np.testing.assert_array_equal(d["row_offsets"], ro)
    np.testing.assert_array_equal(d["col_indices"], co)
    np.testing.assert_array_equal(d["edges"], e)
    np.testing.assert_array_equal(d["tri_slots"], ts)
    np.testing.assert_array_equal(d["spring_slots"], ss)
    np.testing.assert_array_equal(d["pin_slots"], ps)


def test_topology_1024_closed_form_counts():
    w = 1024
    _, tris = scenes.grid(w)
    d = chysx.topology_query(w * w, tris, True, np.arange(w, dtype=np.int32))
    e = 2 * w * (w - 1) + (w - 1) ** 2
    assert d["edges"].shape == (e, 2) == (3141633, 2)
    assert d["row_offsets"][-1] == 2 * e == 6283266
    assert d["tri_slots"].shape == (9 * 2093058,)


def test_topology_rejects_out_of_range():
    with pytest.raises(IndexError):
        chysx.topology_query(4, np.array([[0, 1, 7]], np.int32), False, np.zeros(0, np.int32))


def test_python_surface_errors_match_reference():
    sim = chysx.ClothSimulator()
    # particle_count <= 0 -> step is a silent no-op (cloth_simulator.cu:817-820)
    sim.step(1.0 / 60.0)
    sim.set_external_buffers(0, 0, 10)
    with pytest.raises(RuntimeError):
        sim.step(1.0 / 60.0)  # pos / vel unset (cloth_simulator.cu:863-867)
    sim.set_external_buffers(64, 64, 10)
    with pytest.raises(ValueError):
        sim.step(0.0)
    with pytest.raises(ValueError):
        sim.set_pins(np.zeros((2, 2), np.int32), np.zeros((2, 3), np.float32))
    with pytest.raises(ValueError):
        sim.set_pins(np.zeros(2, np.int32), np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError):
        sim.set_mesh(np.zeros((4, 2), np.int32))
    sim.set_pins(np.array([1, 2], np.int32), np.zeros((2, 3), np.float32))
    assert sim.num_pins() == 2
    with pytest.raises(ValueError):
        sim.update_pin_targets(np.zeros((3, 3), np.float32))
    sim.clear_pins()
    assert sim.num_pins() == 0
    with pytest.raises(ValueError):
        sim.redistribute_mass_area_weighted(1.0, 0, 10)
    with pytest.raises(ValueError):
        sim.set_pcg_tolerance(-1.0)
    sim.set_pcg_iterations(0)
    assert sim.pcg_iterations() == 1  # clamped like the reference (cloth_simulator.h:597-599)
    sim.set_pcg_iterations(50)
    assert sim.pcg_iterations() == 50


def test_material_pod_and_live_view():
    m = chysx.ClothMaterial()
    assert (m.lame_mu, m.lame_lambda, m.bending) == (0.0, 0.0, 0.0)
    assert m.density == pytest.approx(0.05) and m.damping == 0.0
    assert (m.gx, m.gy) == (0.0, 0.0) and m.gz == pytest.approx(-9.81)
    sim = chysx.ClothSimulator()
    m.gy = -9.8
    m.gz = 0.0
    sim.set_material(m)
    m.gy = 1.0  # the simulator holds a copy
    assert sim.material.gy == pytest.approx(-9.8)
    view = sim.material
    view.damping = 0.25  # the property is a live reference
    assert sim.material.damping == pytest.approx(0.25)


def test_mesh_edges_and_counts_without_gpu():
    s = scenes.config1()
    sim = chysx.ClothSimulator()
    sim.set_mesh(s.tris)
    e = sim.edges()
    assert e.shape == (261, 2)
    np.testing.assert_array_equal(e, scenes.edges_of(s.tris))
    assert sim.num_springs() == 0 and sim.num_fem_stretch_triangles() == 0
    with pytest.raises(RuntimeError):
        sim.build_springs_from_current_positions(1e3)  # positions not set


@pytest.mark.parametrize("k", [2, 3, 5, 8])
def test_projection_small_solve_matches_eigen_truncated_pinv(k):
    """The per-instance projection solve (cyclic Jacobi, host twin of the
    device code) against numpy's eigh-based truncated solve."""
    from oracle import warmstart_oracle as W
    rng = np.random.default_rng(k)
    n = 50
    M = rng.normal(size=(n, n))
    A = M @ M.T + n * np.eye(n)
    V = rng.normal(size=(n, k))
    V[:, -1] = V[:, 0] * 0.5 + V[:, 1] * 1e-14  # numerically collinear column
    b = rng.normal(size=n)
    G = V.T @ A @ V
    g = V.T @ b
    c, kept = chysx._chysx_native.ws_small_solve(np.ascontiguousarray(G), g)
    x_ref, _, kept_ref = W.galerkin_projection(A, b, V)
    assert kept == kept_ref == k - 1
    np.testing.assert_allclose(V @ c, x_ref, rtol=1e-8, atol=1e-10)


def test_projection_small_solve_rejects_bad_size():
    with pytest.raises(ValueError):
        chysx._chysx_native.ws_small_solve(np.eye(9), np.ones(9))


@pytest.mark.parametrize("case", ["grid7", "grid20", "random"])
def test_topology_with_bending_bit_identical_to_oracle(case):
    from oracle.oracle import OracleCloth as OC
    if case.startswith("grid"):
        w = int(case[4:])
        pos, tris = scenes.grid(w)
        n = w * w
    else:
        rng = np.random.default_rng(11)
        n = 80
        tris = rng.integers(0, n, size=(200, 3)).astype(np.int32)
        tris[::17, 1] = tris[::17, 0]  # degenerate triangles
        pos = rng.normal(size=(n, 3)).astype(np.float32)
    d = chysx._chysx_native.topology_query_bending(n, tris)
    oc = OC(64)
    oc.set_mesh(tris, n)
    oc.build_stretch(pos, 1e3)
    oc.build_bending(pos, 1e-2)
    oc.set_params((0, -9.8, 0), 0.0, False, False, 0.0, 1)
    oc.build_topology()
    ro, co = oc.pattern()
    dih, _, _, ds = oc.bending(tris.shape[0])
    np.testing.assert_array_equal(d["row_offsets"], ro)
    np.testing.assert_array_equal(d["col_indices"], co)
    np.testing.assert_array_equal(d["dihedrals"], dih)
    np.testing.assert_array_equal(d["dihedral_slots"], ds)


def test_committed_warm_start_parameters_load():
    from chysx import warmstart
    theta, meta = warmstart.load_params()
    assert theta.dtype == np.float32 and theta.size == chysx._chysx_native.ws_num_params()
    assert np.all(np.isfinite(theta))
    assert meta["mode"] == "learned" and meta["pairs"] * meta["width"] ** 2 >= 50000 * 256 ** 2
    assert meta["loss_last"] < meta["loss_first"]
