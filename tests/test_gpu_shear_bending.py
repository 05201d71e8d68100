"""GPU parity of FEM shear and dihedral bending against the FP64 oracle, on
the reference demo's cloth (stretch 5000, shear 500, bending 0.01,
area-lumped mass, damping 2, two pinned corners) flat and crumpled.

Bars as in test_gpu_step.py: pattern and dihedral list bit-exact, rest
angles within 2e-6 rad, Hessian blocks within 2e-6 of the largest block of
their row (FP32 rounding), every solve at true relative residual <= 1e-6,
positions after one step from a shared state within 1e-5 relative.
"""

import numpy as np
import pytest

from chysx import scenes
from chysx.harness import make_sim
from oracle.oracle import oracle_from_scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("seed", [None, 5])
def test_drape_first_step_system_matches_oracle(cuda, seed):
    s = scenes.drape(w=21, seed=seed)
    g = make_sim(s, tol=1e-6, springs_active=False, force_active=False)
    oc, p, v, im, f = oracle_from_scene(s, precision=64, springs_active=False, force_active=False, tol=1e-12,
                                        max_iter=100000)
    assert g.sim.num_fem_shear_triangles() == s.tris.shape[0]
    dih, rest = g.sim.bending_dihedrals()
    odih, orest, _, _ = oc.bending(s.tris.shape[0])
    np.testing.assert_array_equal(dih, odih)
    np.testing.assert_allclose(rest, orest, atol=2e-6)
    if seed is None:
        assert np.abs(orest).max() < 1e-6
    else:
        assert np.abs(orest).max() > 0.1
    g.sim.step(s.dt)
    oc.step(p, v, im, f, s.dt, 1)
    diag, ro, co, vals, rhs, dx = g.sim.debug_dump_last_solve()
    oro, oco = oc.pattern()
    np.testing.assert_array_equal(ro, oro)
    np.testing.assert_array_equal(co, oco)
    od, ov, orhs, odx = oc.system()
    rows = np.repeat(np.arange(s.n), np.diff(ro))
    scale = np.maximum(np.abs(od).max(axis=(1, 2)), np.abs(ov).max(axis=(1, 2)).max(initial=0.0))
    assert np.all(np.abs(diag - od) <= 2e-6 * np.abs(od).max(axis=(1, 2), keepdims=True) + 1e-12)
    rowmax = np.abs(od).max(axis=(1, 2))[rows][:, None, None]
    assert np.all(np.abs(vals - ov) <= 2e-6 * rowmax + 1e-12)
    assert scale.min() > 0
    # The elastic forces cancel at rest; FP32 rounding of each element's
    # strain (eps32 ~ 6e-8) times its force per unit strain (k A |dw| ~ 100 N
    # for stretch k = 5000, A = 8e-4 m^2, |dw| ~ 1/h = 25 /m) over the 12
    # stretch + shear elements of a row bounds the difference by ~7e-5 N
    # (|m g| ~ 8e-4 N per particle).
    np.testing.assert_allclose(rhs, orhs, rtol=1e-5, atol=1e-4)
    assert g.sim.last_solve_stats()["relres"] <= 1e-6
    assert rel(g.positions(), p) <= 1e-5


def bsr_residual(diag, ro, co, vals, b, x):
    n = diag.shape[0]
    y = np.einsum("nij,nj->ni", diag.astype(np.float64), x)
    rows = np.repeat(np.arange(n), np.diff(ro))
    np.add.at(y, rows, np.einsum("kij,kj->ki", vals.astype(np.float64), x[co]))
    r = b.astype(np.float64) - y
    return float(np.sqrt((r * r).sum() / max((b.astype(np.float64) ** 2).sum(), 1e-30)))


def compare_systems(g, oc, n):
    """GPU system of the last solve vs the oracle's, both assembled at the
    same state: blocks within 2e-6 of their row's largest block, rhs within
    1e-4 N plus 1e-5 relative."""
    diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
    od, ov, orhs, _ = oc.system()
    rows = np.repeat(np.arange(n), np.diff(ro))
    rowmax = np.abs(od).max(axis=(1, 2))
    assert np.all(np.abs(diag - od) <= 2e-6 * rowmax[:, None, None] + 1e-12)
    assert np.all(np.abs(vals - ov) <= 2e-6 * rowmax[rows][:, None, None] + 1e-12)
    np.testing.assert_allclose(rhs, orhs, rtol=1e-5, atol=1e-4)
    return od, ro, co, ov, orhs


@pytest.mark.parametrize("seed", [None, 5])
def test_drape_resync(cuda, seed):
    """Every 10 steps the oracle steps from the GPU's state: the two assembled
    systems must agree (deformed states, so every element term is live), the
    GPU's dx must solve its own dumped system to 1e-6 (recomputed here in
    FP64), and positions must agree to 5e-5.  This cloth is stiff
    (k dt^2 / m ~ 2e4 per element, condition number ~ 1e5): its rows hold
    ~6e3 N/m entries that nearly cancel against a ~3e-4 N right-hand side, so
    two solves that both stop at a 1e-6 relative residual may differ by more
    than the 1e-5 kept for the softer workloads."""
    s = scenes.drape(w=33, seed=seed)
    g = make_sim(s, tol=1e-6, springs_active=False, force_active=False)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, springs_active=False, force_active=False, tol=1e-11,
                                        max_iter=100000)
    for k in range(60):
        check = (k % 10) == 0
        if check:
            p0 = g.positions().astype(np.float64)
            v0 = g.velocities().astype(np.float64)
            dx0 = g.sim.get_dx64() if k > 0 else np.zeros((s.n, 3))
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["converged"] and st["relres"] <= 1e-6, (k, st)
        if check:
            oc.set_dx(dx0)
            oc.step(p0, v0, im, f, s.dt, 1)
            compare_systems(g, oc, s.n)
            diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
            assert bsr_residual(diag, ro, co, vals, rhs, g.sim.get_dx64()) <= 1.0001e-6, k
            assert rel(g.positions(), p0) <= 5e-5, k


def test_bending_only_system_at_deformed_state(cuda):
    """Bending alone (no FEM triangles), stiff enough to dominate, on a
    crumpled rest shape and a perturbed current state: the GPU's assembled
    gradient and Gauss-Newton blocks match the oracle's."""
    from oracle.oracle import OracleCloth
    s = scenes.drape(w=12, seed=3)
    g = make_sim(s, tol=1e-8, springs_active=False, force_active=False)
    g.sim.set_mesh(s.tris)  # drops every built element set
    g.sim.build_bending_from_current_positions(2.0)
    rng = np.random.default_rng(0)
    h = 0.8 / 11
    x1 = s.pos.astype(np.float64) + rng.uniform(-0.2 * h, 0.2 * h, s.pos.shape)
    v1 = rng.normal(size=s.pos.shape) * 0.05
    g.set_state(x1, v1)
    g.sim.step(s.dt)
    oc = OracleCloth(64)
    oc.set_mesh(s.tris, s.n)
    oc.build_bending(s.pos.astype(np.float64), 2.0)
    oc.set_pins(s.pin_idx, s.pin_targets, s.pin_k)
    oc.set_params(s.gravity, s.damping, False, False, 1e-12, 100000)
    p = x1.astype(np.float32).astype(np.float64)
    v = v1.astype(np.float32).astype(np.float64)
    oc.step(p, v, s.inv_mass.astype(np.float64), None, s.dt, 1)
    dih, rest = g.sim.bending_dihedrals()
    odih, orest, _, _ = oc.bending(s.tris.shape[0])
    np.testing.assert_array_equal(dih, odih)
    np.testing.assert_allclose(rest, orest, atol=2e-6)
    od, ro, co, ov, orhs = compare_systems(g, oc, s.n)
    # the bending forces are far above the rounding floor here
    grav = (1.0 / s.inv_mass.astype(np.float64))[:, None] * np.array(s.gravity)
    assert np.abs(orhs - (1.0 / s.inv_mass.astype(np.float64))[:, None] * v / s.dt - grav).max() > 1e-2


def test_bending_only_pattern_without_triangles(cuda):
    """Bending without any FEM triangles: the pattern holds only the dihedral
    pairs (the reference adds triangle pairs only for installed FEM sets)."""
    s = scenes.drape(w=9, seed=2)
    g = make_sim(s, tol=1e-6, springs_active=False, force_active=False)
    g.sim.set_mesh(s.tris)  # drops every built element set
    g.sim.build_bending_from_current_positions(0.5)
    g.sim.step(s.dt)
    assert g.sim.num_fem_stretch_triangles() == 0 and g.sim.num_bending_dihedrals() > 0
    _, ro, co, _, _, _ = g.sim.debug_dump_last_solve()
    dih, _ = g.sim.bending_dihedrals()
    pairs = {(int(a), int(b)) for q in dih for a in q for b in q if a != b}
    assert ro[-1] == len(pairs)
