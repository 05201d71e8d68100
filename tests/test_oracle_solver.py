"""Pins the oracle's pattern, slot maps, assembly and PCG:
brute-force pattern builds, closed-form counts, dense library solves,
finite differences of the assembled gradient, the free-fall closed form and
the committed golden record."""

import os

import numpy as np
import pytest

from chysx import scenes
from oracle.oracle import OracleCloth, oracle_from_scene, pcg_solve

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cloth10_golden.npz")


def brute_pattern(n, tris, springs=None):
    pairs = set()
    for t in tris:
        for a in range(3):
            for b in range(3):
                if a != b and t[a] != t[b]:
                    pairs.add((int(t[a]), int(t[b])))
    if springs is not None:
        for a, b in springs:
            if a != b:  # a degenerate triangle yields a self-edge: diagonal only
                pairs.add((int(a), int(b)))
                pairs.add((int(b), int(a)))
    ro = np.zeros(n + 1, np.int64)
    cols = []
    for r, c in sorted(pairs):
        ro[r + 1] += 1
        cols.append(c)
    return np.cumsum(ro), np.array(cols)


def build(tris, n, springs_active=False, pins=()):
    oc = OracleCloth(64)
    oc.set_mesh(tris, n)
    pos = np.zeros((n, 3))
    oc.build_springs(pos, 1.0)
    oc.build_stretch(pos, 1.0)
    oc.set_pins(np.array(pins, np.int32), np.zeros((len(pins), 3)), 1e6)
    oc.set_params(springs_active=springs_active)
    oc.build_topology()
    return oc


@pytest.mark.parametrize("w", [2, 3, 10, 17])
def test_grid_pattern_closed_form_and_brute_force(w):
    pos, tris = scenes.grid(w)
    oc = build(tris, w * w)
    ro, co = oc.pattern()
    e = 2 * w * (w - 1) + (w - 1) ** 2
    assert ro[-1] == 2 * e == co.size
    bro, bco = brute_pattern(w * w, tris)
    np.testing.assert_array_equal(ro, bro)
    np.testing.assert_array_equal(co, bco)
    deg = np.diff(ro)
    assert deg.max() == (6 if w > 2 else 3) and deg.min() >= 2
    for r in range(w * w):  # ascending, no diagonal
        c = co[ro[r]:ro[r + 1]]
        assert np.all(np.diff(c) > 0) and r not in c
    # springs on the triangle edges leave the pattern unchanged
    oc2 = build(tris, w * w, springs_active=True)
    ro2, co2 = oc2.pattern()
    np.testing.assert_array_equal(ro, ro2)
    np.testing.assert_array_equal(co, co2)


@pytest.mark.parametrize("seed", range(3))
def test_random_mesh_pattern_and_slots(seed):
    rng = np.random.default_rng(seed)
    n = 40
    tris = rng.integers(0, n, size=(60, 3)).astype(np.int32)
    tris[0] = [5, 5, 7]  # degenerate triangle: repeated vertex
    oc = build(tris, n, springs_active=True, pins=(3, 11))
    ro, co = oc.pattern()
    edges = oc.edges()
    bro, bco = brute_pattern(n, tris, edges)
    np.testing.assert_array_equal(ro, bro)
    np.testing.assert_array_equal(co, bco)
    ts, ss, ps = oc.slots(tris.shape[0], edges.shape[0], 2)

    def decode(slot, r, c):
        if slot < 0:
            return -slot - 1 == r == c
        return ro[r] <= slot < ro[r + 1] and co[slot] == c

    for t in range(tris.shape[0]):
        for a in range(3):
            for b in range(3):
                assert decode(ts[9 * t + 3 * a + b], tris[t, a], tris[t, b])
    for e, (a, b) in enumerate(edges):
        for q, (r, c) in enumerate([(a, a), (a, b), (b, a), (b, b)]):
            assert decode(ss[4 * e + q], r, c)
    assert list(ps) == [-4, -12]


def test_pattern_out_of_range_raises():
    with pytest.raises(IndexError):
        build(np.array([[0, 1, 9]], np.int32), 5)


def random_spd_bsr(n, rng):
    _, tris = scenes.grid(int(np.sqrt(n)))
    m = int(np.sqrt(n)) ** 2
    ro, co = brute_pattern(m, tris)
    dense = np.zeros((3 * m, 3 * m))
    vals = np.zeros((co.size, 3, 3))
    for r in range(m):
        for q in range(ro[r], ro[r + 1]):
            c = co[q]
            if c > r:
                b = 0.3 * rng.normal(size=(3, 3))
                dense[3 * r:3 * r + 3, 3 * c:3 * c + 3] = b
                dense[3 * c:3 * c + 3, 3 * r:3 * r + 3] = b.T
    dense += np.diag(np.abs(dense).sum(1) + 1.0)  # diagonally dominant SPD
    diag = np.stack([dense[3 * r:3 * r + 3, 3 * r:3 * r + 3] for r in range(m)])
    for r in range(m):
        for q in range(ro[r], ro[r + 1]):
            c = co[q]
            vals[q] = dense[3 * r:3 * r + 3, 3 * c:3 * c + 3]
    return m, dense, diag, ro, co, vals


@pytest.mark.parametrize("seed", range(3))
def test_pcg_matches_dense_solve(seed):
    rng = np.random.default_rng(seed)
    m, dense, diag, ro, co, vals = random_spd_bsr(64, rng)
    b = rng.normal(size=(m, 3))
    x, it, rr = pcg_solve(diag, ro, co, vals, b, tol=1e-12, max_iter=1000)
    ref = np.linalg.solve(dense, b.reshape(-1)).reshape(m, 3)
    np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-10)
    assert rr <= 1e-12 and 0 < it < 200


def test_pcg_single_block_is_exact_in_one_iteration():
    a = np.array([[4.0, 1, 0.5], [1, 3, 0.2], [0.5, 0.2, 2]])
    b = np.array([[1.0, -2.0, 0.5]])
    x, it, rr = pcg_solve(a[None], np.array([0, 0]), np.zeros(0, np.int32), np.zeros((0, 3, 3)), b,
                          tol=1e-12, max_iter=10)
    np.testing.assert_allclose(x[0], np.linalg.solve(a, b[0]), rtol=1e-14)
    assert it <= 1


def test_pcg_fixed_iteration_mode_runs_exactly_max_iter():
    rng = np.random.default_rng(7)
    m, dense, diag, ro, co, vals = random_spd_bsr(64, rng)
    b = rng.normal(size=(m, 3))
    _, it, _ = pcg_solve(diag, ro, co, vals, b, tol=0.0, max_iter=7)
    assert it == 7


def test_assembled_hessian_is_fd_of_assembled_gradient_at_rest():
    """At the rest state no Baraff-Witkin term is filtered, so the assembled
    matrix minus inertia must equal the derivative of the assembled gradient
    (rhs = m v/dt + m g - grad E): checks every slot of the scatter."""
    s = scenes.config1()
    n = s.n
    rest = s.pos.astype(np.float64)

    def system_at(x):
        oc = OracleCloth(64)
        oc.set_mesh(s.tris, n)
        oc.build_springs(rest, s.spring_k)
        oc.build_stretch(rest, s.stretch_k)
        oc.set_pins(s.pin_idx, s.pin_targets, s.pin_k)
        oc.set_params(s.gravity, 0.0, True, False, 0.0, 1)
        p = np.array(x, np.float64)
        v = np.zeros((n, 3))
        oc.step(p, v, s.inv_mass, None, s.dt, 1)
        ro, co = oc.pattern()
        diag, vals, rhs, _ = oc.system()
        return ro, co, diag, vals, rhs

    ro, co, diag, vals, rhs0 = system_at(rest)
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        A[3 * r:3 * r + 3, 3 * r:3 * r + 3] = diag[r]
        for q in range(ro[r], ro[r + 1]):
            A[3 * r:3 * r + 3, 3 * co[q]:3 * co[q] + 3] = vals[q]
    m_dt2 = (1.0 / s.inv_mass) / s.dt ** 2
    A -= np.diag(np.repeat(m_dt2, 3))
    h = 1e-6
    rng = np.random.default_rng(0)
    for j in rng.choice(3 * n, 24, replace=False):
        xp = rest.copy().reshape(-1)
        xm = rest.copy().reshape(-1)
        xp[j] += h
        xm[j] -= h
        # rhs(x) = M/dt^2 (x_tilde - x) - grad E(x) with v = 0:
        # x_tilde - x = dt^2 g  ->  d rhs / dx = -H_E
        col = -(system_at(xp.reshape(n, 3))[4] - system_at(xm.reshape(n, 3))[4]).reshape(-1) / (2 * h)
        np.testing.assert_allclose(A[:, j], col, rtol=1e-5, atol=1e-5 * np.abs(A).max())
    np.testing.assert_allclose(A, A.T, atol=1e-9 * np.abs(A).max())


@pytest.mark.parametrize("prec", [64, 32])
def test_freefall_closed_form(prec):
    s = scenes.freefall(6)
    oc, pos, vel, im, f = oracle_from_scene(s, precision=prec, tol=1e-9)
    oc.step(pos, vel, im, f, s.dt, 10)
    expect = s.pos.astype(np.float64).copy()
    expect[:, 1] += s.dt ** 2 * -9.8 * 10 * 11 / 2
    tol = 1e-12 if prec == 64 else 2e-6
    np.testing.assert_allclose(pos, expect, atol=tol)
    np.testing.assert_allclose(vel[:, 1], -9.8 * 10 * s.dt, rtol=1e-9 if prec == 64 else 1e-4)


def test_step_matches_dense_solve_of_its_system():
    s = scenes.config1()
    oc, pos, vel, im, f = oracle_from_scene(s, precision=64, tol=1e-12)
    oc.step(pos, vel, im, f, s.dt, 3)
    ro, co = oc.pattern()
    diag, vals, rhs, dx = oc.system()
    n = s.n
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        A[3 * r:3 * r + 3, 3 * r:3 * r + 3] = diag[r]
        for q in range(ro[r], ro[r + 1]):
            A[3 * r:3 * r + 3, 3 * co[q]:3 * co[q] + 3] = vals[q]
    ref = np.linalg.solve(A, rhs.reshape(-1)).reshape(n, 3)
    np.testing.assert_allclose(dx, ref, rtol=1e-8, atol=1e-12)
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


def test_reference_mode_fixed_iterations():
    s = scenes.config1()
    oc, pos, vel, im, f = oracle_from_scene(s, precision=32, springs_active=False, force_active=False,
                                            tol=0.0, max_iter=50)
    it, rr, cv = oc.step(pos, vel, im, f, s.dt, 5)
    assert list(it) == [50] * 5 and not cv.any()


def test_oracle_reproduces_golden_record():
    g = np.load(GOLDEN)
    s = scenes.config1()
    oc, pos, vel, im, f = oracle_from_scene(s, precision=64, tol=1e-10, max_iter=100000)
    np.testing.assert_array_equal(oc.edges(), g["edges"])
    oc.step(pos, vel, im, f, s.dt, 1)
    ro, co = oc.pattern()
    np.testing.assert_array_equal(ro, g["row_offsets"])
    np.testing.assert_array_equal(co, g["col_indices"])
    ts, ss, ps = oc.slots(s.tris.shape[0], g["edges"].shape[0], 1)
    np.testing.assert_array_equal(ts, g["tri_slots"])
    np.testing.assert_array_equal(ss, g["spring_slots"])
    diag, vals, rhs, dx = oc.system()
    np.testing.assert_array_equal(diag, g["diag1"])
    np.testing.assert_array_equal(vals, g["values1"])
    np.testing.assert_array_equal(rhs, g["rhs1"])
    np.testing.assert_allclose(pos, g["pos1"], rtol=0, atol=1e-14)
    oc.step(pos, vel, im, f, s.dt, 9)
    np.testing.assert_allclose(pos, g["pos10"], rtol=0, atol=1e-12)


def test_thread_count_does_not_change_results():
    s = scenes.config2(w=40, steps=3)
    out = []
    for threads in (1, 4):
        oc, pos, vel, im, f = oracle_from_scene(s, precision=64, tol=1e-8, threads=threads)
        oc.step(pos, vel, im, f, s.dt, 3)
        out.append(pos.copy())
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("prec", [64, 32])
def test_damping_and_uniform_force_closed_form(prec):
    """A uniform translation keeps every elastic term at rest, so each step
    is dx = dt v + dt^2 (g + f/m) and v' = (dx/dt) exp(-c dt)
    (finalize_step_kernel, cloth_simulator.cu:222-244; compute_x_tilde_kernel
    with a force, :83-106)."""
    s = scenes.freefall(5)
    s.damping = 0.7
    f = np.array([0.02, 0.0, -0.01])
    s.force = np.tile(f, (s.n, 1)).astype(np.float32)
    oc, pos, vel, im, ff = oracle_from_scene(s, precision=prec, tol=1e-10, force_active=True)
    oc.set_params(s.gravity, s.damping, True, True, 1e-10, 10000)
    m = 1.0 / s.inv_mass[0]
    a = np.array(s.gravity) + f / m
    decay = np.exp(-s.damping * s.dt)
    x = np.zeros(3)
    v = np.zeros(3)
    for _ in range(12):
        dx = s.dt * v + s.dt ** 2 * a
        x = x + dx
        v = dx / s.dt * decay
    oc.step(pos, vel, im, ff, s.dt, 12)
    tol = 1e-12 if prec == 64 else 3e-6
    np.testing.assert_allclose(pos - s.pos, np.tile(x, (s.n, 1)), atol=tol)
    np.testing.assert_allclose(vel, np.tile(v, (s.n, 1)), atol=tol * 100)
