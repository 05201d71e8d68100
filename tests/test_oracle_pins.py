"""Pins the oracle's pin projection (an extension: the reference's pins are
penalties, pin_constraint.h:5-15) against what the mathematics fixes:

* the eliminated system is the unpinned system (pin stiffness 0) with the
  pinned displacements d_p = t - x_n moved to the right-hand side
  (b_f - A_fp d_p), couplings to pinned particles removed and pinned rows
  I dx_p = d_p;
* the step's dx solves the reduced system A_ff dx_f = b_f - A_fp d_p (a
  dense numpy solve of the unpinned system);
* a pinned particle ends every step exactly at its target, also when the
  target moves;
* penalty pins tend to it as their stiffness grows, by O(1/k).
"""

import numpy as np

from chysx import scenes
from oracle.oracle import oracle_from_scene


def dense(diag, ro, co, vals):
    n = diag.shape[0]
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        A[3 * r:3 * r + 3, 3 * r:3 * r + 3] = diag[r]
        for q in range(ro[r], ro[r + 1]):
            A[3 * r:3 * r + 3, 3 * co[q]:3 * co[q] + 3] = vals[q]
    return A


def deformed_state(s, seed=0):
    rng = np.random.default_rng(seed)
    h = 1.0 / (s.grid_w - 1)
    p = s.pos.astype(np.float64) + rng.uniform(-0.1 * h, 0.1 * h, s.pos.shape)
    v = rng.normal(scale=0.05, size=s.pos.shape)
    return np.ascontiguousarray(p), np.ascontiguousarray(v)


def pinned_scene():
    s = scenes.config1()
    s.pin_idx = np.array([0, 9, 45], np.int32)
    s.pin_targets = (s.pos[s.pin_idx] + np.array([[0.0, 0.02, 0.0], [0.01, 0.0, -0.01], [0.0, -0.03, 0.02]])
                     ).astype(np.float32)
    return s


def test_eliminated_system_is_the_unpinned_system_reduced():
    s = pinned_scene()
    p, v = deformed_state(s)
    free = make(s, projection=False, pin_k=0.0)
    free.assemble(p, v, s.inv_mass, None, s.dt)
    d0, v0, b0, _ = free.system()
    ro, co = free.pattern()
    proj = make(s, projection=True)
    proj.assemble(p, v, s.inv_mass, None, s.dt)
    d1, v1, b1, _ = proj.system()
    pinned = np.zeros(s.n, bool)
    pinned[s.pin_idx] = True
    dp = np.zeros((s.n, 3))
    dp[s.pin_idx] = s.pin_targets.astype(np.float64) - p[s.pin_idx]
    rows = np.repeat(np.arange(s.n), np.diff(ro))
    couple = pinned[rows] | pinned[co]
    np.testing.assert_array_equal(v1[couple], 0.0)
    np.testing.assert_array_equal(v1[~couple], v0[~couple])
    for i in np.flatnonzero(pinned):
        np.testing.assert_array_equal(d1[i], np.eye(3))
        np.testing.assert_allclose(b1[i], dp[i], rtol=0, atol=0)
    expect = b0.copy()
    np.add.at(expect, rows[~pinned[rows]], -np.einsum("kij,kj->ki", v0[~pinned[rows]], dp[co[~pinned[rows]]]))
    np.testing.assert_allclose(b1[~pinned], expect[~pinned], rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(d1[~pinned], d0[~pinned])


def make(s, projection, pin_k=None, tol=1e-13):
    oc, _, _, _, _ = oracle_from_scene(s, precision=64, tol=tol, max_iter=100000)
    if pin_k is not None:
        oc.set_pins(s.pin_idx, s.pin_targets, pin_k)
    if projection:
        oc.set_pin_projection(True)
    return oc


def test_step_solves_the_reduced_system():
    s = pinned_scene()
    p, v = deformed_state(s, 1)
    free = make(s, projection=False, pin_k=0.0)
    free.assemble(p, v, s.inv_mass, None, s.dt)
    d0, v0, b0, _ = free.system()
    ro, co = free.pattern()
    A = dense(d0, ro, co, v0)
    pinned = np.zeros(3 * s.n, bool)
    for i in s.pin_idx:
        pinned[3 * i:3 * i + 3] = True
    dp = np.zeros(3 * s.n)
    dp[pinned] = (s.pin_targets.astype(np.float64) - p[s.pin_idx]).reshape(-1)
    xf = np.linalg.solve(A[~pinned][:, ~pinned], b0.reshape(-1)[~pinned] - A[~pinned][:, pinned] @ dp[pinned])
    oc = make(s, projection=True)
    pp, vv = p.copy(), v.copy()
    oc.step(pp, vv, s.inv_mass, None, s.dt, 1)
    dx = oc.get_dx().reshape(-1)
    np.testing.assert_allclose(dx[~pinned], xf, rtol=0, atol=1e-10 * np.abs(xf).max())
    np.testing.assert_array_equal(dx[pinned], dp[pinned])


def test_pinned_particles_sit_exactly_at_moving_targets():
    s = pinned_scene()
    oc = make(s, projection=True, tol=1e-8)
    p = s.pos.astype(np.float64).copy()
    v = np.zeros_like(p)
    tgt = s.pin_targets.astype(np.float64)
    for k in range(12):
        if k == 6:
            tgt = tgt + np.array([0.0, 0.05, 0.01])
            oc.set_pins(s.pin_idx, tgt.astype(np.float32), 1e6)
            tgt = tgt.astype(np.float32).astype(np.float64)
        x_prev = p.copy()
        oc.step(p, v, s.inv_mass, None, s.dt, 1)
        np.testing.assert_array_equal(p[s.pin_idx], tgt)
        np.testing.assert_allclose(v[s.pin_idx], (tgt - x_prev[s.pin_idx]) / s.dt, rtol=1e-12)


def test_stiff_penalty_pins_tend_to_the_projection():
    """Penalty pins of stiffness k differ from the projection by O(1/k): ten
    times stiffer, ten times closer (k = 1e6 and 1e7, solved to 1e-14; much
    stiffer pins would let the pin rows dominate |b| and loosen the free
    rows' convergence under a relative stopping test)."""
    s = pinned_scene()
    p0, v0 = deformed_state(s, 2)
    out = {}
    for name, proj, k in [("projection", True, None), ("k6", False, 1e6), ("k7", False, 1e7)]:
        oc = make(s, projection=proj, pin_k=k, tol=1e-14)
        p, v = p0.copy(), v0.copy()
        oc.step(p, v, s.inv_mass, None, s.dt, 3)
        out[name] = p
    scale = np.linalg.norm(out["projection"] - p0)
    d6 = np.linalg.norm(out["k6"] - out["projection"])
    d7 = np.linalg.norm(out["k7"] - out["projection"])
    assert 7.0 <= d6 / d7 <= 13.0, (d6, d7)
    assert d7 <= 1e-4 * scale
