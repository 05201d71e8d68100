"""GPU parity on the reference's special cases, against the FP64 oracle.

* mass_from_inv_mass_kernel (cloth_simulator.cu:66-77): a null inverse-mass
  pointer means 1 kg per particle; w <= 1e-12 (including w = 0) means 1e8 kg;
* degenerate rest shapes: determinant <= 1e-6 gives the identity frame
  (triangle_stretch_constraint.cu:95-99), zero-area triangles, a collapsed
  deformed edge (kMinNorm, :39);
* coincident spring ends (kMinLength = 1e-6, spring_constraint.cu:35).

Bars as in test_gpu_step.py: blocks within 2e-6 of their row's largest
block, every solve at true relative residual <= 1e-6, positions after one
step from a shared state within 1e-5 relative.
"""

import numpy as np
import pytest

from chysx import scenes
from chysx.harness import make_sim
from oracle.oracle import oracle_from_scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def assert_system_matches(g, oc, p0, v0, im, f, dt, rhs_atol):
    """The system the GPU assembled for its last step against the oracle's
    system at the same state."""
    diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
    oc.assemble(p0, v0, im, f, dt)
    od, ov, orhs, _ = oc.system()
    oro, oco = oc.pattern()
    np.testing.assert_array_equal(ro, oro)
    np.testing.assert_array_equal(co, oco)
    rows = np.maximum(np.abs(od).max(axis=(1, 2)), np.zeros(len(od)))
    for r in range(len(od)):
        blk = np.abs(ov[ro[r]:ro[r + 1]]).max() if ro[r + 1] > ro[r] else 0.0
        rows[r] = max(rows[r], blk)
    assert np.all(np.abs(diag - od) <= 2e-6 * rows[:, None, None] + 1e-12)
    rr = np.repeat(np.arange(len(od)), np.diff(ro))
    assert np.all(np.abs(vals - ov) <= 2e-6 * rows[rr][:, None, None] + 1e-12)
    np.testing.assert_allclose(rhs, orhs, rtol=1e-5, atol=rhs_atol)
    return diag


def step_pair(g, oc, im, f, dt, rhs_atol=5e-5):
    """One GPU step and one oracle step (tol 1e-11) from the GPU's state."""
    p0 = g.positions().astype(np.float64)
    v0 = g.velocities().astype(np.float64)
    dx0 = g.sim.get_dx64()
    if dx0.shape[0] != p0.shape[0]:  # before the first step: a zero warm start
        dx0 = np.zeros_like(p0)
    g.sim.step(dt)
    st = g.sim.last_solve_stats()
    assert st["converged"] and st["relres"] <= 1e-6, st
    diag = assert_system_matches(g, oc, p0.copy(), v0.copy(), im, f, dt, rhs_atol)
    oc.set_dx(dx0)
    p, v = p0.copy(), v0.copy()
    oc.step(p, v, im, f, dt, 1)
    assert rel(g.positions(), p) <= 1e-5
    return diag, p0


def test_null_inv_mass_means_one_kilogram(cuda):
    """inv_mass_ptr = 0: every particle weighs 1 kg.  The GPU matches the
    oracle stepped with no inverse mass, and runs bitwise like an explicit
    inverse mass of 1.0 (1/1 = 1 exactly)."""
    import torch
    s = scenes.config1()
    g = make_sim(s, tol=1e-6)
    g.sim.set_external_buffers(g.pos.data_ptr(), g.vel.data_ptr(), s.n, 0, 0)
    oc, _, _, _, _ = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    for k in range(6):
        diag, _ = step_pair(g, oc, None, None, s.dt)
        if k == 0:
            # at rest the elastic blocks lie in the cloth's plane (xz), so the
            # out-of-plane entry of a free row is the inertia 1 kg / dt^2 alone
            assert abs(diag[55, 1, 1] - 3600.0) <= 1e-6 * 3600.0
    one = make_sim(s, tol=1e-6)
    one.inv_mass.copy_(torch.ones(s.n, dtype=torch.float32))
    ref = make_sim(s, tol=1e-6)
    ref.sim.set_external_buffers(ref.pos.data_ptr(), ref.vel.data_ptr(), s.n, 0, 0)
    for _ in range(6):
        one.sim.step(s.dt)
        ref.sim.step(s.dt)
    np.testing.assert_array_equal(one.positions(), ref.positions())


def test_zero_and_tiny_inv_mass_mean_1e8_kilograms(cuda):
    """w = 0 and w = 1e-13 (<= 1e-12) both give 1e8 kg: the diagonal of those
    rows is 1e8 / dt^2 (elastic terms are 1e-8 of it), not 1e13 / dt^2 or
    infinite.  Gravity is off and the light particles start with seeded
    velocities (the heavy ones at rest), so no row dominates |b| and the
    1e-6 stopping test constrains every row; a 1e8 kg particle then hardly
    moves: |dx| <= |grad E| dt^2 / 1e8."""
    s = scenes.config1()
    heavy = [5, 37, 63]
    s.gravity = (0.0, 0.0, 0.0)
    s.inv_mass = s.inv_mass.copy()
    s.inv_mass[5] = 0.0
    s.inv_mass[37] = 0.0
    s.inv_mass[63] = 1e-13
    s.vel = (0.1 * np.random.default_rng(7).normal(size=s.pos.shape)).astype(np.float32)
    s.vel[heavy] = 0.0
    g = make_sim(s, tol=1e-6)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    for k in range(5):
        diag, p0 = step_pair(g, oc, im, f, s.dt)
        for i in heavy:
            np.testing.assert_allclose(np.diag(diag[i]), 1e8 / s.dt ** 2, rtol=1e-6)
        moved = g.positions()[heavy].astype(np.float64) - p0[heavy]
        assert np.abs(moved).max() <= 1e-7, moved


def test_degenerate_triangles_and_coincident_spring_ends(cuda):
    """scenes.degenerate(): a millimetre triangle (rest determinant 2^-21,
    identity frame, area 2^-22), a zero-area collinear triangle, and a
    particle on top of another joined by a rest-length-0 spring (length
    clamped to 1e-6) and a triangle with two coincident corners.  Step by step
    from the GPU's state the systems and positions match the oracle, every
    solve converges and nothing is NaN."""
    s = scenes.degenerate()
    g = make_sim(s, tol=1e-6)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    r, c = s.meta["coincident"], s.meta["coincident_with"]
    for k in range(12):
        # deformed states: FP32 rounding of the spring lengths (eps32 |x| with
        # |x| ~ 1.4 m, k = 1e3, about six springs per row) bounds the rhs
        # difference by ~5e-4 N
        step_pair(g, oc, im, f, s.dt, rhs_atol=5e-4)
        p = g.positions()
        assert np.all(np.isfinite(p)) and np.all(np.isfinite(g.velocities()))
    # the rest-length-0 spring keeps the coincident pair together
    assert np.linalg.norm(p[r] - p[c]) < 1e-3
    # the identity frame of the millimetre triangle acts at rest: its corners
    # are pulled towards unit edge lengths' strain of -1 scaled by the tiny
    # area, so they move, yet stay finite and near the corner
    assert np.linalg.norm(p[list(s.meta["small"])] - p[0]) < 0.05


def test_degenerate_collapsed_edge_in_deformed_state(cuda):
    """Rest shapes built on the regular grid, then particle 7 is moved onto
    particle 8 before the step: the springs between them are shorter than
    1e-6 (clamped) and the triangles sharing that edge have a collapsed u or
    v edge (clamped norm).  The step matches the oracle."""
    s = scenes.config1()
    g = make_sim(s, tol=1e-6)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    g.sim.step(s.dt)  # builds rest data from the undeformed grid, as the oracle did
    p = g.positions().copy()
    p[7] = p[8]
    g.set_state(p, g.velocities())
    for _ in range(4):
        step_pair(g, oc, im, f, s.dt)
        assert np.all(np.isfinite(g.positions()))
