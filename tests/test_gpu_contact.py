"""GPU parity of static-shape contact (planes and oriented boxes, penalty,
lagged IPC friction, Coulomb-cone projection) against the FP64 oracle.

Bars: blocks within 2e-6 of their row's largest block (5e-5 on rows in
friction contact: the friction block alpha = mu k depth f1(|u_t|) has
d f1 / d|u_t| = -1/eps^2 = -1e8 /m, and the FP32 rounding of the tangential
slip u_t = v dt - n (n . v dt), ~eps32 |v dt| ~ 2e-9 m, moves alpha by up to
~2e-5 relative), rhs within 1e-4 N + 1e-5 relative, every solve at true relative residual <= 1e-6, positions
after one step from a shared state within 5e-5 relative (the penalty adds
k / (m / dt^2) ~ 1e3 to the conditioning of contact rows), and a cloth
resting on the ground at the closed-form depth h - m g / k.
"""

import numpy as np
import pytest

from chysx import scenes
from chysx.harness import make_sim
from oracle.oracle import box_sdf, oracle_from_scene

pytestmark = pytest.mark.gpu

H, K = 0.01, 1.0e3
GROUND = 0.3  # plane y = -0.3
BOX = dict(center=(0.4, -0.2, 0.45), half_ext=(0.15, 0.08, 0.2))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def axes(angle):
    c, s = np.cos(angle), np.sin(angle)
    return np.array([c, 0.0, s]), np.array([0.0, 1.0, 0.0]), np.array([-s, 0.0, c])


def add_shapes(sim, oc, mu, angle=0.3):
    ex, ey, ez = axes(angle)
    sim.add_static_plane(np.array([0.0, 1.0, 0.0], np.float32), GROUND)
    sim.add_static_box(np.array(BOX["center"], np.float32), np.array(BOX["half_ext"], np.float32),
                       ex.astype(np.float32), ey.astype(np.float32), ez.astype(np.float32))
    sim.set_static_contact_thickness(H)
    sim.set_static_contact_stiffness(K)
    sim.set_static_contact_friction(mu)
    oc.add_plane((0.0, 1.0, 0.0), GROUND)
    oc.add_box(BOX["center"], BOX["half_ext"], ex.astype(np.float32), ey.astype(np.float32), ez.astype(np.float32))
    oc.set_contact(H, K, mu, 1e-4)


def contact_rows(x):
    """Particles within the contact thickness of the plane or the box."""
    ex, ey, ez = axes(0.3)
    on_box = np.array([box_sdf(BOX["center"], BOX["half_ext"], ex, ey, ez, q)[0] < H for q in x])
    return on_box | (x[:, 1] + GROUND < H)


def compare_systems(g, oc, n, friction_rows=None):
    diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
    od, ov, orhs, _ = oc.system()
    rows = np.repeat(np.arange(n), np.diff(ro))
    rowmax = np.abs(od).max(axis=(1, 2))
    rtol = np.full(n, 2e-6)
    if friction_rows is not None:
        rtol[friction_rows] = 5e-5
    assert np.all(np.abs(diag - od) <= rtol[:, None, None] * rowmax[:, None, None] + 1e-12)
    assert np.all(np.abs(vals - ov) <= 2e-6 * rowmax[rows][:, None, None] + 1e-12)
    np.testing.assert_allclose(rhs, orhs, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("mu", [0.0, 0.4])
def test_drape_on_box_and_ground_resync(cuda, mu):
    s = scenes.drape(w=25)
    g = make_sim(s, tol=1e-6, springs_active=False, force_active=False)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, springs_active=False, force_active=False, tol=1e-11,
                                        max_iter=100000)
    assert g.sim.static_plane_count() == 0
    add_shapes(g.sim, oc, mu)
    assert g.sim.static_plane_count() == 1 and g.sim.static_box_count() == 1
    assert g.sim.static_contact_friction() == pytest.approx(mu)
    box_max = ground_max = 0
    for k in range(90):
        check = (k % 10) == 9
        if check:
            p0 = g.positions().astype(np.float64)
            v0 = g.velocities().astype(np.float64)
            dx0 = g.sim.get_dx64()
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["converged"] and st["relres"] <= 1e-6, (k, st)
        if check:
            oc.set_dx(dx0)
            oc.step(p0, v0, im, f, s.dt, 1)
            compare_systems(g, oc, s.n, contact_rows(p0) if mu > 0 else None)
            assert rel(g.positions(), p0) <= 5e-5, k
            x = g.positions().astype(np.float64)
            ex, ey, ez = axes(0.3)
            box_max = max(box_max, int(sum(box_sdf(BOX["center"], BOX["half_ext"], ex, ey, ez, q)[0] < H for q in x)))
            ground_max = max(ground_max, int((x[:, 1] + GROUND < H).sum()))
    # contact was exercised: the box always; the ground too without friction
    # (with mu = 0.4 the box holds the cloth up)
    assert box_max > 0
    if mu == 0.0:
        assert ground_max > 0


def test_resting_on_ground_closed_form(cuda):
    s = scenes.freefall(w=6)
    g = make_sim(s, tol=1e-9, springs_active=False, force_active=False)
    g.sim.add_static_plane(np.array([0.0, 1.0, 0.0], np.float32), 0.05)
    g.sim.set_static_contact_thickness(H)
    g.sim.set_static_contact_stiffness(5e3)
    mat = g.sim.material
    mat.damping = 5.0
    for _ in range(400):
        g.sim.step(s.dt)
    m = 1.0 / s.inv_mass[0]
    expect = -0.05 + H - m * 9.8 / 5e3
    np.testing.assert_allclose(g.positions()[:, 1], expect, atol=2e-6)


@pytest.mark.parametrize("mode", ["projection", "learned"])
def test_contact_with_warm_start_modes(cuda, mode):
    """The drape falling onto the box with friction, stepped with a warm
    start: contact lives in the assembled operator, which the warm-start
    stages multiply by like any other, so every solve converges and a step
    from the simulator's state matches the oracle's (same bars as without a
    warm start)."""
    from chysx import warmstart
    s = scenes.drape(w=25)
    g = make_sim(s, tol=1e-6, springs_active=False, force_active=False)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, springs_active=False, force_active=False, tol=1e-11,
                                        max_iter=100000)
    add_shapes(g.sim, oc, 0.4)
    if mode == "learned":
        g.sim.ws_set_params(warmstart.load_params()[0])
    g.sim.set_warm_start_mode(mode)
    used = []
    for k in range(50):
        check = k in (29, 49)
        if check:
            p0 = g.positions().astype(np.float64)
            v0 = g.velocities().astype(np.float64)
            dx0 = g.sim.get_dx64()
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["converged"] and st["relres"] <= 1e-6, (mode, k, st)
        if k >= 2:
            w = g.sim.ws_stats()
            used.append(w["projection_used"] if mode == "projection" else w["used"])
        if check:
            oc.set_dx(dx0)
            oc.step(p0, v0, im, f, s.dt, 1)
            assert rel(g.positions(), p0) <= 5e-5, (mode, k)
    assert np.mean(used) > 0.3, (mode, np.mean(used))

