"""Pins the oracle's static-shape contact (collision/static_contact.h:1-312,
static_contact.cu:40-380; friction.cuh:28-36) against what fixes it:

  * box signed distance against a brute-force closest point (outside) and
    the face distances (inside); the normal against the gradient of the
    distance by central differences;
  * the penalty force and its Gauss-Newton block against finite differences
    of E = k/2 max(0, h - d)^2 for planes and boxes, through the assembled
    system;
  * the friction ramp f1 at 0, at eps and past it (C0, value 2/eps at 0);
  * the Coulomb cone: stick cancels the tangential load, slip removes
    exactly mu * f_n along it, no normal load means no projection;
  * a resting cloth on a ground plane settles at the closed-form penetration
    h - m g / k.
"""

import numpy as np
import pytest

from chysx import scenes
from oracle.oracle import OracleCloth, box_sdf

BOX = dict(center=(0.1, -0.2, 0.3), half_ext=(0.5, 0.25, 0.4))


def rot(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q[:, 0], q[:, 1], q[:, 2]


def brute_box(center, half, ex, ey, ez, x):
    R = np.stack([ex, ey, ez], 1)
    q = R.T @ (np.asarray(x) - center)
    d = np.abs(q) - half
    if np.all(d <= 0):
        return float(d.max())
    closest = np.clip(q, -np.asarray(half), half)
    return float(np.linalg.norm(q - closest))


@pytest.mark.parametrize("seed", range(4))
def test_box_sdf_matches_brute_force_and_its_gradient(seed):
    ex, ey, ez = rot(seed)
    rng = np.random.default_rng(100 + seed)
    for _ in range(40):
        x = np.asarray(BOX["center"]) + rng.uniform(-1.0, 1.0, 3)
        d, n = box_sdf(BOX["center"], BOX["half_ext"], ex, ey, ez, x)
        assert d == pytest.approx(brute_box(np.asarray(BOX["center"]), np.asarray(BOX["half_ext"]), ex, ey, ez, x),
                                  abs=1e-12)
        h = 1e-7
        g = np.array([(box_sdf(BOX["center"], BOX["half_ext"], ex, ey, ez, x + h * e)[0] -
                       box_sdf(BOX["center"], BOX["half_ext"], ex, ey, ez, x - h * e)[0]) / (2 * h)
                      for e in np.eye(3)])
        np.testing.assert_allclose(n, g, atol=1e-5)
        assert abs(np.linalg.norm(n) - 1) < 1e-12


def contact_system(x, v, shapes, h=0.01, k=1e4, mu=0.0, eps=1e-4, dt=1.0 / 60.0, mass=0.01):
    """Assembled system of a 3-particle mesh with no elastic elements: only
    inertia and contact act."""
    oc = OracleCloth(64)
    oc.set_mesh(np.array([[0, 1, 2]], np.int32), 3)
    for kind, args in shapes:
        getattr(oc, kind)(*args)
    oc.set_contact(h, k, mu, eps)
    oc.set_params((0.0, -9.8, 0.0), 0.0, False, False, 0.0, 1)
    oc.assemble(np.asarray(x, np.float64), np.asarray(v, np.float64), np.full(3, 1.0 / mass), None, dt)
    diag, _, rhs, _ = oc.system()
    inertia = mass / dt ** 2
    return diag - inertia * np.eye(3), rhs - mass * (np.asarray(v) / dt + np.array([0.0, -9.8, 0.0]))


def penalty_energy(x, shapes, h, k):
    best = 0.0
    for kind, args in shapes:
        if kind == "add_plane":
            n, d0 = np.asarray(args[0], np.float64), args[1]
            d = n @ x + d0
        else:
            d, _ = box_sdf(*args, x)
        best = max(best, h - d)
    return 0.5 * k * best ** 2


@pytest.mark.parametrize("case", ["plane", "box_out", "box_in"])
def test_penalty_force_and_block_match_energy_fd(case):
    ex, ey, ez = rot(7)
    n = np.array([0.2, 1.0, -0.3])
    n /= np.linalg.norm(n)
    if case == "plane":
        shapes = [("add_plane", (n, 0.05))]
        x0 = -0.06 * n + np.array([0.3, 0.0, 0.0]) - (np.array([0.3, 0.0, 0.0]) @ n) * n
    else:
        shapes = [("add_box", (BOX["center"], BOX["half_ext"], ex, ey, ez))]
        c = np.asarray(BOX["center"])
        x0 = c + ex * 0.505 + ey * 0.1 if case == "box_out" else c + ex * 0.49 + ey * 0.05
    h, k = 0.02, 3e3
    X = np.stack([x0, x0 + [5.0, 5.0, 5.0], x0 + [-5.0, 5.0, 5.0]])  # only particle 0 in contact
    V = np.zeros((3, 3))
    H, F = contact_system(X, V, shapes, h=h, k=k)
    assert penalty_energy(x0, shapes, h, k) > 0
    g = np.array([(penalty_energy(x0 + 1e-7 * e, shapes, h, k) - penalty_energy(x0 - 1e-7 * e, shapes, h, k)) / 2e-7
                  for e in np.eye(3)])
    np.testing.assert_allclose(F[0], -g, rtol=1e-6, atol=1e-6 * np.abs(g).max())
    # Gauss-Newton block k n n^T, n the unit outward normal along -F
    nn = F[0] / np.linalg.norm(F[0])
    np.testing.assert_allclose(H[0], k * np.outer(nn, nn), atol=1e-9 * k)
    np.testing.assert_allclose(F[1:], 0.0, atol=1e-12)
    np.testing.assert_allclose(H[1:], 0.0, atol=1e-12)


def test_friction_ramp():
    eps = 1e-3
    # alpha = mu k depth f1(|u_t|): the tangential diagonal of a contact with
    # slip u gives f1(u) * mu * k * depth
    h, k, mu, dt = 0.01, 2e3, 0.5, 0.01
    depth = 0.004
    out = []
    for u in [0.0, 0.5 * eps, eps, 2 * eps, 10 * eps]:
        x = np.array([[0.0, h - depth, 0.0], [5, 5, 5], [6, 5, 5]])
        v = np.zeros((3, 3))
        v[0, 0] = u / dt
        H, _ = contact_system(x, v, [("add_plane", ((0, 1, 0), 0.0))], h=h, k=k, mu=mu, eps=eps, dt=dt)
        out.append(H[0][0, 0] / (mu * k * depth))
    np.testing.assert_allclose(out, [2 / eps, 1.5 / eps, 1 / eps, 1 / (2 * eps), 1 / (10 * eps)], rtol=1e-9)


def test_coulomb_cone_stick_slip_and_release():
    """Particle on the plane y = 0 (thickness h): the force F0 = m(v/dt + g)
    (no elastic terms) is projected on the cone of radius mu * max(0, -F0.n).
    Zero lagged slip isolates the projection."""
    h, k, dt, m = 0.01, 1e4, 0.01, 0.01
    x = np.array([[0.0, h - 0.002, 0.0], [5, 5, 5], [6, 5, 5]])
    v = np.zeros((3, 3))
    v[0] = [0.3, -0.5, 0.0]  # pushes into the plane and sideways
    F0 = m * (v[0] / dt + np.array([0.0, -9.8, 0.0]))
    fn = -F0[1]
    for mu, stick in [(10.0, True), (0.05, False)]:
        oc = OracleCloth(64)
        oc.set_mesh(np.array([[0, 1, 2]], np.int32), 3)
        oc.add_plane((0, 1, 0), 0.0)
        oc.set_contact(h, k, mu, 1e9)  # eps huge: alpha ~ 2 mu k depth / eps -> 0
        oc.set_params((0.0, -9.8, 0.0), 0.0, False, False, 0.0, 1)
        oc.assemble(x, v, np.full(3, 1.0 / m), None, dt)
        _, _, rhs, _ = oc.system()
        tangential = rhs[0] - (rhs[0] @ [0, 1, 0]) * np.array([0, 1, 0])
        if stick:
            np.testing.assert_allclose(tangential, 0.0, atol=1e-9)
        else:
            np.testing.assert_allclose(tangential[0], F0[0] - mu * fn, rtol=1e-9)
        # the normal part is F0_n + k depth
        assert rhs[0][1] == pytest.approx(F0[1] + k * 0.002, rel=1e-9)
    # no normal load (moving away from the plane fast): no projection
    v2 = np.zeros((3, 3))
    v2[0] = [0.3, 2.0, 0.0]
    oc = OracleCloth(64)
    oc.set_mesh(np.array([[0, 1, 2]], np.int32), 3)
    oc.add_plane((0, 1, 0), 0.0)
    oc.set_contact(h, k, 10.0, 1e9)
    oc.set_params((0.0, -9.8, 0.0), 0.0, False, False, 0.0, 1)
    oc.assemble(x, v2, np.full(3, 1.0 / m), None, dt)
    _, _, rhs, _ = oc.system()
    assert rhs[0][0] == pytest.approx(m * v2[0][0] / dt, rel=1e-6)


def test_cloth_rests_on_ground_at_closed_form_depth():
    """A free cloth dropped on the plane y = -0.05 comes to rest with every
    particle at signed distance h - m g / k (penalty balances weight; the
    stretch forces vanish on a flat rest state)."""
    s = scenes.freefall(w=6)
    h, k = 0.01, 5e3
    oc = OracleCloth(64)
    oc.set_mesh(s.tris, s.n)
    pos = s.pos.astype(np.float64)
    oc.build_stretch(pos, 1e3)
    oc.add_plane((0, 1, 0), 0.05)
    oc.set_contact(h, k)
    oc.set_params(s.gravity, 5.0, False, False, 1e-10, 10000)
    v = np.zeros_like(pos)
    oc.step(pos, v, s.inv_mass.astype(np.float64), None, s.dt, 400)
    m = 1.0 / s.inv_mass[0]
    expect = -0.05 + h - m * 9.8 / k
    np.testing.assert_allclose(pos[:, 1], expect, atol=1e-7)
    np.testing.assert_allclose(v, 0.0, atol=1e-6)
