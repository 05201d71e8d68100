"""Pins the oracle's element physics against mathematics that fixes it:
finite differences of the energies the reference defines
(spring_constraint.cu:56-71, triangle_stretch_constraint.cu:116-147),
closed forms, invariances and the PSD filter, and numpy's matrix inverse."""

import numpy as np
import pytest

from oracle.oracle import mat3_inverse, spring_element, triangle_element


def spring_energy(xa, xb, L, k):
    return 0.5 * k * (np.linalg.norm(xa - xb) - L) ** 2


def fd_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        g.flat[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


@pytest.mark.parametrize("seed", range(5))
def test_spring_gradient_matches_energy_fd(seed):
    rng = np.random.default_rng(seed)
    xa, xb = rng.normal(size=3), rng.normal(size=3)
    L = 0.7 * np.linalg.norm(xa - xb) * rng.uniform(0.5, 1.5)
    k = 10 ** rng.uniform(2, 4)
    g, _ = spring_element(xa, xb, L, k)
    fd = fd_grad(lambda x: spring_energy(x, xb, L, k), xa)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6 * k)


@pytest.mark.parametrize("seed", range(5))
def test_spring_hessian_stretched_is_exact(seed):
    rng = np.random.default_rng(seed)
    xa, xb = rng.normal(size=3), rng.normal(size=3)
    L = 0.8 * np.linalg.norm(xa - xb)  # stretched: l >= L
    k = 1e3
    _, H = spring_element(xa, xb, L, k)
    fd = np.stack([fd_grad(lambda x: spring_element(x, xb, L, k)[0][j], xa) for j in range(3)])
    np.testing.assert_allclose(H, fd, rtol=1e-5, atol=1e-5 * k)
    np.testing.assert_allclose(H, H.T, rtol=0, atol=1e-12 * k)


def test_spring_hessian_compressed_is_rank_one_psd():
    xa, xb = np.array([0.3, 0.1, -0.2]), np.array([-0.1, 0.4, 0.5])
    l = np.linalg.norm(xa - xb)
    L, k = 1.5 * l, 1e3
    _, H = spring_element(xa, xb, L, k)
    w, V = np.linalg.eigh(H)
    a = k * L / l
    np.testing.assert_allclose(sorted(w), [0, 0, a], atol=1e-9 * a)
    d = (xa - xb) / l
    assert abs(abs(V[:, -1] @ d) - 1) < 1e-9


def test_spring_closed_form_two_particles():
    # rest length 1 stretched to 2 along x: gradient k*(2-1)*x_hat on a.
    g, H = spring_element(np.array([2.0, 0, 0]), np.zeros(3), 1.0, 100.0)
    np.testing.assert_allclose(g, [100.0, 0, 0])
    # a = k L / l = 50; H = (k - a) I + a xx^T
    np.testing.assert_allclose(H, np.diag([100.0, 50.0, 50.0]))


def tri_energy(xr, x, k):
    # E = k A / 2 (cu^2 + cv^2), cu = |F e_u| - 1 (triangle_energy_kernel)
    _, _, dm, area = triangle_element(xr, xr, k)
    a, b = x[1] - x[0], x[2] - x[0]
    wu = a * dm[0, 0] + b * dm[1, 0]
    wv = a * dm[0, 1] + b * dm[1, 1]
    return 0.5 * k * area * ((np.linalg.norm(wu) - 1) ** 2 + (np.linalg.norm(wv) - 1) ** 2)


def random_tri(rng):
    xr = rng.normal(size=(3, 3))
    return xr


@pytest.mark.parametrize("seed", range(6))
def test_triangle_gradient_matches_energy_fd(seed):
    rng = np.random.default_rng(seed)
    xr = random_tri(rng)
    x = xr + 0.2 * rng.normal(size=(3, 3))
    k = 1e3
    g, _, _, _ = triangle_element(xr, x, k)
    fd = fd_grad(lambda y: tri_energy(xr, y.reshape(3, 3), k), x.reshape(9)).reshape(3, 3)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


@pytest.mark.parametrize("seed", range(6))
def test_triangle_hessian_exact_when_stretched(seed):
    rng = np.random.default_rng(seed)
    xr = random_tri(rng)
    # uniform dilation by 1.3 plus small noise -> cu, cv > 0: no term filtered
    c = xr.mean(0)
    x = c + 1.3 * (xr - c) + 0.01 * rng.normal(size=(3, 3))
    k = 1e3
    _, H, _, _ = triangle_element(xr, x, k)
    Hf = H.transpose(0, 2, 1, 3).reshape(9, 9)
    fd = np.stack([fd_grad(lambda y: triangle_element(xr, y.reshape(3, 3), k)[0].reshape(9)[j], x.reshape(9))
                   for j in range(9)])
    np.testing.assert_allclose(Hf, fd, rtol=1e-4, atol=1e-5 * np.abs(fd).max())


@pytest.mark.parametrize("seed", range(6))
def test_triangle_invariances_and_psd(seed):
    rng = np.random.default_rng(seed)
    xr = random_tri(rng)
    x = xr + 0.3 * rng.normal(size=(3, 3))
    k = 1e3
    g, H, _, _ = triangle_element(xr, x, k)
    Hf = H.transpose(0, 2, 1, 3).reshape(9, 9)
    # translation null space: gradients and block rows sum to zero
    np.testing.assert_allclose(g.sum(0), 0, atol=1e-10 * np.abs(g).max())
    np.testing.assert_allclose(H.sum(1), 0, atol=1e-10 * np.abs(H).max())
    # symmetry and PSD (Baraff-Witkin filtering keeps every element PSD)
    np.testing.assert_allclose(Hf, Hf.T, atol=1e-10 * np.abs(Hf).max())
    assert np.linalg.eigvalsh(Hf).min() > -1e-9 * np.abs(Hf).max()
    # rigid rotation of the current state rotates the gradient
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    g2, H2, _, _ = triangle_element(xr, x @ q.T + rng.normal(size=3), k)
    np.testing.assert_allclose(g2, g @ q.T, atol=1e-9 * np.abs(g).max())
    np.testing.assert_allclose(H2[0, 1], q @ H[0, 1] @ q.T, atol=1e-9 * np.abs(H).max())


def test_triangle_rest_state_and_closed_form_stretch():
    xr = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0]])
    g, H, dm, area = triangle_element(xr, xr, 1e3)
    np.testing.assert_allclose(dm, np.eye(2), atol=1e-15)
    assert area == pytest.approx(0.5)
    np.testing.assert_allclose(g, 0, atol=1e-12)
    s, k = 1.25, 200.0
    x = xr.copy()
    x[1, 0] = s
    g, _, _, _ = triangle_element(xr, x, k)
    keff = 0.5 * k
    np.testing.assert_allclose(g[0], [-keff * (s - 1), 0, 0], atol=1e-12)
    np.testing.assert_allclose(g[1], [keff * (s - 1), 0, 0], atol=1e-12)
    np.testing.assert_allclose(g[2], 0, atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_triangle_rest_frame(seed):
    rng = np.random.default_rng(seed)
    xr = random_tri(rng)
    _, _, dm, area = triangle_element(xr, xr, 1.0)
    ab, ac = xr[1] - xr[0], xr[2] - xr[0]
    assert area == pytest.approx(0.5 * np.linalg.norm(np.cross(ab, ac)), rel=1e-12)
    # Dm^-1 maps the rest edge matrix in the triangle's own 2-D frame to I
    lab = np.linalg.norm(ab)
    uc = ab @ ac / lab
    vc = np.sqrt(ac @ ac - uc * uc)
    np.testing.assert_allclose(dm @ np.array([[lab, uc], [0, vc]]), np.eye(2), atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_mat3_inverse_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(3, 3))
    a = a @ a.T + np.eye(3)
    np.testing.assert_allclose(mat3_inverse(a), np.linalg.inv(a), rtol=1e-12, atol=1e-12)


# ---- the reference's guards on degenerate elements -------------------------

def test_spring_coincident_ends_clamped_length():
    """Coincident ends: the length is clamped to kMinLength = 1e-6
    (spring_constraint.cu:35) and the direction d / l is 0.  Rest length 0:
    l = 1e-6 >= L, a = k L / l = 0, so H = k I and the gradient is 0.  Rest
    length 0.1: l < L, H = a d d^T = 0, gradient 0.  Without the clamp both
    would divide by zero."""
    x = np.array([0.3, -0.2, 0.7])
    g, H = spring_element(x, x.copy(), 0.0, 250.0)
    np.testing.assert_array_equal(g, 0.0)
    np.testing.assert_allclose(H, 250.0 * np.eye(3), rtol=0, atol=0)
    g, H = spring_element(x, x.copy(), 0.1, 250.0)
    np.testing.assert_array_equal(g, 0.0)
    np.testing.assert_array_equal(H, 0.0)


def test_triangle_degenerate_rest_frames():
    """Rest determinant lab * v_c <= 1e-6 gives the identity frame
    (triangle_stretch_constraint.cu:95-99).  A collinear triangle has area 0,
    so its forces and blocks vanish; a sliver 2.5e-6 off its base of 0.2 has
    determinant 5e-7, the identity frame and area 0.2 * 2.5e-6 / 2.  Just
    above the threshold (determinant 2e-6) the frame is the real inverse."""
    col = np.array([[0.0, 0, 0], [0.2, 0, 0], [0.5, 0, 0]])
    x = col + np.array([[0, 0.01, 0], [0.02, 0, 0], [0, 0, 0.03]])
    g, H, dm, area = triangle_element(col, x, 1e3)
    np.testing.assert_array_equal(dm, np.eye(2))
    assert area == 0.0
    np.testing.assert_array_equal(g, 0.0)
    np.testing.assert_array_equal(H, 0.0)
    sl = np.array([[0.0, 0, 0], [0.2, 0, 0], [0.1, 0, -2.5e-6]])
    _, _, dm, area = triangle_element(sl, sl, 1e3)
    np.testing.assert_array_equal(dm, np.eye(2))
    assert area == pytest.approx(0.5 * 0.2 * 2.5e-6, rel=1e-9)
    ok = np.array([[0.0, 0, 0], [0.2, 0, 0], [0.1, 0, -1e-5]])
    _, _, dm, area = triangle_element(ok, ok, 1e3)
    np.testing.assert_allclose(dm @ np.array([[0.2, 0.1], [0.0, 1e-5]]), np.eye(2), atol=1e-9)


def test_triangle_collapsed_edge_clamped_norm():
    """A deformed triangle whose first edge has collapsed (x1 = x0): w_u = 0 is
    clamped to kMinNorm = 1e-6 (triangle_stretch_constraint.cu:39), so its
    direction is 0 and only the v term acts: on the unit right triangle
    (D_m^-1 = I, dw_v = (-1, 0, 1)) the gradient is k A c_v w_v_hat on
    corner 2, its negative on corner 0 and 0 on corner 1."""
    xr = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0]])
    x = np.array([[0.0, 0, 0], [0.0, 0, 0], [0.0, 1.5, 0.2]])
    k = 300.0
    g, H, _, area = triangle_element(xr, x, k)
    wv = x[2] - x[0]
    lv = np.linalg.norm(wv)
    expect = k * area * (lv - 1.0) * wv / lv
    np.testing.assert_allclose(g[2], expect, rtol=1e-12)
    np.testing.assert_allclose(g[0], -expect, rtol=1e-12)
    np.testing.assert_allclose(g[1], 0.0, atol=1e-12)
    assert np.all(np.isfinite(H))
    # the u term is gone from the Hessian: corner 1 couples to nothing
    np.testing.assert_allclose(H[1], 0.0, atol=1e-9)
    np.testing.assert_allclose(H[:, 1], 0.0, atol=1e-9)
