"""Pins the oracle's FEM shear and dihedral bending against mathematics that
fixes them (the reference ships no fixtures for either):

  * shear (triangle_stretch_constraint.cu:50-110 with theta = pi/4,
    cloth_simulator.cu:423-449): D_m^-1 R(theta)^-1 against numpy, gradient
    against finite differences of the reference energy, and a simple-shear
    closed form that the stretch element does not resist to first order;
  * bending (bending_constraint.h:1-129, bending_constraint.cu:50-246):
    a_i = d theta / d x_i against finite differences of theta, the angle of a
    hinge folded by a known angle, translation / rotation invariance, the
    dihedral list of cloth_simulator.cu:451-514 against a brute-force
    restatement, and the widened pattern against a brute-force pair set;
  * the whole assembly with stretch + shear + bending + pins at rest equals
    the derivative of the assembled gradient (every slot of the scatter).
"""

import numpy as np
import pytest

from chysx import scenes
from oracle.oracle import OracleCloth, SHEAR_THETA, dihedral_element, find_dihedrals, triangle_element

C32 = float(np.cos(np.float32(SHEAR_THETA)))
S32 = float(np.sin(np.float32(SHEAR_THETA)))


def fd_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        g.flat[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def tri_energy(dm, area, x, k):
    a, b = x[1] - x[0], x[2] - x[0]
    wu = a * dm[0, 0] + b * dm[1, 0]
    wv = a * dm[0, 1] + b * dm[1, 1]
    return 0.5 * k * area * ((np.linalg.norm(wu) - 1) ** 2 + (np.linalg.norm(wv) - 1) ** 2)


@pytest.mark.parametrize("seed", range(4))
def test_shear_rest_frame_is_rotated_stretch_frame(seed):
    xr = np.random.default_rng(seed).normal(size=(3, 3))
    _, _, dm0, a0 = triangle_element(xr, xr, 1.0, 0.0)
    _, _, dm1, a1 = triangle_element(xr, xr, 1.0, SHEAR_THETA)
    rinv = np.array([[C32, S32], [-S32, C32]])
    np.testing.assert_allclose(dm1, dm0 @ rinv, rtol=1e-6, atol=1e-7 * np.abs(dm0).max())
    assert a1 == a0


@pytest.mark.parametrize("seed", range(5))
def test_shear_gradient_matches_energy_fd(seed):
    rng = np.random.default_rng(seed)
    xr = rng.normal(size=(3, 3))
    x = xr + 0.2 * rng.normal(size=(3, 3))
    k = 500.0
    g, H, dm, area = triangle_element(xr, x, k, SHEAR_THETA)
    fd = fd_grad(lambda y: tri_energy(dm, area, y.reshape(3, 3), k), x.reshape(9)).reshape(3, 3)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    Hf = H.transpose(0, 2, 1, 3).reshape(9, 9)
    np.testing.assert_allclose(Hf, Hf.T, atol=1e-10 * np.abs(Hf).max())
    assert np.linalg.eigvalsh(Hf).min() > -1e-9 * np.abs(Hf).max()


def test_simple_shear_closed_form():
    """Right triangle (0,0), (1,0), (0,1) sheared x' = x + gamma y: the shear
    element's axes are (1,-1)/sqrt2 and (1,1)/sqrt2, so
    E(gamma) = k A / 2 [(|(1-gamma, -1)|/sqrt2 - 1)^2 + (|(1+gamma, 1)|/sqrt2 - 1)^2]
    and only vertex 2 moves: grad_2.x = dE/dgamma.  The stretch element's
    response is second order in gamma."""
    xr = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0]])
    k, A = 300.0, 0.5
    c, s = C32, S32

    def E(gm):
        u = np.array([c * 1 - s * gm, -s])      # F (c, -s)
        v = np.array([s * 1 + c * gm, c])       # F (s, c)
        return 0.5 * k * A * ((np.linalg.norm(u) - 1) ** 2 + (np.linalg.norm(v) - 1) ** 2)

    for gm in [0.01, 0.1, -0.2]:
        x = xr.copy()
        x[2, 0] = gm
        g, _, _, _ = triangle_element(xr, x, k, SHEAR_THETA)
        dE = (E(gm + 1e-6) - E(gm - 1e-6)) / 2e-6
        assert g[2, 0] == pytest.approx(dE, rel=1e-6, abs=1e-9)
        gs, _, _, _ = triangle_element(xr, x, k, 0.0)
        if abs(gm) <= 0.01:
            assert abs(gs[2, 0]) < 0.05 * abs(g[2, 0])


@pytest.mark.parametrize("seed", range(6))
def test_dihedral_gradient_matches_fd_of_angle(seed):
    x = np.random.default_rng(seed).normal(size=(4, 3))
    th, a, ok = dihedral_element(x)
    assert ok
    fd = fd_grad(lambda y: dihedral_element(y.reshape(4, 3))[0], x.reshape(12), h=1e-7).reshape(4, 3)
    np.testing.assert_allclose(a, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())
    # translation and rotation invariance of theta
    np.testing.assert_allclose(a.sum(0), 0.0, atol=1e-10 * np.abs(a).max())
    np.testing.assert_allclose(np.cross(x, a).sum(0), 0.0, atol=1e-9 * np.abs(a).max() * np.abs(x).max())


@pytest.mark.parametrize("alpha", [0.0, 0.3, -0.7, 2.0])
def test_dihedral_angle_of_a_folded_hinge(alpha):
    """Hinge v0 -> v1 along x; wing v2 in the +y half plane, wing v3 rotated
    by alpha about the hinge from the flat position (-y): |theta| = alpha and
    the sign follows alpha."""
    v0, v1 = np.zeros(3), np.array([1.0, 0, 0])
    v2 = np.array([0.3, 1.0, 0.0])
    v3 = np.array([0.6, -np.cos(alpha), np.sin(alpha)])
    th, _, ok = dihedral_element(np.stack([v0, v1, v2, v3]))
    assert ok
    assert abs(abs(th) - abs(alpha)) < 1e-12
    th2, _, _ = dihedral_element(np.stack([v0, v1, v2, v3 * np.array([1, 1, -1])]))
    assert th2 == pytest.approx(-th, abs=1e-12)


def test_degenerate_dihedral_is_invalid():
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, 0, 0], [0.2, 1.0, 0]])  # wing on the hinge line
    _, a, ok = dihedral_element(x)
    assert not ok and not a.any()


def brute_dihedrals(tris):
    third = {}
    for t in tris:
        v = [int(q) for q in t]
        third[(v[0], v[1])] = v[2]
        third[(v[1], v[2])] = v[0]
        third[(v[2], v[0])] = v[1]
    out = []
    for (v0, v1), v2 in third.items():
        if v0 < v1 and (v1, v0) in third:
            out.append((v0, v1, v2, third[(v1, v0)]))
    return sorted(out)


@pytest.mark.parametrize("w", [3, 6, 11])
def test_grid_dihedrals(w):
    _, tris = scenes.grid(w)
    d = find_dihedrals(tris)
    assert [tuple(r) for r in d] == brute_dihedrals(tris)
    n_edges = 2 * w * (w - 1) + (w - 1) ** 2
    assert d.shape[0] == n_edges - 4 * (w - 1)  # every interior edge once


def test_random_mesh_dihedrals_and_pattern():
    rng = np.random.default_rng(7)
    n = 60
    tris = rng.integers(0, n, size=(150, 3)).astype(np.int32)
    tris = tris[(tris[:, 0] != tris[:, 1]) & (tris[:, 1] != tris[:, 2]) & (tris[:, 0] != tris[:, 2])]
    d = find_dihedrals(tris)
    assert [tuple(r) for r in d] == brute_dihedrals(tris)
    pos = rng.normal(size=(n, 3))
    oc = OracleCloth(64)
    oc.set_mesh(tris, n)
    oc.build_stretch(pos, 1e3)
    oc.build_shear(pos, 1e3)
    oc.build_bending(pos, 1e-2)
    oc.set_params((0, -9.8, 0), 0.0, False, False, 0.0, 1)
    oc.build_topology()
    ro, co = oc.pattern()
    pairs = set()
    for t in tris:
        for a in t:
            for b in t:
                if a != b:
                    pairs.add((int(a), int(b)))
    for q in d:
        for a in q:
            for b in q:
                if a != b:
                    pairs.add((int(a), int(b)))
    got = {(r, int(c)) for r in range(n) for c in co[ro[r]:ro[r + 1]]}
    assert got == pairs
    assert all(np.all(np.diff(co[ro[r]:ro[r + 1]]) > 0) for r in range(n))
    # slot round trip for shear and dihedral slots
    _, _, ss, ds = oc.bending(tris.shape[0])
    for t in range(tris.shape[0]):
        for a in range(3):
            for b in range(3):
                sl = ss[t, 3 * a + b]
                r, c = tris[t, a], tris[t, b]
                if r == c:
                    assert sl == -r - 1
                else:
                    assert co[sl] == c and ro[r] <= sl < ro[r + 1]
    for k in range(d.shape[0]):
        for a in range(4):
            for b in range(4):
                sl = ds[k, 4 * a + b]
                r, c = d[k, a], d[k, b]
                if r == c:
                    assert sl == -r - 1
                else:
                    assert co[sl] == c and ro[r] <= sl < ro[r + 1]


def test_grid_pattern_with_bending_closed_form():
    """W x W grid with bending: rows gain the wing-to-wing pairs (for each
    incident triangle, the vertex across its far edge)."""
    w = 9
    pos, tris = scenes.grid(w)
    oc = OracleCloth(64)
    oc.set_mesh(tris, w * w)
    oc.build_stretch(pos, 1e3)
    oc.build_bending(pos, 1e-2)
    oc.set_params((0, -9.8, 0), 0.0, False, False, 0.0, 1)
    oc.build_topology()
    ro, co = oc.pattern()
    d = find_dihedrals(tris)
    pairs = set()
    for t in tris:
        for a in t:
            for b in t:
                if a != b:
                    pairs.add((int(a), int(b)))
    for q in d:
        for a in q:
            for b in q:
                if a != b:
                    pairs.add((int(a), int(b)))
    assert ro[-1] == len(pairs)
    # interior vertex (not within 2 of the border): 6 triangle neighbours and
    # one partner across the far edge of each of its 6 triangles
    r = 4 * w + 4
    assert ro[r + 1] - ro[r] == 12


def test_flat_rest_state_and_assembly_fd_with_shear_and_bending():
    s = scenes.config1()
    n = s.n
    rest = s.pos.astype(np.float64)

    def system_at(x):
        oc = OracleCloth(64)
        oc.set_mesh(s.tris, n)
        oc.build_stretch(rest, s.stretch_k)
        oc.build_shear(rest, 400.0)
        oc.build_bending(rest, 0.05)
        oc.set_pins(s.pin_idx, s.pin_targets, s.pin_k)
        oc.set_params(s.gravity, 0.0, False, False, 0.0, 1)
        p = np.array(x, np.float64)
        oc.step(p, np.zeros((n, 3)), s.inv_mass, None, s.dt, 1)
        ro, co = oc.pattern()
        diag, vals, rhs, _ = oc.system()
        _, rest_angles, _, _ = oc.bending(s.tris.shape[0])
        return ro, co, diag, vals, rhs, rest_angles

    ro, co, diag, vals, rhs0, rest_angles = system_at(rest)
    np.testing.assert_allclose(rest_angles, 0.0, atol=1e-12)  # the grid is flat
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        A[3 * r:3 * r + 3, 3 * r:3 * r + 3] = diag[r]
        for q in range(ro[r], ro[r + 1]):
            A[3 * r:3 * r + 3, 3 * co[q]:3 * co[q] + 3] = vals[q]
    A -= np.diag(np.repeat((1.0 / s.inv_mass) / s.dt ** 2, 3))
    np.testing.assert_allclose(A, A.T, atol=1e-9 * np.abs(A).max())
    h = 1e-6
    rng = np.random.default_rng(1)
    for j in rng.choice(3 * n, 24, replace=False):
        xp = rest.copy().reshape(-1)
        xm = rest.copy().reshape(-1)
        xp[j] += h
        xm[j] -= h
        col = -(system_at(xp.reshape(n, 3))[4] - system_at(xm.reshape(n, 3))[4]).reshape(-1) / (2 * h)
        np.testing.assert_allclose(A[:, j], col, rtol=1e-5, atol=1e-5 * np.abs(A).max())
