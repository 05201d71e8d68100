"""Pins the FP64 reference of the warm-start network: backward against
central finite differences of its forward; mean aggregation and its adjoint
against a dense adjacency product."""

import numpy as np
import pytest

from chysx import scenes
from oracle import warmstart_oracle as W


def small_graph(w=5, seed=0):
    pos, tris = scenes.grid(w)
    import chysx
    d = chysx.topology_query(w * w, tris, False, np.zeros(0, np.int32))
    rng = np.random.default_rng(seed)
    n = w * w
    F = rng.normal(size=(n, W.FIN)) * 0.5
    E = np.zeros((n, 32))
    E[:, :4] = rng.normal(size=(n, 4)) * 0.3
    return n, d["row_offsets"].astype(np.int64), d["col_indices"].astype(np.int64), F, E


def test_mean_aggregation_and_adjoint_dense():
    n, ro, co, F, _ = small_graph(6)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, co[ro[i]:ro[i + 1]]] = 1.0 / (ro[i + 1] - ro[i])
    h = np.random.default_rng(1).normal(size=(n, 7))
    np.testing.assert_allclose(W.mean_agg(h, ro, co), A @ h, atol=1e-14)
    np.testing.assert_allclose(W.mean_agg_T(h, ro, co), A.T @ h, atol=1e-14)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    n, ro, co, F, E = small_graph(5, seed)
    theta = W.init_params(seed, scale=1.0)
    theta += np.random.default_rng(seed + 7).normal(size=theta.size) * 0.05  # nonzero biases
    rng = np.random.default_rng(seed + 3)
    target = rng.normal(size=(n, 3))
    valid = np.ones(n)
    valid[-2:] = 0
    L, g, _ = W.loss_and_grad(theta, F, E, target, valid, ro, co)
    idx = rng.choice(theta.size, 60, replace=False)
    h = 1e-6
    for k in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (W.loss_and_grad(tp, F, E, target, valid, ro, co)[0] -
              W.loss_and_grad(tm, F, E, target, valid, ro, co)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])
    assert W.NPARAM == theta.size == 35104


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.normal(size=(n, n))
    return M @ M.T + n * np.eye(n)


def test_galerkin_projection_exact_when_solution_in_span():
    A = _spd(30, 0)
    rng = np.random.default_rng(1)
    V = rng.normal(size=(30, 4))
    xs = V @ np.array([0.3, -1.2, 2.0, 0.5])
    b = A @ xs
    x, c, kept = W.galerkin_projection(A, b, V)
    assert kept == 4
    np.testing.assert_allclose(x, xs, rtol=1e-10, atol=1e-12)


def test_galerkin_projection_minimises_energy_and_drops_collinear():
    A = _spd(40, 2)
    rng = np.random.default_rng(3)
    b = rng.normal(size=40)
    V = rng.normal(size=(40, 3))
    V = np.concatenate([V, V[:, :1] * 2.0], 1)  # a collinear column
    x, c, kept = W.galerkin_projection(A, b, V)
    assert kept == 3
    phi = lambda y: 0.5 * y @ A @ y - b @ y  # noqa: E731
    for _ in range(20):
        dc = rng.normal(size=4) * 1e-3
        assert phi(V @ (c + dc)) >= phi(x) - 1e-12
    # the residual is A-orthogonal to span(V): V^T (b - A x) = 0
    np.testing.assert_allclose(V.T @ (b - A @ x), 0.0, atol=1e-9)


def test_linesearch_loss_gradient_matches_finite_differences():
    A = _spd(25, 4)
    rng = np.random.default_rng(5)
    e = rng.normal(size=25)
    r = A @ e
    d = rng.normal(size=25)

    def L_of(dd):
        return W.linesearch_loss_grad(dd, r, A @ dd, e @ r)[0]

    L, g = W.linesearch_loss_grad(d, r, A @ d, e @ r)
    h = 1e-6
    fd = np.array([(L_of(d + h * np.eye(25)[i]) - L_of(d - h * np.eye(25)[i])) / (2 * h) for i in range(25)])
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-9)
    assert 0.0 <= L <= 1.0


def test_linesearch_loss_limits():
    A = _spd(20, 6)
    rng = np.random.default_rng(7)
    e = rng.normal(size=20)
    r = A @ e
    L, g = W.linesearch_loss_grad(3.0 * e, r, A @ (3.0 * e), e @ r)
    assert abs(L) < 1e-12 and np.abs(g).max() < 1e-10
    # a direction A-orthogonal to e leaves all of the error
    d = rng.normal(size=20)
    d -= (d @ r) / (e @ r) * e
    L, _ = W.linesearch_loss_grad(d, r, A @ d, e @ r)
    assert abs(L - 1.0) < 1e-12
