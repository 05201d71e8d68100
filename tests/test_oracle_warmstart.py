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
