"""GPU checks of the learned warm start.

The network runs with BF16 weights, inputs and activations on MFMA and FP32
accumulation, so it is compared with the FP64 reference
(oracle/warmstart_oracle.py) evaluated on the same BF16-rounded weights and
inputs; the remaining difference is the BF16 rounding of the activations:
  * forward:  |out_gpu - out_ref| / |out_ref| <= 3e-2 (first 3 columns);
  * gradient: cosine(g_gpu, g_ref) >= 0.995, |g_gpu - g_ref| / |g_ref| <= 0.1,
    per-matrix relative error <= 0.15;
  * loss within 2e-2 relative.
A wrong MFMA lane layout or a transposed operand gives O(1) errors.
"""

import os

import numpy as np
import pytest

import chysx
from chysx import scenes
from chysx.harness import make_sim
from oracle import warmstart_oracle as W

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def quantize_params(theta):
    p = W.unpack(theta)
    for k in p:
        if k.startswith("W"):
            p[k] = bf16(p[k]).astype(np.float64)
    return W.pack(p)


def graph(w=20, seed=0):
    _, tris = scenes.grid(w)
    d = chysx.topology_query(w * w, tris, False, np.zeros(0, np.int32))
    rng = np.random.default_rng(seed)
    n = w * w
    F = np.zeros((n, 32), np.float32)
    F[:, :15] = rng.normal(size=(n, 15)) * 0.7
    E = np.zeros((n, 32), np.float32)
    E[:, :4] = rng.normal(size=(n, 4)) * 0.3
    return n, d["row_offsets"], d["col_indices"], bf16(F), bf16(E)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_fp64_reference(cuda, seed):
    n, ro, co, F, E = graph(21, seed)
    theta = W.init_params(seed).astype(np.float32)
    theta += np.random.default_rng(seed).normal(size=theta.size).astype(np.float32) * 0.02
    out = chysx._chysx_native.wsnet_forward(theta, ro, co, F, E)
    ref = W.forward(quantize_params(theta.astype(np.float64)), F.astype(np.float64), E.astype(np.float64),
                    ro.astype(np.int64), co.astype(np.int64))
    assert rel(out[:, :3], ref[:, :3]) <= 3e-2
    assert rel(out, ref) <= 3e-2


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_fp64_reference(cuda, seed):
    n, ro, co, F, E = graph(19, seed)
    rng = np.random.default_rng(seed + 10)
    theta = W.init_params(seed + 5).astype(np.float32)
    theta += rng.normal(size=theta.size).astype(np.float32) * 0.02
    target = rng.normal(size=(n, 3)).astype(np.float32)
    loss, g = chysx._chysx_native.wsnet_loss_grad(theta, ro, co, F, E, target)
    L, gr, _ = W.loss_and_grad(quantize_params(theta.astype(np.float64)), F.astype(np.float64),
                               E.astype(np.float64), target.astype(np.float64), np.ones(n),
                               ro.astype(np.int64), co.astype(np.int64))
    assert abs(loss - L) <= 2e-2 * L
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos >= 0.995, cos
    assert rel(g, gr) <= 0.1
    o = 0
    for name, shp in W.SHAPES:
        k = int(np.prod(shp))
        if name.startswith("W"):
            assert rel(g[o:o + k], gr[o:o + k]) <= 0.15, name
        o += k


def test_untrained_network_falls_back_to_previous_guess(cuda):
    """A random network predicts worse guesses than the previous step's dx:
    the solver must keep the previous guess, so the run is bitwise the plain
    solver's."""
    s = scenes.config2(w=40, steps=6)
    runs = []
    for mode in ["previous", "learned"]:
        g = make_sim(s, tol=1e-6)
        g.sim.ws_init_params(3)
        g.sim.set_warm_start_mode(mode)
        used = []
        for k in range(6):
            g.sim.step(s.dt)
            if mode == "learned" and k >= 2:  # the network runs once two solutions exist
                st = g.sim.ws_stats()
                used.append(st["used"])
                assert st["r0_previous"] > 0 and st["r0_learned"] > 0
        runs.append((g.positions(), used))
    if not any(runs[1][1]):
        np.testing.assert_array_equal(runs[0][0], runs[1][0])


def test_online_training_reduces_loss(cuda):
    s = scenes.config2(w=48, steps=60, seed=9)
    g = make_sim(s, tol=1e-6)
    g.sim.ws_init_params(1)
    g.sim.ws_set_training(True, 2e-3)
    losses = []
    for k in range(60):
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["converged"]
        losses.append(g.sim.ws_stats()["loss"])
    assert np.mean(losses[-10:]) < 0.5 * np.mean(losses[2:12]), (losses[:12], losses[-10:])
