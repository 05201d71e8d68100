"""GPU checks of the learned warm start.

The network runs with BF16 weights, inputs and activations on MFMA and FP32
accumulation (at inference each layer is one fused kernel: neighbour mean
inside the operand load; the simulator folds the decoder into the last
layer), so it is compared with the FP64 reference
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
    # the simulator's path: decoder folded into the last fused layer, the
    # three direction columns only, from the same BF16-rounded last layer
    d = chysx._chysx_native.wsnet_forward(theta, ro, co, F, E, direction_only=True)
    assert d.shape == (n, 3)
    assert rel(d, ref[:, :3]) <= 3e-2
    assert rel(d, out[:, :3]) <= 1e-5


def test_forward_matches_fp64_reference_wide_and_empty_rows(cuda):
    """Rows with up to 12 neighbours (the grid's second ring: more neighbour
    slots than one gather batch of the fused layer holds) and rows with none
    (their neighbour mean is zero): the fused layers and the folded decoder
    still match the FP64 reference."""
    w = 17
    n = w * w
    rows = []
    for i in range(w):
        for j in range(w):
            nb = []
            if i < w - 1:  # the last grid row is isolated
                for di in range(-2, 3):
                    for dj in range(-2, 3):
                        if 0 < abs(di) + abs(dj) <= 2 and 0 <= i + di < w - 1 and 0 <= j + dj < w:
                            nb.append((i + di) * w + j + dj)
            rows.append(sorted(nb))
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(r) for r in rows])
    co = np.array([c for r in rows for c in r], np.int32)
    assert np.diff(ro).max() == 12 and np.diff(ro).min() == 0
    rng = np.random.default_rng(7)
    F = np.zeros((n, 32), np.float32)
    F[:, :15] = rng.normal(size=(n, 15)) * 0.7
    E = np.zeros((n, 32), np.float32)
    E[:, :4] = rng.normal(size=(n, 4)) * 0.3
    F, E = bf16(F), bf16(E)
    theta = W.init_params(3).astype(np.float32)
    theta += np.random.default_rng(3).normal(size=theta.size).astype(np.float32) * 0.02
    out = chysx._chysx_native.wsnet_forward(theta, ro, co, F, E)
    ref = W.forward(quantize_params(theta.astype(np.float64)), F.astype(np.float64), E.astype(np.float64),
                    ro.astype(np.int64), co.astype(np.int64))
    assert rel(out, ref) <= 3e-2
    d = chysx._chysx_native.wsnet_forward(theta, ro, co, F, E, direction_only=True)
    assert rel(d, ref[:, :3]) <= 3e-2
    assert rel(d, out[:, :3]) <= 1e-5


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


def _system(dump):
    import scipy.sparse as sp
    diag, ro, co, vals, rhs, _ = dump
    n = diag.shape[0]
    A = sp.bsr_matrix((vals.astype(np.float64), co, ro), shape=(3 * n, 3 * n)) + \
        sp.block_diag(list(diag.astype(np.float64)), format="bsr")
    return A.tocsr(), rhs.astype(np.float64).reshape(-1)


def test_projection_matches_numpy_galerkin(cuda):
    """The GPU projection stage against an independent FP64 computation on the
    solved system (debug dump) and the FP32-rounded history of converged dx:
    kept directions, acceptance, and the relative residuals of dx_prev and of
    the projected guess (predicted on the GPU from Gram sums)."""
    s = scenes.config3(w=40, steps=10)
    g = make_sim(s, tol=1e-6)
    g.sim.set_warm_start_mode("projection")
    g.sim.ws_set_history(4)
    hist = []
    checked = 0
    for k in range(10):
        g.sim.step(s.dt)
        st = g.sim.ws_stats()
        if len(hist) >= 2:
            A, b = _system(g.sim.debug_dump_last_solve())
            V = np.stack([h.reshape(-1).astype(np.float64) for h in hist[::-1][:4]], 1)
            x, c, kept = W.galerkin_projection(A, b, V)
            nb = np.linalg.norm(b)
            r_prev = np.linalg.norm(b - A @ V[:, 0]) / nb
            r_proj = np.linalg.norm(b - A @ x) / nb
            phi = lambda y: 0.5 * y @ (A @ y) - b @ y  # noqa: E731
            assert st["projection_dims"] == kept
            assert st["projection_used"] == int(phi(x) < phi(V[:, 0]))
            assert abs(st["r0_previous"] - r_prev) <= 1e-6 * r_prev, (k, st, r_prev)
            assert abs(st["r0_projection"] - r_proj) <= 1e-5 * r_proj, (k, st, r_proj)
            checked += 1
        else:
            assert st["projection_used"] == 0 and st["r0_projection"] == 0.0
        hist.append(g.sim.get_dx64().astype(np.float32))
    assert checked == 8


def test_projection_cuts_iterations_same_solution(cuda):
    """Fewer iterations, and every step still the oracle's step: at steps 15,
    25 and 35 the oracle steps from the GPU's state (tol 1e-11) and the GPU's
    next positions must match to 1e-5.  (Whole trajectories of this wind cloth
    are not compared across warm starts: it amplifies 1e-6-level solve
    differences chaotically.)"""
    from oracle.oracle import oracle_from_scene
    s = scenes.config3(w=64, steps=40)
    its = {}
    for mode in ["previous", "projection"]:
        g = make_sim(s, tol=1e-6)
        g.sim.set_warm_start_mode(mode)
        oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
        it = []
        for k in range(40):
            check = mode == "projection" and k in (15, 25, 35)
            if check:
                p0 = g.positions().astype(np.float64)
                v0 = g.velocities().astype(np.float64)
                dx0 = g.sim.get_dx64()
            g.sim.step(s.dt)
            st = g.sim.last_solve_stats()
            assert st["converged"], (mode, k, st)
            it.append(st["iters"])
            if check:
                oc.set_dx(dx0)
                oc.step(p0, v0, im, f, s.dt, 1)
                assert rel(g.positions(), p0) <= 1e-5, k
        its[mode] = float(np.mean(it[10:]))
    assert its["projection"] < 0.9 * its["previous"], its


@pytest.mark.parametrize("base,mode", [("previous", "learned"), ("projection", "projection_learned")])
def test_zero_network_leaves_base_guess_unchanged(cuda, base, mode):
    """A network with zero output proposes no direction: each learned mode
    must run bitwise like its base mode."""
    s = scenes.config2(w=40, steps=8)
    runs = []
    for m in [base, mode]:
        g = make_sim(s, tol=1e-6)
        g.sim.ws_set_params(np.zeros(chysx._chysx_native.ws_num_params(), np.float32))
        g.sim.set_warm_start_mode(m)
        for k in range(8):
            g.sim.step(s.dt)
            if m == mode:
                assert g.sim.ws_stats()["used"] == 0
        runs.append(g.positions())
    np.testing.assert_array_equal(runs[0], runs[1])


def test_learned_direction_applied_and_steps_stay_exact(cuda):
    """A random network's direction lowers the energy whenever it is not
    A-orthogonal to the error, so it is applied; every solve still converges
    and each checked step matches the oracle's step from the same state."""
    from oracle.oracle import oracle_from_scene
    s = scenes.config2(w=40, steps=12)
    g = make_sim(s, tol=1e-6)
    g.sim.ws_init_params(3)
    g.sim.set_warm_start_mode("learned")
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    used = []
    for k in range(12):
        check = k in (5, 10)
        if check:
            p0 = g.positions().astype(np.float64)
            v0 = g.velocities().astype(np.float64)
            dx0 = g.sim.get_dx64()
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["converged"]
        st = g.sim.ws_stats()
        if k >= 2:  # the network runs once two solutions exist
            used.append(st["used"])
            assert st["r0_previous"] > 0 and st["r0_learned"] > 0 and st["projection_used"] == 0
        if check:
            oc.set_dx(dx0)
            oc.step(p0, v0, im, f, s.dt, 1)
            assert rel(g.positions(), p0) <= 1e-5, k
    assert np.mean(used) > 0.5


def test_fused_inference_matches_training_path(cuda):
    """Inference runs the features and the encoder in one kernel, the layers
    on the compact edge features and the decoder in the last layer's epilogue;
    training (here with a zero learning rate, so the parameters stay put) runs
    the features kernel, separate GEMMs and the 32 edge-feature columns.  From
    the same state both give the same direction: the predicted residual after
    the line search agrees to FP32 rounding of the decoder, and so do the
    steps.  A fused kernel that read the wrong rows or columns of its tiles
    would give O(1) differences."""
    from chysx import warmstart
    theta, _ = warmstart.load_params()
    s = scenes.config3(w=72, steps=6)
    sims = []
    for train in (False, True):
        g = make_sim(s, tol=1e-6)
        g.sim.ws_set_params(theta)
        g.sim.set_warm_start_mode("learned")
        if train:
            g.sim.ws_set_training(True, 0.0)
        sims.append(g)
    for k in range(3):
        for g in sims:
            g.sim.step(s.dt)
            assert g.sim.last_solve_stats()["converged"]
    a, b = sims[0].sim.ws_stats(), sims[1].sim.ws_stats()
    assert a["used"] == 1 and b["used"] == 1, (a, b)
    assert a["r0_learned"] > 0
    assert abs(a["r0_learned"] - b["r0_learned"]) <= 1e-4 * a["r0_learned"], (a, b)
    assert rel(sims[0].positions(), sims[1].positions()) <= 1e-6
    np.testing.assert_array_equal(sims[0].sim.ws_get_params(), sims[1].sim.ws_get_params())


@pytest.mark.parametrize("mode", ["learned", "projection_learned"])
def test_warm_start_on_bending_cloth_matches_oracle(cuda, mode):
    """The reference demo's crumpled drape (stretch, shear and dihedral
    bending: 9-float blocks and rows of up to 12 neighbours, more slots than
    one gather batch of the fused layers), stepped with the committed
    parameters: every solve converges, the direction runs, and a step from
    the simulator's state matches the FP64 oracle's step from that state."""
    from chysx import warmstart
    from oracle.oracle import oracle_from_scene
    theta, _ = warmstart.load_params(warmstart.weights_for(mode))
    s = scenes.drape(w=24, seed=5)
    g = make_sim(s, tol=1e-6, springs_active=False, force_active=False)
    g.sim.ws_set_params(theta)
    g.sim.set_warm_start_mode(mode)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, springs_active=False, force_active=False, tol=1e-11,
                                        max_iter=100000)
    used = []
    for k in range(10):
        check = k == 7
        if check:
            p0 = g.positions().astype(np.float64)
            v0 = g.velocities().astype(np.float64)
            dx0 = g.sim.get_dx64()
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["converged"], (mode, k)
        if k >= 2:
            used.append(g.sim.ws_stats()["used"])
        if check:
            oc.set_dx(dx0)
            oc.step(p0, v0, im, f, s.dt, 1)
            assert rel(g.positions(), p0) <= 1e-5, (mode, k)
    assert np.mean(used) > 0.3, (mode, used)


def test_residual_acceptance_falls_back_when_residual_rises(cuda):
    """With the residual rule the network's direction is kept only when the
    predicted residual drops below the previous step's guess; otherwise the
    solve starts from dx_prev.  The trained network, 40 steps of a wind cloth:
    every kept direction lowered the residual, steps fell back, and the
    solves all converge."""
    s = scenes.config3(w=64, steps=40)
    from chysx import warmstart
    theta, _ = warmstart.load_params()
    g = make_sim(s, tol=1e-6)
    g.sim.ws_set_params(theta)
    g.sim.set_warm_start_mode("learned")
    with pytest.raises(ValueError):
        g.sim.ws_set_acceptance("sometimes")
    g.sim.ws_set_acceptance("residual")
    used = []
    for k in range(40):
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["converged"]
        st = g.sim.ws_stats()
        if k >= 2:
            used.append(st["used"])
            if st["used"]:
                assert st["r0_learned"] < st["r0_previous"], (k, st)
    assert np.mean(used) < 1.0


def test_online_training_reduces_loss(cuda):
    """Online training from dx_prev: the line-search loss (share of the A-norm
    error of dx_prev the direction leaves) starts near 1 for a fresh network
    and falls well below it."""
    s = scenes.config2(w=48, steps=120, seed=9)
    g = make_sim(s, tol=1e-6)
    g.sim.ws_init_params(1)
    g.sim.set_warm_start_mode("learned")
    g.sim.ws_set_training(True, 2e-3)
    losses = []
    for k in range(120):
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["converged"]
        if k >= 2:
            losses.append(g.sim.ws_stats()["loss"])
    losses = np.array(losses)
    assert np.all((losses >= -1e-6) & (losses <= 1 + 1e-6))
    assert losses[0] > 0.8
    assert np.mean(losses[-20:]) < 0.6 * np.mean(losses[:5]), (losses[:5], losses[-20:])


def test_trained_network_gains_over_previous_guess(cuda):
    """The committed parameters (trained on the GPU from dx_prev over at least
    50000 x 256^2 cloth rows) on a held-out 128x128 cloth of the training
    distribution: the learned warm start needs fewer PCG iterations than the
    previous-step guess, and the projection fewer still."""
    from chysx import warmstart
    theta, meta = warmstart.load_params()
    assert meta["mode"] == "learned" and meta["pairs"] * meta["width"] ** 2 >= 50000 * 256 ** 2
    scene = scenes.config5_train(count=1, w=128, seed=12345)[0]
    its = {}
    for mode in ["previous", "learned", "projection"]:
        g = make_sim(scene, tol=1e-6)
        g.sim.ws_set_params(theta)
        g.sim.set_warm_start_mode(mode)
        it = []
        for k in range(40):
            g.sim.step(scene.dt)
            st = g.sim.last_solve_stats()
            assert st["converged"], (mode, k)
            it.append(st["iters"])
        its[mode] = float(np.mean(it[10:]))
    assert its["learned"] <= 0.95 * its["previous"], its
    assert its["projection"] <= 0.85 * its["previous"], its


@pytest.mark.parametrize("mode", ["projection", "learned", "projection_learned"])
def test_batch_warm_start_modes_match_oracle_and_single_cloths(cuda, mode):
    """An uneven batch (config-4 recipe, W in 8..40) stepped in each warm-start
    mode with the committed parameters: every cloth converges, its step from
    the batch's state matches the FP64 oracle (the warm start only moves the
    initial guess), the warm start ran, and the batch is bitwise equal to the
    same cloths stepped one by one in the same mode (every stage is row-local
    or summed per cloth in a fixed order)."""
    from chysx import warmstart
    from chysx.harness import make_batch
    from oracle.oracle import oracle_from_scene
    theta, _ = warmstart.load_params()
    sc = scenes.config4(count=10, wmin=8, wmax=40, steps=8, seed=21)
    b = make_batch(sc, tol=1e-6)
    b.batch.ws_set_params(theta)
    b.batch.set_warm_start_mode(mode)
    singles = []
    for s in sc:
        g = make_sim(s, tol=1e-6, force_active=False)
        g.sim.ws_set_params(theta)
        g.sim.set_warm_start_mode(mode)
        singles.append(g)
    used = []
    for k in range(6):
        if k == 5:
            pos0 = b.pos.cpu().numpy().astype(np.float64)
            vel0 = b.vel.cpu().numpy().astype(np.float64)
        b.batch.step(1.0 / 60.0)
        for g in singles:
            g.sim.step(1.0 / 60.0)
        for c in range(len(sc)):
            st = b.batch.stats(c)
            assert st["converged"] and st["relres"] <= 1e-6, (mode, k, c, st)
            assert st["iters"] == singles[c].sim.last_solve_stats()["iters"], (mode, k, c)
            if k >= 2:
                w = b.batch.ws_stats(c)
                used.append(w["projection_used"] if mode == "projection" else w["used"])
    assert np.mean(used) > 0.3, (mode, np.mean(used))
    for c, s in enumerate(sc):
        np.testing.assert_array_equal(b.positions(c), singles[c].positions())
        a = b.offsets[c]
        oc, _, _, im, _ = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000, force_active=False)
        p = np.ascontiguousarray(pos0[a:a + s.n])
        v = np.ascontiguousarray(vel0[a:a + s.n])
        oc.step(p, v, im, None, 1.0 / 60.0, 1)
        assert rel(b.positions(c), p) <= 1e-5, (mode, c)
