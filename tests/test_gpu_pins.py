"""GPU pin projection (an extension: pins as constraints) against the FP64
oracle's restatement of it (tests/test_oracle_pins.py pins the oracle).

* the assembled system is the oracle's eliminated system: pinned rows are
  the identity with right-hand side target - x_n, couplings to pinned
  particles are zero (exactly), other blocks within 2e-6 of their row's
  largest;
* a pinned particle ends every step bitwise at its (FP32) target, also when
  the targets move;
* one step from a shared state matches the oracle's to 1e-5 relative, on the
  packed symmetric storage and on the general one;
* a batch with pin projection matches single simulators bitwise.
"""

import numpy as np
import pytest

from chysx import scenes
from chysx.harness import make_batch, make_sim
from oracle.oracle import oracle_from_scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def pinned_scene(w=24):
    s = scenes.config3(w=w, steps=20)
    top = (w - 1) * w
    s.pin_idx = np.array([top, top + w // 2, top + w - 1, w // 2], np.int32)
    s.pin_targets = (s.pos[s.pin_idx] + np.array([[0.0, 0.0, 0.0], [0.0, 0.01, 0.0], [0.02, 0.0, 0.0],
                                                  [0.0, -0.05, 0.0]])).astype(np.float32)
    return s


@pytest.mark.parametrize("symmetric", [1, 0])
def test_pin_projection_matches_oracle_and_sits_on_targets(cuda, symmetric):
    s = pinned_scene()
    g = make_sim(s, tol=1e-6)
    g.sim.set_autotune(False)
    g.sim.set_launch_config(dict(spmv_symmetric=symmetric, spmv_packed=symmetric))
    g.sim.set_pin_projection(True)
    assert g.sim.pin_projection()
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    oc.set_pin_projection(True)
    tgt = s.pin_targets.copy()
    for k in range(16):
        if k == 8:  # the targets move
            tgt = (tgt + np.array([0.0, 0.02, -0.01], np.float32)).astype(np.float32)
            g.sim.update_pin_targets(tgt)
            oc.set_pins(s.pin_idx, tgt, s.pin_k)
            oc.set_pin_projection(True)
        p0 = g.positions().astype(np.float64)
        v0 = g.velocities().astype(np.float64)
        dx0 = g.sim.get_dx64() if k > 0 else np.zeros((s.n, 3))
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["converged"] and st["relres"] <= 1e-6, (k, st)
        pos = g.positions()
        np.testing.assert_array_equal(pos[s.pin_idx], tgt)
        if k in (0, 8, 15):
            diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
            oc.assemble(p0, v0, im, f, s.dt)
            od, ov, orhs, _ = oc.system()
            pinned = np.zeros(s.n, bool)
            pinned[s.pin_idx] = True
            rows = np.repeat(np.arange(s.n), np.diff(ro))
            couple = pinned[rows] | pinned[co]
            assert np.all(vals[couple] == 0.0)
            for i in s.pin_idx:
                np.testing.assert_array_equal(diag[i], np.eye(3, dtype=np.float32))
            np.testing.assert_allclose(rhs[pinned], orhs[pinned], rtol=1e-6, atol=1e-9)
            scale = np.abs(od).max(axis=(1, 2))
            assert np.all(np.abs(diag - od) <= 2e-6 * scale[:, None, None] + 1e-12)
            assert np.all(np.abs(vals - ov) <= 2e-6 * scale[rows][:, None, None] + 1e-12)
            np.testing.assert_allclose(rhs, orhs, rtol=1e-5, atol=5e-5)
        oc.set_dx(dx0)
        p, v = p0.copy(), v0.copy()
        oc.step(p, v, im, f, s.dt, 1)
        np.testing.assert_array_equal(p[s.pin_idx], tgt.astype(np.float64))
        assert rel(pos, p) <= 1e-5, k


def test_pin_projection_batch_matches_single_cloths(cuda):
    sc = scenes.config4(count=5, wmin=8, wmax=30, steps=4, seed=14)
    b = make_batch(sc, tol=1e-6)
    b.batch.set_pin_projection(True)
    singles = []
    for s in sc:
        g = make_sim(s, tol=1e-6, force_active=False)
        g.sim.set_pin_projection(True)
        singles.append(g)
    for _ in range(4):
        b.batch.step(1.0 / 60.0)
        for g in singles:
            g.sim.step(1.0 / 60.0)
    for k, (s, g) in enumerate(zip(sc, singles)):
        assert b.batch.stats(k)["converged"]
        np.testing.assert_array_equal(b.positions(k), g.positions())
        np.testing.assert_array_equal(b.positions(k)[s.pin_idx], s.pin_targets)


@pytest.mark.parametrize("mode", ["projection", "learned", "projection_learned"])
def test_pin_projection_with_warm_start_modes(cuda, mode):
    """The warm-start stages run after assembly has started every pinned row
    at its target displacement and may move it (the projection combines past
    dx, the learned direction has a component there); the pinned rows' own
    equations (I dx = d) bring them back within the tolerance, and finalize
    puts a projected pin exactly on its target.  Every solve converges, the
    pins sit bitwise on their (moving) targets, and a step matches the
    oracle's from the same state and initial guess."""
    from chysx import warmstart
    s = pinned_scene()
    g = make_sim(s, tol=1e-6)
    g.sim.set_pin_projection(True)
    if mode != "projection":
        g.sim.ws_set_params(warmstart.load_params(warmstart.weights_for(mode))[0])
    g.sim.set_warm_start_mode(mode)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    oc.set_pin_projection(True)
    tgt = s.pin_targets.copy()
    used = []
    for k in range(12):
        if k == 6:  # the targets move
            tgt = (tgt + np.array([0.0, 0.02, -0.01], np.float32)).astype(np.float32)
            g.sim.update_pin_targets(tgt)
            oc.set_pins(s.pin_idx, tgt, s.pin_k)
            oc.set_pin_projection(True)
        p0 = g.positions().astype(np.float64)
        v0 = g.velocities().astype(np.float64)
        dx0 = g.sim.get_dx64() if k > 0 else np.zeros((s.n, 3))
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["converged"] and st["relres"] <= 1e-6, (mode, k, st)
        if k >= 2:
            w = g.sim.ws_stats()
            used.append(w["projection_used"] if mode == "projection" else w["used"])
        pos = g.positions()
        np.testing.assert_array_equal(pos[s.pin_idx], tgt)
        if k in (5, 6, 11):
            oc.set_dx(dx0)
            p, v = p0.copy(), v0.copy()
            oc.step(p, v, im, f, s.dt, 1)
            assert rel(pos, p) <= 1e-5, (mode, k)
    assert np.mean(used) > 0.3, (mode, used)

