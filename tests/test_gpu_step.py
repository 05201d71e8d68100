"""GPU parity of the implicit cloth step against the FP64 oracle.

Tolerances (FP32 assembly / SpMV / vectors, FP64 iterate and reductions):
  * pattern, column indices, slot-derived layout: bit-exact;
  * Hessian blocks: |d| <= 2e-6 * max|block| (FP32 rounding of ~20 terms);
  * every solve: true relative residual <= 1e-6, recomputed here in FP64 from
    the dumped system;
  * positions after one step from a shared state: <= 1e-5 relative;
  * trajectories over short horizons: <= 1e-5 relative.  Long trajectories of
    these configurations amplify any perturbation (the FP64 oracle at 1e-6
    vs at 1e-12 already differs by ~1e-4 after a few hundred steps), so long
    runs are checked step by step from re-synchronised states.
"""

import os

import numpy as np
import pytest

from chysx import scenes
from chysx.harness import make_batch, make_sim
from oracle.oracle import oracle_from_scene

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cloth10_golden.npz")
CONVERGED = 2
STOPPED = 3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def dense_system(diag, ro, co, vals):
    n = diag.shape[0]
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        A[3 * r:3 * r + 3, 3 * r:3 * r + 3] = diag[r]
        for q in range(ro[r], ro[r + 1]):
            A[3 * r:3 * r + 3, 3 * co[q]:3 * co[q] + 3] = vals[q]
    return A


def bsr_residual(diag, ro, co, vals, b, x):
    """||b - A x|| / ||b|| in FP64 with a row loop (any size)."""
    n = diag.shape[0]
    y = np.einsum("nij,nj->ni", diag.astype(np.float64), x)
    rows = np.repeat(np.arange(n), np.diff(ro))
    contrib = np.einsum("kij,kj->ki", vals.astype(np.float64), x[co])
    np.add.at(y, rows, contrib)
    r = b.astype(np.float64) - y
    return float(np.sqrt((r * r).sum() / max((b.astype(np.float64) ** 2).sum(), 1e-30)))


def test_cfg1_first_step_matches_oracle(cuda):
    s = scenes.config1()
    g = make_sim(s, tol=1e-6)
    g.sim.step(s.dt)
    st = g.sim.last_solve_stats()
    assert st["status"] == CONVERGED and st["relres"] <= 1e-6
    diag, ro, co, vals, rhs, dx = g.sim.debug_dump_last_solve()
    gold = np.load(GOLDEN)
    np.testing.assert_array_equal(ro, gold["row_offsets"])
    np.testing.assert_array_equal(co, gold["col_indices"])
    od, ov = gold["diag1"], gold["values1"]
    tol_d = 2e-6 * np.abs(od).max(axis=(1, 2), keepdims=True) + 1e-12
    assert np.all(np.abs(diag - od) <= tol_d)
    tol_v = 2e-6 * np.abs(ov).max(axis=(1, 2), keepdims=True) + 1e-12
    assert np.all(np.abs(vals - ov) <= tol_v)
    # At rest the stretch forces cancel; their FP32 rounding leaves about
    # eps32 * keff * |dw| * (triangles per row) ~ 1e-5 N per particle
    # (|m g| = 0.098 N), which bounds the rhs difference.
    np.testing.assert_allclose(rhs, gold["rhs1"], rtol=1e-6, atol=5e-5)
    x = g.sim.get_dx64()
    assert bsr_residual(diag, ro, co, vals, rhs, x) <= 1.0001e-6
    assert rel(x, gold["dx1"]) <= 1e-3
    assert rel(g.positions(), gold["pos1"]) <= 1e-6


def test_cfg1_every_step_converges_and_short_trajectory(cuda):
    s = scenes.config1()
    g = make_sim(s, tol=1e-6)
    gold = np.load(GOLDEN)
    iters = []
    for k in range(1, 201):
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["status"] == CONVERGED and st["relres"] <= 1e-6, (k, st)
        iters.append(st["iters"])
        if k == 10:
            assert rel(g.positions(), gold["pos10"]) <= 1e-5
    assert 10 <= np.mean(iters) <= 500


def test_freefall_closed_form(cuda):
    s = scenes.freefall(8)
    g = make_sim(s, tol=1e-9, max_iter=1000)
    for _ in range(10):
        g.sim.step(s.dt)
    expect = s.pos.astype(np.float64).copy()
    expect[:, 1] += s.dt ** 2 * -9.8 * 10 * 11 / 2
    np.testing.assert_allclose(g.positions(), expect, atol=2e-6)
    np.testing.assert_allclose(g.velocities()[:, 1], -9.8 * 10 * s.dt, rtol=1e-4)


def resync_check(s, steps, every, tol=1e-6, pos_tol=1e-5, dx_tol=2e-3):
    g = make_sim(s, tol=tol)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000)
    worst = 0.0
    for k in range(steps):
        check = (k % every) == 0
        if check:
            p0 = g.positions().astype(np.float64)
            v0 = g.velocities().astype(np.float64)
            dx0 = g.sim.get_dx64() if k > 0 else np.zeros((s.n, 3))
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["status"] == CONVERGED and st["relres"] <= tol, (k, st)
        if check:
            oc.set_dx(dx0)
            oc.step(p0, v0, im, f, s.dt, 1)
            e = rel(g.positions(), p0)
            worst = max(worst, e)
            assert e <= pos_tol, (k, e)
            assert rel(g.sim.get_dx64(), oc.get_dx()) <= dx_tol, k
    return g, worst


def test_cfg2_jittered_per_element_k_resync(cuda):
    s = scenes.config2(w=64, steps=100)
    resync_check(s, 100, 10)


def test_cfg3_wind_resync(cuda):
    s = scenes.config3(w=48, steps=40)
    resync_check(s, 40, 5)


def test_cfg2_full_size_resync_and_residual(cuda):
    s = scenes.config2(w=256, steps=30)
    g, _ = resync_check(s, 30, 15)
    diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
    assert ro[-1] == 2 * 195585
    assert bsr_residual(diag, ro, co, vals, rhs, g.sim.get_dx64()) <= 1.0001e-6


@pytest.mark.parametrize("iters", [50, 49, 1])
def test_reference_mode_runs_fixed_iterations(cuda, iters):
    """Exactly max_iter iterations.  An odd count ends on an update pass that
    deferred its x += alpha p, so the solve flushes that term before it
    stops; the iterate must still match the oracle's."""
    s = scenes.config1()
    g = make_sim(s, tol=0.0, max_iter=iters, springs_active=False, force_active=False)
    oc, pos, vel, im, f = oracle_from_scene(s, precision=64, springs_active=False, force_active=False,
                                            tol=0.0, max_iter=iters)
    for _ in range(5):
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["iters"] == iters and st["status"] == STOPPED
    oc.step(pos, vel, im, f, s.dt, 5)
    assert rel(g.positions(), pos) <= 1e-5


def test_determinism_bitwise(cuda):
    s = scenes.config2(w=64, steps=20)
    runs = []
    for _ in range(2):
        g = make_sim(s, tol=1e-6)
        for _ in range(20):
            g.sim.step(s.dt)
        runs.append(g.positions())
    np.testing.assert_array_equal(runs[0], runs[1])


def test_update_pin_targets_and_mass_redistribution(cuda):
    import torch
    s = scenes.config1()
    g = make_sim(s, tol=1e-6)
    g.sim.step(s.dt)
    tgt = s.pin_targets + np.array([[0.0, 0.05, 0.0]], np.float32)
    g.sim.update_pin_targets(tgt)
    for _ in range(5):
        g.sim.step(s.dt)
    # stiff penalty pin follows its target
    assert np.linalg.norm(g.positions()[0] - tgt[0]) < 1e-3
    # area-lumped mass: interior 1/(density*area_share)
    inv = torch.zeros(s.n, dtype=torch.float32, device="cuda:0")
    g.sim.redistribute_mass_area_weighted(0.3, inv.data_ptr(), s.n)
    got = inv.cpu().numpy()
    expect = 1.0 / (0.3 * scenes.vertex_areas(g.positions(), s.tris))  # from the current positions
    np.testing.assert_allclose(got, expect, rtol=1e-5)


def test_batch_matches_single_simulators_bitwise(cuda):
    sc = [scenes.config2(w=20, steps=3, seed=11), scenes.config1(), scenes.config3(w=33, steps=3)]
    for s in sc:
        s.force = None
    b = make_batch(sc, tol=1e-6)
    singles = [make_sim(s, tol=1e-6, force_active=False) for s in sc]
    for _ in range(3):
        b.batch.step(1.0 / 60.0)
        for g in singles:
            g.sim.step(1.0 / 60.0)
    every = b.batch.stats_all()
    assert b.batch.num_cloths() == len(sc) and every["iters"].shape == (len(sc),)
    for k, g in enumerate(singles):
        st = b.batch.stats(k)
        assert st["status"] == CONVERGED
        assert st["iters"] == g.sim.last_solve_stats()["iters"]
        for key in ("iters", "restarts", "status", "relres", "converged"):
            assert every[key][k] == st[key], (k, key)
        np.testing.assert_array_equal(b.positions(k), g.positions())


def test_damping_and_uniform_force_closed_form(cuda):
    s = scenes.freefall(9)
    s.damping = 0.7
    f = np.array([0.02, 0.0, -0.01])
    s.force = np.tile(f, (s.n, 1)).astype(np.float32)
    g = make_sim(s, tol=1e-9, max_iter=1000)
    a = np.array(s.gravity) + f * s.inv_mass[0]
    decay = np.exp(-s.damping * s.dt)
    x = np.zeros(3)
    v = np.zeros(3)
    for _ in range(12):
        g.sim.step(s.dt)
        dx = s.dt * v + s.dt ** 2 * a
        x = x + dx
        v = dx / s.dt * decay
    np.testing.assert_allclose(g.positions() - s.pos, np.tile(x, (s.n, 1)), atol=3e-6)
    np.testing.assert_allclose(g.velocities(), np.tile(v, (s.n, 1)), atol=3e-4)


def test_launch_shapes_do_not_change_results(cuda):
    """Workgroup shapes, SpMV variants and where the solver state is advanced
    (inside every SpMV workgroup, or in a kernel of its own) are speed knobs
    only: every inner product is summed per 16-slice group, then over groups
    in one order, and the same scalar step follows.  The PCG passes' own
    workgroup shape is fixed (16 waves, one group each)."""
    s = scenes.config2(w=72, steps=5)
    g = make_sim(s, tol=1e-6)
    with pytest.raises(ValueError):
        g.sim.set_launch_config(dict(spmv_waves=4))
    runs = []
    for cfg in [dict(spmv_unroll=1, assemble_waves=2),
                dict(spmv_unroll=4, assemble_waves=8),
                dict(spmv_unroll=2, finalize_waves=16),
                dict(spmv_unroll=4, spmv_branch_free=1),
                dict(spmv_unroll=1, spmv_branch_free=1, assemble_waves=1),
                dict(spmv_unroll=2, pcg_split_advance=1),
                dict(spmv_unroll=1, spmv_branch_free=1, pcg_split_advance=1)]:
        g = make_sim(s, tol=1e-6)
        g.sim.set_autotune(False)
        g.sim.set_launch_config(cfg)
        for _ in range(5):
            g.sim.step(s.dt)
        runs.append((g.positions(), g.sim.get_dx64()))
    for p, x in runs[1:]:
        np.testing.assert_array_equal(p, runs[0][0])
        np.testing.assert_array_equal(x, runs[0][1])


def test_autotuner_picks_shapes_without_changing_results(cuda):
    """The autotuner (problems of 32768 padded rows or more) times the SpMV
    pass variants, the assembly's workgroup widths and the two places the PCG
    state can be advanced, and keeps the fastest of each; stepping with it
    gives bitwise the positions and iterate of stepping without it."""
    s = scenes.config2(w=192, steps=4)
    runs = []
    for tune in (True, False):
        g = make_sim(s, tol=1e-6)
        g.sim.set_autotune(tune)
        for _ in range(3):
            g.sim.step(s.dt)
            assert g.sim.last_solve_stats()["converged"]
        runs.append((g.positions(), g.sim.get_dx64()))
        kinds = {t[0] for t in g.sim.tune_log()}
        if tune:
            assert {"spmv", "spmv_branch_free", "assemble", "iteration_advance_in_pass",
                    "iteration_advance_kernel"} <= kinds, kinds
            assert all(t[3] > 0.0 for t in g.sim.tune_log())
        else:
            assert not kinds
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_batch_uneven_cloths_each_match_oracle(cuda):
    """Uneven batch (config-4 recipe, smaller sizes): every cloth converges on
    its own and each cloth's step matches the oracle from the same state."""
    sc = scenes.config4(count=24, wmin=8, wmax=40, steps=4, seed=5)
    b = make_batch(sc, tol=1e-6)
    for _ in range(3):
        b.batch.step(1.0 / 60.0)
    pos0 = b.pos.cpu().numpy().astype(np.float64)
    vel0 = b.vel.cpu().numpy().astype(np.float64)
    b.batch.step(1.0 / 60.0)
    for k in range(0, len(sc), 5):
        s = sc[k]
        st = b.batch.stats(k)
        assert st["converged"] and st["relres"] <= 1e-6
        a = b.offsets[k]
        oc, _, _, im, _ = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000, force_active=False)
        p = np.ascontiguousarray(pos0[a:a + s.n])
        v = np.ascontiguousarray(vel0[a:a + s.n])
        oc.step(p, v, im, None, 1.0 / 60.0, 1)
        assert rel(b.positions(k), p) <= 1e-5, k


def test_symmetric_spmv_mode_matches_default(cuda):
    """The symmetric SpMV pass reads each lower block as the transpose of its
    upper mirror.  Both halves are assembled independently in FP32, so results
    agree to rounding: every solve converges and positions match the default
    mode to 1e-6 relative over a few steps; the oracle step check holds."""
    s = scenes.config2(w=72, steps=6)
    runs = []
    for sym in [0, 1]:
        g = make_sim(s, tol=1e-6)
        g.sim.set_autotune(False)
        g.sim.set_launch_config(dict(spmv_symmetric=sym))
        assert g.sim.launch_config()["spmv_symmetric"] == sym
        for _ in range(6):
            g.sim.step(s.dt)
            st = g.sim.last_solve_stats()
            assert st["status"] == CONVERGED and st["relres"] <= 1e-6
        diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
        assert bsr_residual(diag, ro, co, vals, rhs, g.sim.get_dx64()) <= 1.0001e-6
        runs.append(g.positions())
    assert rel(runs[1], runs[0].astype(np.float64)) <= 1e-6


def test_packed_blocks_bitwise_equal_to_full_blocks(cuda):
    """Spring and stretch blocks are symmetric by construction (the element
    formulas fill (b, a) from (a, b)), so the packed upper-triangle storage
    holds the same operator: positions, iterations and dumped systems agree
    bitwise with 9-float symmetric storage.  With bending (a_i a_j^T blocks
    are not symmetric) storage stays at 9 floats."""
    s = scenes.config2(w=72, steps=6)
    runs = []
    for packed in [0, 1]:
        g = make_sim(s, tol=1e-6)
        g.sim.set_autotune(False)
        g.sim.set_launch_config(dict(spmv_symmetric=1, spmv_packed=packed))
        assert g.sim.launch_config()["spmv_packed"] == packed
        iters = []
        for _ in range(6):
            g.sim.step(s.dt)
            st = g.sim.last_solve_stats()
            assert st["status"] == CONVERGED and st["relres"] <= 1e-6
            iters.append(st["iters"])
        assert g.sim.block_floats() == (6 if packed else 9)
        runs.append((g.positions(), g.velocities(), g.sim.get_dx64(), iters, g.sim.debug_dump_last_solve()))
    (p0, v0, d0, i0, s0), (p1, v1, d1, i1, s1) = runs
    assert i0 == i1
    assert np.array_equal(p0, p1) and np.array_equal(v0, v1) and np.array_equal(d0, d1)
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)

    d = scenes.drape(w=20)
    g = make_sim(d, tol=1e-6, springs_active=False, force_active=False)
    g.sim.set_launch_config(dict(spmv_symmetric=1, spmv_packed=1))
    g.sim.step(d.dt)
    assert g.sim.block_floats() == 9
    assert g.sim.last_solve_stats()["status"] == CONVERGED


def test_index16_bitwise_equal_and_fallback(cuda):
    """16-bit column / mirror offsets are the same indices: bitwise-equal
    results.  A mesh whose numbering puts neighbours more than 32767 rows
    apart falls back to 32-bit indices and still solves."""
    s = scenes.config2(w=72, steps=4)
    runs = []
    for i16 in [0, 1]:
        g = make_sim(s, tol=1e-6)
        g.sim.set_autotune(False)
        g.sim.set_launch_config(dict(spmv_index16=i16))
        for _ in range(4):
            g.sim.step(s.dt)
            assert g.sim.last_solve_stats()["status"] == CONVERGED
        assert g.sim.index_bits() == (16 if i16 else 32)
        runs.append((g.positions(), g.sim.get_dx64()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])

    # 200 x 200 wind cloth with a random particle numbering: neighbours far apart
    base = scenes.config3(w=200, steps=2)
    perm = np.random.default_rng(3).permutation(base.n)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(base.n)
    sc = scenes.Scene(**{**base.__dict__})
    sc.pos = np.ascontiguousarray(base.pos[perm])
    sc.vel = np.ascontiguousarray(base.vel[perm])
    sc.inv_mass = np.ascontiguousarray(base.inv_mass[perm])
    sc.force = np.ascontiguousarray(base.force[perm])
    sc.tris = inv[base.tris].astype(np.int32)
    sc.pin_idx = inv[base.pin_idx].astype(np.int32)
    g = make_sim(sc, tol=1e-6)
    ref = make_sim(base, tol=1e-6)
    for _ in range(2):
        g.sim.step(sc.dt)
        ref.sim.step(base.dt)
        assert g.sim.last_solve_stats()["status"] == CONVERGED
    assert g.sim.index_bits() == 32 and ref.sim.index_bits() == 16
    # the same cloth, numbered differently: the same motion to rounding
    assert rel(g.positions()[inv], ref.positions().astype(np.float64)) <= 1e-5


def test_storage_switches_keep_padding_zero(cuda):
    """Switching between storages reuses the value buffers with another
    layout; the branch-free SpMV multiplies padding entries, so they must be
    zero after a switch.  A simulator that switches general -> packed ->
    general -> packed matches one that never switches (rounding level)."""
    s = scenes.config3(w=128, steps=6)
    a = make_sim(s, tol=1e-6)
    b = make_sim(s, tol=1e-6)
    for g in (a, b):
        g.sim.set_autotune(False)
        g.sim.set_launch_config(dict(spmv_branch_free=1))
    for mode in [(1, 1), (0, 0), (1, 1), (0, 0), (1, 1)]:
        a.sim.set_launch_config(dict(spmv_symmetric=mode[0], spmv_packed=mode[1]))
        a.sim.step(s.dt)
        b.sim.step(s.dt)
        for g in (a, b):
            st = g.sim.last_solve_stats()
            assert st["status"] == CONVERGED and st["relres"] <= 1e-6, (mode, st)
    assert rel(a.positions(), b.positions().astype(np.float64)) <= 1e-6


def test_reference_mode_step_returns_before_its_stream_finishes(cuda):
    """tol = 0 (the reference's fixed 50 iterations): step() enqueues the solve
    on the caller's stream and returns, like the reference after its graph
    launch (pcg_solver.cu:882-894).  The stream is held busy by a spin kernel
    queued before the step, so it is still running when step() returns.
    Results are bitwise those of the polled (blocking) step, which waits for
    its solve to end; the statistics wait for the stream."""
    import torch
    s = scenes.config3(w=96, steps=6)
    stream = torch.cuda.Stream()
    runs = []
    for blocking in (False, True):
        g = make_sim(s, tol=0.0, max_iter=50, springs_active=False, force_active=False)
        g.sim.set_blocking_step(blocking)
        g.sim.step(s.dt, stream.cuda_stream)  # builds topology (synchronous)
        torch.cuda.synchronize()
        pending = []
        for _ in range(5):
            with torch.cuda.stream(stream):
                torch.cuda._sleep(20_000_000)
            g.sim.step(s.dt, stream.cuda_stream)
            pending.append(not stream.query())
            st = g.sim.last_solve_stats()  # waits for the step
            assert st["iters"] == 50 and st["status"] == STOPPED and st["relres"] > 0
        stream.synchronize()
        runs.append((g.positions(), g.velocities(), g.sim.get_dx64(), st["relres"]))
        if not blocking:
            assert all(pending), pending
    (p0, v0, x0, r0), (p1, v1, x1, r1) = runs
    assert np.array_equal(p0, p1) and np.array_equal(v0, v1) and np.array_equal(x0, x1) and r0 == r1


def test_reference_mode_step_graph_replays_and_recaptures(cuda):
    """In reference mode (tol = 0) the whole step, assembly to finalize, is one
    graph, replayed while nothing it was captured with changes: a new dt
    captures it again and new pin targets reach it through its parameters.
    Every step is bitwise the polled step's.  A fixed solve longer than the
    graph's bound (600 iterations) replays the chunk graphs instead."""
    s = scenes.config3(w=64, steps=8)
    out = []
    for blocking in (False, True):
        g = make_sim(s, tol=0.0, max_iter=50)
        g.sim.set_blocking_step(blocking)
        tgt = s.pin_targets.copy()
        for k in range(8):
            if k == 6:
                tgt = (tgt + np.array([0.0, 0.01, 0.0], np.float32)).astype(np.float32)
                g.sim.update_pin_targets(tgt)
            g.sim.step(s.dt if k < 4 else 0.5 * s.dt)
            st = g.sim.last_solve_stats()
            assert st["iters"] == 50 and st["status"] == STOPPED, (k, st)
        out.append((g.positions(), g.velocities(), g.sim.get_dx64()))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    g = make_sim(s, tol=0.0, max_iter=600)
    for _ in range(2):
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["iters"] == 600


def test_reference_mode_batch_enqueues_and_matches_blocking(cuda):
    """A batch whose cloths all run fixed iteration counts is enqueued the same
    way; its results are bitwise those of the blocking step."""
    sc = scenes.config4(count=6, wmin=8, wmax=30, steps=3, seed=8)
    out = []
    for blocking in (False, True):
        b = make_batch(sc, tol=0.0, max_iter=30)
        b.batch.set_blocking_step(blocking)
        for _ in range(3):
            b.batch.step(1.0 / 60.0)
        assert all(b.batch.stats(k)["iters"] == 30 for k in range(len(sc)))
        out.append(b.pos.cpu().numpy())
    np.testing.assert_array_equal(out[0], out[1])


def test_plain_spmv_matches_numpy_bsr_product(cuda):
    """The plain SpMV (the reference's one-thread-per-block-row formulation)
    against a numpy FP64 product with the dumped system, on the general
    storage it reads; with the symmetric storages it refuses."""
    import torch
    s = scenes.config2(w=40, steps=2)
    g = make_sim(s, tol=1e-6)
    g.sim.set_autotune(False)
    x = torch.randn(s.n, 3, dtype=torch.float32, device="cuda:0", generator=torch.Generator("cuda:0").manual_seed(4))
    y = torch.zeros_like(x)
    g.sim.step(s.dt)
    with pytest.raises(ValueError):
        g.sim.spmv(x.data_ptr(), y.data_ptr())
    g.sim.set_launch_config(dict(spmv_symmetric=0, spmv_packed=0))
    g.sim.step(s.dt)
    g.sim.spmv(x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    diag, ro, co, vals, _, _ = g.sim.debug_dump_last_solve()
    xh = x.cpu().numpy().astype(np.float64)
    ref = np.einsum("nij,nj->ni", diag.astype(np.float64), xh)
    rows = np.repeat(np.arange(s.n), np.diff(ro))
    np.add.at(ref, rows, np.einsum("kij,kj->ki", vals.astype(np.float64), xh[co]))
    got = y.cpu().numpy()
    scale = np.abs(diag).max(axis=(1, 2)) * np.abs(xh).max()
    assert np.all(np.abs(got - ref) <= 1e-6 * scale[:, None] * 8), np.abs(got - ref).max()
    assert rel(got, ref) <= 1e-6


@pytest.mark.parametrize("name,horizon", [("cfg1", 20), ("cfg2_64", 60), ("cfg3_64", 15)])
def test_trajectory_within_1e5_of_fp64_oracle_up_to_horizon(cuda, name, horizon):
    """Positions along the whole trajectory from rest, with no re-synchronisation,
    stay within 1e-5 relative of the FP64 oracle solved to 1e-12, up to a
    per-configuration horizon.  The horizon is set by the stopping tolerance,
    not by FP32: the oracle itself, in FP64 at tolerance 1e-6, first departs by
    more than 1e-5 at step 71 (config 1), 289 (config 2 at 64^2) and 30
    (config 3 at 64^2); its FP32-matrix variant at steps 27, 80 and 21.  The
    horizons here sit below those."""
    s = {"cfg1": scenes.config1, "cfg2_64": lambda: scenes.config2(w=64, steps=horizon),
         "cfg3_64": lambda: scenes.config3(w=64, steps=horizon)}[name]()
    g = make_sim(s, tol=1e-6)
    oc, p, v, im, f = oracle_from_scene(s, precision=64, tol=1e-12, max_iter=100000, threads=8)
    worst = 0.0
    for k in range(horizon):
        g.sim.step(s.dt)
        assert g.sim.last_solve_stats()["converged"]
        oc.step(p, v, im, f, s.dt, 1)
        worst = max(worst, rel(g.positions(), p))
        assert worst <= 1e-5, (k + 1, worst)
