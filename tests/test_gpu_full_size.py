"""The 1024 x 1024 wind cloth (the benchmark workload) at full size, checked
through properties that do not depend on size:

  * the pattern has the closed-form count nnz = 2 (2 W (W - 1) + (W - 1)^2);
  * every solve converges, and the true relative residual of the FP64
    iterate, recomputed in numpy FP64 from the dumped system, is <= 1e-6;
  * the dumped operator is symmetric (block (j, i) is block (i, j)^T);
  * one FP32 step from a shared state matches the FP64 oracle (solved to
    1e-11) to 1e-5 relative in positions, the bar of the small configs.
"""

import os

import numpy as np
import pytest

from chysx import scenes
from chysx.harness import make_sim
from oracle.oracle import oracle_from_scene

pytestmark = pytest.mark.gpu

W = 1024


def bsr_residual(diag, ro, co, vals, b, x):
    """||b - A x|| / ||b|| in FP64 (bincount per component: any size)."""
    n = diag.shape[0]
    y = np.einsum("nij,nj->ni", diag.astype(np.float64), x)
    rows = np.repeat(np.arange(n), np.diff(ro))
    contrib = np.einsum("kij,kj->ki", vals.astype(np.float64), x[co])
    for c in range(3):
        y[:, c] += np.bincount(rows, weights=contrib[:, c], minlength=n)
    r = b.astype(np.float64) - y
    return float(np.sqrt((r * r).sum() / max((b.astype(np.float64) ** 2).sum(), 1e-30)))


def test_cloth1024_full_size_properties(cuda):
    s = scenes.config3(w=W, steps=4)
    g = make_sim(s, tol=1e-6)
    for _ in range(3):
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        assert st["converged"] and st["relres"] <= 1e-6, st
    diag, ro, co, vals, rhs, _ = g.sim.debug_dump_last_solve()
    assert int(ro[-1]) == 2 * (2 * W * (W - 1) + (W - 1) ** 2)
    assert np.all(np.diff(ro) >= 0)
    assert bsr_residual(diag, ro, co, vals, rhs, g.sim.get_dx64()) <= 1.0001e-6

    n = s.n
    rows = np.repeat(np.arange(n), np.diff(ro)).astype(np.int64)
    keys = rows * n + co
    t = np.searchsorted(keys, co.astype(np.int64) * n + rows)
    assert np.array_equal(keys[t], co.astype(np.int64) * n + rows)
    assert np.array_equal(vals[t], vals.transpose(0, 2, 1))
    del vals, diag

    # one step from the GPU's state against the FP64 oracle
    threads = min(16, os.cpu_count() or 1)
    oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=1e-11, max_iter=100000, threads=threads)
    p0 = g.positions().astype(np.float64)
    v0 = g.velocities().astype(np.float64)
    oc.set_dx(g.sim.get_dx64())
    its, relres, conv = oc.step(p0, v0, im, f, s.dt, 1)
    assert conv[0]
    g.sim.step(s.dt)
    assert g.sim.last_solve_stats()["converged"]
    gp = g.positions().astype(np.float64)
    assert np.linalg.norm(gp - p0) / np.linalg.norm(p0) <= 1e-5


def test_batch512_every_cloth_converges_and_runs_are_bitwise_equal(cuda):
    """The 512-cloth batch (3.2 M particles, 3376 workgroups per pass): every
    cloth converges to 1e-6 in every step, and two runs from the same state
    are bitwise equal.  Each SpMV pass folds the previous partials into the
    solver state in every workgroup while some of them already write this
    pass's partials; at this size the workgroups of one cloth do not all run
    at once, so a pass that wrote a slot it also reads would diverge here."""
    import torch
    from chysx.harness import make_batch

    sc = scenes.config4(steps=3)
    runs = []
    for _ in range(2):
        b = make_batch(sc, tol=1e-6, max_iter=20000)
        iters = []
        for _ in range(3):
            b.batch.step(sc[0].dt)
            st = [b.batch.stats(k) for k in range(len(sc))]
            assert all(x["converged"] and x["relres"] <= 1e-6 for x in st)
            iters.append([x["iters"] for x in st])
        torch.cuda.synchronize()
        runs.append((b.pos.cpu().numpy(), iters))
        del b
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]
