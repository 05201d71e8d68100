"""Cloth-step benchmark on one MI355X.

Workload (default): the 1024x1024 cloth -- 1 m square, top edge pinned,
gravity plus a seeded constant wind, springs and Baraff-Witkin stretch
(k = 1e3), m = 0.01 kg, dt = 1/60, every step solved by block-Jacobi PCG to a
1e-6 true relative residual, warm-started from the previous step's dx.
Inputs are resident in HBM before the timed region.

Prints one JSON line: steps/s (value), PCG iterations per step, the roofline
of the dominant kernel (the PCG SpMV: direction update + A p + p.Ap) from HIP
events on the launch stream, and a CPU baseline (the FP64 oracle stepping the
same state on the host cores, bounded sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "ai-phyiscs_amd")]

METRIC = "cloth steps/s + PCG iters/step to 1e-6 residual, learned vs plain warm start"
PEAK_HBM_GBS = 8000.0


def kernel_bytes(n_rows: int, nnz: int):
    """Algorithmic HBM bytes per launch (FP32 blocks of 36 B, 4 B columns,
    12 B float3, 24 B double3)."""
    return {
        # plain y = A x: values+cols 40/block; diag 36 + row offset 4 + x 12 + y 12 per row (+4)
        "spmv": 40.0 * nnz + 64.0 * n_rows + 4.0,
        # PCG SpMV: values+cols 40/block; diag 36, z 12, p_prev 12 reads, p 12, Ap 12 writes per row
        "pcg_spmv": 40.0 * nnz + 84.0 * n_rows,
        # PCG update: p 12, Ap 12, x 24, r 12, M^-1 24 reads; x 24, r 12, z 12 writes per row
        "pcg_update": 132.0 * n_rows,
        # reference formulation of one PCG iteration: B_spmv + 180 per row
        "pcg_iteration_ref": 40.0 * nnz + 244.0 * n_rows + 4.0,
    }


def spmv_layout_bytes(n_rows: int, nnz: int, storage: str, index_bits: int) -> float:
    """Bytes the SpMV pass must move on a storage: general streams every block
    (36 B); the symmetric storages stream the nnz/2 upper blocks (36 B, or
    24 B packed) and re-read lower blocks through their mirrors (L2), with a
    column index per entry and a mirror index per lower entry (4 B, or 2 B as
    16-bit offsets); diagonal 36 B (24 B packed); z, p_prev read and p, Ap
    written, 12 B each."""
    ib = index_bits / 8.0
    vectors = 48.0 * n_rows
    if storage == "general":
        return 36.0 * nnz + ib * nnz + 36.0 * n_rows + vectors
    if storage == "symmetric":
        return 18.0 * nnz + ib * nnz + ib * nnz / 2 + 36.0 * n_rows + vectors
    return 12.0 * nnz + ib * nnz + ib * nnz / 2 + 24.0 * n_rows + vectors


def batch_roofline(sc, iters, ms_per_step):
    """HBM roofline of the batch step: every cloth's PCG iterations times the
    bytes one iteration of that cloth moves on the default storage (packed
    symmetric blocks, 16-bit offsets: SpMV pass 15 nnz + 72 N, update pass
    72 N or 144 N alternately), plus each solve's start (residual pass 18 nnz
    + 120 N) and the assembly (upper blocks 12 nnz, diagonal, inverse and
    rhs 60 N written; positions, velocities, masses 28 N and element data
    read), over the measured time per step."""
    import numpy as np
    from chysx import scenes
    total = 0.0
    for s, it in zip(sc, iters):
        n = s.n
        e = scenes.edges_of(s.tris).shape[0]
        nnz = 2 * e
        t = s.tris.shape[0]
        per_iter = 15.0 * nnz + 72.0 * n + 108.0 * n
        start = 18.0 * nnz + 120.0 * n
        asm = 12.0 * nnz + 88.0 * n + t * (12 + 16 + 4 + 4) + e * (8 + 4 + 4) + 8.0 * 3 * t + 8.0 * 2 * e
        total += it * per_iter + start + asm
    achieved = total / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "whole batch step (assembly, PCG start, every cloth's iterations)",
            "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
            "bytes_per_step": total, "ms_per_step": round(ms_per_step, 3),
            "iters_weighted_by_rows": round(float(np.average(iters, weights=[s.n for s in sc])), 2)}


def bench_batch(args):
    """512 cloths, W ~ U{16..128}, per-cloth k, mass and pins, stepped
    together: one launch per kernel, per-cloth PCG state and convergence."""
    import numpy as np
    import torch
    from chysx import scenes
    from chysx.harness import make_batch

    sc = scenes.config4(steps=args.warmup + args.steps)
    t0 = time.perf_counter()
    b = make_batch(sc, tol=args.tol, max_iter=20000)
    if args.split_advance is not None:
        b.batch.set_split_advance(args.split_advance)
    b.batch.step(sc[0].dt)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    for _ in range(args.warmup - 1):
        b.batch.step(sc[0].dt)
    torch.cuda.synchronize()
    cpu_state = None
    if not args.no_cpu:
        cpu_state = (b.pos.cpu().numpy().astype(np.float64), b.vel.cpu().numpy().astype(np.float64))
    it_mean, it_max = [], []
    it_sum = np.zeros(len(sc))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.batch.step(sc[0].dt)
        st = b.batch.stats_all()  # every cloth, one call
        if not st["converged"].all():
            raise RuntimeError("a cloth did not converge")
        its = st["iters"]
        it_sum += its.astype(np.float64)
        it_mean.append(np.mean(its))
        it_max.append(np.max(its))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_total = int(sum(s.n for s in sc))
    roofline = batch_roofline(sc, it_sum / args.steps, 1e3 * elapsed / args.steps)
    ws_cmp = None
    if args.ws_steps > 0:
        del b
        ws_cmp = batch_warm_start_comparison(args, sc, make_batch)
        b = None

    cpu = None
    if cpu_state is not None:
        from oracle.oracle import oracle_from_scene
        threads = args.cpu_threads
        pos_all, vel_all = cpu_state
        secs_total, rows_done, its = 0.0, 0, 0
        order = np.random.default_rng(0).permutation(len(sc))[:48]
        for k in order:
            s = sc[k]
            a = int(np.cumsum([0] + [q.n for q in sc])[k])
            oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=args.tol, max_iter=20000, threads=threads,
                                                force_active=False)
            oc.build_topology()
            p0 = np.ascontiguousarray(pos_all[a:a + s.n])
            v0 = np.ascontiguousarray(vel_all[a:a + s.n])
            secs, it = oc.time_steps(p0, v0, im, None, s.dt, 1)
            secs_total += secs
            rows_done += s.n
            its += it
        est = secs_total * n_total / rows_done
        cpu = {"value": round(1.0 / est, 5), "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle stepping 48 randomly chosen cloths of the batch (one step each, cold start, "
                         f"{rows_done} particles, {secs_total:.2f} s), scaled by particle count to all 512 cloths",
               "host_cpu": host_cpu_model(), "host_nproc": os.cpu_count()}
    out = {
        "metric": METRIC,
        "value": round(args.steps / elapsed, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch512: 512 cloths W~U{16..128}, k~10^U(2,4), m~U(0.005,0.02) kg, pins in "
                               "{corner, two corners, top edge}, dt=1/60, PCG to 1e-6 per cloth",
                   "particles": n_total, "cloths": len(sc), "parallelism": "1 GPU"},
        "pcg_iters_per_step_mean": round(float(np.mean(it_mean)), 2),
        "pcg_iters_per_step_max": round(float(np.mean(it_max)), 2),
        "cloth_iterations": batch_tail(sc, it_sum / args.steps),
        "warm_start_comparison": ws_cmp,
        "setup_s": round(setup_s, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def warm_start_comparison(args, scene, make_sim):
    """Same workload and step window with each warm start: previous-step dx
    (the reference's), dx_prev plus the learned direction (committed trained
    parameters), Galerkin projection onto the last 6 solutions, and the
    projection plus the learned direction."""
    import numpy as np
    import torch
    from chysx import warmstart as ws
    out = {"warmup": args.warmup, "steps": args.ws_steps, "accept": args.ws_accept}
    pos = {}
    modes = ("previous", "learned", "projection", "projection_learned")
    for mode in modes:
        path = args.ws_weights or ws.weights_for(mode)
        theta = ws.load_params(path)[0] if os.path.exists(path) else None
        if mode.endswith("learned") and theta is None:
            out[mode] = None
            continue
        g = make_sim(scene, tol=args.tol, max_iter=20000)
        if theta is not None:
            g.sim.ws_set_params(theta)
        g.sim.set_warm_start_mode(mode)
        g.sim.ws_set_acceptance(args.ws_accept)
        for _ in range(args.warmup):
            g.sim.step(scene.dt)
        torch.cuda.synchronize()
        its, used, proj = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.ws_steps):
            g.sim.step(scene.dt)
            st = g.sim.last_solve_stats()
            if not st["converged"]:
                raise RuntimeError(f"PCG did not converge ({mode}): {st}")
            its.append(st["iters"])
            w = g.sim.ws_stats()
            used.append(w["used"])
            proj.append(w["projection_used"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        pos[mode] = g.positions()
        out[mode] = {"steps_per_s": round(args.ws_steps / el, 3), "pcg_iters_per_step": round(float(np.mean(its)), 2),
                     "projection_used_frac": round(float(np.mean(proj)), 3),
                     "learned_used_frac": round(float(np.mean(used)), 3),
                     "weights": os.path.relpath(path, ROOT) if mode.endswith("learned") else None}
        del g
    for mode in modes[1:]:
        if out.get(mode):
            out[mode]["iters_ratio_vs_previous"] = round(out["previous"]["pcg_iters_per_step"] /
                                                         out[mode]["pcg_iters_per_step"], 3)
            out[mode]["speedup_vs_previous"] = round(out[mode]["steps_per_s"] / out["previous"]["steps_per_s"], 3)
            out[mode]["max_pos_diff_vs_previous_m"] = float(np.abs(pos[mode] - pos["previous"]).max())
    return out


def batch_tail(sc, its):
    """Where the batch's rounds go: percentiles of the cloths' mean PCG
    iterations per step, and the cloths, 1024-row groups and rows still
    iterating after r rounds (the passes launch every group and a finished
    group's workgroups exit at once), with the share of small cloths (at most
    4096 rows) among them."""
    import numpy as np
    rows = np.array([s.n for s in sc])
    groups = np.ceil(rows / 1024.0)
    out = {"percentiles_0_25_50_75_90_100": [round(float(x), 1) for x in np.percentile(its, [0, 25, 50, 75, 90, 100])],
           "iterating_after_round": {}}
    w = np.array([s.grid_w for s in sc])
    out["mean_iterations_by_width"] = {f"{lo}-{hi}": round(float(its[(w >= lo) & (w <= hi)].mean()), 1)
                                       for lo, hi in ((16, 40), (41, 64), (65, 96), (97, 128))
                                       if ((w >= lo) & (w <= hi)).any()}
    for r in (0, 50, 100, 150, 200, 250, 300, 350):
        live = its > r
        out["iterating_after_round"][str(r)] = {"cloths": int(live.sum()), "groups": int(groups[live].sum()),
                                                "rows": int(rows[live].sum()),
                                                "small_cloths": int((live & (rows <= 4096)).sum())}
    return out


def batch_warm_start_comparison(args, sc, make_batch):
    """The batch stepped with each warm start (fresh batch per mode)."""
    import numpy as np
    import torch
    from chysx import warmstart as ws
    out = {"warmup": args.warmup, "steps": args.ws_steps, "accept": args.ws_accept}
    modes = ("previous", "learned", "projection", "projection_learned")
    for mode in modes:
        path = args.ws_weights or ws.weights_for(mode)
        theta = ws.load_params(path)[0] if os.path.exists(path) else None
        if mode.endswith("learned") and theta is None:
            out[mode] = None
            continue
        b = make_batch(sc, tol=args.tol, max_iter=20000)
        if theta is not None:
            b.batch.ws_set_params(theta)
        b.batch.set_warm_start_mode(mode)
        b.batch.ws_set_acceptance(args.ws_accept)
        for _ in range(args.warmup):
            b.batch.step(sc[0].dt)
        torch.cuda.synchronize()
        it_mean, it_max, used = [], [], []
        el = 0.0
        for _ in range(args.ws_steps):
            # each step timed to its end on the device; the per-cloth
            # warm-start records are read outside the timed span
            t0 = time.perf_counter()
            b.batch.step(sc[0].dt)
            st = b.batch.stats_all()
            torch.cuda.synchronize()
            el += time.perf_counter() - t0
            if not st["converged"].all():
                raise RuntimeError(f"a cloth did not converge ({mode})")
            it_mean.append(np.mean(st["iters"]))
            it_max.append(np.max(st["iters"]))
            used.append(np.mean([b.batch.ws_stats(k)["used"] for k in range(len(sc))]))
        out[mode] = {"steps_per_s": round(args.ws_steps / el, 3),
                     "pcg_iters_per_step_mean": round(float(np.mean(it_mean)), 2),
                     "pcg_iters_per_step_max": round(float(np.mean(it_max)), 2),
                     "learned_used_frac": round(float(np.mean(used)), 3)}
        del b
    for mode in modes[1:]:
        if out.get(mode):
            out[mode]["speedup_vs_previous"] = round(out[mode]["steps_per_s"] / out["previous"]["steps_per_s"], 3)
    return out


def gbs(b, ms):
    return b / (ms * 1e-3) / 1e9


def storage_of(g):
    cfg = g.sim.launch_config()
    return ("general" if not cfg["spmv_symmetric"] else
            "symmetric_packed" if g.sim.block_floats() == 6 else "symmetric")


def time_storages(g, s, n, nnz, args, stream, names):
    """The SpMV pass (with and without the state advance) and a whole
    iteration on each matrix storage in `names` (same matrix values,
    re-assembled in that storage; speed only), and the plain SpMV on the
    general storage.  Leaves the simulator in its own storage."""
    cfg0 = g.sim.launch_config()
    table = {"general": (0, 0), "symmetric": (1, 0), "symmetric_packed": (1, 1)}
    storage, t_spmv = {}, None
    for name in names:
        sym, packed = table[name]
        g.sim.set_launch_config(dict(spmv_symmetric=sym, spmv_packed=packed))
        g.sim.time_kernel(3, 1, s.dt, stream)  # one assembly into this storage
        t_sp = g.sim.time_kernel(1, args.kernel_reps, s.dt, stream)
        t_na = g.sim.time_kernel(5, args.kernel_reps, s.dt, stream)
        t_it = g.sim.time_kernel(4, args.kernel_reps, s.dt, stream)
        lb = spmv_layout_bytes(n, nnz, name, g.sim.index_bits())
        storage[name] = {"block_floats": g.sim.block_floats(), "index_bits": g.sim.index_bits(),
                         "layout_bytes": lb, "ms": round(t_sp, 5),
                         "GB/s_layout": round(gbs(lb, t_sp), 1),
                         "frac_layout": round(gbs(lb, t_sp) / PEAK_HBM_GBS, 4),
                         "GB/s_ref_formulation": round(gbs(kernel_bytes(n, nnz)["pcg_spmv"], t_sp), 1),
                         # the same pass without the solver-state advance it runs first
                         "ms_no_advance": round(t_na, 5),
                         "frac_layout_no_advance": round(gbs(lb, t_na) / PEAK_HBM_GBS, 4),
                         "iteration_ms": round(t_it, 5)}
        if name == "general":
            t_spmv = g.sim.time_kernel(0, args.kernel_reps, s.dt, stream)  # plain SpMV reads general storage
    g.sim.set_launch_config(dict(spmv_symmetric=cfg0["spmv_symmetric"], spmv_packed=cfg0["spmv_packed"]))
    g.sim.time_kernel(3, 1, s.dt, stream)
    # where the solver state is advanced: inside every SpMV workgroup (0) or
    # in a kernel of its own before the pass (1); -1 (default) picks by size
    mode = storage_of(g)
    adv = {}
    for split in (0, 1):
        g.sim.set_launch_config(dict(pcg_split_advance=split))
        adv[f"split_{split}"] = {"spmv_pass_ms": round(g.sim.time_kernel(1, args.kernel_reps, s.dt, stream), 5),
                                 "iteration_ms": round(g.sim.time_kernel(4, args.kernel_reps, s.dt, stream), 5)}
    g.sim.set_launch_config(dict(pcg_split_advance=cfg0["pcg_split_advance"]))
    storage[mode]["advance_variants"] = adv
    return storage, t_spmv


def spmv_roofline(storage, mode, kb, traffic):
    st = storage[mode]
    lb, ms = st["layout_bytes"], st["ms"]
    return {
        "bound": "hbm",
        "kernel": f"pcg_spmv_kernel, {mode} storage, {st['index_bits']}-bit indices "
                  "(solver-state advance; p = z + beta p_prev; Ap = A p; p.Ap)",
        "achieved": round(gbs(lb, ms), 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(gbs(lb, ms) / PEAK_HBM_GBS, 4),
        "traffic": traffic.get({"symmetric_packed": "pcg_spmv_kernel", "symmetric": "pcg_spmv_kernel_symmetric",
                                "general": "pcg_spmv_kernel_general"}[mode]),
        "bytes_per_launch": lb,
        "ms_per_launch": round(ms, 5),
        "ref_formulation_bytes": kb["pcg_spmv"],
        "ref_formulation_GB/s": round(gbs(kb["pcg_spmv"], ms), 1),
    }


def hbm_kernels(args, make_sim):
    """PCG kernels of the wind cloth at --hbm-size (default 2048^2: the packed
    SpMV pass touches ~680 MB and an iteration ~930 MB, several times the
    256 MB Infinity Cache), on the default storage and on the general one."""
    import numpy as np
    import torch
    from chysx import scenes
    w = args.hbm_size
    s = scenes.config3(w=w, steps=2)
    g = make_sim(s, tol=args.tol, max_iter=20000)
    g.sim.step(s.dt)
    g.sim.step(s.dt)
    st = g.sim.last_solve_stats()
    torch.cuda.synchronize()
    n, nnz = s.n, g.sim.num_off_diag_blocks()
    kb = kernel_bytes(n, nnz)
    stream = torch.cuda.current_stream().cuda_stream
    mode = storage_of(g)
    t_up = g.sim.time_kernel(2, args.kernel_reps, s.dt, stream)
    storage, t_spmv = time_storages(g, s, n, nnz, args, stream, (mode, "general"))
    iter_bytes = storage[mode]["layout_bytes"] + 0.5 * (72.0 + 144.0) * n  # SpMV pass + mean update pass
    out = {
        "workload": f"cloth_{w}x{w}_wind, system of its second step (PCG {st['iters']} iterations)",
        "particles": n, "off_diag_blocks": nnz,
        "pcg_spmv_storage": storage,
        "spmv_plain": {"bytes": kb["spmv"], "ms": round(t_spmv, 5), "GB/s": round(gbs(kb["spmv"], t_spmv), 1),
                       "frac": round(gbs(kb["spmv"], t_spmv) / PEAK_HBM_GBS, 4)},
        "pcg_update_adding_two_terms": {"bytes": 144.0 * n, "ms": round(t_up, 5),
                                        "GB/s": round(gbs(144.0 * n, t_up), 1),
                                        "frac": round(gbs(144.0 * n, t_up) / PEAK_HBM_GBS, 4)},
        "pcg_iteration": {"layout_bytes": iter_bytes, "ms": round(storage[mode]["iteration_ms"], 5),
                          "GB/s_layout": round(gbs(iter_bytes, storage[mode]["iteration_ms"]), 1),
                          "frac_layout": round(gbs(iter_bytes, storage[mode]["iteration_ms"]) / PEAK_HBM_GBS, 4),
                          "ref_bytes": kb["pcg_iteration_ref"],
                          "GB/s_ref_formulation": round(gbs(kb["pcg_iteration_ref"],
                                                            storage[mode]["iteration_ms"]), 1)},
        "spmv_pass_roofline": spmv_roofline(storage, mode, kb, read_traffic(w)),
    }
    out["spmv_pass_roofline"]["working_set"] = (f"{w}x{w}: {storage[mode]['layout_bytes'] / 1e6:.0f} MB per "
                                                "pass, beyond the 256 MB Infinity Cache")
    del g
    return out


def copy_ceiling(args):
    """Hand-written copy and triad kernels over --copy-gib GiB buffers: the
    bandwidth a streaming kernel reaches here, beside the 8 TB/s peak."""
    import torch
    from chysx import _chysx_native as nat
    stream = torch.cuda.current_stream().cuda_stream
    nbytes = int(args.copy_gib * (1 << 30))
    out = {"bytes_per_buffer": nbytes}
    best = 0.0
    for kind, name in enumerate(["copy", "triad", "copy_nontemporal", "triad_nontemporal"]):
        ms, moved = nat.stream_bench(kind, nbytes, 20, stream)
        r = gbs(moved, ms)
        out[name] = {"ms": round(ms, 5), "bytes": moved, "GB/s": round(r, 1), "frac": round(r / PEAK_HBM_GBS, 4)}
        best = max(best, r)
    out["best_GB/s"] = round(best, 1)
    out["best_frac_of_peak"] = round(best / PEAK_HBM_GBS, 4)
    return out


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, s, g, cpu_state):
    """The FP64 oracle (C++ restatement of the reference step, OpenMP) stepping
    the state the GPU timed region starts from, on --cpu-threads host threads
    and on 1 thread; then the GPU re-stepped from that state against the
    oracle's result."""
    import numpy as np
    from oracle.oracle import oracle_from_scene
    p0, v0, dx0 = cpu_state
    runs = {}
    for threads in (args.cpu_threads, 1):
        oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=args.tol, max_iter=20000, threads=threads)
        oc.build_topology()
        oc.set_dx(dx0)
        p, v = p0.copy(), v0.copy()
        secs, it = oc.time_steps(p, v, im, f, s.dt, args.cpu_steps)  # p, v advance in place
        runs[threads] = (secs, it, p)
    secs, it, p_cpu = runs[args.cpu_threads]
    secs1, _, _ = runs[1]
    cpu = {"value": round(args.cpu_steps / secs, 5), "unit": "steps/s", "cores": args.cpu_threads, "kind": "port",
           "sample": f"FP64 oracle (C++ restatement of the reference step, OpenMP) stepping the "
                     f"{args.size}x{args.size} cloth {args.cpu_steps} step(s) from the state the GPU timed "
                     f"region starts from; {it} PCG iterations, {secs:.2f} s on {args.cpu_threads} threads",
           "pcg_iters": it,
           "one_thread": {"value": round(args.cpu_steps / secs1, 5), "unit": "steps/s", "cores": 1,
                          "seconds": round(secs1, 3)},
           "host_cpu": host_cpu_model(), "host_nproc": os.cpu_count(),
           "threads_note": "the box's CPU share for one GPU is 16 threads, so the multi-thread run uses 16 of "
                           "the host's cores, not all of them"}
    # FP32 GPU step vs the FP64 oracle from the same state (both to the same tolerance)
    g.set_state(p0, v0)
    g.sim.set_warm_start(dx0)
    for _ in range(args.cpu_steps):
        g.sim.step(s.dt)
    gp = g.positions().astype(np.float64)
    d_gpu, d_cpu = gp - p0, p_cpu - p0
    fp_err = {"steps": args.cpu_steps, "tol": args.tol,
              "max_abs_pos_diff_m": float(np.abs(gp - p_cpu).max()),
              "rel_pos_diff": float(np.linalg.norm(gp - p_cpu) / np.linalg.norm(p_cpu)),
              "rel_displacement_diff": float(np.linalg.norm(d_gpu - d_cpu) / max(np.linalg.norm(d_cpu), 1e-30))}
    return cpu, fp_err


def reference_mode(args, scene, make_sim):
    """The reference's own stopping rule on the same cloth: a fixed 50 PCG
    iterations per step, no convergence test (pcg_solver.cu:542-592).  Each
    step is one captured graph launched without a host wait; the window is
    timed from the first launch to the stream's end."""
    import torch
    g = make_sim(scene, tol=0.0, max_iter=50)
    for _ in range(5):
        g.sim.step(scene.dt)
    torch.cuda.synchronize()
    steps = 20
    t0 = time.perf_counter()
    for _ in range(steps):
        g.sim.step(scene.dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = g.sim.last_solve_stats()
    return {"pcg_iters_per_step": st["iters"], "steps": steps, "steps_per_s": round(steps / el, 3),
            "ms_per_step": round(1e3 * el / steps, 3), "relres_after_50": float(st["relres"])}


def full_run(args, make_sim):
    """The whole 300-step scenario from rest: PCG iterations of every step,
    so the timed window's iteration count can be set against the run's."""
    import numpy as np
    import torch
    from chysx import scenes
    s = scenes.config3(w=args.size)
    g = make_sim(s, tol=args.tol, max_iter=20000)
    its = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(s.steps):
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        if not st["converged"]:
            raise RuntimeError(f"PCG did not converge in the full run: {st}")
        its.append(int(st["iters"]))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    its = np.array(its)
    w0, w1 = args.warmup, args.warmup + args.steps
    return {"steps": int(s.steps), "steps_per_s_incl_setup_step": round(s.steps / el, 3),
            "pcg_iters_mean": round(float(its.mean()), 2), "pcg_iters_min": int(its.min()),
            "pcg_iters_max": int(its.max()),
            "pcg_iters_timed_window_mean": round(float(its[w0:w1].mean()), 2) if w1 <= len(its) else None,
            "pcg_iters_per_step": its.tolist()}


def read_traffic(size):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (recorded in profiles/traffic_cloth<size>.md as
    `traffic_bytes_per_launch <kernel> <bytes>` lines)."""
    path = os.path.join(ROOT, "profiles", f"traffic_cloth{size}.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            line = line.strip().strip("`")
            if line.startswith("traffic_bytes_per_launch "):
                _, name, val = line.split()
                out[name] = float(val)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["cloth1024", "batch512"], default="cloth1024",
                    help="cloth1024: one 1024x1024 cloth (default); batch512: 512 uneven cloths in one launch")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-steps", type=int, default=1, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--ws-steps", type=int, default=40,
                    help="timed steps per warm-start mode in the projection / learned comparison (0: skip)")
    ap.add_argument("--ws-weights", default=None, help="learned warm-start parameters (.npz)")
    ap.add_argument("--ws-accept", choices=["energy", "residual"], default="energy",
                    help="acceptance rule of the learned direction in warm_start_comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (known bytes) for counter calibration")
    ap.add_argument("--hbm-size", type=int, default=2048,
                    help="grid width of the cloth whose PCG kernels give the HBM roofline (its matrix and "
                         "vectors exceed the 256 MB Infinity Cache); 0 skips it")
    ap.add_argument("--copy-gib", type=float, default=1.0,
                    help="GiB per buffer of the hand-written copy / triad HBM probe (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (the box's CPU share for one GPU); a 1-thread run "
                         "is reported beside it")
    ap.add_argument("--warm-start", default="previous",
                    choices=["previous", "learned", "projection", "projection_learned"],
                    help="warm start of the timed run (previous: the reference's)")
    ap.add_argument("--split-advance", type=int, default=None, choices=[-1, 0, 1],
                    help="where the PCG state is advanced: -1 by problem size (default), 0 inside every SpMV "
                         "workgroup, 1 in a kernel of its own")
    ap.add_argument("--full-run", action="store_true", default=True,
                    help="also step the whole scenario (300 steps) from rest and report iterations per step")
    ap.add_argument("--no-full-run", dest="full_run", action="store_false")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the cloth step has no multi-GPU path (it would "
              "shard by row bands with a one-row halo per PCG iteration).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from chysx import scenes
    from chysx.harness import make_sim

    dev = torch.device("cuda:0")
    if args.workload == "batch512":
        bench_batch(args)
        return
    s = scenes.config3(w=args.size, steps=args.warmup + args.steps)
    t0 = time.perf_counter()
    g = make_sim(s, tol=args.tol, max_iter=20000)
    if args.split_advance is not None:
        g.sim.set_launch_config(dict(pcg_split_advance=args.split_advance))
    if args.warm_start != "previous":
        from chysx import warmstart as ws
        if args.warm_start.endswith("learned"):
            g.sim.ws_set_params(ws.load_params(args.ws_weights or ws.weights_for(args.warm_start))[0])
        g.sim.set_warm_start_mode(args.warm_start)
    g.sim.step(s.dt)  # builds topology + first solve
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    for _ in range(args.warmup - 1):
        g.sim.step(s.dt)
    torch.cuda.synchronize()
    # CPU-baseline state: the state the timed region starts from
    cpu_state = None
    if not args.no_cpu:
        cpu_state = (g.positions().astype(np.float64), g.velocities().astype(np.float64), g.sim.get_dx64())

    iters, restarts, relres = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.sim.step(s.dt)
        st = g.sim.last_solve_stats()
        iters.append(st["iters"])
        restarts.append(st["restarts"])
        relres.append(st["relres"])
        if not st["converged"]:
            raise RuntimeError(f"PCG did not converge: {st}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed

    n = s.n
    nnz = g.sim.num_off_diag_blocks()
    kb = kernel_bytes(n, nnz)
    # Kernel timings with HIP events on the launch stream (null stream).
    stream = torch.cuda.current_stream().cuda_stream
    mode0 = storage_of(g)
    t_pcg_update = g.sim.time_kernel(2, args.kernel_reps, s.dt, stream)
    storage, t_spmv = time_storages(g, s, n, nnz, args, stream, ("general", "symmetric", "symmetric_packed"))
    extra_kernels = {
        "working_set": f"{args.size}x{args.size}: the SpMV pass touches {storage[mode0]['layout_bytes'] / 1e6:.0f} MB, "
                       "an iteration about 1.4x that; within the 256 MB Infinity Cache, so these rates are "
                       "effective (cache-resident), not HBM",
        "spmv_plain": {"bytes": kb["spmv"], "ms": round(t_spmv, 5),
                       "GB/s": round(gbs(kb["spmv"], t_spmv), 1),
                       "frac": round(gbs(kb["spmv"], t_spmv) / PEAK_HBM_GBS, 4)},
        "pcg_update": {"bytes": kb["pcg_update"], "ms": round(t_pcg_update, 5),
                       "GB/s": round(gbs(kb["pcg_update"], t_pcg_update), 1)},
        "pcg_iteration": {"ref_bytes": kb["pcg_iteration_ref"], "ms": round(storage[mode0]["iteration_ms"], 5),
                          "GB/s_ref_formulation": round(gbs(kb["pcg_iteration_ref"],
                                                            storage[mode0]["iteration_ms"]), 1)},
        "pcg_spmv_storage": storage,
        "spmv_pass_roofline": spmv_roofline(storage, mode0, kb, read_traffic(args.size)),
    }
    launch_cfg = g.sim.launch_config()
    tune = [list(t) for t in g.sim.tune_log()]

    cpu = None
    fp_err = None
    if cpu_state is not None:
        cpu, fp_err = cpu_baseline(args, s, g, cpu_state)
    del g

    full = full_run(args, make_sim) if args.full_run else None
    ref_mode = reference_mode(args, s, make_sim) if args.full_run else None

    ws_cmp = None
    if args.ws_steps > 0:
        ws_cmp = warm_start_comparison(args, s, make_sim)

    # The HBM roofline: the same kernels on a cloth whose matrix and vectors
    # exceed the Infinity Cache.
    hbm = hbm_kernels(args, make_sim) if args.hbm_size > 0 else None
    copy = copy_ceiling(args) if args.copy_gib > 0 else None
    if hbm is not None:
        roofline = hbm["spmv_pass_roofline"]
        if copy is not None:
            roofline["copy_ceiling_GB/s"] = copy["best_GB/s"]
            roofline["frac_of_copy_ceiling"] = round(roofline["achieved"] / copy["best_GB/s"], 4)
    else:
        roofline = extra_kernels["spmv_pass_roofline"]

    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "f32 assembly/matrix/SpMV/vectors; f64 iterate dx, inner products, alpha/beta, test",
        "data": "synthetic",
        "config": {
            "workload": f"cloth_{args.size}x{args.size}_wind: 1 m grid, top edge pinned, gravity + seeded wind, "
                        f"springs + stretch k=1e3, m=0.01 kg, dt=1/60, PCG to {args.tol:g} true relative residual",
            "particles": n,
            "off_diag_blocks": nnz,
            "warm_start": "previous-step dx (plain)" if args.warm_start == "previous" else args.warm_start,
            "parallelism": "1 GPU",
        },
        "pcg_iters_per_step": round(float(np.mean(iters)), 2),
        "pcg_restarts_per_step": round(float(np.mean(restarts)), 3),
        "max_relres": float(np.max(relres)),
        "warm_start_comparison": ws_cmp,
        "setup_s": round(setup_s, 2),
        "roofline": roofline,
        "hbm_copy_ceiling": copy,
        "kernels": extra_kernels,
        "kernels_hbm": hbm,
        "full_run": full,
        "reference_mode": ref_mode,
        "cpu_baseline": cpu,
        "fp32_vs_fp64_oracle": fp_err,
        "launch_config": launch_cfg,
        "autotune": tune,
    }
    if args.calibrate_copy:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)  # 256 MiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
