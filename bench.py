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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.batch.step(sc[0].dt)
        st = [b.batch.stats(k) for k in range(len(sc))]
        if not all(x["converged"] for x in st):
            raise RuntimeError("a cloth did not converge")
        its = [x["iters"] for x in st]
        it_mean.append(np.mean(its))
        it_max.append(np.max(its))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_total = int(sum(s.n for s in sc))
    ws_cmp = None
    if args.ws_steps > 0:
        del b
        ws_cmp = batch_warm_start_comparison(args, sc, make_batch)
        b = None

    cpu = None
    if cpu_state is not None:
        from oracle.oracle import oracle_from_scene
        threads = min(16, os.cpu_count() or 1)
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
                         f"{rows_done} particles, {secs_total:.2f} s), scaled by particle count to all 512 cloths"}
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
        "warm_start_comparison": ws_cmp,
        "setup_s": round(setup_s, 2),
        "roofline": None,
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
    path = args.ws_weights or ws.DEFAULT_WEIGHTS
    theta, meta = (None, None)
    if os.path.exists(path):
        theta, meta = ws.load_params(path)
    out = {"warmup": args.warmup, "steps": args.ws_steps, "weights": os.path.relpath(path, ROOT) if theta is not None
           else None}
    if meta is not None:
        out["trained_pairs"] = meta.get("pairs")
        out["train_seconds"] = meta.get("train_seconds")
    pos = {}
    modes = ("previous", "learned", "projection", "projection_learned")
    for mode in modes:
        if mode.endswith("learned") and theta is None:
            out[mode] = None
            continue
        g = make_sim(scene, tol=args.tol, max_iter=20000)
        if theta is not None:
            g.sim.ws_set_params(theta)
        g.sim.set_warm_start_mode(mode)
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
                     "learned_used_frac": round(float(np.mean(used)), 3)}
        del g
    for mode in modes[1:]:
        if out.get(mode):
            out[mode]["iters_ratio_vs_previous"] = round(out["previous"]["pcg_iters_per_step"] /
                                                         out[mode]["pcg_iters_per_step"], 3)
            out[mode]["speedup_vs_previous"] = round(out[mode]["steps_per_s"] / out["previous"]["steps_per_s"], 3)
            out[mode]["max_pos_diff_vs_previous_m"] = float(np.abs(pos[mode] - pos["previous"]).max())
    return out


def batch_warm_start_comparison(args, sc, make_batch):
    """The batch stepped with each warm start (fresh batch per mode)."""
    import numpy as np
    import torch
    from chysx import warmstart as ws
    path = args.ws_weights or ws.DEFAULT_WEIGHTS
    theta = ws.load_params(path)[0] if os.path.exists(path) else None
    out = {"warmup": args.warmup, "steps": args.ws_steps,
           "weights": os.path.relpath(path, ROOT) if theta is not None else None}
    modes = ("previous", "learned", "projection", "projection_learned")
    for mode in modes:
        if mode.endswith("learned") and theta is None:
            out[mode] = None
            continue
        b = make_batch(sc, tol=args.tol, max_iter=20000)
        if theta is not None:
            b.batch.ws_set_params(theta)
        b.batch.set_warm_start_mode(mode)
        for _ in range(args.warmup):
            b.batch.step(sc[0].dt)
        torch.cuda.synchronize()
        it_mean, it_max, used = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.ws_steps):
            b.batch.step(sc[0].dt)
            st = [b.batch.stats(k) for k in range(len(sc))]
            if not all(x["converged"] for x in st):
                raise RuntimeError(f"a cloth did not converge ({mode})")
            its = [x["iters"] for x in st]
            it_mean.append(np.mean(its))
            it_max.append(np.max(its))
            used.append(np.mean([b.batch.ws_stats(k)["used"] for k in range(len(sc))]))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[mode] = {"steps_per_s": round(args.ws_steps / el, 3),
                     "pcg_iters_per_step_mean": round(float(np.mean(it_mean)), 2),
                     "pcg_iters_per_step_max": round(float(np.mean(it_max)), 2),
                     "learned_used_frac": round(float(np.mean(used)), 3)}
        del b
    for mode in modes[1:]:
        if out.get(mode):
            out[mode]["speedup_vs_previous"] = round(out[mode]["steps_per_s"] / out["previous"]["steps_per_s"], 3)
    return out


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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (known bytes) for counter calibration")
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
    _, ro, _, _, _, _ = g.sim.debug_dump_last_solve()
    nnz = int(ro[-1])
    kb = kernel_bytes(n, nnz)
    # Kernel timings with HIP events on the launch stream (null stream).
    stream = torch.cuda.current_stream().cuda_stream
    cfg0 = g.sim.launch_config()
    mode0 = ("general" if not cfg0["spmv_symmetric"] else
             "symmetric_packed" if g.sim.block_floats() == 6 else "symmetric")
    t_pcg_update = g.sim.time_kernel(2, args.kernel_reps, s.dt, stream)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9  # noqa: E731
    # The SpMV pass and a whole iteration on each matrix storage (same
    # matrix values, re-assembled in that storage; speed only).
    storage = {}
    for name, sym, packed in [("general", 0, 0), ("symmetric", 1, 0), ("symmetric_packed", 1, 1)]:
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
                         "GB/s_ref_formulation": round(gbs(kb["pcg_spmv"], t_sp), 1),
                         # the same pass without the solver-state advance it runs first
                         "ms_no_advance": round(t_na, 5),
                         "frac_layout_no_advance": round(gbs(lb, t_na) / PEAK_HBM_GBS, 4),
                         "iteration_ms": round(t_it, 5)}
        if name == "general":
            t_spmv = g.sim.time_kernel(0, args.kernel_reps, s.dt, stream)  # plain SpMV reads general storage
    g.sim.set_launch_config(dict(spmv_symmetric=cfg0["spmv_symmetric"], spmv_packed=cfg0["spmv_packed"]))
    g.sim.time_kernel(3, 1, s.dt, stream)
    t_pcg_spmv = storage[mode0]["ms"]
    t_iter = storage[mode0]["iteration_ms"]
    lb0 = storage[mode0]["layout_bytes"]

    traffic = read_traffic(args.size)
    roofline = {
        "bound": "hbm",
        "kernel": f"pcg_spmv_kernel, {mode0} storage, {storage[mode0]['index_bits']}-bit indices "
                  "(solver-state advance; p = z + beta p_prev; Ap = A p; p.Ap)",
        "achieved": round(gbs(lb0, t_pcg_spmv), 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(gbs(lb0, t_pcg_spmv) / PEAK_HBM_GBS, 4),
        "traffic": traffic.get({"symmetric_packed": "pcg_spmv_kernel", "symmetric": "pcg_spmv_kernel_symmetric",
                                "general": "pcg_spmv_kernel_general"}[mode0]),
        "bytes_per_launch": lb0,
        "ms_per_launch": round(t_pcg_spmv, 5),
        "ref_formulation_bytes": kb["pcg_spmv"],
        "ref_formulation_GB/s": round(gbs(kb["pcg_spmv"], t_pcg_spmv), 1),
    }
    extra_kernels = {
        "spmv_plain": {"bytes": kb["spmv"], "ms": round(t_spmv, 5),
                       "GB/s": round(gbs(kb["spmv"], t_spmv), 1),
                       "frac": round(gbs(kb["spmv"], t_spmv) / PEAK_HBM_GBS, 4)},
        "pcg_update": {"bytes": kb["pcg_update"], "ms": round(t_pcg_update, 5),
                       "GB/s": round(gbs(kb["pcg_update"], t_pcg_update), 1)},
        "pcg_iteration": {"ref_bytes": kb["pcg_iteration_ref"], "ms": round(t_iter, 5),
                          "GB/s_ref_formulation": round(gbs(kb["pcg_iteration_ref"], t_iter), 1)},
        "pcg_spmv_storage": storage,
    }

    ws_cmp = None
    if args.ws_steps > 0:
        ws_cmp = warm_start_comparison(args, s, make_sim)

    cpu = None
    fp_err = None
    if cpu_state is not None:
        from oracle.oracle import oracle_from_scene
        threads = min(16, os.cpu_count() or 1)
        oc, _, _, im, f = oracle_from_scene(s, precision=64, tol=args.tol, max_iter=20000, threads=threads)
        p0, v0, dx0 = cpu_state
        p_start, v_start = p0.copy(), v0.copy()
        oc.build_topology()
        oc.set_dx(dx0)
        secs, it = oc.time_steps(p0, v0, im, f, s.dt, args.cpu_steps)  # p0, v0 advance in place
        cpu = {"value": round(args.cpu_steps / secs, 5), "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle (C++ restatement of the reference step, OpenMP) stepping the "
                         f"{args.size}x{args.size} cloth {args.cpu_steps} step(s) from the state the GPU timed "
                         f"region starts from; {it} PCG iterations, {secs:.2f} s",
               "pcg_iters": it}
        # FP32 GPU step vs the FP64 oracle from the same state (both to the same tolerance)
        g.set_state(p_start, v_start)
        g.sim.set_warm_start(dx0)
        for _ in range(args.cpu_steps):
            g.sim.step(s.dt)
        gp = g.positions().astype(np.float64)
        d_gpu, d_cpu = gp - p_start, p0 - p_start
        fp_err = {"steps": args.cpu_steps, "tol": args.tol,
                  "max_abs_pos_diff_m": float(np.abs(gp - p0).max()),
                  "rel_pos_diff": float(np.linalg.norm(gp - p0) / np.linalg.norm(p0)),
                  "rel_displacement_diff": float(np.linalg.norm(d_gpu - d_cpu) / max(np.linalg.norm(d_cpu), 1e-30))}

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
            "warm_start": "previous-step dx (plain)",
            "parallelism": "1 GPU",
        },
        "pcg_iters_per_step": round(float(np.mean(iters)), 2),
        "pcg_restarts_per_step": round(float(np.mean(restarts)), 3),
        "max_relres": float(np.max(relres)),
        "warm_start_comparison": ws_cmp,
        "setup_s": round(setup_s, 2),
        "roofline": roofline,
        "kernels": extra_kernels,
        "cpu_baseline": cpu,
        "fp32_vs_fp64_oracle": fp_err,
        "launch_config": g.sim.launch_config(),
        "autotune": [list(t) for t in g.sim.tune_log()],
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
