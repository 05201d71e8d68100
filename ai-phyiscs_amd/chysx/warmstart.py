"""Learned PCG warm start: training driver, parameter files and evaluation.

The network itself lives in the library (BF16 MFMA forward and backward
kernels, Adam on the GPU).  Training pairs (state, converged dx) come from
stepping random W x W cloths (random pins, wind, stiffness and mass; W = 256
by default, 1024 for the committed parameters) with training on: each step
of a batch of cloths gives one pair per cloth and one Adam step over the
whole batch on the line-search loss (the share of the
base guess's A-norm error the network's direction leaves).  The base guess
is the previous step's dx for the 'learned' mode and the Galerkin
projection for 'projection_learned'; a parameter file is trained for one of
them.  The solution each step converges to does not depend on the warm
start.
"""

from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional

import numpy as np

from . import _chysx_native
from . import scenes as _scenes
from .harness import make_batch

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "warmstart_cloth.npz")
# parameters trained on top of the projection (mode projection_learned)
PROJECTION_WEIGHTS = os.path.join(DATA_DIR, "warmstart_cloth_projection.npz")


def weights_for(mode: str) -> str:
    """The committed parameter file for a warm-start mode: the one trained on
    top of the projection for projection_learned when it exists, else the one
    trained from the previous step's dx."""
    if mode == "projection_learned" and os.path.exists(PROJECTION_WEIGHTS):
        return PROJECTION_WEIGHTS
    return DEFAULT_WEIGHTS


def num_params() -> int:
    return int(_chysx_native.ws_num_params())


def save_params(path: str, theta: np.ndarray, meta: Dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, theta=np.asarray(theta, np.float32), meta=np.array(json.dumps(meta)))


def load_params(path: str = DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        theta = z["theta"].astype(np.float32)
        meta = json.loads(str(z["meta"]))
    if theta.size != num_params():
        raise ValueError(f"{path}: {theta.size} parameters, the network has {num_params()}")
    return theta, meta


def train(pairs: int = 50000, cloths: int = 16, width: int = 256, episode: int = 120, lr: float = 1e-3,
          lr_final: float = 1e-4, seed: int = 5, tol: float = 1e-6, log_every: int = 50,
          theta0: Optional[np.ndarray] = None, time_budget_s: Optional[float] = None, verbose: bool = True,
          mode: str = "learned"):
    """Online training on fresh batches of random cloths.  Returns (theta,
    meta) where meta records the pairs seen, the loss curve and timing."""
    import torch
    rng_seed = seed
    theta = theta0
    seen = 0
    losses = []
    t0 = time.time()
    steps_total = int(np.ceil(pairs / cloths))
    step_global = 0
    episode_idx = 0
    while seen < pairs:
        sc = _scenes.config5_train(count=cloths, w=width, seed=rng_seed + 1000 * episode_idx)
        b = make_batch(sc, tol=tol, force_active=True)
        if theta is None:
            b.batch.ws_init_params(seed)
        else:
            b.batch.ws_set_params(theta)
        b.batch.set_warm_start_mode(mode)
        for k in range(episode):
            frac = min(step_global / max(steps_total - 1, 1), 1.0)
            cur_lr = lr * (lr_final / lr) ** frac
            b.batch.ws_set_training(True, cur_lr)
            b.batch.step(sc[0].dt)
            st0 = b.batch.ws_stats(0)
            if k >= 2:  # the network trains once two solutions exist
                seen += cloths
                step_global += 1
                losses.append(float(st0["loss"]))
                if verbose and step_global % log_every == 0:
                    print(f"[warmstart] pairs {seen:6d}  loss {np.mean(losses[-log_every:]):.4e}  "
                          f"lr {cur_lr:.2e}  {time.time() - t0:6.1f}s", flush=True)
            if seen >= pairs:
                break
            if time_budget_s is not None and time.time() - t0 > time_budget_s:
                break
        theta = b.batch.ws_get_params()
        del b
        torch.cuda.synchronize()
        episode_idx += 1
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            break
    meta = {"mode": mode, "pairs": int(seen), "cloths_per_batch": cloths, "width": width, "episode_steps": episode,
            "lr": lr, "lr_final": lr_final, "seed": seed, "tol": tol, "train_seconds": round(time.time() - t0, 1),
            "loss_first": float(np.mean(losses[:20])) if losses else None,
            "loss_last": float(np.mean(losses[-20:])) if losses else None,
            "loss_curve": [float(np.mean(losses[i:i + log_every])) for i in range(0, len(losses), log_every)]}
    return theta, meta


MODES = ("previous", "learned", "projection", "projection_learned")


def compare(sim_factory, steps: int, warmup: int, theta: np.ndarray, modes=MODES, accept: str = "energy"):
    """Run the same workload with each warm start; returns per-mode mean
    iterations, steps/s, share of steps using each stage and final positions.
    `accept`: the network direction's acceptance rule ('energy' or
    'residual', see ClothSimulator.ws_set_acceptance)."""
    import torch
    out = {}
    for mode in modes:
        g = sim_factory()
        g.sim.ws_set_params(theta)
        g.sim.set_warm_start_mode(mode)
        g.sim.ws_set_acceptance(accept)
        for _ in range(warmup):
            g.sim.step(g.scene.dt)
        torch.cuda.synchronize()
        its, used, proj = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            g.sim.step(g.scene.dt)
            st = g.sim.last_solve_stats()
            if not st["converged"]:
                raise RuntimeError(f"PCG did not converge ({mode}): {st}")
            its.append(st["iters"])
            ws = g.sim.ws_stats()
            used.append(ws["used"])
            proj.append(ws["projection_used"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[mode] = {"iters": float(np.mean(its)), "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps,
                     "learned_used_frac": float(np.mean(used)), "projection_used_frac": float(np.mean(proj)),
                     "pos": g.positions()}
    return out


def _main(argv=None):
    import argparse
    import torch  # noqa: F401  (one HIP runtime: torch first)
    from . import scenes as sc
    from .harness import make_sim

    ap = argparse.ArgumentParser(description="Train / evaluate the learned PCG warm start.")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=50000)
    t.add_argument("--cloths", type=int, default=16)
    t.add_argument("--width", type=int, default=256)
    t.add_argument("--episode", type=int, default=120)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-final", type=float, default=1e-4)
    t.add_argument("--seed", type=int, default=5)
    t.add_argument("--time-budget", type=float, default=None)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--mode", choices=["learned", "projection_learned"], default="learned")
    t.add_argument("--init", default=None, help="continue from this parameter file (same mode)")
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--width", type=int, default=1024)
    e.add_argument("--scene", choices=["wind", "train"], default="wind")
    e.add_argument("--steps", type=int, default=60)
    e.add_argument("--warmup", type=int, default=10)
    e.add_argument("--accept", choices=["energy", "residual"], default="energy")
    a = ap.parse_args(argv)
    if a.cmd == "train":
        theta0, meta0 = load_params(a.init) if a.init else (None, None)
        if meta0 is not None and meta0.get("mode") != a.mode:
            raise SystemExit(f"--init {a.init} was trained for mode {meta0.get('mode')}, not {a.mode}")
        theta, meta = train(a.pairs, a.cloths, a.width, a.episode, a.lr, a.lr_final, a.seed, theta0=theta0,
                            time_budget_s=a.time_budget, mode=a.mode)
        if meta0 is not None:
            # the pairs this file has seen in all: the starting file's and this run's
            meta["init"] = {k: meta0[k] for k in ("pairs", "cloths_per_batch", "width", "lr", "lr_final", "seed",
                                                  "train_seconds") if k in meta0}
            meta["pairs_this_run"] = meta["pairs"]
            meta["pairs"] = int(meta0["pairs"]) + int(meta["pairs"])
        save_params(a.out, theta, meta)
        print(json.dumps({k: v for k, v in meta.items() if k != "loss_curve"}), flush=True)
    else:
        theta, _ = load_params(a.weights)
        if a.scene == "wind":
            scene = sc.config3(w=a.width)
        else:
            scene = sc.config5_train(count=1, w=a.width, seed=12345)[0]
        r = compare(lambda: make_sim(scene, tol=1e-6), a.steps, a.warmup, theta, accept=a.accept)
        out = {m: {k: round(v, 4) for k, v in r[m].items() if k != "pos"} for m in r}
        out["scene"] = scene.name
        out["accept"] = a.accept
        for m in MODES[1:]:
            out[m]["iter_ratio_vs_previous"] = round(r["previous"]["iters"] / max(r[m]["iters"], 1e-9), 4)
            out[m]["time_ratio_vs_previous"] = round(r[m]["steps_per_s"] / r["previous"]["steps_per_s"], 4)
            out[m]["max_pos_diff_m"] = float(np.abs(r["previous"]["pos"] - r[m]["pos"]).max())
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    _main()
