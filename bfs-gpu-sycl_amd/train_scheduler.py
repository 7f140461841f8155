#!/usr/bin/env python3
"""Trains the per-level scheduler MLP on the GPU and writes sched_mlp.json.

1. For each training graph (R-MAT scales 16..24 with edge factors 8..32 and
   two grids, none of them an evaluation workload), generate it in HBM and
   record, for every level of searches from its roots, the scheduler inputs
   and the device time of every step kind from the same saved state
   (bfsx_sched_record).
2. Fit the 8 -> 32 -> 32 -> 4 network with the engine's GPU
   forward/backward/Adam kernels (bfsx_mlp_train).  Targets are log2 of each
   kind's step time relative to LLB's on the same level (the scheduler runs
   the argmin, which only needs the order); rows are weighted by how much
   time the level can cost, so that the slow levels of large graphs decide
   the fit.
3. Evaluate learned vs fixed LLB (harmonic GTEPS) on evaluation workloads
   with their own roots, and save weights plus metadata.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import bfsx  # noqa: E402
import workloads  # noqa: E402

FEATURES = ["log2((1+frontier vertices)/(1+n))/32", "log2((1+frontier edges)/(1+nnz))/32",
            "log2((1+unvisited vertices)/(1+n))/32", "log2((1+unvisited edges)/(1+nnz))/32",
            "log2(1+frontier edges/frontier vertices)/16", "log2(1+nnz/n)/16", "log2(1+n)/32",
            "1 if the last step was bottom-up"]
TRAIN_SETS = ",".join(workloads.TRAIN_WORKLOADS)


def log(*a):
    print(*a, flush=True)


def evaluate(eng, roots, reps=3):
    out = {}
    for s in ("llb", "learned"):
        eng.run_batch(roots, s)  # warm-up
        runs = [[st["ms"] for st in eng.run_batch(roots, s)[1]] for _ in range(reps)]
        _, stats = eng.run_batch(roots, s)
        edges = np.array([st["edges_count"] for st in stats], np.float64)
        out[s] = float(bfsx.harmonic_gteps(edges, np.mean(runs, axis=0)))
        if s == "learned":
            out["learned_steps"] = {k: int(sum(st["steps"][k] for st in stats))
                                    for k in bfsx.STEP_NAMES}
    out["speedup"] = out["learned"] / out["llb"]
    return out


def targets_and_weights(T, M):
    """Relative log2 targets (LLB = 0) and per-row weights.

    A row's weight is log2(1 + best step time / 8 us), at least 1/8: levels
    that take milliseconds count more than levels that take microseconds.
    """
    Tr = np.where(M > 0, T - T[:, :1], 0.0).astype(np.float32)
    best = np.where(M > 0, T, np.inf).min(axis=1)
    w = np.maximum(np.log2(1.0 + 2.0 ** best / 8.0), 0.125)
    return Tr, (M * w[:, None]).astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "sched_mlp.json"))
    ap.add_argument("--datasets", default=TRAIN_SETS)
    ap.add_argument("--eval", default="rmat19,grid4096,rmat22")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--epsilon", type=float, default=0.25)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--eval-roots", type=int, default=64)
    ap.add_argument("--save-data", default="")
    args = ap.parse_args()

    eng = bfsx.Engine(0)
    Xs, Ts, Ms, tags = [], [], [], []
    names = args.datasets.split(",")
    for k, name in enumerate(names):
        t0 = time.time()
        w, nnz = workloads.build(eng, name)
        roots = workloads.roots(eng, w)
        X, T, M = eng.sched_record(roots, repeats=args.repeats, epsilon=args.epsilon,
                                   seed=args.seed + k)
        Xs.append(X); Ts.append(T); Ms.append(M); tags += [name] * len(X)
        us = np.where(M > 0, 2.0 ** T, np.nan)
        log(f"{name}: nnz={nnz} roots={len(roots)} rows={len(X)} in {time.time() - t0:.1f}s; "
            f"total us per kind (llb,glb,slb,pull) = {np.nansum(us, axis=0).round(0).tolist()}, "
            f"per-level best = {np.nansum(np.nanmin(us, axis=1)):.0f}")
    X = np.concatenate(Xs); T = np.concatenate(Ts); M = np.concatenate(Ms)
    if args.save_data:
        np.savez(args.save_data, X=X, T=T, M=M, tags=np.array(tags))
    Tr, Mw = targets_and_weights(T, M)
    W0 = bfsx.sched_init_weights(args.seed)
    t0 = time.time()
    W, curve = eng.mlp_train(X, Tr, Mw, W0, args.epochs, args.lr)
    log(f"trained {len(X)} rows, {args.epochs} epochs in {time.time() - t0:.1f}s: "
        f"loss {curve[0]:.3f} -> {curve[-1]:.4f}")
    pred = eng.mlp_forward(X, W)
    big = np.where(M > 0, pred, np.inf).argmin(axis=1)
    best = np.where(M > 0, T, np.inf).argmin(axis=1)
    log(f"picks the fastest kind on {np.mean(big == best):.3f} of recorded levels")
    us = np.where(M > 0, 2.0 ** T, np.inf)
    rows = np.arange(len(X))
    for name in names:
        sel = np.array(tags) == name
        log(f"  {name}: recorded us  llb {us[sel, 0].sum():.0f}  picked "
            f"{us[rows[sel], big[sel]].sum():.0f}  best {us[rows[sel], best[sel]].sum():.0f}")

    eng.set_scheduler_weights(W)
    evals = {}
    for name in [e for e in args.eval.split(",") if e]:
        w, _ = workloads.build(eng, name)
        roots = workloads.roots(eng, w)[:args.eval_roots]
        evals[name] = evaluate(eng, roots)
        log(f"eval {name}: {evals[name]}")
    meta = {
        "format": "bfsx scheduler MLP 8-32-32-4 ReLU, float32, row-major W1,b1,W2,b2,W3,b3",
        "inputs": FEATURES,
        "outputs": "log2 of the step time of LLB, GLB, SLB, PULL relative to LLB; the smallest "
                   "allowed one is run",
        "train_workloads": names,
        "rows": int(len(X)),
        "epochs": args.epochs,
        "lr": args.lr,
        "init_seed": args.seed,
        "epsilon": args.epsilon,
        "repeats": args.repeats,
        "loss_first": float(curve[0]),
        "loss_final": float(curve[-1]),
        "fastest_pick_rate": float(np.mean(big == best)),
        "eval": evals,
        "device": eng.device_info()["arch"],
        "weights": [float(x) for x in W],
    }
    with open(args.out, "w") as f:
        json.dump(meta, f, indent=1)
    log(f"wrote {args.out}")
    eng.close()


if __name__ == "__main__":
    main()
