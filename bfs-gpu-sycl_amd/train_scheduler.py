#!/usr/bin/env python3
"""Trains the per-level scheduler MLP on the GPU and writes sched_mlp.json.

1. For each training workload (R-MAT scale 19, the 4096 x 4096 grid, R-MAT
   scale 22), generate the graph in HBM and record, for every level of
   searches from training roots, the scheduler inputs and the device time of
   every step kind from the same saved state (bfsx_sched_record).  Training
   roots use a different seed from the workloads' evaluation roots.
2. Fit the 8 -> 32 -> 32 -> 4 network to log2(step microseconds) with the
   engine's GPU forward/backward/Adam kernels (bfsx_mlp_train).
3. Evaluate learned vs fixed LLB (harmonic GTEPS) on each workload's own
   roots, and save weights plus metadata.
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

FEATURES = ["log2(1+frontier vertices)/32", "log2(1+frontier edges)/32",
            "log2(1+unvisited vertices)/32", "log2(1+unvisited edges)/32", "level/64",
            "previous kind & 1", "previous kind >> 1", "log2(1+previous step us)/16"]


def log(*a):
    print(*a, flush=True)


def init_weights(seed):
    from oracle import mlp_ref
    return mlp_ref.init_weights(seed)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "sched_mlp.json"))
    ap.add_argument("--datasets", default="rmat19,grid4096,rmat22")
    ap.add_argument("--roots", type=int, default=32, help="training roots per R-MAT workload")
    ap.add_argument("--grid-roots", type=int, default=6)
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
        n_train = args.grid_roots if w["kind"] == "grid" else args.roots
        roots = eng.pick_roots(n_train, w["seed_roots"] + 100, 1 if w["kind"] == "grid" else 0)
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
    W0 = init_weights(args.seed)
    t0 = time.time()
    W, curve = eng.mlp_train(X, T, M, W0, args.epochs, args.lr)
    log(f"trained {len(X)} rows, {args.epochs} epochs in {time.time() - t0:.1f}s: "
        f"loss {curve[0]:.3f} -> {curve[-1]:.4f}")
    pred = eng.mlp_forward(X, W)
    big = np.where(M > 0, pred, np.inf).argmin(axis=1)
    best = np.where(M > 0, T, np.inf).argmin(axis=1)
    log(f"picks the fastest kind on {np.mean(big == best):.3f} of recorded levels")

    eng.set_scheduler_weights(W)
    evals = {}
    for name in names:
        w, _ = workloads.build(eng, name)
        roots = workloads.roots(eng, w)[:args.eval_roots]
        evals[name] = evaluate(eng, roots)
        log(f"eval {name}: {evals[name]}")
    meta = {
        "format": "bfsx scheduler MLP 8-32-32-4 ReLU, float32, row-major W1,b1,W2,b2,W3,b3",
        "inputs": FEATURES,
        "outputs": "log2(device microseconds) of LLB, GLB, SLB, PULL; the smallest is run",
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
