#!/usr/bin/env python3
"""Benchmark: GTEPS (harmonic mean over 64 roots) of the MI355X BFS engine.

Workload (default): R-MAT scale 26, edge factor 16, a/b/c/d = .57/.19/.19/.05,
seeded permutation, deduplicated and symmetrized, CSC with uint64 col_ptr and
uint32 row_idx, built directly in HBM; 64 seeded roots of degree >= 1.

A step is one sweep of BFS from all 64 roots with the headline strategy.  Per
root, the time is the level loop only (graph resident, levels not copied);
each root's time is the mean of its timed runs after dropping the first 10 %
(the reference's avgExecutionTime90f).  value = 64 / sum_r(t_r / E_r), with
E_r the reference's edgesCount (sum of out-degree over reached vertices).

Also reported: fixed LLB on the same roots (the learned scheduler's
baseline), the dominant step kernel against the HBM roofline, and the
reference CPU BFS timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "bfs-gpu-sycl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def per_root_times(runs):
    """runs: [n_runs][n_roots] ms -> per-root mean after dropping the first 10 %."""
    a = np.asarray(runs, np.float64)
    skip = int(a.shape[0] * 0.1)
    return a[skip:].mean(axis=0)


def sweep(eng, roots, strategy, n):
    out, st = [], None
    for _ in range(n):
        _, st = eng.run_batch(roots, strategy)
        out.append([s["ms"] for s in st])
    return out, st


def algorithmic_bytes(F, E_F, D, w_ptr, L_b=4):
    """Top-down useful bytes of one level: one row_idx read and one visited
    probe per edge, col_ptr read twice and a queue read per frontier vertex,
    a queue write and a level write per discovered vertex."""
    return 5.0 * E_F + F * (2.0 * w_ptr + 4.0) + D * (4.0 + L_b)


def pull_bytes(n, UV, S, D, w_ptr, L_b=4):
    """Bottom-up useful bytes of one level: the visited bitmap read, the
    current frontier bitmap read once and the next one written (n/8 each),
    col_ptr twice per unvisited candidate, one row_idx read and one frontier
    probe (charged 1 byte, as the top-down model charges a visited probe) per
    adjacency entry examined, a level write per discovered vertex."""
    return 3.0 * n / 8.0 + UV * 2.0 * w_ptr + 5.0 * S + D * L_b


KERNEL_NAMES = {"llb": "k_llb", "slb": "k_llb (strided)", "glb": "k_glb_expand + scan kernels",
                "pull": "k_pull2"}


def roofline_leg(eng, roots, strategy, info):
    """Per-level kernel times (HIP events around each step's main kernel on
    the engine stream) and level histograms -> algorithmic bytes / time of
    the dominant step kind's kernel.

    `achieved` charges each launch the bytes its own algorithm must move:
    algorithmic_bytes for a top-down step, pull_bytes (with the adjacency
    entries the launch examined) for a bottom-up one; `frac` = achieved / HBM
    peak.  `achieved_topdown_equiv` charges every launch the top-down bytes of
    its level whatever the step kind (the fixed denominator that makes GTEPS
    comparable across kinds): a bottom-up launch skips most of those edges,
    so that figure can exceed the peak and is no bandwidth.  with_traffic
    adds the measured HBM bytes per launch."""
    import bfsx
    w_ptr = info["col_ptr_bytes"]
    n = float(info["n_cols"])
    live = n - float(info["n_isolated"])  # vertices a bottom-up step may test
    pull = bfsx.STEP_NAMES.index("pull")
    per_kind_bytes = np.zeros(4)   # top-down bytes of the levels
    per_kind_own = np.zeros(4)     # the kind's own model
    per_kind_us = np.zeros(4)
    per_kind_step_us = np.zeros(4)
    per_kind_launches = np.zeros(4)
    for r in roots:
        eng.run(int(r), strategy, levels=False, step_times=True)
        tr = eng.last_trace()
        counts, edges = eng.level_histogram()
        reached = np.cumsum(np.asarray(counts, np.float64))
        for rec in tr:
            L = rec["level"]
            F = float(counts[L])
            E = float(edges[L])
            D = float(counts[L + 1]) if L + 1 < len(counts) else 0.0
            k = rec["choice"]
            td = algorithmic_bytes(F, E, D, w_ptr)
            per_kind_bytes[k] += td
            if k == pull:
                per_kind_own[k] += pull_bytes(n, live - reached[L], float(rec["pull_scanned"]),
                                              D, w_ptr)
            else:
                per_kind_own[k] += td
            per_kind_us[k] += rec["kernel_us"]
            per_kind_step_us[k] += rec["us"]
            per_kind_launches[k] += 1
    k = int(np.argmax(per_kind_us))
    own = per_kind_own[k] / (per_kind_us[k] * 1e-6) / 1e9
    td = per_kind_bytes[k] / (per_kind_us[k] * 1e-6) / 1e9
    name = bfsx.STEP_NAMES[k]
    launches = max(per_kind_launches[k], 1)
    return {
        "kernel": KERNEL_NAMES[name],
        "step": name,
        "bound": "hbm",
        "achieved": round(own, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(own / HBM_PEAK_GBS, 4),
        "traffic": None,
        "launches": int(per_kind_launches[k]),
        "avg_launch_us": round(per_kind_us[k] / launches, 2),
        "avg_step_us": round(per_kind_step_us[k] / launches, 2),
        "alg_bytes_per_launch": round(per_kind_own[k] / launches, 1),
        "achieved_topdown_equiv": round(td, 1),
        "topdown_bytes_per_launch": round(per_kind_bytes[k] / launches, 1),
        "kind_time_share": round(float(per_kind_us[k] / per_kind_us.sum()), 4),
    }


def with_traffic(rl, workload):
    """Attaches the measured HBM bytes per launch and, from them, the
    kernel's actual HBM throughput (traffic / average launch time)."""
    rl["traffic"] = traffic_from_profiles(workload, rl["step"])
    if rl["traffic"]:
        hbm = rl["traffic"] / (rl["avg_launch_us"] * 1e-6) / 1e9
        rl["traffic_over_alg"] = round(rl["traffic"] / rl["alg_bytes_per_launch"], 2)
        rl["fabric_achieved"] = round(hbm, 1)
        rl["fabric_frac"] = round(hbm / HBM_PEAK_GBS, 4)
    return rl


def traffic_from_profiles(workload, step):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    scaled by the copy-kernel calibration: the `hbm_bytes_per_launch:` line of
    profiles/traffic_<step>_<workload>.md, if that summary is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{step}_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip().strip("`")
        if line.startswith("hbm_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def host_cpu():
    """The host the CPU legs run on: logical CPUs, the ones this process may
    use, and the model name from /proc/cpuinfo."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
            "model": model}


class one_core:
    """Pins the calling thread (the one that runs the CPU BFS through ctypes)
    to a single CPU for the duration: the reference's CPU BFS is
    single-threaded (its OpenMP is disabled, src/run_bfs_gpu.hpp:14-15)."""

    def __enter__(self):
        self.old = os.sched_getaffinity(0)
        self.cpu = max(self.old)
        os.sched_setaffinity(0, {self.cpu})
        return self

    def __exit__(self, *a):
        os.sched_setaffinity(0, self.old)


def cpu_bfs_times(cp, ri, roots, gpu_levels_of, budget_s=None):
    """The reference's run_bfs_cpu.hpp (oracle/_ref) or, where that is not
    built, the oracle's restatement of it, on one pinned core: per root
    (edgesCount, seconds).  The GPU levels of every root timed are compared
    bit-exact.  Stops after budget_s seconds (at least one root)."""
    from oracle import cpu
    kind = "reference" if cpu.ref_available() else "port"
    ref = cpu.RefGraph(cp, ri) if kind == "reference" else None
    edges, secs = [], []
    t_start = time.time()
    with one_core() as pin:
        for r in roots:
            if kind == "reference":
                lev, ec, ms = ref.run(int(r))
            else:
                t0 = time.perf_counter()
                lev, ec, _ = cpu.bfs_levelsync(cp, ri, int(r))
                ms = (time.perf_counter() - t0) * 1e3
            want = gpu_levels_of(int(r))
            if not np.array_equal(lev, want):
                raise AssertionError(f"GPU levels differ from the CPU BFS at root {r}")
            edges.append(ec)
            secs.append(ms * 1e-3)
            if budget_s is not None and time.time() - t_start > budget_s:
                break
        core = pin.cpu
    if ref is not None:
        ref.close()
    return kind, core, np.array(edges, np.float64), np.array(secs, np.float64)


def cpu_baseline(eng, roots, gpu_levels_of, budget_s=30.0):
    """CPU BFS on a bounded sample of the headline workload's roots."""
    cp, ri = eng.download_csc()
    kind, core, edges, secs = cpu_bfs_times(cp, ri, roots, gpu_levels_of, budget_s)
    for e, t in zip(edges, secs):
        log(f"cpu baseline: {t * 1e3:.0f} ms, {e / t / 1e9:.4f} GTEPS")
    value = len(edges) / float(np.sum(secs / edges)) / 1e9
    return {"value": round(value, 5), "unit": "GTEPS", "cores": 1, "kind": kind,
            "sample": f"{len(edges)} of {len(roots)} roots (about {budget_s:.0f} s of CPU work), "
                      f"single-threaded run_bfs_cpu pinned to CPU {core}, harmonic mean; "
                      f"GPU levels bit-exact on these roots",
            "roots_checked": int(len(edges)), "host": host_cpu()}


def small_configs(eng, names=("rmat16", "rmat19"), runs=3):
    """R-MAT scales 16 and 19 (the CPU reference's case and the first GPU
    case), all 64 roots: harmonic GTEPS of every strategy on the GPU (first
    10 % of runs dropped) and of the CPU BFS on one core; the GPU levels of
    every root are compared with the CPU's."""
    import bfsx
    import workloads
    out = {}
    for name in names:
        w, nnz = workloads.build(eng, name)
        roots = workloads.roots(eng, w)
        cp, ri = eng.download_csc()
        res = {"vertices": int(len(cp) - 1), "nnz": int(nnz), "roots": int(len(roots)), "gpu": {}}
        for s in ("llb", "glb", "slb", "pull", "learned", "beamer"):
            if s == "learned":
                if not bfsx.have_scheduler_weights():
                    continue
                eng.load_scheduler()
            sweep(eng, roots, s, 1)
            r_runs, st = sweep(eng, roots, s, runs)
            e = np.array([x["edges_count"] for x in st], np.float64)
            res["gpu"][s] = round(float(bfsx.harmonic_gteps(e, per_root_times(r_runs))), 3)
        eng.set_scheduler_weights(None)
        kind, core, edges, secs = cpu_bfs_times(cp, ri, roots,
                                                lambda r: eng.run(r, "llb")[0])
        cpu_g = len(edges) / float(np.sum(secs / edges)) / 1e9
        res["cpu"] = {"value": round(cpu_g, 5), "unit": "GTEPS", "cores": 1, "kind": kind,
                      "roots": int(len(edges)), "cpu_ms_total": round(float(secs.sum()) * 1e3, 1)}
        log(f"{name}: gpu {res['gpu']} cpu {cpu_g:.4f} GTEPS over {len(edges)} roots")
        out[name] = res
    return out


def bfsx_weights():
    import bfsx
    return bfsx.have_scheduler_weights()


def sequential_searches(eng, sources, off, args):
    """The batch's searches one at a time, each on its own graph, learned
    schedule: the 256 roots on R-MAT scale 22 alone, each small graph's root
    on that graph alone (same seeds, so the same vertex ids up to the
    packing offset).  Aggregate = all edges / all search time (the mean of
    the timed runs per search).  Rebuilds the graphs, so it runs after the
    batch."""
    import workloads
    n = max(1, min(args.steps, 3))
    edges, ms = 0.0, 0.0
    eng.load_scheduler()
    try:
        workloads.build(eng, "rmat22")
        big = np.asarray(sources[:256], np.uint32)
        _, st = eng.run_batch(big, "learned")  # warm-up
        runs, st = sweep(eng, big, "learned", n)
        edges += float(sum(x["edges_count"] for x in st))
        ms += float(np.sum(per_root_times(runs)))
        a, b, c = workloads.RMAT_ABC
        for i, (scale, ef) in enumerate(workloads.small_graph_shapes(len(off) - 2)):
            eng.generate_rmat(scale, ef, a, b, c, 4000 + i, 5000 + i, 4)
            eng.order_adjacency(workloads.ORDER)
            r = np.array([int(sources[256 + i]) - int(off[i + 1])], np.uint32)
            eng.run_batch(r, "learned")
            runs, st = sweep(eng, r, "learned", n)
            edges += float(st[0]["edges_count"])
            ms += float(np.sum(per_root_times(runs)))
    finally:
        eng.set_scheduler_weights(None)
    out = {"gteps": round(edges / (ms * 1e-3) / 1e9, 2), "ms_total": round(ms, 3),
           "searches": int(len(sources)), "edges": edges,
           "graphs": "R-MAT 22 alone for its 256 roots, each small graph alone for its root"}
    log(f"sequential learned searches: {out}")
    return out


def bench_batch(eng, args):
    """R-MAT scale 22 with 256 searches in one bitset batch, packed with 64
    small R-MAT graphs searched in the same launches; aggregate GTEPS =
    sum over all searches of edgesCount / batch time."""
    import workloads
    t0 = time.time()
    sources, bits, off, n_bits = workloads.build_batch(eng)
    info = eng.info()
    log(f"built batch: V={info['n_cols']} nnz={info['nnz']} graphs={len(off) - 1} "
        f"in {time.time() - t0:.1f}s")
    # the batch's launch parameters tuned on this batch (outside the timed
    # region): relabelling, direction divisor, hub threshold, levels per
    # captured launch
    tuned, tb, ta = eng.autotune_multi(sources, bits, passes=1)
    log(f"batch autotune: {tb:.1f} -> {ta:.1f} GTEPS", {k: tuned[k] for k in tuned if k.startswith("ms_")})
    for _ in range(args.warmup):
        eng.run_multi(sources, bits, n_bits, levels=False)
    times, st = [], None
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        _, st = eng.run_multi(sources, bits, n_bits, levels=False)
        times.append(st["ms"])
    wall = time.perf_counter() - t_wall
    t = float(np.mean(times[int(len(times) * 0.1):]))
    value = st["edges_count"] / (t * 1e-3) / 1e9
    seq = sequential_searches(eng, sources, off, args) if bfsx_weights() else None
    return {
        "metric": "aggregate GTEPS of one batched launch: 256 searches on R-MAT scale 22 "
                  "plus 64 packed R-MAT graphs (scales 12-18), uint8 levels",
        "value": round(value, 2), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / max(args.steps, 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "uint32",
        "data": "synthetic (graphs generated on the GPU from fixed seeds)",
        "config": {"workload": "batch22", "vertices": info["n_cols"], "nnz": info["nnz"],
                   "searches": int(len(sources)), "bits": n_bits, "graphs": int(len(off) - 1),
                   "adjacency_order": workloads.ORDER,
                   "levels": st["n_levels"], "push_levels": st["push_levels"],
                   "pull_levels": st["pull_levels"], "batch_ms": round(t, 3),
                   "edges_all_searches": st["edges_count"]},
        "tuned": {k: v for k, v in tuned.items() if k.startswith("ms_")},
        "sequential_learned": seq,
        "speedup_vs_sequential": round(value / seq["gteps"], 3) if seq else None,
    }


def compare_sweep(eng, roots, strategy, n):
    """Harmonic GTEPS of `strategy` on the same roots (one warm-up sweep, n
    timed sweeps, first 10 % dropped per root)."""
    import bfsx
    sweep(eng, roots, strategy, 1)
    runs, st = sweep(eng, roots, strategy, n)
    e = np.array([x["edges_count"] for x in st], np.float64)
    return float(bfsx.harmonic_gteps(e, per_root_times(runs))), st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat26")
    ap.add_argument("--strategy", default="auto",
                    help="headline strategy: auto (learned when trained weights exist, else llb)")
    ap.add_argument("--roots", type=int, default=0, help="limit the number of roots (0 = all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the R-MAT 16 / 19 legs (GPU strategies and CPU on all 64 roots)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuned fixed-LLB leg")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--compare-steps", type=int, default=3,
                    help="timed sweeps of each comparison strategy (fixed LLB, tuned LLB, "
                         "direction-optimising rule)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the streaming copy probe over this many bytes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(the search would shard by roots or by destination slices)",
              file=sys.stderr)
        sys.exit(2)

    import bfsx
    import workloads

    eng = bfsx.Engine(0)
    dev = eng.device_info()
    log("device", dev)
    if args.workload == "batch22":
        print(json.dumps(bench_batch(eng, args)), flush=True)
        eng.close()
        return
    if args.copy_probe:
        # plain streaming copy of known size: calibrates the HBM byte counters
        # (FETCH_SIZE / WRITE_SIZE) and gives the achievable copy bandwidth
        ms = eng.stream_copy(args.copy_probe, repeats=10)
        print(json.dumps({"probe": "k_stream_copy", "bytes": args.copy_probe, "ms": ms,
                          "GBps_read_plus_write": 2 * args.copy_probe / (ms * 1e-3) / 1e9}),
              flush=True)
        eng.close()
        return
    # achievable streaming bandwidth on this GPU, unprofiled: a 1 GiB copy
    copy_bytes = 1 << 30
    copy_ms = eng.stream_copy(copy_bytes, repeats=20)
    copy_probe = {"kernel": "k_stream_copy", "bytes": copy_bytes, "ms": round(copy_ms, 4),
                  "GBps_read_plus_write": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1)}
    log("copy probe", copy_probe)
    small = None
    if not args.no_small and args.workload == "rmat26":
        small = small_configs(eng) if not args.no_cpu else None
    t0 = time.time()
    w, nnz = workloads.build(eng, args.workload)
    info = eng.info()
    log(f"built {args.workload}: V={info['n_cols']} nnz={nnz} in {time.time() - t0:.1f}s")
    roots = workloads.roots(eng, w)
    if args.roots:
        roots = roots[:args.roots]
    strategy = args.strategy
    if strategy == "auto":
        strategy = "learned" if bfsx.have_scheduler_weights() else "llb"
    sched_meta = None
    if strategy == "learned":
        sched_meta = eng.load_scheduler()
    log("strategy", strategy, "roots", len(roots))

    # warm-up, then the timed steps (each a sweep over all roots)
    sweep(eng, roots, strategy, args.warmup)
    t_wall = time.perf_counter()
    runs, stats = sweep(eng, roots, strategy, args.steps)
    wall = time.perf_counter() - t_wall
    edges = np.array([s["edges_count"] for s in stats], np.float64)
    t_root = per_root_times(runs)
    value = bfsx.harmonic_gteps(edges, t_root)
    log(f"{strategy}: {value:.2f} GTEPS (harmonic, {len(roots)} roots)")

    result = {
        "metric": "GTEPS (harmonic mean, 64 roots) on 1 MI355X: learned vs fixed LLB; HBM %peak",
        "value": round(float(value), 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / max(args.steps, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (graph generated on the GPU from fixed seeds)",
        "config": {"workload": args.workload, "kind": w["kind"],
                   **({"scale": w["scale"], "edge_factor": w["edge_factor"]}
                      if w["kind"] == "rmat" else {"side": w["side"]}),
                   "vertices": info["n_cols"], "nnz": info["nnz"],
                   "col_ptr_bytes": info["col_ptr_bytes"], "roots": len(roots),
                   "strategy": strategy, "adjacency_order": w["order"],
                   "mean_edges_per_root": float(edges.mean())},
        "graph500_gteps": round(float(value) / 2, 3),
        "copy_probe": copy_probe,
    }
    if strategy != "llb":
        n = args.compare_steps
        llb, _ = compare_sweep(eng, roots, "llb", n)
        result["fixed_llb_gteps"] = round(llb, 3)
        judge = llb
        if not args.no_tune:
            # the judge with its launch shape tuned on 8 of the roots
            default = eng.params
            tuned, before, after = eng.autotune(roots[:8], "llb", passes=1)
            llb_t, _ = compare_sweep(eng, roots, "llb", n)
            eng.set_params(**default)
            result["fixed_llb_tuned"] = {"gteps": round(llb_t, 3), "params": tuned,
                                         "tune_roots": 8}
            judge = max(llb, llb_t)
        result["speedup_vs_fixed_llb"] = round(float(value / judge), 3)
        beam, st_b = compare_sweep(eng, roots, "beamer", n)
        result["beamer_gteps"] = round(beam, 3)
        result["speedup_vs_beamer"] = round(float(value / beam), 3)
        result["steps_used"] = {k: int(sum(s["steps"][k] for s in stats)) for k in bfsx.STEP_NAMES}
        result["steps_used_beamer"] = {k: int(sum(s["steps"][k] for s in st_b))
                                       for k in bfsx.STEP_NAMES}
        if sched_meta is not None:
            result["scheduler"] = {"train_workloads": sched_meta["train_workloads"],
                                   "rows": sched_meta["rows"],
                                   "loss_final": round(sched_meta["loss_final"], 4)}
    if not args.no_roofline:
        result["roofline"] = with_traffic(roofline_leg(eng, roots, strategy, info), args.workload)
        if strategy != "llb":
            # the top-down explore kernel on its own (fixed LLB run)
            result["explore_roofline_llb"] = with_traffic(roofline_leg(eng, roots, "llb", info),
                                                          args.workload)
        # whole searches against the top-down algorithmic bytes (the same
        # denominator for every step kind, like TEPS): above 1 means faster
        # than a top-down search streaming at the HBM peak could be
        reached = np.array([s["reached"] for s in stats], np.float64)
        b_alg = 5.0 * edges + reached * (2.0 * info["col_ptr_bytes"] + 8.0 + 4.0)
        ach = float(b_alg.sum() / (t_root.sum() * 1e-3) / 1e9)
        result["search_topdown_equiv"] = {"achieved": round(ach, 1), "unit": "GB/s",
                                          "ratio_to_peak": round(ach / HBM_PEAK_GBS, 4),
                                          "bytes_per_root": round(float(b_alg.mean()), 1)}
    if not args.no_cpu:
        lev_cache = {}

        def gpu_levels_of(r):
            if r not in lev_cache:
                lev_cache[r] = eng.run(r, strategy)[0]
            return lev_cache[r]

        result["cpu_baseline"] = cpu_baseline(eng, roots, gpu_levels_of, args.cpu_budget)
    if small is not None:
        result["small_configs"] = small
    eng.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
