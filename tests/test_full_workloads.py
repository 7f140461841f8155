"""The benchmark configurations at full size, levels bit-exact against the
oracle (16 host threads run the oracle's FIFO BFS while the GPU searches):

* R-MAT scale 26, edge factor 16 (2.1 G stored edges, uint64 col_ptr), all
  64 roots, learned schedule and fixed LLB;
* the 4096 x 4096 grid (5 % of edges dropped, 0.01 % long-range edges), 64
  roots of the largest component, every step kind and the learned schedule
  in the host loop, fixed LLB/SLB and the learned schedule in the captured
  hipGraph loop;
* the batched launch: 256 searches on R-MAT scale 22 and 64 packed R-MAT
  graphs of scales 12-18 in the same bitset batch, uint8 levels.

Each takes under two minutes on one MI355X and up to ~30 GB of host memory.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bfsx
import workloads
from oracle import cpu

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

POOL = 16  # host threads for the oracle (the GPU box's CPU share)


def say(*a):
    print(*a, flush=True)


def check_roots(engine, cp, ri, roots, modes):
    """modes: [(strategy, loop)], loop False (host loop), True (captured
    hipGraph loop) or "device" (device loop); the oracle runs 16 roots at a
    time."""
    with ThreadPoolExecutor(POOL) as pool:
        for lo in range(0, len(roots), POOL):
            batch = [int(r) for r in roots[lo:lo + POOL]]
            want = list(pool.map(lambda r: cpu.bfs_queue(cp, ri, r), batch))
            for r, (lev_w, ec_w, _) in zip(batch, want):
                for s, loop in modes:
                    lev, st = engine.run(r, s, graph_loop=loop is True,
                                         device_loop=loop == "device")
                    bad = np.flatnonzero(lev != lev_w)
                    assert bad.size == 0, (s, loop, r, bad[:5], lev[bad[:5]], lev_w[bad[:5]])
                    assert st["edges_count"] == ec_w, (s, loop, r)
                    assert st["max_level"] == int(lev_w.max()), (s, loop, r)
            say("roots", lo, "to", lo + len(batch) - 1, "exact")


def test_rmat26_all_64_roots(engine):
    w, nnz = workloads.build(engine, "rmat26")
    roots = workloads.roots(engine, w)
    assert len(roots) == 64
    cp, ri = engine.download_csc()
    assert cp.dtype == np.uint64 and len(ri) == nnz and nnz > 2 * 10 ** 9
    say("rmat26", nnz)
    engine.load_scheduler()
    try:
        check_roots(engine, cp, ri, roots, [("learned", False), ("llb", False),
                                            ("learned", "device")])
    finally:
        engine.set_scheduler_weights(None)


def test_grid4096_all_64_roots(engine):
    w, nnz = workloads.build(engine, "grid4096")
    roots = workloads.roots(engine, w)
    assert len(roots) == 64
    cp, ri = engine.download_csc()
    say("grid4096", nnz)
    lev0 = cpu.bfs_queue(cp, ri, int(roots[0]))[0]
    assert (lev0 >= 0).sum() * 2 > len(lev0) and (lev0[roots] >= 0).all()
    assert lev0.max() > 300  # many thin levels
    engine.load_scheduler()
    try:
        check_roots(engine, cp, ri, roots,
                    [(s, False) for s in ("llb", "slb", "glb", "pull", "learned")] +
                    [(s, True) for s in ("llb", "slb", "learned")] +
                    [(s, "device") for s in ("llb", "learned")])
    finally:
        engine.set_scheduler_weights(None)


def test_batch22_with_64_graphs(engine):
    sources, bits, off, n_bits = workloads.build_batch(engine)
    assert len(sources) == 320 and n_bits == 256 and len(off) == 66
    cp, ri = engine.download_csc()
    say("batch", len(cp) - 1, len(ri))
    lev, st = engine.run_multi(sources, bits, n_bits)
    assert lev.shape == (256, len(cp) - 1)

    def lane(b):
        src = sources[bits == b]
        return cpu.msbfs_u8(cp, ri, src, np.zeros(len(src), np.uint16), 1)[0]

    deg = np.diff(cp.astype(np.int64))
    edges = 0
    with ThreadPoolExecutor(POOL) as pool:
        for lo in range(0, 256, POOL):
            for b, want in zip(range(lo, lo + POOL), pool.map(lane, range(lo, lo + POOL))):
                bad = np.flatnonzero(lev[b] != want)
                assert bad.size == 0, (b, bad[:5], lev[b][bad[:5]], want[bad[:5]])
                edges += int(deg[want != 0xFF].sum())
    assert st["edges_count"] == edges
    # each packed graph's search stays inside its graph and equals a plain BFS
    for g in range(64):
        a, z = int(off[g + 1]), int(off[g + 2])
        r = int(sources[256 + g])
        want = cpu.bfs_queue(cp, ri, r)[0]
        w8 = np.where(want < 0, 255, want).astype(np.uint8)
        assert (lev[g][a:z] == w8[a:z]).all(), g
    say("batch exact:", st)
