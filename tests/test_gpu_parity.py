"""GPU parity: levels from the HIP kernels (through the C ABI) are bit-exact
against the CPU oracle for every root and every step kind; the GPU graph
builder is bit-exact against the oracle's restatement of the recipe."""
import numpy as np
import pytest

import bfsx
import workloads
from graphs import (binary_tree, csc_from_edges, path_graph, random_graph, star_graph,
                    symmetrize)
from oracle import cpu, mlp_ref

pytestmark = pytest.mark.gpu

STEPS = ["llb", "slb", "glb", "pull"]


def say(*a):
    print(*a, flush=True)


def assert_levels(eng, cp, ri, roots, strategies, n_rows=None, parents=False):
    for r in roots:
        want, ec, _ = cpu.bfs_queue(cp, ri, int(r), n_rows=n_rows)
        for s in strategies:
            out = eng.run(int(r), s, parents=parents)
            lev, st = out[0], out[1]
            bad = np.flatnonzero(lev != want)
            assert bad.size == 0, (s, int(r), bad[:10], lev[bad[:10]], want[bad[:10]])
            assert st["edges_count"] == ec, (s, int(r))
            assert st["reached"] == int((want >= 0).sum())
            assert st["max_level"] == int(want.max())
            if parents:
                assert cpu.check_parents(cp, ri, int(r), want, out[2], n_rows=n_rows) == 0, s


# ---------------------------------------------------------------------------
# graph builder
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scale,ptr", [(8, 4), (11, 4), (13, 8), (15, 4)])
def test_rmat_builder_matches_oracle(engine, scale, ptr):
    say("rmat builder", scale, ptr)
    nnz = engine.generate_rmat(scale, 16, 0.57, 0.19, 0.19, 1000 + scale, 2000 + scale, ptr)
    cp, ri = engine.download_csc()
    wcp, wri = cpu.rmat_csc(scale, 16, 0.57, 0.19, 0.19, 1000 + scale, 2000 + scale, ptr)
    assert nnz == len(wri)
    assert cp.dtype == wcp.dtype
    assert (cp == wcp).all() and (ri == wri).all()
    info = engine.info()
    deg = np.diff(cp.astype(np.int64))
    assert info["max_degree"] == deg.max()
    assert info["n_isolated"] == int((deg == 0).sum())
    assert info["flags"] & bfsx.GRAPH_SYMMETRIC


def test_grid_builder_matches_oracle(engine):
    say("grid builder")
    nnz = engine.generate_grid(64, 0.05, 7, 77, 4)
    cp, ri = engine.download_csc()
    wcp, wri = cpu.grid_csc(64, 0.05, 7, 77, 4)
    assert nnz == len(wri) and (cp == wcp).all() and (ri == wri).all()


def test_builder_matches_committed_checksums(engine):
    """The GPU builder's R-MAT 16, R-MAT 19 and 256 x 256 grid (id-ordered
    lists) have the nnz and SHA-256 digests committed in
    tests/golden/generated_checksums.json."""
    import json
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_generated_checksums as mk
    want = json.load(open(os.path.join(here, "golden", "generated_checksums.json")))
    for name, wl, over in mk.GRAPHS:
        workloads.build(engine, wl, order=None, **over)
        got = mk.digest(*engine.download_csc())
        for k in ("nnz", "vertices", "col_ptr_bytes", "col_ptr_sha256", "row_idx_sha256"):
            assert got[k] == want[name][k], (name, k)


def test_roots_match_oracle(engine):
    w, _ = workloads.build(engine, "rmat16")
    roots = workloads.roots(engine, w)
    cp, ri = engine.download_csc()
    deg = np.diff(cp.astype(np.int64))
    want = cpu.pick_roots((deg > 0).astype(np.uint8), 64, w["seed_roots"])
    assert len(roots) == 64 and (roots == want).all()


# ---------------------------------------------------------------------------
# parity on the workloads
# ---------------------------------------------------------------------------
def test_rmat16_all_roots_all_steps(engine):
    w, _ = workloads.build(engine, "rmat16")
    roots = workloads.roots(engine, w)
    cp, ri = engine.download_csc()
    say("rmat16", len(ri))
    assert len(roots) == 64
    engine.load_scheduler()
    try:
        assert_levels(engine, cp, ri, roots, STEPS + ["learned"])
    finally:
        engine.set_scheduler_weights(None)


def test_rmat19_all_roots_all_steps(engine):
    w, _ = workloads.build(engine, "rmat19")
    roots = workloads.roots(engine, w)
    cp, ri = engine.download_csc()
    say("rmat19", len(ri))
    assert len(roots) == 64
    engine.load_scheduler()
    try:
        assert_levels(engine, cp, ri, roots, STEPS + ["learned"])
    finally:
        engine.set_scheduler_weights(None)


def test_rmat_uint64_col_ptr(engine):
    engine.generate_rmat(14, 16, 0.57, 0.19, 0.19, 5, 6, 8)
    cp, ri = engine.download_csc()
    assert cp.dtype == np.uint64
    roots = engine.pick_roots(8, 7)
    assert_levels(engine, cp, ri, roots, STEPS)


def test_parents_are_valid(engine):
    engine.generate_rmat(14, 8, 0.57, 0.19, 0.19, 15, 16, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(6, 17)
    assert_levels(engine, cp, ri, roots, STEPS, parents=True)


def test_run_batch_equals_single_runs(engine):
    engine.generate_rmat(13, 16, 0.57, 0.19, 0.19, 21, 22, 4)
    roots = engine.pick_roots(5, 23)
    lev, stats = engine.run_batch(roots, "glb", levels=True)
    for i, r in enumerate(roots):
        one, st = engine.run(int(r), "llb")
        assert (lev[i] == one).all()
        assert stats[i]["edges_count"] == st["edges_count"]


# ---------------------------------------------------------------------------
# launch shapes and edge cases
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("wg", [64, 128, 512, 1024])
def test_llb_work_group_sizes(engine, wg):
    engine.generate_rmat(13, 16, 0.57, 0.19, 0.19, 31, 32, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(4, 33)
    old = engine.params
    try:
        engine.set_params(llb_wg=wg, hub_degree=0)
        assert_levels(engine, cp, ri, roots, ["llb", "slb"])
        engine.set_params(hub_degree=64)
        assert_levels(engine, cp, ri, roots, ["llb", "slb"])
    finally:
        engine.set_params(**old)


@pytest.mark.parametrize("ratio", [0, 16, 0xFFFFFFFF])
def test_mark_and_atomic_discovery_modes(engine, ratio):
    """Every level marked+compacted (0), every level exact-once (never), mixed."""
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 71, 72, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(6, 73)
    old = engine.params
    try:
        engine.set_params(mark_ratio_q10=ratio)
        assert_levels(engine, cp, ri, roots, STEPS, parents=True)
        r = int(roots[0])
        want, _, _ = cpu.bfs_queue(cp, ri, r)
        hist = np.bincount(want[want >= 0])
        for s in STEPS:
            engine.run(r, s, trace=True)
            assert [t["frontier_vertices"] for t in engine.last_trace()] == list(hist)
    finally:
        engine.set_params(**old)


@pytest.mark.parametrize("bucket", [0, 512, 0xFFFFFFFF])
def test_glb_bucketed_discovery(engine, bucket):
    """GLB levels that mark use bucketed discovery always (0), when most
    edges lead to unvisited vertices (512) or never; the next frontier
    bitmap it writes feeds bottom-up steps of the learned schedule."""
    engine.generate_rmat(16, 16, 0.57, 0.19, 0.19, 91, 92, 8)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(8, 93)
    old = engine.params
    try:
        engine.set_params(mark_ratio_q10=0, bucket_ratio_q10=bucket)
        assert_levels(engine, cp, ri, roots, ["glb"])
        engine.load_scheduler()
        assert_levels(engine, cp, ri, roots, ["learned"])
        engine.set_params(mark_ratio_q10=16)
        assert_levels(engine, cp, ri, roots, ["glb", "learned"])
    finally:
        engine.set_scheduler_weights(None)
        engine.set_params(**old)


@pytest.mark.parametrize("glb_wg,chunk,bpc", [(128, 256, 1), (256, 512, 4), (1024, 65536, 2)])
def test_glb_shapes(engine, glb_wg, chunk, bpc):
    engine.generate_rmat(13, 16, 0.57, 0.19, 0.19, 41, 42, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(4, 43)
    old = engine.params
    try:
        engine.set_params(glb_wg=glb_wg, glb_chunk=chunk, glb_blocks_per_cu=bpc)
        assert_levels(engine, cp, ri, roots, ["glb"])
    finally:
        engine.set_params(**old)


@pytest.mark.parametrize("ptr", [4, 8])
def test_hub_first_adjacency_order(engine, ptr):
    """Lists re-sorted by (neighbour degree desc, id asc) hold the same
    neighbours; every step kind still gives the oracle's levels and valid
    parents; ordering by id restores the builder's arrays bit-exact."""
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 81, 82, ptr)
    cp, ri = engine.download_csc()
    engine.order_adjacency("hub_first")
    assert engine.info()["flags"] & bfsx.GRAPH_HUB_FIRST
    cp2, ri2 = engine.download_csc()
    assert (cp2 == cp).all()
    deg = np.diff(cp.astype(np.int64))
    col = np.repeat(np.arange(len(deg)), deg)
    assert (ri2[np.lexsort((ri2, col))] == ri).all()  # same multiset per list
    same = col[1:] == col[:-1]
    a, b = ri2[:-1][same], ri2[1:][same]
    assert ((deg[a] > deg[b]) | ((deg[a] == deg[b]) & (a < b))).all()
    roots = engine.pick_roots(6, 83)
    assert_levels(engine, cp, ri, roots, STEPS, parents=True)
    engine.load_scheduler()
    assert_levels(engine, cp, ri, roots, ["learned"])
    engine.set_scheduler_weights(None)
    engine.order_adjacency("id")
    assert not engine.info()["flags"] & bfsx.GRAPH_HUB_FIRST
    cp3, ri3 = engine.download_csc()
    assert (cp3 == cp).all() and (ri3 == ri).all()


@pytest.mark.parametrize("ptr", [4, 8])
def test_pull_hub_first_copy(engine, ptr):
    """A hub-first copy read only by bottom-up steps: the resident lists stay
    in id order, every step kind (bottom-up ones through the copy) gives the
    oracle's levels and valid parents, the multi-source search too; a new
    graph, or re-sorting the lists, drops the copy."""
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 85, 86, ptr)
    cp, ri = engine.download_csc()
    engine.order_adjacency("pull_hub_first")
    info = engine.info()
    flags = info["flags"]
    assert flags & bfsx.GRAPH_PULL_COPY and not flags & bfsx.GRAPH_HUB_FIRST
    # first neighbours among the 128 top-degree vertices are probed in the
    # hub bitmap (one vertex in 256, whole waves)
    assert info["pull_hubs"] == 128
    cp2, ri2 = engine.download_csc()
    assert (cp2 == cp).all() and (ri2 == ri).all()
    roots = engine.pick_roots(6, 87)
    assert_levels(engine, cp, ri, roots, STEPS, parents=True)
    engine.load_scheduler()
    try:
        assert_levels(engine, cp, ri, roots, ["learned"])
    finally:
        engine.set_scheduler_weights(None)
    many = engine.pick_roots(128, 88)
    lev, st = engine.run_multi(many, np.arange(128), 128)
    assert st["pull_levels"] > 0
    assert (lev == cpu.msbfs_u8(cp, ri, many, np.arange(128), 128)).all()
    engine.order_adjacency("id")
    assert not engine.info()["flags"] & bfsx.GRAPH_PULL_COPY
    assert engine.info()["pull_hubs"] == 0
    engine.order_adjacency("pull_hub_first")
    engine.generate_rmat(12, 8, 0.57, 0.19, 0.19, 85, 86, ptr)
    assert not engine.info()["flags"] & bfsx.GRAPH_PULL_COPY
    assert engine.info()["pull_hubs"] == 0
    cp, ri = engine.download_csc()
    assert_levels(engine, cp, ri, engine.pick_roots(3, 89), ["pull"])


def test_pull_scale24_hub_first_copy(engine):
    """R-MAT scale 24 (16.8 M vertices: the LDS frontier summary is one bit
    per 64 vertices, coarser than on the smaller graphs) through bottom-up
    steps only, reading the hub-first copy: levels bit-exact."""
    engine.generate_rmat(24, 16, 0.57, 0.19, 0.19, 91, 92, 8)
    engine.order_adjacency("pull_hub_first")
    cp, ri = engine.download_csc()
    for r in engine.pick_roots(2, 93):
        lev, st = engine.run(int(r), "pull")
        want, ec, _ = cpu.bfs_queue(cp, ri, int(r))
        assert np.array_equal(lev, want) and st["edges_count"] == ec, int(r)


def test_llb_scale24_hot_copy(engine):
    """R-MAT scale 24 (2^24 vertices, the smallest graph that gets the hot
    copy): fixed LLB and SLB, whose heavy levels after half the edges are
    visited read the top-down copy with the 2^20 top-degree neighbours named
    by index and probe them in the hot bitmap; levels bit-exact and parents
    valid, with the default rule and with every level marking."""
    say("rmat24 hot copy")
    engine.generate_rmat(24, 16, 0.57, 0.19, 0.19, 95, 96, 8)
    engine.order_adjacency("pull_hub_first")
    assert engine.info()["hot_vertices"] == 1 << 20
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(2, 97)
    old = engine.params
    try:
        for ratio in (old["mark_ratio_q10"], 0):
            engine.set_params(mark_ratio_q10=ratio)
            assert_levels(engine, cp, ri, roots, ["llb", "slb"], parents=True)
            say("rmat24 hot copy, mark ratio", ratio, "exact")
    finally:
        engine.set_params(**old)
    engine.order_adjacency("id")
    assert engine.info()["hot_vertices"] == 0


def test_directed_graph_with_fewer_rows(engine):
    """Directed input, duplicates, self loops, n_rows < n_cols (reference quirk)."""
    rng = np.random.default_rng(9)
    n = 5000
    cp, ri = random_graph(rng, n, 30000, directed=True)
    n_rows = n - 3
    engine.upload_csc(cp, ri, n_rows=n_rows, symmetric=False)
    assert_levels(engine, cp, ri, rng.integers(0, n, 6), ["llb", "slb", "glb"], n_rows=n_rows)
    with pytest.raises(bfsx.BfsxError) as e:
        engine.run(0, "pull")
    assert e.value.status == 7


def test_closed_form_graphs(engine):
    for cp, ri, root in [(path_graph(3000)[0], path_graph(3000)[1], 1500),
                         (star_graph(70000)[0], star_graph(70000)[1], 5),
                         (binary_tree(100000)[0], binary_tree(100000)[1], 0)]:
        engine.upload_csc(cp, ri, symmetric=True)
        assert_levels(engine, cp, ri, [root], STEPS)


def test_tiny_and_empty_graphs(engine):
    cp, ri = csc_from_edges(1, [], [])
    engine.upload_csc(cp, ri, symmetric=True)
    lev, st = engine.run(0, "llb")
    assert list(lev) == [0] and st["edges_count"] == 0 and st["n_levels"] == 1
    cp, ri = csc_from_edges(100, [], [])
    engine.upload_csc(cp, ri, symmetric=True)
    for s in STEPS:
        lev, st = engine.run(42, s)
        assert lev[42] == 0 and (np.delete(lev, 42) == -1).all()
    cp, ri = csc_from_edges(4, *symmetrize([0, 1], [1, 2]))
    engine.upload_csc(cp, ri, symmetric=True)
    for s in STEPS:
        lev, _ = engine.run(3, s)  # isolated root
        assert list(lev) == [-1, -1, -1, 0]


def test_bad_arguments(engine):
    cp, ri = path_graph(10)
    engine.upload_csc(cp, ri, symmetric=True)
    with pytest.raises(bfsx.BfsxError) as e:
        engine.run(10, "llb")
    assert e.value.status == 1
    with pytest.raises(bfsx.BfsxError):
        engine.run(0, 9)
    with pytest.raises(bfsx.BfsxError):
        engine.set_params(llb_wg=96)
    bad = cp.copy()
    bad[3] = 0
    with pytest.raises(bfsx.BfsxError):
        engine.upload_csc(bad, ri)
    with pytest.raises(bfsx.BfsxError):
        engine.upload_csc(cp, np.full_like(ri, 10))


def test_trace_is_consistent(engine):
    engine.generate_rmat(14, 16, 0.57, 0.19, 0.19, 51, 52, 4)
    cp, ri = engine.download_csc()
    r = int(engine.pick_roots(1, 53)[0])
    want, ec, _ = cpu.bfs_queue(cp, ri, r)
    hist = np.bincount(want[want >= 0])
    deg = np.diff(cp.astype(np.int64))
    for s in STEPS:
        lev, st = engine.run(r, s, trace=True)
        tr = engine.last_trace()
        assert len(tr) == st["n_levels"] == len(hist)
        for k, rec in enumerate(tr):
            assert rec["level"] == k and rec["choice"] == bfsx.STRATEGIES[s]
            assert rec["frontier_vertices"] == hist[k]
            assert rec["frontier_edges"] == deg[want == k].sum()
            assert rec["discovered"] == (hist[k + 1] if k + 1 < len(hist) else 0)
            assert rec["us"] > 0


def test_frontier_edges_with_saturated_degrees(engine):
    """Vertices of 65535 and more neighbours (their 16-bit degree saturates,
    and the degree sums of new frontiers then read col_ptr) discovered by
    every step kind, with the hub-first copy (first-neighbour data of the
    two-phase bottom-up step) and without: the per-level frontier edge
    counts of the trace equal the oracle's degree sums.  A root joined to
    300 middle vertices, each joined to a hub of 70000 leaves, and an
    isolated vertex."""
    n_leaves, n_mid = 70000, 300
    hub, root = 0, 1
    mids = np.arange(2, 2 + n_mid)
    leaves = np.arange(2 + n_mid, 2 + n_mid + n_leaves)
    n = 2 + n_mid + n_leaves + 1
    src = np.concatenate([np.full(n_mid, root), mids, np.full(n_leaves, hub)])
    dst = np.concatenate([mids, np.full(n_mid, hub), leaves])
    cp, ri = csc_from_edges(n, *symmetrize(src, dst))
    deg = np.diff(cp.astype(np.int64))
    assert deg[hub] >= 65535
    engine.upload_csc(cp, ri, symmetric=True)
    for order in ("id", "pull_hub_first"):
        engine.order_adjacency(order)
        for r in (root, int(leaves[5]), int(mids[7])):
            want, ec, _ = cpu.bfs_queue(cp, ri, r)
            hist = np.bincount(want[want >= 0])
            for s in STEPS + ["learned"]:
                lev, st = engine.run(r, s, trace=True)
                assert np.array_equal(lev, want) and st["edges_count"] == ec, (order, r, s)
                tr = engine.last_trace()
                assert [t["frontier_edges"] for t in tr] == [
                    int(deg[want == k].sum()) for k in range(len(hist))], (order, r, s)


@pytest.mark.parametrize("drop", [0.05, 0.3])
def test_largest_component_roots(engine, drop):
    """Seeded roots of the largest connected component equal the oracle's
    seeded choice among the members of that component, labelled by scipy's
    connected_components (0.3 of the grid's edges dropped leaves many
    components)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    engine.generate_grid(256, drop, 6, 99, 4)
    roots = engine.pick_roots(16, 5, mode=1)
    assert len(roots) == 16
    cp, ri = engine.download_csc()
    n = len(cp) - 1
    A = csr_matrix((np.ones(len(ri), np.int8), ri.astype(np.int64), cp.astype(np.int64)), (n, n))
    nc, labels = connected_components(A, directed=False)
    sizes = np.bincount(labels)
    assert sizes.max() * 2 > n
    member = (labels == np.argmax(sizes)).astype(np.uint8)
    want = cpu.pick_roots(member, 16, 5)
    assert (roots == want).all(), (roots, want)
    if drop > 0.2:
        assert nc > 100


@pytest.mark.parametrize("sym", [True, False])
def test_top_degree_roots(engine, sym):
    """Mode 2: the k vertices of largest degree, ties to the smaller id --
    scripts/rootNodes.py:133-137 (networkx degree, stable sort) -- checked
    against networkx on a symmetric graph and a numpy restatement on a
    directed one with many ties."""
    import networkx as nx
    from scipy.sparse import csr_matrix
    if sym:
        engine.generate_rmat(13, 8, 0.57, 0.19, 0.19, 181, 182, 8)
        cp, ri = engine.download_csc()
    else:
        rng = np.random.default_rng(4)
        cp, ri = random_graph(rng, 3000, 9000, directed=True)
        engine.upload_csc(cp, ri, n_rows=2990, symmetric=False)
    n = len(cp) - 1
    deg = np.diff(cp.astype(np.int64))
    for k in (1, 5, 64):
        got = engine.pick_roots(k, 0, mode=2)
        rows = n if sym else 2990
        want = np.lexsort((np.arange(rows), -deg[:rows]))[:k]
        assert (got == want).all(), (k, got[:8], want[:8])
        if sym:
            A = csr_matrix((np.ones(len(ri)), ri.astype(np.int64), cp.astype(np.int64)), (n, n))
            G = nx.from_scipy_sparse_array(A)
            top = [v for v, _ in sorted(G.degree, key=lambda x: x[1], reverse=True)[:k]]
            assert list(got) == top


def test_level_histogram_and_step_times(engine):
    engine.generate_rmat(14, 16, 0.57, 0.19, 0.19, 61, 62, 8)
    cp, ri = engine.download_csc()
    r = int(engine.pick_roots(1, 63)[0])
    want, _, _ = cpu.bfs_queue(cp, ri, r)
    deg = np.diff(cp.astype(np.int64))
    for s in STEPS:
        engine.run(r, s, step_times=True, levels=False)
        tr = engine.last_trace()
        counts, edges = engine.level_histogram()
        hist = np.bincount(want[want >= 0])
        assert (counts == hist).all()
        assert (edges == np.array([deg[want == k].sum() for k in range(len(hist))])).all()
        assert [t["frontier_vertices"] for t in tr] == list(hist)
        assert all(t["us"] > 0 for t in tr)
        # the main kernel's events sit inside the step's
        assert all(0 < t["kernel_us"] <= t["us"] + 1.0 for t in tr), tr
        if s != "pull":
            assert all(t["pull_scanned"] == 0 for t in tr)
            continue
        # a bottom-up level reads every list of a candidate that stays
        # undiscovered, at least one entry of each discovered one, and never
        # more than the candidates' lists
        later = lambda k: (want > k) | (want < 0)
        for t in tr:
            k = t["level"]
            d = int(hist[k + 1]) if k + 1 < len(hist) else 0
            lo = int(deg[later(k + 1)].sum()) + d
            hi = int(deg[later(k)].sum())
            assert lo <= t["pull_scanned"] <= hi, (k, lo, t["pull_scanned"], hi)


def test_short_searches_then_long_ones(engine):
    """Searches of one level (isolated root) and of many levels alternate on
    one context, with and without per-level traces, in the host loop and the
    captured loop: each equals the oracle (the level-end words the host polls
    must never carry a tag an earlier search already left there)."""
    cp, ri = path_graph(300)
    cp = np.concatenate([cp, [cp[-1]]]).astype(cp.dtype)  # vertex 300: isolated
    engine.upload_csc(cp, ri, symmetric=True)
    for trace in (False, True, False):
        for s in ("llb", "slb", "glb", "pull"):
            for r in (300, 0, 300, 150, 300, 300, 7):
                want, _, _ = cpu.bfs_queue(cp, ri, r)
                lev, st = engine.run(r, s, trace=trace)
                assert np.array_equal(lev, want), (trace, s, r)
                assert st["n_levels"] == int(want.max()) + 1
                if s in ("llb", "slb") and not trace:
                    lev, _ = engine.run(r, s, graph_loop=True)
                    assert np.array_equal(lev, want), ("loop", s, r)


def test_back_to_back_runs_into_pinned_buffers(engine):
    """Searches issued back to back, levels copied into pinned host buffers
    (DMA writes to host memory interleaved with the level loop's host-mapped
    level-end words): every run equals the oracle.  The level-end words are
    tagged with their sequence number because such writes can reach the host
    out of order; untagged, about 1 run in 200 of this loop ended a level
    early, having read a new sequence number next to the previous level's
    frontier size."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    hip.hipHostFree.argtypes = [C.c_void_p]
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 91, 1091, 4)
    cp, ri = engine.download_csc()
    n = len(cp) - 1
    root = int(np.argmax(np.diff(cp.astype(np.int64))))
    want, _, _ = cpu.bfs_queue(cp, ri, root)
    lib = bfsx.lib()
    bufs = []
    try:
        for _ in range(8):
            p = C.c_void_p()
            assert hip.hipHostMalloc(C.byref(p), 4 * n, 0) == 0
            bufs.append(p)
        views = [np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int32)), (n,)) for p in bufs]
        st = bfsx.RunStats()
        for s in ("slb", "llb", "glb"):
            sid = bfsx.strategy_id(s)
            for i in range(480):
                rc = lib.bfsx_run(engine._h, root, sid, 0, bufs[i % 8], None, C.byref(st))
                assert rc == 0
                if i % 8 == 7:
                    for v in views:
                        bad = np.flatnonzero(v != want)
                        assert bad.size == 0, (s, i, bad.size, bad[:5])
    finally:
        for p in bufs:
            hip.hipHostFree(p)


def test_stream_copy(engine):
    ms = engine.stream_copy(1 << 26, 4)
    assert ms > 0


# ---------------------------------------------------------------------------
# captured level loop (hipGraph, device-side termination)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("K", [1, 8, 64])
def test_graph_loop_parity(engine, K):
    old = engine.params
    try:
        engine.set_params(graph_levels=K)
        cases = []
        engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 111, 112, 4)
        cases.append((engine.download_csc(), engine.pick_roots(4, 113)))
        engine.generate_grid(200, 0.05, 3, 114, 4)
        cases.append((engine.download_csc(), engine.pick_roots(3, 115, 1)))
        cp, ri = path_graph(2500)
        engine.upload_csc(cp, ri, symmetric=True)
        cases.append(((cp, ri), np.array([3, 1250], np.uint32)))
        for (cp, ri), roots in cases:
            engine.upload_csc(cp, ri, symmetric=True)
            for r in roots:
                want, ec, _ = cpu.bfs_queue(cp, ri, int(r))
                for s in ("llb", "slb"):
                    lev, st = engine.run(int(r), s, graph_loop=True)
                    assert np.array_equal(lev, want), (K, s, int(r))
                    assert st["edges_count"] == ec
                    assert st["n_levels"] == int(want.max()) + 1 == st["steps"][s]
    finally:
        engine.set_params(**old)


def test_graph_loop_learned_and_mark_modes(engine):
    from oracle import mlp_ref
    engine.generate_rmat(14, 16, 0.57, 0.19, 0.19, 121, 122, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(4, 123)
    old = engine.params
    try:
        for ratio in (0, 64, 0xFFFFFFFF):
            engine.set_params(mark_ratio_q10=ratio)
            for seed in range(3):
                engine.set_scheduler_weights(mlp_ref.init_weights(200 + seed) * 3)
                for r in roots:
                    want, ec, _ = cpu.bfs_queue(cp, ri, int(r))
                    lev, st = engine.run(int(r), "learned", graph_loop=True)
                    assert np.array_equal(lev, want) and st["edges_count"] == ec
                    assert st["steps"]["glb"] == st["steps"]["pull"] == 0
                    assert st["steps"]["llb"] + st["steps"]["slb"] == int(want.max()) + 1
    finally:
        engine.set_scheduler_weights(None)
        engine.set_params(**old)


# ---------------------------------------------------------------------------
# device loop (cooperative launch of thin levels)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ratio", [16, 0, 0xFFFFFFFF])
def test_device_loop_parity(engine, ratio):
    """Runs of thin levels in one cooperative launch (mark threshold as set,
    never thin, always thin), fixed kinds and the learned schedule: levels,
    edge counts and per-kind level counts match the oracle and the host
    loop, on an R-MAT graph, a grid with hundreds of levels, a path and a
    directed graph with n_rows < n_cols."""
    old = engine.params
    try:
        engine.set_params(mark_ratio_q10=ratio)
        cases = []
        engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 221, 222, 4)
        engine.order_adjacency("pull_hub_first")
        cases.append((engine.download_csc(), engine.pick_roots(4, 223), True, None))
        engine.generate_grid(300, 0.05, 5, 224, 8)
        cases.append((engine.download_csc(), engine.pick_roots(3, 225, 1), True, None))
        cp, ri = path_graph(3000)
        cases.append(((cp, ri), np.array([5, 1500], np.uint32), True, None))
        rng = np.random.default_rng(12)
        dcp, dri = random_graph(rng, 5000, 30000, directed=True)
        cases.append(((dcp, dri), rng.integers(0, 5000, 3), False, 4990))
        for (cp, ri), roots, sym, n_rows in cases:
            engine.upload_csc(cp, ri, n_rows=n_rows, symmetric=sym)
            if sym:
                engine.order_adjacency("pull_hub_first")
            for r in roots:
                want, ec, _ = cpu.bfs_queue(cp, ri, int(r), n_rows=n_rows)
                for s in ("llb", "slb", "glb") + (("pull",) if sym else ()):
                    lev, st = engine.run(int(r), s, device_loop=True)
                    assert np.array_equal(lev, want), (ratio, s, int(r))
                    assert st["edges_count"] == ec
                    assert st["n_levels"] == int(want.max()) + 1 == sum(st["steps"].values())
                    assert st["steps"][s] == st["n_levels"]
                if sym:
                    for seed in range(2):
                        engine.set_scheduler_weights(mlp_ref.init_weights(240 + seed) * 3)
                        lev, st = engine.run(int(r), "learned", device_loop=True)
                        assert np.array_equal(lev, want), ("learned", seed, int(r))
                        assert st["n_levels"] == int(want.max()) + 1 == sum(st["steps"].values())
                    engine.load_scheduler()
                    lev, st = engine.run(int(r), "learned", device_loop=True)
                    assert np.array_equal(lev, want) and st["edges_count"] == ec
                    engine.set_scheduler_weights(None)
    finally:
        engine.set_scheduler_weights(None)
        engine.set_params(**old)


@pytest.mark.parametrize("ratio", [4, 0, 0xFFFFFFFF])
def test_speculative_steps(engine, ratio):
    """The learned host loop launches the next level's bottom-up or thin
    top-down step before it has read the current level's end; a step whose
    guard fails (another kind picked, an empty or a marking frontier) exits
    without a write.  Levels, edge counts and per-kind level counts with
    speculation match the oracle and the same searches without it, under
    random scheduler weights (frequent guesses that fail) and the trained
    ones, on an R-MAT graph, a grid, a path and a directed graph; searches
    with parents or per-level traces between them stay exact."""
    old = engine.params
    try:
        engine.set_params(mark_ratio_q10=ratio)
        cases = []
        engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 251, 252, 4)
        engine.order_adjacency("pull_hub_first")
        cases.append((engine.download_csc(), engine.pick_roots(3, 253), True, None))
        engine.generate_grid(300, 0.05, 5, 254, 8)
        cases.append((engine.download_csc(), engine.pick_roots(2, 255, 1), True, None))
        cases.append((path_graph(2000), np.array([7, 1000], np.uint32), True, None))
        rng = np.random.default_rng(13)
        dcp, dri = random_graph(rng, 5000, 30000, directed=True)
        cases.append(((dcp, dri), rng.integers(0, 5000, 2), False, 4990))
        for ci, ((cp, ri), roots, sym, n_rows) in enumerate(cases):
            say("speculation", ratio, "case", ci)
            engine.upload_csc(cp, ri, n_rows=n_rows, symmetric=sym)
            if sym:
                engine.order_adjacency("pull_hub_first")
            for r in roots:
                want, ec, _ = cpu.bfs_queue(cp, ri, int(r), n_rows=n_rows)
                for seed in (None, 260, 261, 262):
                    if seed is None:
                        engine.load_scheduler()
                    else:
                        engine.set_scheduler_weights(mlp_ref.init_weights(seed) * 3)
                    steps = []
                    for spec in (1, 0, 1):
                        engine.set_params(speculate=spec)
                        lev, st = engine.run(int(r), "learned")
                        assert np.array_equal(lev, want), (ratio, ci, seed, spec, int(r))
                        assert st["edges_count"] == ec
                        assert st["n_levels"] == int(want.max()) + 1 == sum(st["steps"].values())
                        steps.append(st["steps"])
                    assert steps[0] == steps[1] == steps[2], (ratio, ci, seed, int(r))
                    engine.set_params(speculate=1)
                    lev, st, par = engine.run(int(r), "learned", parents=True)
                    assert np.array_equal(lev, want) and st["steps"] == steps[0]
                    assert cpu.check_parents(cp, ri, int(r), want, par, n_rows=n_rows) == 0
                    lev, st = engine.run(int(r), "learned", trace=True)[:2]
                    assert np.array_equal(lev, want) and st["steps"] == steps[0]
    finally:
        engine.set_scheduler_weights(None)
        engine.set_params(**old)


def test_device_loop_block_counts_and_tiny_graphs(engine):
    old = engine.params
    try:
        engine.generate_grid(200, 0.05, 3, 231, 4)
        cp, ri = engine.download_csc()
        roots = engine.pick_roots(2, 232, 1)
        for groups in (1, 7, 256, 4096):
            engine.set_params(coop_groups=groups, llb_wg=1024 if groups == 7 else 256)
            for r in roots:
                want, _, _ = cpu.bfs_queue(cp, ri, int(r))
                assert np.array_equal(engine.run(int(r), "slb", device_loop=True)[0], want)
        engine.set_params(**old)
        c1, r1 = csc_from_edges(1, [], [])
        engine.upload_csc(c1, r1, symmetric=True)
        lev, st = engine.run(0, "llb", device_loop=True)
        assert list(lev) == [0] and st["n_levels"] == 1
        c4, r4 = csc_from_edges(4, *symmetrize([0, 1], [1, 2]))
        engine.upload_csc(c4, r4, symmetric=True)
        assert list(engine.run(3, "llb", device_loop=True)[0]) == [-1, -1, -1, 0]
        assert list(engine.run(0, "slb", device_loop=True)[0]) == [0, 1, 2, -1]
        for bad in (dict(trace=True), dict(parents=True), dict(graph_loop=True)):
            with pytest.raises(bfsx.BfsxError) as e:
                engine.run(0, "llb", device_loop=True, **bad)
            assert e.value.status == 7
        with pytest.raises(bfsx.BfsxError):
            engine.run(0, "beamer", device_loop=True)
    finally:
        engine.set_params(**old)


def test_graph_loop_unsupported_combinations(engine):
    engine.generate_rmat(10, 8, 0.57, 0.19, 0.19, 1, 2, 4)
    for s in ("glb", "pull"):
        with pytest.raises(bfsx.BfsxError) as e:
            engine.run(0, s, graph_loop=True)
        assert e.value.status == 7
    with pytest.raises(bfsx.BfsxError):
        engine.run(0, "llb", graph_loop=True, parents=True)
    with pytest.raises(bfsx.BfsxError):
        engine.run(0, "llb", graph_loop=True, step_times=True)


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("strategy", ["llb", "glb"])
def test_autotune_installs_valid_parameters(engine, strategy):
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 161, 162, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(6, 163)
    old = engine.params
    try:
        p, before, after = engine.autotune(roots[:3], strategy, passes=1)
        assert after >= before > 0
        assert engine.params == p
        assert_levels(engine, cp, ri, roots[3:], [strategy])
    finally:
        engine.set_params(**old)
