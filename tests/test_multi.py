"""Batched multi-source search (bitsets of up to 256 searches, uint8 levels)
against the oracle's per-bit multi-source BFS; packed multi-graph CSC against
the per-graph recipe."""
import numpy as np
import pytest

import bfsx
from graphs import path_graph, random_graph
from oracle import cpu

pytestmark = pytest.mark.gpu


def check(eng, cp, ri, sources, bits, n_bits, n_rows=None):
    lev, st = eng.run_multi(sources, bits, n_bits)
    want = cpu.msbfs_u8(cp, ri, sources, bits, n_bits, n_rows=n_rows)
    bad = np.argwhere(lev != want)
    assert bad.size == 0, (n_bits, bad[:5], lev[tuple(bad[0])], want[tuple(bad[0])])
    deg = np.diff(cp.astype(np.int64))
    rows = len(deg) if n_rows is None else n_rows
    reached = (want[:, :rows] != 0xFF)
    assert st["edges_count"] == int((reached * deg[:rows]).sum())
    return st


@pytest.mark.parametrize("n_bits", [1, 17, 32, 33, 64, 100, 128, 200, 256])
def test_multi_source_matches_oracle(engine, n_bits):
    engine.generate_rmat(14, 16, 0.57, 0.19, 0.19, 131, 132, 4)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(n_bits, 133)
    st = check(engine, cp, ri, roots, np.arange(n_bits), n_bits)
    assert st["n_levels"] >= 3


def test_multi_source_push_and_pull_levels(engine):
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 141, 142, 8)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(256, 143)
    st = check(engine, cp, ri, roots, np.arange(256), 256)
    assert st["push_levels"] > 0 and st["pull_levels"] > 0


def test_shared_bits_and_duplicate_sources(engine):
    engine.generate_rmat(13, 8, 0.57, 0.19, 0.19, 151, 152, 4)
    cp, ri = engine.download_csc()
    rng = np.random.default_rng(0)
    roots = engine.pick_roots(40, 153)
    sources = np.concatenate([roots, roots[:5]])
    bits = np.concatenate([rng.integers(0, 9, 40), [1, 2, 3, 4, 5]])
    check(engine, cp, ri, sources, bits, 9)


def test_directed_graph_runs_top_down(engine):
    rng = np.random.default_rng(3)
    cp, ri = random_graph(rng, 3000, 20000, directed=True)
    engine.upload_csc(cp, ri, n_rows=2990, symmetric=False)
    src = rng.integers(0, 3000, 70)
    st = check(engine, cp, ri, src, np.arange(70), 70, n_rows=2990)
    assert st["pull_levels"] == 0


def test_level_overflow_is_an_error(engine):
    cp, ri = path_graph(400)
    engine.upload_csc(cp, ri, symmetric=True)
    with pytest.raises(bfsx.BfsxError) as e:
        engine.run_multi([0], [0], 1)
    assert e.value.status == 6
    lev, _ = engine.run_multi([200], [0], 1)  # deepest level 200 fits
    assert lev[0, 0] == 200 and lev[0, 399] == 199


def test_packed_graphs_match_recipe_and_run_in_one_batch(engine):
    scales, efs = [10, 12, 9, 11], [8, 16, 32, 12]
    sg = [4000 + i for i in range(4)]
    sp = [5000 + i for i in range(4)]
    off, nnz = engine.generate_rmat_union(scales, efs, sg, sp, 4)
    cp, ri = engine.download_csc()
    assert list(off) == list(np.concatenate([[0], np.cumsum([1 << s for s in scales])]))
    total = 0
    for g in range(4):
        gcp, gri = cpu.rmat_csc(scales[g], efs[g], 0.57, 0.19, 0.19, sg[g], sp[g])
        a, b = int(off[g]), int(off[g + 1])
        sub = cp[a:b + 1].astype(np.int64) - int(cp[a])
        assert (sub == gcp).all()
        assert (ri[cp[a]:cp[b]] == gri + a).all()
        total += len(gri)
    assert nnz == total
    # one root per graph, each in bit g, plus extra searches on graph 1
    rng = np.random.default_rng(1)
    roots = [int(off[g]) + int(np.argmax(np.diff(cp[off[g]:off[g + 1] + 1].astype(np.int64))))
             for g in range(4)]
    extra = list(rng.integers(int(off[1]), int(off[2]), 20))
    sources = roots + extra
    bits = list(range(4)) + list(range(4, 24))
    lev, _ = engine.run_multi(sources, bits, 24)
    for g in range(4):
        want, _, _ = cpu.bfs_queue(cp, ri, roots[g])
        w8 = np.where(want < 0, 255, want).astype(np.uint8)
        assert (lev[g] == w8).all()
        assert (lev[g][:off[g]] == 255).all() and (lev[g][off[g + 1]:] == 255).all()


@pytest.mark.parametrize("relabel,K,divisor,hub", [(0, 8, 14, 256), (1, 2, 14, 256),
                                                    (1, 64, 4, 64), (1, 4, 32, 1024),
                                                    (0, 16, 1, 32)])
def test_batch_shapes_match_oracle(engine, relabel, K, divisor, hub):
    """The batch on the graph as stored and on its degree-relabelled copy,
    with 2..64 levels per captured hipGraph launch, different direction
    divisors and hub thresholds: levels and edge counts equal the oracle's."""
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 201, 202, 4)
    engine.order_adjacency("pull_hub_first")
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(200, 203)
    old = engine.params
    try:
        engine.set_params(ms_relabel=relabel, ms_graph_levels=K, ms_push_divisor=divisor,
                          ms_pull_hub=hub)
        st = check(engine, cp, ri, roots, np.arange(200), 200)
        if divisor >= 14:
            assert st["push_levels"] > 0 and st["pull_levels"] > 0
        # the captured loop replayed, then another bitset width
        check(engine, cp, ri, roots, np.arange(200), 200)
        check(engine, cp, ri, roots[:33], np.arange(33), 33)
    finally:
        engine.set_params(**old)


def test_batch_autotune_installs_valid_parameters(engine):
    engine.generate_rmat(15, 16, 0.57, 0.19, 0.19, 211, 212, 8)
    cp, ri = engine.download_csc()
    roots = engine.pick_roots(128, 213)
    old = engine.params
    try:
        p, before, after = engine.autotune_multi(roots, np.arange(128), passes=1)
        assert after >= before > 0 and engine.params == p
        check(engine, cp, ri, roots, np.arange(128), 128)
    finally:
        engine.set_params(**old)
