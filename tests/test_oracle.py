"""Pins the CPU oracle (oracle/) before anything is compared against it.

The reference has no tests or golden level vectors for this path, so the
oracle is pinned by (1) the reference's own src/run_bfs_cpu.hpp compiled from
its sources into oracle/_ref when the reference is present, (2) closed forms,
(3) brute force and scipy's shortest paths on small graphs, (4) BFS
invariants, and (5) published Philox-4x32-10 known-answer vectors for the
generator's random streams.  Planted mistakes (in the oracle's output, on the
CPU) must be caught by the invariant checker.
"""
import numpy as np
import pytest

from oracle import cpu
from graphs import (binary_tree, brute_levels, csc_from_edges, edges_count, full_grid,
                    invariant_violations, path_graph, random_graph, scipy_levels, star_graph,
                    symmetrize)


# ---------------------------------------------------------------------------
# Philox-4x32-10 known-answer vectors (Random123 kat_vectors, philox4x32_10)
# ---------------------------------------------------------------------------
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert list(cpu.philox4x32_10(ctr, key)) == want


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
BFS_FUNCS = [cpu.bfs_levelsync, cpu.bfs_queue]


@pytest.mark.parametrize("bfs", BFS_FUNCS)
def test_path_closed_form(bfs):
    n = 97
    cp, ri = path_graph(n)
    for r in (0, 13, 96):
        lev, ec, _ = bfs(cp, ri, r)
        assert (lev == np.abs(np.arange(n) - r)).all()
        assert ec == len(ri)  # every vertex reached: all stored edges explored


@pytest.mark.parametrize("bfs", BFS_FUNCS)
def test_star_closed_form(bfs):
    n = 50
    cp, ri = star_graph(n)
    lev, ec, it = bfs(cp, ri, 0)
    assert lev[0] == 0 and (lev[1:] == 1).all()
    lev, ec, it = bfs(cp, ri, 7)
    want = np.full(n, 2)
    want[0], want[7] = 1, 0
    assert (lev == want).all()
    assert it == 3  # levels 0,1,2 expanded


@pytest.mark.parametrize("bfs", BFS_FUNCS)
def test_binary_tree_closed_form(bfs):
    n = 1000
    cp, ri = binary_tree(n)
    lev, _, _ = bfs(cp, ri, 0)
    assert (lev == np.floor(np.log2(np.arange(n) + 1)).astype(np.int32)).all()


@pytest.mark.parametrize("bfs", BFS_FUNCS)
def test_grid_manhattan(bfs):
    side = 23
    cp, ri = full_grid(side)
    r = 5 * side + 17
    lev, _, _ = bfs(cp, ri, r)
    rr, cc = np.divmod(np.arange(side * side), side)
    assert (lev == np.abs(rr - 5) + np.abs(cc - 17)).all()


@pytest.mark.parametrize("bfs", BFS_FUNCS)
def test_disconnected_and_isolated(bfs):
    # two paths 0-1-2 and 3-4, vertex 5 isolated
    cp, ri = csc_from_edges(6, *symmetrize([0, 1, 3], [1, 2, 4]))
    lev, ec, _ = bfs(cp, ri, 1)
    assert list(lev) == [1, 0, 1, -1, -1, -1]
    assert ec == 4
    lev, ec, it = bfs(cp, ri, 5)
    assert list(lev) == [-1, -1, -1, -1, -1, 0]
    assert ec == 0 and it == 1


@pytest.mark.parametrize("bfs", BFS_FUNCS)
def test_self_loops_and_duplicates_do_not_change_levels(bfs):
    rng = np.random.default_rng(5)
    n = 200
    s = rng.integers(0, n, 800)
    d = rng.integers(0, n, 800)
    cp0, ri0 = csc_from_edges(n, s, d)
    s2 = np.concatenate([s, s, np.arange(n)])
    d2 = np.concatenate([d, d, np.arange(n)])
    cp1, ri1 = csc_from_edges(n, s2, d2)
    for r in range(0, n, 37):
        assert (bfs(cp0, ri0, r)[0] == bfs(cp1, ri1, r)[0]).all()


def test_rows_below_cols_quirk():
    # vertices >= n_rows get levels but are never expanded (run_bfs_cpu.hpp:31)
    cp, ri = csc_from_edges(4, [0, 1, 2], [1, 2, 3])
    for bfs in BFS_FUNCS:
        lev, ec, _ = bfs(cp, ri, 0, n_rows=2)
        assert list(lev) == [0, 1, 2, -1]
        assert ec == 2


def test_uint64_col_ptr_same_as_uint32():
    rng = np.random.default_rng(11)
    cp, ri = random_graph(rng, 300, 2000)
    for r in (0, 100, 299):
        a = cpu.bfs_levelsync(cp, ri, r)
        b = cpu.bfs_levelsync(cp.astype(np.uint64), ri, r)
        assert (a[0] == b[0]).all() and a[1:] == b[1:]


def test_bad_arguments_raise():
    cp, ri = path_graph(5)
    with pytest.raises(ValueError):
        cpu.bfs_levelsync(cp, ri, 5)
    with pytest.raises(ValueError):
        cpu.bfs_queue(cp, ri, 99)


# ---------------------------------------------------------------------------
# brute force, scipy, invariants, reference
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(40))
def test_oracle_vs_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 48))
    m = int(rng.integers(0, 4 * n + 1))
    directed = bool(seed % 2)
    cp, ri = random_graph(rng, n, m, directed=directed)
    n_rows = n if seed % 3 else max(1, n - int(rng.integers(0, 3)))
    for r in range(n):
        want = brute_levels(cp, ri, r, n_rows)
        for bfs in BFS_FUNCS:
            lev, ec, it = bfs(cp, ri, r, n_rows=n_rows)
            assert (lev == want).all(), (seed, r, bfs.__name__)
            assert ec == edges_count(cp, want, n_rows)
        a = cpu.bfs_levelsync(cp, ri, r, n_rows=n_rows)
        b = cpu.bfs_queue(cp, ri, r, n_rows=n_rows)
        assert a[2] == b[2]


@pytest.mark.parametrize("seed", range(6))
def test_oracle_vs_scipy_shortest_path(seed):
    rng = np.random.default_rng(100 + seed)
    n = 3000
    cp, ri = random_graph(rng, n, 9000, directed=bool(seed % 2))
    for r in rng.integers(0, n, 4):
        want = scipy_levels(cp, ri, int(r))
        lev, _, _ = cpu.bfs_queue(cp, ri, int(r))
        assert (lev == want).all()
        assert invariant_violations(cp, ri, int(r), lev) == []


@pytest.mark.skipif(not cpu.ref_available(), reason="oracle/_ref not built (no reference here)")
@pytest.mark.parametrize("seed", range(8))
def test_oracle_vs_reference_run_bfs_cpu(seed):
    """The restatement equals the reference's own run_bfs_cpu.hpp, compiled from source."""
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(2, 5000))
    cp, ri = random_graph(rng, n, int(rng.integers(0, 8 * n)), directed=bool(seed % 2))
    n_rows = n if seed % 4 else n - 1
    g = cpu.RefGraph(cp, ri, n_rows=n_rows)
    for r in rng.integers(0, n, 5):
        lev_ref, ec_ref, _ = g.run(int(r))
        lev, ec, _ = cpu.bfs_levelsync(cp, ri, int(r), n_rows=n_rows)
        assert (lev == lev_ref).all()
        assert ec == ec_ref
    g.close()


@pytest.mark.skipif(not cpu.ref_available(), reason="oracle/_ref not built (no reference here)")
def test_oracle_vs_reference_on_rmat():
    cp, ri = cpu.rmat_csc(13, 16, 0.57, 0.19, 0.19, 1001, 2001)
    g = cpu.RefGraph(cp, ri)
    deg = np.diff(cp.astype(np.int64))
    for r in cpu.pick_roots((deg > 0).astype(np.uint8), 16, 3001):
        lev_ref, ec_ref, _ = g.run(int(r))
        lev, ec, _ = cpu.bfs_queue(cp, ri, int(r))
        assert (lev == lev_ref).all() and ec == ec_ref


def test_planted_mistakes_are_caught():
    """Each plausible mistake in a level array breaks an invariant or a closed form."""
    cp, ri = cpu.rmat_csc(11, 8, 0.57, 0.19, 0.19, 7, 8)
    deg = np.diff(cp.astype(np.int64))
    r = int(np.argmax(deg))
    lev, _, _ = cpu.bfs_queue(cp, ri, r)
    assert invariant_violations(cp, ri, r, lev) == []
    reached = np.flatnonzero(lev > 0)
    deep = np.flatnonzero(lev == lev.max())
    plants = []
    x = lev.copy(); x[reached[3]] += 1; plants.append(x)          # off by one (too deep)
    x = lev.copy(); x[deep[0]] -= 1; plants.append(x)             # off by one (too shallow)
    x = lev.copy(); x[r] = 1; plants.append(x)                    # root not at 0
    x = lev.copy(); x[reached[5]] = -1; plants.append(x)          # dropped vertex
    x = lev.copy(); x[lev > 0] += 1; plants.append(x)             # level index shifted
    for bad in plants:
        assert invariant_violations(cp, ri, r, bad) != []


# ---------------------------------------------------------------------------
# multi-source and parents
# ---------------------------------------------------------------------------
def test_msbfs_is_min_over_sources():
    rng = np.random.default_rng(3)
    n = 500
    cp, ri = random_graph(rng, n, 1500, directed=False)
    sources = rng.integers(0, n, 12).astype(np.uint32)
    bits = (np.arange(12) % 5).astype(np.uint16)
    out = cpu.msbfs_u8(cp, ri, sources, bits, 5)
    for b in range(5):
        per = [cpu.bfs_queue(cp, ri, int(s))[0] for s, bb in zip(sources, bits) if bb == b]
        per = np.stack([np.where(p < 0, 10 ** 6, p) for p in per])
        want = per.min(axis=0)
        want = np.where(want >= 10 ** 6, 255, want).astype(np.uint8)
        assert (out[b] == want).all()


def test_msbfs_overflow_detected():
    cp, ri = path_graph(300)
    with pytest.raises(OverflowError):
        cpu.msbfs_u8(cp, ri, [0], [0], 1)


def test_check_parents():
    cp, ri = binary_tree(31)
    lev, _, _ = cpu.bfs_queue(cp, ri, 0)
    par = np.array([0] + [(i - 1) // 2 for i in range(1, 31)], np.uint32)
    assert cpu.check_parents(cp, ri, 0, lev, par) == 0
    bad = par.copy(); bad[10] = 3          # not a neighbour
    assert cpu.check_parents(cp, ri, 0, lev, bad) == 1
    bad = par.copy(); bad[7] = 7           # wrong level
    assert cpu.check_parents(cp, ri, 0, lev, bad) == 1


@pytest.mark.parametrize("shuffle", [False, True])
def test_check_parents_against_brute_force(shuffle):
    """The checker (binary search in ascending lists, a scan in others)
    against a set of edges, on a random graph with planted wrong parents."""
    rng = np.random.default_rng(5)
    cp, ri = random_graph(rng, 400, 3000)
    if shuffle:  # lists out of order: the scanning path
        ri = ri.copy()
        for u in range(400):
            seg = ri[cp[u]:cp[u + 1]]
            ri[cp[u]:cp[u + 1]] = seg[::-1]
    root = int(np.argmax(np.diff(cp.astype(np.int64))))
    lev, _, _ = cpu.bfs_queue(cp, ri, root)
    edges = {(u, int(v)) for u in range(400) for v in ri[cp[u]:cp[u + 1]]}
    par = np.full(400, 0xFFFFFFFF, np.uint32)
    for v in range(400):
        if lev[v] > 0:
            par[v] = next(u for u in range(400) if lev[u] == lev[v] - 1 and (u, v) in edges)
    par[root] = root
    assert cpu.check_parents(cp, ri, root, lev, par) == 0
    bad = par.copy()
    want = 0
    for v in rng.choice(np.flatnonzero(lev > 0), 40, replace=False):
        bad[v] = rng.integers(0, 400)
    for v in range(400):
        if lev[v] > 0:
            u = int(bad[v])
            want += not (lev[u] == lev[v] - 1 and (u, v) in edges)
    assert want > 0 and cpu.check_parents(cp, ri, root, lev, bad) == want


# ---------------------------------------------------------------------------
# graph recipe restatement
# ---------------------------------------------------------------------------
def test_rmat_recipe_properties():
    cp, ri = cpu.rmat_csc(12, 16, 0.57, 0.19, 0.19, 1001, 2001)
    n = 1 << 12
    cp64 = cp.astype(np.int64)
    assert cp64[0] == 0 and (np.diff(cp64) >= 0).all() and cp64[-1] == len(ri)
    src = np.repeat(np.arange(n), np.diff(cp64))
    assert (src != ri).all()                                   # no self loops
    key = src * n + ri
    assert (np.diff(key) > 0).all()                            # sorted, no duplicates
    assert (np.sort(ri.astype(np.int64) * n + src) == key).all()  # symmetric
    assert len(ri) <= 2 * 16 * n


def test_rmat_quadrant_frequencies():
    src, dst = cpu.rmat_edges(1, 200000, 0.57, 0.19, 0.19, 99)
    q = np.bincount(src * 2 + dst, minlength=4) / 200000
    assert np.allclose(q, [0.57, 0.19, 0.19, 0.05], atol=0.005)


def test_permutation_is_bijection_and_seeded():
    p = cpu.permutation(1000, 5)
    assert (np.sort(p) == np.arange(1000)).all()
    assert (p == cpu.permutation(1000, 5)).all()
    assert (p != cpu.permutation(1000, 6)).any()


def test_grid_recipe():
    side = 16
    cp, ri = cpu.grid_csc(side, 0.0, 0, 1)
    cpf, rif = full_grid(side)
    assert (cp == cpf).all() and (ri == rif).all()
    cp2, ri2 = cpu.grid_csc(side, 0.3, 5, 1)
    assert len(ri2) < len(ri) + 10


def test_pick_roots_seeded_and_eligible():
    elig = np.zeros(1000, np.uint8)
    elig[::3] = 1
    r = cpu.pick_roots(elig, 20, 42)
    assert len(r) == 20 and len(set(r.tolist())) == 20
    assert (elig[r] == 1).all()
    assert (r == cpu.pick_roots(elig, 20, 42)).all()


# ---------------------------------------------------------------------------
# the synthetic-graph recipe against its committed checksums
# ---------------------------------------------------------------------------
def test_generated_graphs_match_committed_checksums():
    """R-MAT 16, R-MAT 19 and a 256 x 256 grid built by the oracle's
    restatement of the recipe have the nnz and SHA-256 digests committed in
    tests/golden/generated_checksums.json (made by
    tests/golden/make_generated_checksums.py)."""
    import json
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_generated_checksums as mk
    want = json.load(open(os.path.join(here, "golden", "generated_checksums.json")))
    assert set(want) == {g[0] for g in mk.GRAPHS}
    for name in want:
        _, cp, ri = mk.build(name)
        got = mk.digest(cp, ri)
        for k in ("nnz", "vertices", "col_ptr_bytes", "col_ptr_sha256", "row_idx_sha256"):
            assert got[k] == want[name][k], (name, k)
