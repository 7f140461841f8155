"""Small graphs with known BFS answers, and independent checkers.

Closed forms: path (|i - r|), star, complete binary tree (floor(log2(i+1))
from root 0), full grid (Manhattan distance), disconnected parts (-1).
Independent checkers: reachability by repeated boolean adjacency products
(no queue, no level loop) and scipy.sparse.csgraph's unweighted shortest
paths, plus the BFS invariants every correct level array satisfies.
"""
import numpy as np


def csc_from_edges(n, src, dst, n_rows=None, dtype=np.uint32, sort=True):
    """CSC in the reference layout: col_ptr over sources, row_idx destinations.
    Duplicates and self loops are kept as given."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.lexsort((dst, src)) if sort else np.argsort(src, kind="stable")
    src, dst = src[order], dst[order]
    counts = np.bincount(src, minlength=n) if len(src) else np.zeros(n, np.int64)
    col_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=col_ptr[1:])
    return col_ptr.astype(dtype), dst.astype(np.uint32)


def symmetrize(src, dst):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    return np.concatenate([src, dst]), np.concatenate([dst, src])


def path_graph(n):
    s = np.arange(n - 1)
    return csc_from_edges(n, *symmetrize(s, s + 1))


def star_graph(n):
    leaves = np.arange(1, n)
    return csc_from_edges(n, *symmetrize(np.zeros(n - 1, np.int64), leaves))


def binary_tree(n):
    child = np.arange(1, n)
    return csc_from_edges(n, *symmetrize((child - 1) // 2, child))


def full_grid(side):
    ids = np.arange(side * side).reshape(side, side)
    h = (ids[:, :-1].ravel(), ids[:, 1:].ravel())
    v = (ids[:-1, :].ravel(), ids[1:, :].ravel())
    s = np.concatenate([h[0], v[0]])
    d = np.concatenate([h[1], v[1]])
    return csc_from_edges(side * side, *symmetrize(s, d))


def random_graph(rng, n, m, directed=True, loops=True, dups=True):
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    if not loops:
        keep = s != d
        s, d = s[keep], d[keep]
    if not directed:
        s, d = symmetrize(s, d)
    if not dups and len(s):
        key = np.unique(s * n + d)
        s, d = key // n, key % n
    return csc_from_edges(n, s, d)


def brute_levels(col_ptr, row_idx, root, n_rows=None):
    """Levels by repeated boolean adjacency products (tiny graphs only)."""
    n = len(col_ptr) - 1
    n_rows = n if n_rows is None else n_rows
    A = np.zeros((n, n), bool)
    for u in range(n_rows):
        A[u, row_idx[col_ptr[u]:col_ptr[u + 1]]] = True
    lev = np.full(n, -1, np.int32)
    lev[root] = 0
    reach = np.zeros(n, bool)
    reach[root] = True
    frontier = reach.copy()
    k = 0
    while frontier.any():
        k += 1
        nxt = (A[frontier].any(axis=0)) & ~reach
        lev[nxt] = k
        reach |= nxt
        frontier = nxt
    return lev


def scipy_levels(col_ptr, row_idx, root):
    """Unweighted shortest-path hop counts from scipy.sparse.csgraph."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import shortest_path
    n = len(col_ptr) - 1
    data = np.ones(len(row_idx), np.float64)
    A = sp.csr_matrix((data, np.asarray(row_idx, np.int64), np.asarray(col_ptr, np.int64)),
                      shape=(n, n))
    d = shortest_path(A, method="D", directed=True, unweighted=True, indices=[root])[0]
    out = np.where(np.isinf(d), -1, d).astype(np.int32)
    return out


def edges_count(col_ptr, levels, n_rows=None):
    """Closed form of the reference's edgesCount: degree sum over reached rows."""
    n = len(col_ptr) - 1
    n_rows = n if n_rows is None else n_rows
    deg = np.diff(np.asarray(col_ptr, np.int64))[:n_rows]
    return int(deg[np.asarray(levels[:n_rows]) >= 0].sum())


def invariant_violations(col_ptr, row_idx, root, levels, n_rows=None):
    """Returns a list of violated BFS invariants (empty when levels are valid):
    root at 0; every edge from a reached row u reaches v with level <= l(u)+1;
    every reached v != root has an in-neighbour row at level l(v)-1;
    unreached vertices have -1 and nothing else is negative."""
    n = len(col_ptr) - 1
    n_rows = n if n_rows is None else n_rows
    lev = np.asarray(levels, np.int64)
    bad = []
    if lev[root] != 0:
        bad.append("root level")
    if (lev < -1).any():
        bad.append("level below -1")
    cp = np.asarray(col_ptr, np.int64)
    src = np.repeat(np.arange(n), np.diff(cp))
    dst = np.asarray(row_idx, np.int64)
    keep = src < n_rows
    src, dst = src[keep], dst[keep]
    ru = lev[src] >= 0
    if ((lev[dst[ru]] < 0) | (lev[dst[ru]] > lev[src[ru]] + 1)).any():
        bad.append("edge skips a level")
    has_parent = np.zeros(n, bool)
    ok = ru & (lev[dst] == lev[src] + 1)
    has_parent[dst[ok]] = True
    reached = (lev > 0)
    if (reached & ~has_parent).any():
        bad.append("reached vertex without a parent one level up")
    return bad
