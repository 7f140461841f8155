"""ctypes wrapper over the CPU oracle.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, only as the checker or the timed CPU baseline;
the product library never loads it.

* ``liboracle.so`` (bfs_oracle.c, graphgen_ref.c): the C restatement of
  src/run_bfs_cpu.hpp:5-85 and the synthetic-graph recipe.
* ``_ref/libref_cpu.so`` (optional): the reference's own run_bfs_cpu.hpp
  compiled from its sources by ``make -C oracle ref``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

P = C.c_void_p
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
U16P = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        L.oracle_bfs_levelsync.restype = C.c_int64
        L.oracle_bfs_levelsync.argtypes = [C.c_uint32, C.c_uint32, P, C.c_int, U32P,
                                           C.c_uint32, I32P, C.POINTER(C.c_uint64)]
        L.oracle_bfs_queue.restype = C.c_int64
        L.oracle_bfs_queue.argtypes = L.oracle_bfs_levelsync.argtypes
        L.oracle_msbfs_u8.restype = C.c_int
        L.oracle_msbfs_u8.argtypes = [C.c_uint32, C.c_uint32, P, C.c_int, U32P, U32P, U16P,
                                      C.c_uint32, C.c_uint32, U8P]
        L.oracle_check_parents.restype = C.c_uint64
        L.oracle_check_parents.argtypes = [C.c_uint32, C.c_uint32, P, C.c_int, U32P,
                                           C.c_uint32, I32P, U32P]
        L.oracle_philox4x32_10.restype = None
        L.oracle_philox4x32_10.argtypes = [U32P, U32P, U32P]
        L.oracle_rmat_edges.restype = None
        L.oracle_rmat_edges.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint64, U32P, U32P]
        L.oracle_permutation.restype = None
        L.oracle_permutation.argtypes = [C.c_uint32, C.c_uint64, U32P]
        L.oracle_build_symmetric_csc.restype = C.c_uint64
        L.oracle_build_symmetric_csc.argtypes = [C.c_uint32, C.c_uint64, U32P, U32P,
                                                 C.c_int, P, U32P]
        L.oracle_rmat_csc.restype = C.c_uint64
        L.oracle_rmat_csc.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, P, U32P]
        L.oracle_grid_csc.restype = C.c_uint64
        L.oracle_grid_csc.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_int, P, U32P]
        L.oracle_pick_roots.restype = C.c_uint32
        L.oracle_pick_roots.argtypes = [C.c_uint32, U8P, C.c_uint32, C.c_uint64,
                                        C.c_uint64, U32P]
        _LIB = L
    return _LIB


def _ptr(col_ptr):
    col_ptr = np.ascontiguousarray(col_ptr)
    if col_ptr.dtype not in (np.uint32, np.uint64):
        raise TypeError("col_ptr must be uint32 or uint64")
    return col_ptr, col_ptr.ctypes.data_as(P), col_ptr.dtype.itemsize


def bfs_levelsync(col_ptr, row_idx, root, n_rows=None):
    """Restatement of run_bfs_cpu: returns (levels int32[n_cols], edgesCount, iterations)."""
    col_ptr, p, w = _ptr(col_ptr)
    n_cols = len(col_ptr) - 1
    n_rows = n_cols if n_rows is None else n_rows
    lev = np.empty(n_cols, np.int32)
    ec = C.c_uint64(0)
    it = lib().oracle_bfs_levelsync(n_rows, n_cols, p, w, np.ascontiguousarray(row_idx, np.uint32),
                                    root, lev, C.byref(ec))
    if it < 0:
        raise ValueError("oracle_bfs_levelsync: bad arguments")
    return lev, ec.value, it


def bfs_queue(col_ptr, row_idx, root, n_rows=None):
    """Exact FIFO BFS: returns (levels int32[n_cols], edgesCount, iterations)."""
    col_ptr, p, w = _ptr(col_ptr)
    n_cols = len(col_ptr) - 1
    n_rows = n_cols if n_rows is None else n_rows
    lev = np.empty(n_cols, np.int32)
    ec = C.c_uint64(0)
    it = lib().oracle_bfs_queue(n_rows, n_cols, p, w, np.ascontiguousarray(row_idx, np.uint32),
                                root, lev, C.byref(ec))
    if it < 0:
        raise ValueError("oracle_bfs_queue: bad arguments")
    return lev, ec.value, it


def msbfs_u8(col_ptr, row_idx, sources, bits, n_bits, n_rows=None):
    """Multi-source BFS per bit lane: uint8 levels [n_bits, n_cols], 0xFF unreached."""
    col_ptr, p, w = _ptr(col_ptr)
    n_cols = len(col_ptr) - 1
    n_rows = n_cols if n_rows is None else n_rows
    out = np.empty((n_bits, n_cols), np.uint8)
    src = np.ascontiguousarray(sources, np.uint32)
    b = np.ascontiguousarray(bits, np.uint16)
    rc = lib().oracle_msbfs_u8(n_rows, n_cols, p, w, np.ascontiguousarray(row_idx, np.uint32),
                               src, b, len(src), n_bits, out.reshape(-1))
    if rc == -2:
        raise OverflowError("level above 254")
    if rc != 0:
        raise ValueError("oracle_msbfs_u8: bad arguments")
    return out


def check_parents(col_ptr, row_idx, root, levels, parent, n_rows=None):
    col_ptr, p, w = _ptr(col_ptr)
    n_cols = len(col_ptr) - 1
    n_rows = n_cols if n_rows is None else n_rows
    return lib().oracle_check_parents(n_rows, n_cols, p, w, np.ascontiguousarray(row_idx, np.uint32),
                                      root, np.ascontiguousarray(levels, np.int32),
                                      np.ascontiguousarray(parent, np.uint32))


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def rmat_thresholds(a, b, c):
    """Cumulative quadrant probabilities scaled to 2^32 (same rule as the engine)."""
    s = 4294967296.0
    return (min(int(a * s), 0xFFFFFFFF), min(int((a + b) * s), 0xFFFFFFFF),
            min(int((a + b + c) * s), 0xFFFFFFFF))


def rmat_edges(scale, n_edges, a, b, c, seed):
    tA, tAB, tABC = rmat_thresholds(a, b, c)
    src = np.empty(n_edges, np.uint32)
    dst = np.empty(n_edges, np.uint32)
    lib().oracle_rmat_edges(scale, n_edges, tA, tAB, tABC, seed, src, dst)
    return src, dst


def permutation(n, seed):
    out = np.empty(n, np.uint32)
    lib().oracle_permutation(n, seed, out)
    return out


def symmetric_csc(n, u, v, ptr_bytes=4):
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    cp = np.empty(n + 1, np.uint64 if ptr_bytes == 8 else np.uint32)
    ri = np.empty(max(2 * len(u), 1), np.uint32)
    nnz = lib().oracle_build_symmetric_csc(n, len(u), u, v, ptr_bytes, cp.ctypes.data_as(P), ri)
    return cp, ri[:nnz].copy()


def rmat_csc(scale, edge_factor, a, b, c, seed_graph, seed_perm, ptr_bytes=4):
    n = 1 << scale
    m = edge_factor * n
    tA, tAB, tABC = rmat_thresholds(a, b, c)
    cp = np.empty(n + 1, np.uint64 if ptr_bytes == 8 else np.uint32)
    ri = np.empty(2 * m, np.uint32)
    nnz = lib().oracle_rmat_csc(scale, m, tA, tAB, tABC, seed_graph, seed_perm, ptr_bytes,
                                cp.ctypes.data_as(P), ri)
    return cp, ri[:nnz].copy()


def grid_csc(side, drop_prob, n_long, seed, ptr_bytes=4):
    t_drop = min(int(drop_prob * 4294967296.0), 0xFFFFFFFF)
    V = side * side
    G = 2 * side * (side - 1)
    cp = np.empty(V + 1, np.uint64 if ptr_bytes == 8 else np.uint32)
    ri = np.empty(2 * (G + n_long), np.uint32)
    nnz = lib().oracle_grid_csc(side, t_drop, n_long, seed, ptr_bytes, cp.ctypes.data_as(P), ri)
    return cp, ri[:nnz].copy()


def pick_roots(eligible, n_roots, seed, max_tries=None):
    eligible = np.ascontiguousarray(eligible, np.uint8)
    V = len(eligible)
    out = np.empty(n_roots, np.uint32)
    if max_tries is None:
        max_tries = 64 * V + 1024
    got = lib().oracle_pick_roots(V, eligible, n_roots, seed, max_tries, out)
    return out[:got].copy()


# --- the reference's own run_bfs_cpu.hpp, compiled from its sources -------------

def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_cpu.so"))


def ref_lib():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_cpu.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref not built (make -C oracle ref)")
        L = C.CDLL(path)
        L.ref_graph_create.restype = P
        L.ref_graph_create.argtypes = [C.c_uint32, C.c_uint32, U32P, U32P, C.c_uint64]
        L.ref_graph_destroy.restype = None
        L.ref_graph_destroy.argtypes = [P]
        L.ref_run_bfs_cpu.restype = C.c_int
        L.ref_run_bfs_cpu.argtypes = [P, C.c_uint32, I32P, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_double)]
        _REF = L
    return _REF


class RefGraph:
    """The reference's CPU BFS over one graph (uint32 col_ptr, as its vectorT holds)."""

    def __init__(self, col_ptr, row_idx, n_rows=None):
        cp = np.asarray(col_ptr)
        if cp.dtype == np.uint64:
            if int(cp[-1]) > 0xFFFFFFFF:
                raise ValueError("reference CPU BFS holds uint32 col_ptr only")
            cp = cp.astype(np.uint32)
        self.col_ptr = np.ascontiguousarray(cp, np.uint32)
        self.row_idx = np.ascontiguousarray(row_idx, np.uint32)
        self.n_cols = len(self.col_ptr) - 1
        self.n_rows = self.n_cols if n_rows is None else n_rows
        self.h = ref_lib().ref_graph_create(self.n_rows, self.n_cols, self.col_ptr,
                                            self.row_idx, len(self.row_idx))

    def run(self, root):
        lev = np.empty(self.n_cols, np.int32)
        ec = C.c_uint64(0)
        ms = C.c_double(0)
        rc = ref_lib().ref_run_bfs_cpu(self.h, root, lev, C.byref(ec), C.byref(ms))
        if rc != 0:
            raise ValueError("ref_run_bfs_cpu failed")
        return lev, ec.value, ms.value

    def close(self):
        if self.h:
            ref_lib().ref_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
