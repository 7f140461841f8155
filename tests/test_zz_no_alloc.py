"""bfsx_run_noalloc, the allocation-free search, through the C ABI: levels
equal bfsx_run's (and the oracle's) on two calls in a row, the second on a
smaller graph, for 4- and 8-byte col_ptr and n_rows < n_cols; the call
leaves the device's free memory unchanged; sizes beyond the fixed scratch
array and out-of-range adjacency entries are errors, not faults.

The graph and level buffers are the caller's, allocated here with the HIP
runtime (the entry point itself allocates nothing)."""
import ctypes as C
import os

import numpy as np
import pytest

import bfsx
from graphs import random_graph
from oracle import cpu

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.skip("HIP runtime library not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    h.hipDeviceSynchronize.argtypes = []
    return h


class Dev:
    """A device buffer owned by the test (the caller of the entry point)."""

    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(int(nbytes), 4)) == 0
        self.nbytes = int(nbytes)

    @classmethod
    def of(cls, h, a):
        a = np.ascontiguousarray(a)
        d = cls(h, a.nbytes)
        if a.nbytes:
            assert h.hipMemcpy(d.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return d

    def get(self, dtype, n):
        out = np.empty(n, dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = None


def free_bytes(h):
    f, t = C.c_size_t(0), C.c_size_t(0)
    assert h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def noalloc(h, d_cp, w, d_ri, n_cols, n_rows, nnz, root, d_lev):
    nl = C.c_uint32(0)
    rc = bfsx.lib().bfsx_run_noalloc(0, None, d_cp.p, w, d_ri.p, n_cols, n_rows, nnz, root,
                                     d_lev.p, C.byref(nl))
    return rc, nl.value


def check_graph(engine, h, cp, ri, roots, n_rows=None):
    n_cols = len(cp) - 1
    n_rows = n_cols if n_rows is None else n_rows
    d_cp, d_ri, d_lev = Dev.of(h, cp), Dev.of(h, ri), Dev(h, 4 * n_cols)
    try:
        for r in roots:
            r = int(r)
            print("noalloc", n_cols, len(ri), cp.itemsize, r, flush=True)
            want, st = engine.run(r, "llb")
            rc, nl = noalloc(h, d_cp, cp.itemsize, d_ri, n_cols, n_rows, len(ri), r, d_lev)
            assert rc == 0, bfsx.lib().bfsx_last_error(None)
            got = d_lev.get(np.int32, n_cols)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, (r, bad[:5], got[bad[:5]], want[bad[:5]])
            assert nl == st["n_levels"]
        lev, _, _ = cpu.bfs_queue(cp, ri, int(roots[0]), n_rows=n_rows)
        noalloc(h, d_cp, cp.itemsize, d_ri, n_cols, n_rows, len(ri), int(roots[0]), d_lev)
        assert np.array_equal(d_lev.get(np.int32, n_cols), lev)
        # no device allocation inside the call (its scratch array is set up)
        before = free_bytes(h)
        rc, _ = noalloc(h, d_cp, cp.itemsize, d_ri, n_cols, n_rows, len(ri), int(roots[-1]), d_lev)
        assert rc == 0 and free_bytes(h) == before
    finally:
        for d in (d_cp, d_ri, d_lev):
            d.free()


def test_noalloc_equals_bfsx_run_on_two_calls(engine):
    h = hip()
    # first call: R-MAT scale 18 (4-byte col_ptr)
    engine.generate_rmat(18, 16, 0.57, 0.19, 0.19, 191, 192, 4)
    cp, ri = engine.download_csc()
    check_graph(engine, h, cp, ri, engine.pick_roots(4, 193))
    # second call, smaller: R-MAT scale 12 with 8-byte col_ptr
    engine.generate_rmat(12, 8, 0.57, 0.19, 0.19, 194, 195, 8)
    cp, ri = engine.download_csc()
    assert cp.dtype == np.uint64
    check_graph(engine, h, cp, ri, engine.pick_roots(3, 196))
    # directed, duplicates and self loops, n_rows < n_cols
    rng = np.random.default_rng(11)
    cp, ri = random_graph(rng, 4000, 25000, directed=True)
    engine.upload_csc(cp, ri, n_rows=3990, symmetric=False)
    check_graph(engine, h, cp, ri, rng.integers(0, 4000, 3), n_rows=3990)


def test_noalloc_limits_and_bad_entries(engine):
    h = hip()
    lib = bfsx.lib()
    d_cp, d_ri, d_lev = Dev(h, 64), Dev(h, 64), Dev(h, 64)
    try:
        nl = C.c_uint32(0)
        big = bfsx.lib().bfsx_run_noalloc
        rc = big(0, None, d_cp.p, 4, d_ri.p, (1 << 22) + 1, 8, 8, 0, d_lev.p, C.byref(nl))
        assert rc == 1 and b"BFSX_NOALLOC_MAX_VERTICES" in lib.bfsx_last_error(None)
        rc = big(0, None, d_cp.p, 4, d_ri.p, 8, 9, 8, 0, d_lev.p, C.byref(nl))
        assert rc == 1  # n_rows > n_cols
        rc = big(0, None, d_cp.p, 4, d_ri.p, 8, 8, 8, 8, d_lev.p, C.byref(nl))
        assert rc == 1  # root out of range
        rc = big(0, None, None, 4, d_ri.p, 8, 8, 8, 0, d_lev.p, C.byref(nl))
        assert rc == 1  # null pointer
    finally:
        for d in (d_cp, d_ri, d_lev):
            d.free()
    # a path 0-1-2-3 whose list of vertex 1 holds an entry past the graph:
    # skipped by the kernel and reported
    cp = np.array([0, 1, 4, 6, 7], np.uint32)
    ri = np.array([1, 0, 2, 40, 1, 3, 2], np.uint32)
    d_cp, d_ri, d_lev = Dev.of(h, cp), Dev.of(h, ri), Dev(h, 16)
    try:
        rc, _ = noalloc(h, d_cp, 4, d_ri, 4, 4, len(ri), 0, d_lev)
        assert rc == 1 and b"out of range" in lib.bfsx_last_error(None)
        assert list(d_lev.get(np.int32, 4)) == [0, 1, 2, 3]
        ok = np.array([1, 0, 2, 1, 1, 3, 2], np.uint32)
        d_ok = Dev.of(h, ok)
        rc, nl = noalloc(h, d_cp, 4, d_ok, 4, 4, len(ok), 3, d_lev)
        d_ok.free()
        assert rc == 0 and nl == 4 and list(d_lev.get(np.int32, 4)) == [3, 2, 1, 0]
    finally:
        for d in (d_cp, d_ri, d_lev):
            d.free()
