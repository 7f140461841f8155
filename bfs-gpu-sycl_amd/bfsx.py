"""Python binding of the MI355X BFS engine (libbfsx.so, C ABI in include/bfsx.h).

The engine is native: every call goes through the C ABI into HIP kernels for
gfx950.  There is no Python or CPU fallback; if libbfsx.so is missing or no
device is visible, the calls raise.

    from bfsx import Engine
    eng = Engine()
    eng.generate_rmat(19, 16, seed_graph=1002, seed_perm=2002)
    roots = eng.pick_roots(64, seed=3002)
    levels, stats = eng.run(int(roots[0]), "llb")
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbfsx.so")

LLB, GLB, SLB, PULL, LEARNED, BEAMER = 0, 1, 2, 3, 4, 5
STRATEGIES = {"llb": LLB, "local": LLB, "glb": GLB, "global": GLB, "slb": SLB,
              "local_strided": SLB, "pull": PULL, "learned": LEARNED, "beamer": BEAMER}
STEP_NAMES = ["llb", "glb", "slb", "pull"]
GRAPH_SYMMETRIC = 1
GRAPH_HUB_FIRST = 2
GRAPH_PULL_COPY = 4
RUN_GRAPH_LOOP = 1
RUN_TRACE = 2
RUN_STEP_TIMES = 4
RUN_DEVICE_LOOP = 8

STATUS = {0: "ok", 1: "invalid argument", 2: "device error", 3: "out of memory",
          4: "invalid state", 5: "i/o error", 6: "overflow", 7: "unsupported"}


class BfsxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status
        self.msg = msg


class GraphInfo(C.Structure):
    _fields_ = [("n_rows", C.c_uint32), ("n_cols", C.c_uint32), ("nnz", C.c_uint64),
                ("col_ptr_bytes", C.c_int32), ("flags", C.c_uint32),
                ("max_degree", C.c_uint32), ("n_isolated", C.c_uint32),
                ("pull_hubs", C.c_uint32), ("hot_vertices", C.c_uint32)]


class RunStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("edges_count", C.c_uint64), ("n_levels", C.c_uint32),
                ("max_level", C.c_uint32), ("reached", C.c_uint64),
                ("steps", C.c_uint32 * 4)]

    def as_dict(self):
        return {"ms": self.ms, "edges_count": self.edges_count, "n_levels": self.n_levels,
                "max_level": self.max_level, "reached": self.reached,
                "steps": {STEP_NAMES[i]: self.steps[i] for i in range(4)}}


class LevelRecord(C.Structure):
    _fields_ = [("level", C.c_uint32), ("choice", C.c_uint32),
                ("frontier_vertices", C.c_uint64), ("frontier_edges", C.c_uint64),
                ("unvisited_vertices", C.c_uint64), ("unvisited_edges", C.c_uint64),
                ("discovered", C.c_uint64), ("us", C.c_double),
                ("pull_scanned", C.c_uint64), ("kernel_us", C.c_double),
                ("sched_evaluated", C.c_uint32), ("sched_in", C.c_float * 8),
                ("sched_out", C.c_float * 4)]


class Params(C.Structure):
    _fields_ = [("llb_wg", C.c_uint32), ("glb_wg", C.c_uint32),
                ("glb_blocks_per_cu", C.c_uint32), ("glb_chunk", C.c_uint32),
                ("hub_degree", C.c_uint32), ("pull_wg", C.c_uint32),
                ("graph_levels", C.c_uint32), ("mark_ratio_q10", C.c_uint32),
                ("bucket_ratio_q10", C.c_uint32), ("ms_relabel", C.c_uint32),
                ("ms_push_divisor", C.c_uint32), ("ms_pull_hub", C.c_uint32),
                ("ms_graph_levels", C.c_uint32), ("coop_groups", C.c_uint32),
                ("speculate", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("edges_count", C.c_uint64), ("n_levels", C.c_uint32),
                ("push_levels", C.c_uint32), ("pull_levels", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


P = C.c_void_p
_lib = None

# every symbol include/bfsx.h declares, with (restype, argtypes)
SIGNATURES = {
    "bfsx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "bfsx_destroy": (None, [P]),
    "bfsx_last_error": (C.c_char_p, [P]),
    "bfsx_status_string": (C.c_char_p, [C.c_int]),
    "bfsx_device_info": (C.c_int, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                   C.POINTER(C.c_uint64)]),
    "bfsx_upload_csc": (C.c_int, [P, C.c_uint32, C.c_uint32, P, C.c_int, P, C.c_uint64,
                                  C.c_uint32]),
    "bfsx_generate_rmat": (C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_uint64)]),
    "bfsx_generate_grid": (C.c_int, [P, C.c_uint32, C.c_double, C.c_uint32, C.c_uint64, C.c_int,
                                     C.POINTER(C.c_uint64)]),
    "bfsx_graph_info_get": (C.c_int, [P, C.POINTER(GraphInfo)]),
    "bfsx_download_csc": (C.c_int, [P, P, P]),
    "bfsx_order_adjacency": (C.c_int, [P, C.c_int]),
    "bfsx_pick_roots": (C.c_int, [P, C.c_uint32, C.c_uint64, C.c_int, P,
                                  C.POINTER(C.c_uint32)]),
    "bfsx_run": (C.c_int, [P, C.c_uint32, C.c_int, C.c_uint32, P, P, C.POINTER(RunStats)]),
    "bfsx_run_batch": (C.c_int, [P, P, C.c_uint32, C.c_int, C.c_uint32, P, P]),
    "bfsx_last_trace": (C.c_int, [P, P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "bfsx_level_histogram": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "bfsx_stream_copy": (C.c_int, [P, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "bfsx_run_multi": (C.c_int, [P, P, P, C.c_uint32, C.c_uint32, P, C.POINTER(BatchStats)]),
    "bfsx_generate_rmat_union": (C.c_int, [P, C.c_uint32, P, P, P, P, C.c_int, P,
                                           C.POINTER(C.c_uint64)]),
    "bfsx_sched_set_weights": (C.c_int, [P, P, C.c_uint32]),
    "bfsx_mlp_forward": (C.c_int, [P, P, C.c_uint32, P, P]),
    "bfsx_mlp_loss_grad": (C.c_int, [P, P, P, P, C.c_uint32, P, P, P]),
    "bfsx_mlp_train": (C.c_int, [P, P, P, P, C.c_uint32, P, C.c_uint32, C.c_float, P]),
    "bfsx_sched_record": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, P, P, P,
                                    C.c_uint32, C.POINTER(C.c_uint32)]),
    "bfsx_autotune": (C.c_int, [P, P, C.c_uint32, C.c_int, C.c_uint32, C.POINTER(Params),
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bfsx_autotune_multi": (C.c_int, [P, P, P, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.POINTER(Params), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "bfsx_get_params": (C.c_int, [P, C.POINTER(Params)]),
    "bfsx_set_params": (C.c_int, [P, C.POINTER(Params)]),
    "bfsx_run_noalloc": (C.c_int, [C.c_int, P, P, C.c_int, P, C.c_uint32, C.c_uint32,
                                   C.c_uint64, C.c_uint32, P, C.POINTER(C.c_uint32)]),
}


def lib():
    """Loads libbfsx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(P) if a is not None else None


def strategy_id(s):
    if isinstance(s, str):
        return STRATEGIES[s.lower()]
    return int(s)


class Engine:
    """One bfsx context: a device, a stream and a resident graph."""

    def __init__(self, device=0):
        self._h = P()
        rc = lib().bfsx_create(device, C.byref(self._h))
        if rc != 0:
            raise BfsxError(rc, "bfsx_create failed (no usable HIP device?)")

    def _check(self, rc):
        if rc != 0:
            raise BfsxError(rc, lib().bfsx_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().bfsx_destroy(self._h)
            self._h = P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device / graph -------------------------------------------------
    def device_info(self):
        name = C.create_string_buffer(128)
        cus = C.c_int(0)
        hbm = C.c_uint64(0)
        self._check(lib().bfsx_device_info(self._h, name, 128, C.byref(cus), C.byref(hbm)))
        return {"arch": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value}

    def upload_csc(self, col_ptr, row_idx, n_rows=None, symmetric=False):
        col_ptr = np.ascontiguousarray(col_ptr)
        if col_ptr.dtype not in (np.uint32, np.uint64):
            raise TypeError("col_ptr must be uint32 or uint64")
        row_idx = np.ascontiguousarray(row_idx, np.uint32)
        n_cols = len(col_ptr) - 1
        n_rows = n_cols if n_rows is None else n_rows
        self._check(lib().bfsx_upload_csc(self._h, n_cols, n_rows, _ptr(col_ptr),
                                          col_ptr.dtype.itemsize, _ptr(row_idx), len(row_idx),
                                          GRAPH_SYMMETRIC if symmetric else 0))

    def generate_rmat(self, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed_graph=1,
                      seed_perm=2, ptr_bytes=4):
        nnz = C.c_uint64(0)
        self._check(lib().bfsx_generate_rmat(self._h, scale, edge_factor, a, b, c, seed_graph,
                                             seed_perm, ptr_bytes, C.byref(nnz)))
        return nnz.value

    def generate_grid(self, side, drop_prob=0.05, n_long=0, seed=1, ptr_bytes=4):
        nnz = C.c_uint64(0)
        self._check(lib().bfsx_generate_grid(self._h, side, drop_prob, n_long, seed, ptr_bytes,
                                             C.byref(nnz)))
        return nnz.value

    def info(self):
        g = GraphInfo()
        self._check(lib().bfsx_graph_info_get(self._h, C.byref(g)))
        return {k: getattr(g, k) for k, _ in g._fields_}

    def download_csc(self):
        g = self.info()
        cp = np.empty(g["n_cols"] + 1, np.uint64 if g["col_ptr_bytes"] == 8 else np.uint32)
        ri = np.empty(max(g["nnz"], 1), np.uint32)
        self._check(lib().bfsx_download_csc(self._h, _ptr(cp), _ptr(ri)))
        return cp, ri[:g["nnz"]]

    def order_adjacency(self, order="hub_first"):
        """Re-sorts the resident adjacency lists: "id" (ascending neighbour
        ids) or "hub_first" (neighbour degree descending, ties by id); or,
        with "pull_hub_first", keeps them and builds a hub-first copy that
        only bottom-up steps read."""
        code = {"id": 0, "hub_first": 1, "pull_hub_first": 2}[order]
        self._check(lib().bfsx_order_adjacency(self._h, code))

    def pick_roots(self, n, seed, mode=0):
        out = np.zeros(n, np.uint32)
        got = C.c_uint32(0)
        self._check(lib().bfsx_pick_roots(self._h, n, seed, mode, _ptr(out), C.byref(got)))
        return out[:got.value].copy()

    # ---- search ---------------------------------------------------------
    def run(self, root, strategy="llb", trace=False, parents=False, levels=True,
            graph_loop=False, step_times=False, device_loop=False):
        g = self.info()
        lev = np.empty(g["n_cols"], np.int32) if levels else None
        par = np.empty(g["n_cols"], np.uint32) if parents else None
        st = RunStats()
        flags = ((RUN_TRACE if trace else 0) | (RUN_GRAPH_LOOP if graph_loop else 0)
                 | (RUN_STEP_TIMES if step_times else 0)
                 | (RUN_DEVICE_LOOP if device_loop else 0))
        self._check(lib().bfsx_run(self._h, int(root), strategy_id(strategy), flags, _ptr(lev),
                                   _ptr(par), C.byref(st)))
        out = [lev, st.as_dict()]
        if parents:
            out.append(par)
        return tuple(out)

    def run_batch(self, roots, strategy="llb", levels=False, trace=False, graph_loop=False,
                  step_times=False, device_loop=False):
        roots = np.ascontiguousarray(roots, np.uint32)
        g = self.info()
        lev = np.empty((len(roots), g["n_cols"]), np.int32) if levels else None
        st = (RunStats * max(len(roots), 1))()
        flags = ((RUN_TRACE if trace else 0) | (RUN_GRAPH_LOOP if graph_loop else 0)
                 | (RUN_STEP_TIMES if step_times else 0)
                 | (RUN_DEVICE_LOOP if device_loop else 0))
        self._check(lib().bfsx_run_batch(self._h, _ptr(roots), len(roots), strategy_id(strategy),
                                         flags, _ptr(lev), C.cast(st, P)))
        return lev, [st[i].as_dict() for i in range(len(roots))]

    def last_trace(self):
        n = C.c_uint32(0)
        self._check(lib().bfsx_last_trace(self._h, None, 0, C.byref(n)))
        recs = (LevelRecord * max(n.value, 1))()
        self._check(lib().bfsx_last_trace(self._h, C.cast(recs, P), n.value, C.byref(n)))
        def val(v):
            return list(v) if isinstance(v, C.Array) else v
        return [{k: val(getattr(recs[i], k)) for k, _ in LevelRecord._fields_}
                for i in range(n.value)]

    def level_histogram(self, max_levels=1024):
        """(counts, edges) per level of the last run; grows the table if needed."""
        while True:
            counts = np.zeros(max_levels, np.uint64)
            edges = np.zeros(max_levels, np.uint64)
            n = C.c_uint32(0)
            rc = lib().bfsx_level_histogram(self._h, _ptr(counts), _ptr(edges), max_levels,
                                            C.byref(n))
            if rc == 6 and n.value > max_levels:
                max_levels = n.value
                continue
            self._check(rc)
            return counts[:n.value].copy(), edges[:n.value].copy()

    def stream_copy(self, nbytes, repeats=10):
        ms = C.c_double(0)
        self._check(lib().bfsx_stream_copy(self._h, nbytes, repeats, C.byref(ms)))
        return ms.value

    def autotune(self, roots, strategy="llb", passes=1):
        """Tunes the launch shape for `strategy` on the resident graph; returns
        (params dict, gteps_before, gteps_after).  The params stay installed."""
        roots = np.ascontiguousarray(roots, np.uint32)
        p = Params()
        before, after = C.c_double(0), C.c_double(0)
        self._check(lib().bfsx_autotune(self._h, _ptr(roots), len(roots), strategy_id(strategy),
                                        passes, C.byref(p), C.byref(before), C.byref(after)))
        return p.as_dict(), before.value, after.value

    def autotune_multi(self, sources, bits, passes=1):
        """Tunes the batched search's parameters on these searches; returns
        (params dict, aggregate GTEPS before, after).  The params stay installed."""
        src = np.ascontiguousarray(sources, np.uint32)
        b = np.ascontiguousarray(bits, np.uint16)
        p = Params()
        before, after = C.c_double(0), C.c_double(0)
        self._check(lib().bfsx_autotune_multi(self._h, _ptr(src), _ptr(b), len(src),
                                              int(b.max()) + 1 if len(b) else 1, passes,
                                              C.byref(p), C.byref(before), C.byref(after)))
        return p.as_dict(), before.value, after.value

    # ---- batched searches -----------------------------------------------
    def run_multi(self, sources, bits, n_bits, levels=True):
        """Searches bit b from every sources[i] with bits[i] == b, all in one
        batch; returns (uint8 levels [n_bits, n_cols] or None, stats)."""
        src = np.ascontiguousarray(sources, np.uint32)
        bit = np.ascontiguousarray(bits, np.uint16)
        g = self.info()
        lev = np.empty((n_bits, g["n_cols"]), np.uint8) if levels else None
        st = BatchStats()
        self._check(lib().bfsx_run_multi(self._h, _ptr(src), _ptr(bit), len(src), n_bits,
                                         _ptr(lev), C.byref(st)))
        return lev, st.as_dict()

    def generate_rmat_union(self, scales, edge_factors, seeds_graph, seeds_perm, ptr_bytes=4):
        sc = np.ascontiguousarray(scales, np.int32)
        ef = np.ascontiguousarray(edge_factors, np.int32)
        sg = np.ascontiguousarray(seeds_graph, np.uint64)
        sp = np.ascontiguousarray(seeds_perm, np.uint64)
        off = np.zeros(len(sc) + 1, np.uint32)
        nnz = C.c_uint64(0)
        self._check(lib().bfsx_generate_rmat_union(self._h, len(sc), _ptr(sc), _ptr(ef), _ptr(sg),
                                                   _ptr(sp), ptr_bytes, _ptr(off), C.byref(nnz)))
        return off, nnz.value

    # ---- learned scheduler ----------------------------------------------
    def set_scheduler_weights(self, w):
        if w is None:
            self._check(lib().bfsx_sched_set_weights(self._h, None, 0))
            return
        w = np.ascontiguousarray(w, np.float32)
        self._check(lib().bfsx_sched_set_weights(self._h, _ptr(w), len(w)))

    def load_scheduler(self, path=None):
        """Loads the committed trained weights; returns the metadata dict."""
        meta = load_scheduler_file(path)
        self.set_scheduler_weights(np.asarray(meta["weights"], np.float32))
        return meta

    def mlp_forward(self, X, W):
        X = np.ascontiguousarray(X, np.float32)
        W = np.ascontiguousarray(W, np.float32)
        Y = np.empty((len(X), 4), np.float32)
        self._check(lib().bfsx_mlp_forward(self._h, _ptr(X), len(X), _ptr(W), _ptr(Y)))
        return Y

    def mlp_loss_grad(self, X, T, M, W):
        X, T, M, W = (np.ascontiguousarray(a, np.float32) for a in (X, T, M, W))
        g = np.empty(SCHED_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(lib().bfsx_mlp_loss_grad(self._h, _ptr(X), _ptr(T), _ptr(M), len(X), _ptr(W),
                                             _ptr(g), _ptr(loss)))
        return float(loss[0]), g

    def mlp_train(self, X, T, M, W, epochs, lr):
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        W = np.array(W, np.float32, copy=True)
        curve = np.empty(max(epochs, 1), np.float32)
        self._check(lib().bfsx_mlp_train(self._h, _ptr(X), _ptr(T), _ptr(M), len(X), _ptr(W),
                                         epochs, lr, _ptr(curve)))
        return W, curve[:epochs]

    def sched_record(self, roots, repeats=2, epsilon=0.0, seed=0, max_rows=1 << 20):
        roots = np.ascontiguousarray(roots, np.uint32)
        X = np.empty((max_rows, 8), np.float32)
        T = np.empty((max_rows, 4), np.float32)
        M = np.empty((max_rows, 4), np.float32)
        n = C.c_uint32(0)
        self._check(lib().bfsx_sched_record(self._h, _ptr(roots), len(roots), repeats, epsilon,
                                            seed, _ptr(X), _ptr(T), _ptr(M), max_rows,
                                            C.byref(n)))
        k = n.value
        return X[:k].copy(), T[:k].copy(), M[:k].copy()

    @property
    def params(self):
        p = Params()
        self._check(lib().bfsx_get_params(self._h, C.byref(p)))
        return p.as_dict()

    def set_params(self, **kw):
        p = Params()
        self._check(lib().bfsx_get_params(self._h, C.byref(p)))
        for k, v in kw.items():
            setattr(p, k, int(v))
        self._check(lib().bfsx_set_params(self._h, C.byref(p)))


SCHED_PARAMS = 1476
SCHED_FILE = os.path.join(HERE, "sched_mlp.json")


def have_scheduler_weights(path=None):
    return os.path.exists(path or SCHED_FILE)


def load_scheduler_file(path=None):
    import json
    with open(path or SCHED_FILE) as f:
        meta = json.load(f)
    if len(meta["weights"]) != SCHED_PARAMS:
        raise ValueError("scheduler file holds the wrong number of parameters")
    return meta


# Layer shapes of the scheduler network, in parameter order (row-major):
# W1[32][8], b1[32], W2[32][32], b2[32], W3[4][32], b3[4].
SCHED_LAYERS = [("W1", (32, 8)), ("b1", (32,)), ("W2", (32, 32)), ("b2", (32,)),
                ("W3", (4, 32)), ("b3", (4,))]


def sched_init_weights(seed):
    """Seeded starting point for scheduler training: Kaiming-uniform weights
    (bound sqrt(6 / fan_in)), zero biases, float32 in the engine's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SCHED_LAYERS:
        if name.startswith("W"):
            bound = np.sqrt(6.0 / shape[1])
            parts.append(rng.uniform(-bound, bound, shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    w = np.concatenate(parts).astype(np.float32)
    assert w.size == SCHED_PARAMS
    return w


def harmonic_gteps(edges, ms):
    """Harmonic mean over roots of edges/time, in GTEPS: n / sum(t_r / E_r)."""
    edges = np.asarray(edges, np.float64)
    ms = np.asarray(ms, np.float64)
    return len(edges) / np.sum((ms * 1e-3) / edges) / 1e9
