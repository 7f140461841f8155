"""Synthetic workloads with fixed seeds (graph = 1000+k, permutation = 2000+k,
roots = 3000+k for workload k; multi-graph member i uses 4000+i)."""

RMAT_ABC = (0.57, 0.19, 0.19)  # d = 0.05

WORKLOADS = {
    # R-MAT scale 16, edge factor 16: the CPU reference's case
    "rmat16": dict(kind="rmat", scale=16, edge_factor=16, seed_graph=1001, seed_perm=2001,
                   seed_roots=3001, n_roots=64, ptr_bytes=4),
    # R-MAT scale 19, edge factor 16: LLB / GLB / SLB on one GPU
    "rmat19": dict(kind="rmat", scale=19, edge_factor=16, seed_graph=1002, seed_perm=2002,
                   seed_roots=3002, n_roots=64, ptr_bytes=4),
    # 4096 x 4096 grid, 5 % of edges dropped, 0.01 % long-range edges
    "grid4096": dict(kind="grid", side=4096, drop_prob=0.05,
                     n_long=round(1e-4 * 2 * 4096 * 4095), seed_graph=1003,
                     seed_roots=3003, n_roots=64, ptr_bytes=4),
    # R-MAT scale 22 for the 256-root bitset multi-source search
    "rmat22": dict(kind="rmat", scale=22, edge_factor=16, seed_graph=1004, seed_perm=2004,
                   seed_roots=3004, n_roots=256, ptr_bytes=4),
    # R-MAT scale 26, edge factor 16, uint64 col_ptr: learned vs fixed LLB
    "rmat26": dict(kind="rmat", scale=26, edge_factor=16, seed_graph=1005, seed_perm=2005,
                   seed_roots=3005, n_roots=64, ptr_bytes=8),
}


# Graphs the per-level scheduler is trained on (seeds 7000+k / 7100+k /
# 7200+k); none of them is an evaluation workload above.
TRAIN_WORKLOADS = {}
for _k, (_scale, _ef) in enumerate([(16, 16), (18, 16), (20, 8), (20, 16), (20, 32), (22, 16),
                                     (23, 16), (24, 16), (25, 16)]):
    TRAIN_WORKLOADS[f"train_rmat{_scale}_ef{_ef}"] = dict(
        kind="rmat", scale=_scale, edge_factor=_ef, seed_graph=7000 + _k, seed_perm=7100 + _k,
        seed_roots=7200 + _k, n_roots=16, ptr_bytes=4)
for _k, _side in enumerate([1024, 2048, 3072, 4096]):
    TRAIN_WORKLOADS[f"train_grid{_side}"] = dict(
        kind="grid", side=_side, drop_prob=0.05, n_long=round(1e-4 * 2 * _side * (_side - 1)),
        seed_graph=7050 + _k, seed_roots=7250 + _k, n_roots=4, ptr_bytes=4)


# Adjacency order the workloads are searched in (bfsx_order_adjacency):
# lists in id order for top-down steps, a hub-first copy for bottom-up ones.
ORDER = "pull_hub_first"


def build(engine, name, order=ORDER, **override):
    """Generates workload `name` in HBM and orders its adjacency lists
    (`order` = "pull_hub_first", "hub_first", "id" or None to keep the
    builder's id order);
    returns (config dict, nnz)."""
    w = dict(WORKLOADS[name] if name in WORKLOADS else TRAIN_WORKLOADS[name], **override)
    if w["kind"] == "rmat":
        a, b, c = RMAT_ABC
        nnz = engine.generate_rmat(w["scale"], w["edge_factor"], a, b, c, w["seed_graph"],
                                   w["seed_perm"], w["ptr_bytes"])
    else:
        nnz = engine.generate_grid(w["side"], w["drop_prob"], w["n_long"], w["seed_graph"],
                                   w["ptr_bytes"])
    if order and order != "id":
        engine.order_adjacency(order)
    w["order"] = order or "id"
    return w, nnz


def roots(engine, w):
    mode = 1 if w["kind"] == "grid" else 0
    return engine.pick_roots(w["n_roots"], w["seed_roots"], mode)


def small_graph_shapes(n_graphs=64):
    """Scale ~ U{12..18} and edge factor ~ U{8..32} for packed graph i, seeded 4000+i."""
    import numpy as np
    out = []
    for i in range(n_graphs):
        rng = np.random.default_rng(4000 + i)
        out.append((int(rng.integers(12, 19)), int(rng.integers(8, 33))))
    return out


def build_batch(engine, big="rmat22", n_small=64, n_big_roots=256, order=ORDER):
    """The batched workload: R-MAT scale 22 (256 roots, bits 0..255) packed in
    one CSC with 64 small R-MAT graphs, graph i searched from one root in bit i,
    adjacency ordered as in build().  Returns (sources, bits, offsets, n_bits)."""
    import numpy as np
    w = WORKLOADS[big]
    a, b, c = RMAT_ABC
    # the big graph alone first, to draw its roots (it sits at offset 0 below)
    engine.generate_rmat(w["scale"], w["edge_factor"], a, b, c, w["seed_graph"], w["seed_perm"],
                         w["ptr_bytes"])
    big_roots = engine.pick_roots(n_big_roots, w["seed_roots"])
    shapes = small_graph_shapes(n_small)
    scales = [w["scale"]] + [s for s, _ in shapes]
    efs = [w["edge_factor"]] + [e for _, e in shapes]
    sg = [w["seed_graph"]] + [4000 + i for i in range(n_small)]
    sp = [w["seed_perm"]] + [5000 + i for i in range(n_small)]
    off, _ = engine.generate_rmat_union(scales, efs, sg, sp, w["ptr_bytes"])
    cp, _ = engine.download_csc()
    deg = np.diff(cp.astype(np.int64))
    small_roots = []
    for i in range(n_small):
        lo, hi = int(off[i + 1]), int(off[i + 2])
        rng = np.random.default_rng(6000 + i)
        cand = lo + rng.permutation(hi - lo)
        small_roots.append(int(cand[np.argmax(deg[cand] > 0)]))
    if order and order != "id":
        engine.order_adjacency(order)
    sources = np.concatenate([big_roots, small_roots]).astype(np.uint32)
    bits = np.concatenate([np.arange(len(big_roots)), np.arange(n_small)]).astype(np.uint16)
    return sources, bits, off, max(len(big_roots), n_small)
