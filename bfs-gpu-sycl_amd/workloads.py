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


def build(engine, name, **override):
    """Generates workload `name` in HBM; returns (config dict, nnz)."""
    w = dict(WORKLOADS[name], **override)
    if w["kind"] == "rmat":
        a, b, c = RMAT_ABC
        nnz = engine.generate_rmat(w["scale"], w["edge_factor"], a, b, c, w["seed_graph"],
                                   w["seed_perm"], w["ptr_bytes"])
    else:
        nnz = engine.generate_grid(w["side"], w["drop_prob"], w["n_long"], w["seed_graph"],
                                   w["ptr_bytes"])
    return w, nnz


def roots(engine, w):
    mode = 1 if w["kind"] == "grid" else 0
    return engine.pick_roots(w["n_roots"], w["seed_roots"], mode)
