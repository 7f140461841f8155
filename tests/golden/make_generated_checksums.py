#!/usr/bin/env python3
"""Writes tests/golden/generated_checksums.json: nnz and SHA-256 digests of
the CSC arrays (col_ptr, then row_idx, little-endian, as stored) of the
synthetic graphs the workloads are built from, computed with the oracle's
restatement of the recipe (oracle/graphgen_ref.c: Philox-4x32-10 R-MAT edges,
seeded vertex permutation, self-loops and duplicates dropped, symmetrized;
the grid with dropped and long-range edges).

The recipe is this project's own (the reference's networkit generator is
third-party and not installed), so these digests are what pins it: a change
to the recipe in the oracle and the GPU builder together changes them.
Run from the repository root after `make -C oracle`:
    python tests/golden/make_generated_checksums.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "bfs-gpu-sycl_amd")]

import workloads  # noqa: E402
from oracle import cpu  # noqa: E402

# (name, workload, overrides): the CPU and first GPU configurations and a
# small grid with the grid workload's recipe
GRAPHS = [("rmat16", "rmat16", {}), ("rmat19", "rmat19", {}),
          ("grid256", "grid4096", {"side": 256, "n_long": round(1e-4 * 2 * 256 * 255)})]


def digest(cp, ri):
    return {"nnz": int(len(ri)), "vertices": int(len(cp) - 1),
            "col_ptr_bytes": int(cp.dtype.itemsize),
            "col_ptr_sha256": hashlib.sha256(cp.astype(cp.dtype.newbyteorder("<")).tobytes()).hexdigest(),
            "row_idx_sha256": hashlib.sha256(ri.astype("<u4").tobytes()).hexdigest()}


def build(name):
    _, wl, over = next(g for g in GRAPHS if g[0] == name)
    w = dict(workloads.WORKLOADS[wl], **over)
    a, b, c = workloads.RMAT_ABC
    if w["kind"] == "rmat":
        cp, ri = cpu.rmat_csc(w["scale"], w["edge_factor"], a, b, c, w["seed_graph"],
                              w["seed_perm"], w["ptr_bytes"])
    else:
        cp, ri = cpu.grid_csc(w["side"], w["drop_prob"], w["n_long"], w["seed_graph"],
                              w["ptr_bytes"])
    return w, cp, ri


def main():
    out = {}
    for name, _, _ in GRAPHS:
        w, cp, ri = build(name)
        out[name] = dict(digest(cp, ri), recipe={k: v for k, v in w.items() if k != "n_roots"})
        print(name, out[name]["nnz"], flush=True)
    with open(os.path.join(HERE, "generated_checksums.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
