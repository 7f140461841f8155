#!/usr/bin/env python3
"""Writes the loader fixtures from the reference's shipped dataset.

The reference ships rmat-19-16 as binary CSC data only (its *-inds.bin files
are missing), under dataset/rmat-19-16-nnz-csc-{1,2,3,4}/.  This script copies
the four small meta files (20 bytes per partition) and the first 4097 offsets
of the one-partition indptr file, and records SHA-256 digests and summary
values of the full files.  It reads data files only; nothing of the
reference is executed.  Run from anywhere: python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import shutil

import numpy as np

REF = "/root/reference/dataset"
HERE = os.path.dirname(os.path.abspath(__file__))
NAME = "rmat-19-16-nnz"


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def main():
    summary = {"dataset": NAME, "partitions": {}}
    for P in (1, 2, 3, 4):
        d = os.path.join(REF, f"{NAME}-csc-{P}")
        meta = os.path.join(d, f"{NAME}-csc-meta.bin")
        shutil.copyfile(meta, os.path.join(HERE, f"{NAME}-csc-{P}-meta.bin"))
        raw = open(meta, "rb").read()
        recs = []
        for k in range(len(raw) // 20):
            rows, cols = np.frombuffer(raw[20 * k:20 * k + 8], "<u4")
            nnz = np.frombuffer(raw[20 * k + 8:20 * k + 16], "<u8")[0]
            start = np.frombuffer(raw[20 * k + 16:20 * k + 20], "<u4")[0]
            ip = np.fromfile(os.path.join(d, f"{NAME}-csc-{k}-indptr.bin"), "<u4")
            recs.append({"rows": int(rows), "cols": int(cols), "nnz": int(nnz),
                         "startRow": int(start), "indptr_len": int(len(ip)),
                         "indptr_last": int(ip[-1]),
                         "indptr_sha256": sha(os.path.join(d, f"{NAME}-csc-{k}-indptr.bin"))})
        summary["partitions"][str(P)] = {"meta_sha256": sha(meta), "records": recs}
    ip = np.fromfile(os.path.join(REF, f"{NAME}-csc-1", f"{NAME}-csc-0-indptr.bin"), "<u4")
    ip[:4097].astype("<u4").tofile(os.path.join(HERE, f"{NAME}-csc-1-indptr-prefix.bin"))
    deg = np.diff(ip.astype(np.int64))
    summary["one_partition"] = {"indptr_1": int(ip[1]), "deg_1": int(deg[1]),
                                "zero_degree_columns": int((deg == 0).sum()),
                                "max_degree": int(deg.max()), "argmax_degree": int(deg.argmax()),
                                "prefix_len": 4097}
    with open(os.path.join(HERE, f"{NAME}-summary.json"), "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary["one_partition"]))


if __name__ == "__main__":
    main()
