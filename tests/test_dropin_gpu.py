"""The reference's own host program on the MI355X engine: src/host.cpp,
src/run_bfs_cpu.hpp and src/graph_load.hpp patched to call the engine's
GPURun (oracle/Makefile `dropin`, built where the reference sources are
present and shipped as binaries in oracle/_ref).  The reference's own check
runs: its CPU BFS against the GPU levels of every run (src/host.cpp:172-188),
which prints [SUCCESS] or [FAILURE]; the JSON record it merges is checked
against the oracle here too.  One binary per strategy macro
(USE_GLOBAL_LOAD_BALANCE / USE_STRIDED_LOCAL_LOAD_BALANCE)."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import cpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "bfs-gpu-sycl_amd", "bfs_1.gpu")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "bfs_1_ref_{}.gpu")
STRATEGIES = ["local", "global", "local_strided"]

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not all(os.path.exists(REF_BIN.format(s)) for s in STRATEGIES),
                                 reason="drop-in binaries not built (make -C oracle dropin)")]


def run(exe, cwd, *args):
    print("run", os.path.basename(exe), *args, flush=True)
    p = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, PWD=str(cwd)))
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    return p.stdout


@pytest.fixture(scope="module")
def dataset(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    run(CLI, d, "--generate=rmat:15:16:91", "--write_dataset=rm15", "--partitions=1",
        "--root=0", "--output=gen.json")
    cp = np.fromfile(d / "dataset" / "rm15-csc-1" / "rm15-csc-0-indptr.bin", "<u4")
    ri = np.fromfile(d / "dataset" / "rm15-csc-1" / "rm15-csc-0-inds.bin", "<u4")
    return d, cp, ri


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_reference_host_validates_engine_levels(dataset, strategy):
    d, cp, ri = dataset
    deg = np.diff(cp.astype(np.int64))
    for root in (int(np.argmax(deg)), int(np.flatnonzero(deg)[7]), 1):
        out = d / f"ref_{strategy}_{root}.json"
        log = run(REF_BIN.format(strategy), d, "--dataset=rm15", f"--root={root}",
                  "--num_runs=5", f"--output={out}")
        # the reference's own CPU-vs-GPU comparison of every run
        assert "[SUCCESS]" in log and "[FAILURE]" not in log, log[-2000:]
        rec = json.load(open(out))[f"rm15_5_{strategy}"]["1"]
        want, ec, _ = cpu.bfs_queue(cp, ri, root)
        assert rec["valid"] is True
        assert rec["edgesCount"] == ec
        assert rec["startVertex"] == root
        assert rec["LevelsGPU"] == np.bincount(want[want >= 0]).tolist()
        raw = rec["rawExecutionTimeAll"]
        assert len(raw) == 5 and all(t > 0 for t in raw)
        assert rec["avgExecutionTime90f"] == pytest.approx(np.mean(raw[5 // 10:]))
