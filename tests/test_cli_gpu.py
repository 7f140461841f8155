"""The bfs_1.gpu CLI end to end on the GPU: datasets in the reference's
binary layout (1-4 partitions), the reference's JSON record, merging of the
output file, every strategy."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import cpu

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "bfs-gpu-sycl_amd", "bfs_1.gpu")

KEYS = {"rawExecutionTimeAll", "avgExecutionTime", "avgExecutionTimeFiltered",
        "minExecutionTimeFiltered", "avgExecutionTime90f", "minExecutionTime90f",
        "totalExecutionTime", "edgesCount", "valid", "startVertex", "numLevels", "avgMTEPS",
        "avgMTEPS90f", "avgGTEPS90f", "avgMTEPSFilter", "maxMTEPSFilter", "edgesCoverage",
        "LevelsGPU"}


def cli(cwd, *args):
    print("cli", *args, flush=True)
    p = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, PWD=str(cwd)))
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


@pytest.fixture(scope="module")
def dataset(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    for P in (1, 2, 3, 4):
        cli(d, "--generate=rmat:14:16:77", f"--write_dataset=rm14", f"--partitions={P}",
            "--root=0", "--output=gen.json")
    cp = np.fromfile(d / "dataset" / "rm14-csc-1" / "rm14-csc-0-indptr.bin", "<u4")
    ri = np.fromfile(d / "dataset" / "rm14-csc-1" / "rm14-csc-0-inds.bin", "<u4")
    return d, cp, ri


@pytest.mark.parametrize("strategy", ["local", "global", "local_strided", "pull", "learned"])
def test_cli_record_matches_oracle(dataset, strategy):
    d, cp, ri = dataset
    deg = np.diff(cp.astype(np.int64))
    root = int(np.argmax(deg))
    want, ec, _ = cpu.bfs_queue(cp, ri, root)
    out = d / f"out_{strategy}.json"
    log = cli(d, "--dataset=rm14", f"--root={root}", "--num_runs=4", f"--strategy={strategy}",
              f"--output={out}")
    assert "[SUCCESS]" in log
    rec = json.load(open(out))[f"rm14_4_{strategy}"]["1"]
    assert KEYS <= set(rec)
    assert rec["valid"] is True
    assert rec["edgesCount"] == ec
    assert rec["LevelsGPU"] == np.bincount(want[want >= 0]).tolist()
    assert rec["numLevels"] == int(want.max()) + 1
    assert rec["startVertex"] == root
    assert len(rec["rawExecutionTimeAll"]) == 4
    assert float(rec["avgGTEPS90f"]) > 0
    assert rec["edgesCoverage"] in (0.0, 100.0)


@pytest.mark.parametrize("P", [2, 3, 4])
def test_cli_partitioned_datasets(dataset, P):
    d, cp, ri = dataset
    out = d / "parts.json"
    cli(d, "--dataset=rm14", "--root=5", f"--partitions={P}", f"--output={out}")
    cli(d, "--dataset=rm14", "--root=5", "--partitions=1", f"--output={out}")
    rec = json.load(open(out))
    assert rec["rm14_1_local"]["1"]["valid"] is True
    want, ec, _ = cpu.bfs_queue(cp, ri, 5)
    assert rec["rm14_1_local"]["1"]["edgesCount"] == ec


def test_cli_merges_output_file(dataset):
    d, _, _ = dataset
    out = d / "merge.json"
    out.write_text(json.dumps({"other": {"8": {"x": 1}}}))
    cli(d, "--dataset=rm14", "--root=3", f"--output={out}")
    cli(d, "--dataset=rm14", "--root=3", "--strategy=global", f"--output={out}")
    rec = json.load(open(out))
    assert rec["other"] == {"8": {"x": 1}}
    assert {"rm14_1_local", "rm14_1_global"} <= set(rec)
