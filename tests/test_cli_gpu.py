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
    log = cli(d, "--dataset=rm14", f"--root={root}", "--num_runs=12", f"--strategy={strategy}",
              f"--output={out}")
    assert "[SUCCESS]" in log
    rec = json.load(open(out))[f"rm14_12_{strategy}"]["1"]
    assert KEYS <= set(rec)
    assert rec["valid"] is True
    assert rec["edgesCount"] == ec
    assert rec["LevelsGPU"] == np.bincount(want[want >= 0]).tolist()
    assert rec["numLevels"] == int(want.max()) + 1
    assert rec["startVertex"] == root
    assert rec["edgesCoverage"] in (0.0, 100.0)
    check_timing_statistics(rec, 12)


def check_timing_statistics(rec, n):
    """The 7 timing keys and the throughput figures follow from the raw run
    times as in src/run_bfs_gpu.hpp:756-790, src/functions.hpp:63-108 and
    src/host.cpp:193-203: mean, 1.5-sigma filter, first 10 % dropped."""
    raw = np.asarray(rec["rawExecutionTimeAll"], np.float64)
    assert len(raw) == n and (raw > 0).all()
    mean, sd = raw.mean(), raw.std()  # population deviation, as the reference
    filt = raw[np.abs(raw - mean) <= 1.5 * sd]
    f90 = raw[n // 10:]
    approx = lambda x: pytest.approx(float(x), rel=1e-12, abs=1e-12)
    assert rec["avgExecutionTime"] == approx(mean)
    assert rec["totalExecutionTime"] == approx(raw.sum())
    assert rec["avgExecutionTimeFiltered"] == approx(filt.mean())
    assert rec["minExecutionTimeFiltered"] == approx(filt.min())
    assert rec["avgExecutionTime90f"] == approx(f90.mean())
    assert rec["minExecutionTime90f"] == approx(f90.min())
    e = rec["edgesCount"]
    assert rec["avgMTEPS"] == approx(e / (1e6 * mean * 1e-3))
    assert rec["avgMTEPS90f"] == approx(e / (1e6 * f90.mean() * 1e-3))
    assert rec["avgGTEPS90f"] == f"{e / (1e9 * f90.mean() * 1e-3):.2f}"
    assert rec["avgMTEPSFilter"] == approx(e / (1e6 * filt.mean() * 1e-3))
    assert rec["maxMTEPSFilter"] == approx(e / (1e6 * filt.min() * 1e-3))


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_cli_partitioned_datasets(dataset, P):
    """A P-partition dataset (destination-row slices, src/graph_load.hpp:
    113-129) is concatenated onto the one GPU; each partition count writes
    its own record file, checked against the oracle on the whole graph."""
    d, cp, ri = dataset
    deg = np.diff(cp.astype(np.int64))
    root = int(np.flatnonzero(deg)[5])
    out = d / f"parts_{P}.json"
    log = cli(d, "--dataset=rm14", f"--root={root}", f"--partitions={P}", "--num_runs=3",
              f"--output={out}")
    assert "[SUCCESS]" in log
    rec = json.load(open(out))["rm14_3_local"]["1"]
    want, ec, _ = cpu.bfs_queue(cp, ri, root)
    assert rec["valid"] is True
    assert rec["edgesCount"] == ec
    assert rec["LevelsGPU"] == np.bincount(want[want >= 0]).tolist()
    check_timing_statistics(rec, 3)


def test_cli_merges_output_file(dataset):
    d, _, _ = dataset
    out = d / "merge.json"
    out.write_text(json.dumps({"other": {"8": {"x": 1}}}))
    cli(d, "--dataset=rm14", "--root=3", f"--output={out}")
    cli(d, "--dataset=rm14", "--root=3", "--strategy=global", f"--output={out}")
    rec = json.load(open(out))
    assert rec["other"] == {"8": {"x": 1}}
    assert {"rm14_1_local", "rm14_1_global"} <= set(rec)
