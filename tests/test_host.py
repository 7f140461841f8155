"""Host-side C++ (loader, writer, JSON, CLI parser, level certificate) on the
CPU, through tests/host_driver.cpp; loader pinned by the reference's shipped
dataset fixtures (tests/golden)."""
import json
import os
import subprocess

import numpy as np
import pytest

from graphs import binary_tree, invariant_violations, random_graph
from oracle import cpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "bfs-gpu-sycl_amd", "host")
GOLD = os.path.join(ROOT, "tests", "golden")
NAME = "rmat-19-16-nnz"
CLI = os.path.join(ROOT, "bfs-gpu-sycl_amd", "bfs_1.gpu")


@pytest.fixture(scope="session")
def driver(tmp_path_factory):
    out = tmp_path_factory.mktemp("drv") / "host_driver"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", HOST, "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host_driver.cpp"),
                    "-o", str(out)], check=True)
    return str(out)


def run(driver, *args, ok=True, raw=False):
    p = subprocess.run([driver, *map(str, args)], capture_output=True, text=True, timeout=300)
    if ok:
        assert p.returncode == 0, p.stderr
        if raw:
            return p.stdout
        return json.loads(p.stdout) if p.stdout.strip().startswith(("{", "[")) else p.stdout
    return p


def write_csc_bin(path, cp, ri):
    with open(path, "wb") as f:
        np.array([len(cp) - 1, len(ri)], np.uint64).tofile(f)
        np.asarray(cp, np.uint64).tofile(f)
        np.asarray(ri, np.uint32).tofile(f)


def read_csc_bin(path):
    raw = open(path, "rb").read()
    n, nnz = np.frombuffer(raw[:16], np.uint64)
    cp = np.frombuffer(raw[16:16 + 8 * (int(n) + 1)], np.uint64)
    ri = np.frombuffer(raw[16 + 8 * (int(n) + 1):], np.uint32)
    assert len(ri) == nnz
    return cp, ri


SUMMARY = json.load(open(os.path.join(GOLD, f"{NAME}-summary.json")))


# ---------------------------------------------------------------------------
# the reference's shipped dataset
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_meta_fixtures_decode(driver, P):
    recs = run(driver, "meta", os.path.join(GOLD, f"{NAME}-csc-{P}-meta.bin"))
    want = SUMMARY["partitions"][str(P)]["records"]
    assert [(r["rows"], r["cols"], r["nnz"], r["startRow"]) for r in recs] == \
        [(w["rows"], w["cols"], w["nnz"], w["startRow"]) for w in want]
    assert all(r["cols"] == 524177 for r in recs)
    assert sum(r["rows"] for r in recs) == 524176          # round8 drops the last vertex
    assert sum(r["nnz"] for r in recs) == 8388607
    for k, r in enumerate(recs):
        assert r["nnz"] == want[k]["indptr_last"]          # meta nnz = end of indptr
        assert r["startRow"] % 8 == 0                      # round8 partition breaks
        assert r["startRow"] == sum(x["rows"] for x in recs[:k])
        assert want[k]["indptr_len"] == 524178


def test_indptr_prefix_fixture():
    ip = np.fromfile(os.path.join(GOLD, f"{NAME}-csc-1-indptr-prefix.bin"), "<u4")
    one = SUMMARY["one_partition"]
    assert len(ip) == one["prefix_len"] and ip[0] == 0
    assert ip[1] == one["indptr_1"] == 42514 and ip[2] - ip[1] == one["deg_1"] == 18688
    assert (np.diff(ip.astype(np.int64)) >= 0).all()


def test_loader_reads_the_prefix_as_a_graph(driver, tmp_path):
    """The fixture prefix (4096 columns) loads through the reference layout."""
    ip = np.fromfile(os.path.join(GOLD, f"{NAME}-csc-1-indptr-prefix.bin"), "<u4")
    n, nnz = len(ip) - 1, int(ip[-1])
    d = tmp_path / "dataset" / "pre-csc-1"
    d.mkdir(parents=True)
    np.array([n, n], "<u4").tofile(d / "pre-csc-meta.bin")
    with open(d / "pre-csc-meta.bin", "ab") as f:
        np.array([nnz], "<u8").tofile(f)
        np.array([0], "<u4").tofile(f)
    ip.tofile(d / "pre-csc-0-indptr.bin")
    rng = np.random.default_rng(0)
    inds = rng.integers(0, n, nnz).astype("<u4")
    inds.tofile(d / "pre-csc-0-inds.bin")
    out = run(driver, "load", tmp_path, "pre", 1)
    assert out == {"rows": n, "cols": n, "multi": False, "wide": False, "nnz": nnz}
    cp, ri = read_csc_bin(tmp_path / "loaded.bin")
    assert (cp == ip).all() and (ri == inds).all()


# ---------------------------------------------------------------------------
# writer <-> loader, nnzSplit breaks
# ---------------------------------------------------------------------------
def numpy_nnz_split(row_counts, P):
    """scripts/generator.py:22-40 restated with numpy's own arange/searchsorted."""
    nnz = np.cumsum(row_counts)
    total = nnz[-1]
    ideal = np.arange(0, total, total / P)
    br = [*nnz.searchsorted(ideal), len(row_counts)]
    return [int(x) + 4 & ~7 for x in br]


@pytest.mark.parametrize("P", [1, 2, 3, 4, 7])
def test_nnz_split_breaks_match_numpy(driver, tmp_path, P):
    rng = np.random.default_rng(P)
    for n in (1000, 4099, 65536):
        counts = rng.zipf(1.7, n).astype(np.uint64) % 5000
        counts.tofile(tmp_path / "c.bin")
        got = run(driver, "breaks", P, tmp_path / "c.bin")
        assert got == numpy_nnz_split(counts.astype(np.int64), P), (n, P)


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_write_then_load_roundtrip(driver, tmp_path, P):
    cp, ri = cpu.rmat_csc(11, 16, 0.57, 0.19, 0.19, 5, 6)
    write_csc_bin(tmp_path / "g.bin", cp, ri)
    br = run(driver, "write", tmp_path, "g", P, tmp_path / "g.bin")
    deg_in = np.bincount(ri, minlength=len(cp) - 1)
    assert br == numpy_nnz_split(deg_in, P)
    out = run(driver, "load", tmp_path, "g", P)
    assert out["cols"] == len(cp) - 1 and out["nnz"] == len(ri) and out["multi"] == (P > 1)
    cp2, ri2 = read_csc_bin(tmp_path / "loaded.bin")
    assert (cp2 == cp).all() and (ri2 == ri).all()
    if P == 1:
        d = tmp_path / "dataset" / "g-csc-1"
        assert (np.fromfile(d / "g-csc-0-indptr.bin", "<u4") == cp).all()
        assert (np.fromfile(d / "g-csc-0-inds.bin", "<u4") == ri).all()
        assert os.path.getsize(d / "g-csc-meta.bin") == 20


def test_loader_errors(driver, tmp_path):
    p = run(driver, "load", tmp_path, "nothing", 1, ok=False)
    assert p.returncode == 1 and "File not found" in p.stderr
    cp, ri = cpu.rmat_csc(8, 8, 0.57, 0.19, 0.19, 1, 2)
    write_csc_bin(tmp_path / "g.bin", cp, ri)
    run(driver, "write", tmp_path, "g", 1, tmp_path / "g.bin")
    d = tmp_path / "dataset" / "g-csc-1"
    ip = d / "g-csc-0-indptr.bin"
    good = ip.read_bytes()
    ip.write_bytes(b"")
    assert "File size is 0" in run(driver, "load", tmp_path, "g", 1, ok=False).stderr
    ip.write_bytes(good[:-4])
    assert "indptr size" in run(driver, "load", tmp_path, "g", 1, ok=False).stderr
    ip.write_bytes(good)
    inds = d / "g-csc-0-inds.bin"
    inds.write_bytes(inds.read_bytes()[:-4])
    assert "inds size" in run(driver, "load", tmp_path, "g", 1, ok=False).stderr
    meta = d / "g-csc-meta.bin"
    meta.write_bytes(meta.read_bytes()[:19])
    assert "partial record" in run(driver, "load", tmp_path, "g", 1, ok=False).stderr


# ---------------------------------------------------------------------------
# JSON, CLI parser, certificate
# ---------------------------------------------------------------------------
def test_json_roundtrip(driver, tmp_path):
    obj = {"b": [1, -2, 3.5, 1e-07, 123456789012345678, True, False, None],
           "a": {"x": "q\"uo\\te\né中", "rawExecutionTimeAll": [0.1, 2.25]},
           "avgGTEPS90f": "12.34", "empty": {}, "list": []}
    (tmp_path / "in.json").write_text(json.dumps(obj))
    out = run(driver, "json", tmp_path / "in.json", raw=True)
    assert json.loads(out) == obj
    assert out == json.dumps(obj, indent=4, sort_keys=True, ensure_ascii=False)


def test_cli_argument_parser(driver):
    p = run(driver, "args", "--num_runs=5", "--dataset", "foo", "--unknown=3", "--root")
    assert p == {"num_runs": "5", "dataset": "foo", "root": "1", "output": "data.json"}


@pytest.mark.parametrize("seed,n", [(0, 1), (1, 3), (2, 10), (3, 12), (4, 37)])
def test_timing_record_matches_numpy(driver, seed, n):
    """GPURun's 7 timing keys (src/run_bfs_gpu.hpp:756-790) from raw run
    times: mean, total, the 1.5-sigma filter (population deviation,
    src/functions.hpp:63-94) and the first-10 %-dropped runs (:96-108),
    restated in numpy; outliers are planted so that the filter drops some."""
    rng = np.random.default_rng(seed)
    raw = rng.uniform(1.0, 2.0, n)
    if n >= 10:
        raw[rng.integers(0, n, 2)] = 50.0
    out = run(driver, "stats", *[repr(float(x)) for x in raw], raw=True)
    rec = json.loads(out.strip().splitlines()[-1])
    mean, sd = raw.mean(), raw.std()
    filt = raw[np.abs(raw - mean) <= 1.5 * sd]
    f90 = raw[int(n * 0.1):]
    if n >= 10:
        assert len(filt) < n and len(f90) < n
    approx = lambda x: pytest.approx(float(x), rel=1e-12)
    assert rec["rawExecutionTimeAll"] == [approx(x) for x in raw]
    assert rec["avgExecutionTime"] == approx(mean)
    assert rec["totalExecutionTime"] == approx(raw.sum())
    assert rec["avgExecutionTimeFiltered"] == approx(filt.mean())
    assert rec["minExecutionTimeFiltered"] == approx(filt.min())
    assert rec["avgExecutionTime90f"] == approx(f90.mean())
    assert rec["minExecutionTime90f"] == approx(f90.min())
    assert "Average (90%) Time:" in out


def test_certificate_accepts_exact_levels_and_rejects_mistakes(driver, tmp_path):
    rng = np.random.default_rng(4)
    for directed in (False, True):
        cp, ri = random_graph(rng, 400, 1600, directed=directed)
        write_csc_bin(tmp_path / "g.bin", cp, ri)
        for r in (0, 17, 399):
            lev, ec, _ = cpu.bfs_queue(cp, ri, r)
            lev.astype(np.int32).tofile(tmp_path / "l.bin")
            out = run(driver, "validate", tmp_path / "g.bin", r, tmp_path / "l.bin")
            assert out["ok"] and out["edges"] == ec
            reached = np.flatnonzero(lev > 0)
            if len(reached) < 3:
                continue
            for bad in (lev + (np.arange(len(lev)) == reached[1]),
                        np.where(np.arange(len(lev)) == reached[2], -1, lev)):
                bad.astype(np.int32).tofile(tmp_path / "l.bin")
                out = run(driver, "validate", tmp_path / "g.bin", r, tmp_path / "l.bin")
                assert not out["ok"]
                assert invariant_violations(cp, ri, r, bad) != []


# ---------------------------------------------------------------------------
# the CLI binary without a GPU
# ---------------------------------------------------------------------------
def test_cli_reports_missing_dataset(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("bfs_1.gpu not built")
    p = subprocess.run([CLI, "--dataset=none", "--root=0"], cwd=tmp_path, capture_output=True,
                       text=True, env=dict(os.environ, PWD=str(tmp_path)), timeout=120)
    assert p.returncode == 1 and "File not found" in p.stderr


def test_cli_without_device_fails_cleanly(driver, tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("bfs_1.gpu not built")
    cp, ri = binary_tree(100)
    write_csc_bin(tmp_path / "g.bin", cp, ri)
    run(driver, "write", tmp_path, "t", 1, tmp_path / "g.bin")
    env = dict(os.environ, PWD=str(tmp_path), HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    p = subprocess.run([CLI, "--dataset=t", "--root=0"], cwd=tmp_path, capture_output=True,
                       text=True, env=env, timeout=120)
    assert p.returncode == 1 and "bfsx_create" in p.stderr
    p = subprocess.run([CLI, "--dataset=t", "--root=100"], cwd=tmp_path, capture_output=True,
                       text=True, env=env, timeout=120)
    assert p.returncode == 1 and "root out of range" in p.stderr
