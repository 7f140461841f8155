"""The host-side C++ (loader, dataset writer, nnz split, JSON, CLI parser,
timing record, level certificate) and the CPU oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on the CPU: any
out-of-bounds access, leak, signed overflow or misaligned load aborts the
driver with a sanitizer report.  (GPU code is not built with sanitizers.)"""
import os
import subprocess

import numpy as np
import pytest

from graphs import random_graph
from oracle import cpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "bfs-gpu-sycl_amd", "host")
GOLD = os.path.join(ROOT, "tests", "golden")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def sanitized(cmd, **kw):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=ENV, **kw)
    report = p.stderr
    assert "ERROR: AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert "LeakSanitizer" not in report, report[-4000:]
    return p


@pytest.fixture(scope="module")
def host_asan(tmp_path_factory):
    out = tmp_path_factory.mktemp("asan") / "host_driver_asan"
    subprocess.run(["g++", "-std=c++17", "-Wall", *SAN, "-I", HOST, "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host_driver.cpp"),
                    "-o", str(out)], check=True)
    return str(out)


def write_csc_bin(path, cp, ri):
    with open(path, "wb") as f:
        np.array([len(cp) - 1, len(ri)], np.uint64).tofile(f)
        np.asarray(cp, np.uint64).tofile(f)
        np.asarray(ri, np.uint32).tofile(f)


def test_host_code_under_asan_ubsan(host_asan, tmp_path):
    # the reference's shipped meta records
    for P in (1, 2, 3, 4):
        meta = os.path.join(GOLD, f"rmat-19-16-nnz-csc-{P}-meta.bin")
        p = sanitized([host_asan, "meta", meta])
        assert p.returncode == 0, p.stderr
    rng = np.random.default_rng(5)
    cp, ri = random_graph(rng, 3000, 20000, directed=True)
    src = tmp_path / "in.bin"
    write_csc_bin(src, cp, ri)
    for P in (1, 3, 4):
        d = tmp_path / f"ds{P}"
        d.mkdir()
        assert sanitized([host_asan, "write", str(d), "g", str(P), str(src)]).returncode == 0
        p = sanitized([host_asan, "load", str(d), "g", str(P)])
        assert p.returncode == 0, p.stderr
    counts = tmp_path / "counts.bin"
    np.diff(cp.astype(np.uint64)).astype(np.uint64).tofile(counts)
    assert sanitized([host_asan, "breaks", "4", str(counts)]).returncode == 0
    j = tmp_path / "x.json"
    j.write_text('{"a": [1, 2.5, "s\\u00e9", {"b": null, "c": true}], "d": -7}')
    assert sanitized([host_asan, "json", str(j)]).returncode == 0
    assert sanitized([host_asan, "args", "--root=3", "--dataset", "x", "--bogus=1"]).returncode == 0
    assert sanitized([host_asan, "stats", "1.5", "2.5", "9.0", "1.25"]).returncode == 0
    sym_cp, sym_ri = random_graph(rng, 2000, 9000, directed=False)
    g = tmp_path / "sym.bin"
    write_csc_bin(g, sym_cp, sym_ri)
    lev, _, _ = cpu.bfs_queue(sym_cp, sym_ri, 7)
    lf = tmp_path / "lev.bin"
    lev.astype(np.int32).tofile(lf)
    p = sanitized([host_asan, "validate", str(g), "7", str(lf)])
    assert p.returncode == 0 and '"ok":true' in p.stdout.replace(" ", ""), p.stdout
    lev[lev == 2] = 3  # a planted mistake is reported, not crashed on
    lev.astype(np.int32).tofile(lf)
    p = sanitized([host_asan, "validate", str(g), "7", str(lf)])
    assert p.returncode == 0 and '"ok":false' in p.stdout.replace(" ", ""), p.stdout
    # loader errors stay exceptions (exit 1), not memory errors
    p = sanitized([host_asan, "load", str(tmp_path / "missing"), "g", "1"])
    assert p.returncode == 1


def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_asan"
    subprocess.run(["gcc", "-std=c11", "-Wall", *SAN, os.path.join(ROOT, "tests", "oracle_driver.c"),
                    os.path.join(ROOT, "oracle", "bfs_oracle.c"),
                    os.path.join(ROOT, "oracle", "graphgen_ref.c"), "-o", str(exe)], check=True)
    p = sanitized([str(exe)])
    assert p.returncode == 0 and "oracle ok" in p.stdout, p.stdout + p.stderr[-3000:]
