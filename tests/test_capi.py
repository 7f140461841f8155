"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and fails with status codes (never aborts) when no device is
visible or arguments are null."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import bfsx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bfsx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bfsx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = bfsx.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(bfsx.SIGNATURES)


def test_status_strings():
    L = bfsx.lib()
    for code, text in bfsx.STATUS.items():
        assert L.bfsx_status_string(code).decode() == text


def test_null_context_is_an_argument_error():
    L = bfsx.lib()
    assert L.bfsx_run(None, 0, 0, 0, None, None, None) == 1
    assert L.bfsx_upload_csc(None, 0, 0, None, 4, None, 0, 0) == 1
    assert L.bfsx_create(0, None) == 1
    L.bfsx_destroy(None)  # no-op


def test_create_without_device_returns_error_not_abort():
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import bfsx, ctypes as C\n"
        "h = C.c_void_p()\n"
        "rc = bfsx.lib().bfsx_create(0, C.byref(h))\n"
        "print(rc)\n" % os.path.join(ROOT, "bfs-gpu-sycl_amd")
    )
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    p = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip().splitlines()[-1] == "2"  # BFSX_ERR_DEVICE


def test_engine_raises_without_device():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and _has_gpu():
        pytest.skip("a GPU is visible")
    with pytest.raises(bfsx.BfsxError) as e:
        bfsx.Engine(0)
    assert e.value.status == 2


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def header_defines():
    out = {}
    for m in re.finditer(r"#define\s+(BFSX_\w+)\s+(\d+)u?\b", open(HEADER).read()):
        out[m.group(1)] = int(m.group(2))
    return out


def test_binding_constants_match_the_header():
    d = header_defines()
    assert bfsx.GRAPH_HUB_FIRST == d["BFSX_GRAPH_HUB_FIRST"]
    assert bfsx.GRAPH_PULL_COPY == d["BFSX_GRAPH_PULL_COPY"]
    assert bfsx.STEP_NAMES.index("llb") == d["BFSX_LLB"]
    assert bfsx.STEP_NAMES.index("pull") == d["BFSX_PULL"]
    assert bfsx.SCHED_PARAMS == d["BFSX_SCHED_PARAMS"]
    # the order names the binding passes to bfsx_order_adjacency
    import inspect
    src = inspect.getsource(bfsx.Engine.order_adjacency)
    codes = dict(re.findall(r'"(\w+)": (\d+)', src))
    assert int(codes["id"]) == d["BFSX_ORDER_ID"]
    assert int(codes["hub_first"]) == d["BFSX_ORDER_HUB_FIRST"]
    assert int(codes["pull_hub_first"]) == d["BFSX_ORDER_PULL_HUB_FIRST"]


def test_bench_reads_the_committed_traffic_summaries():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.traffic_from_profiles("rmat26", "pull")
    assert t is not None and 1e8 < t < 1e10
    rl = bench.with_traffic({"step": "pull", "avg_launch_us": 200.0,
                             "alg_bytes_per_launch": 2.5e8}, "rmat26")
    assert rl["traffic"] == t
    assert abs(rl["fabric_achieved"] - t / 200e-6 / 1e9) < 0.1
    assert abs(rl["fabric_frac"] - rl["fabric_achieved"] / bench.HBM_PEAK_GBS) < 1e-3
    assert abs(rl["traffic_over_alg"] - t / 2.5e8) < 0.01
    assert bench.traffic_from_profiles("rmat26", "nonexistent") is None
