"""The reference's own src/host.cpp, patched by
bfs-gpu-sycl_amd/host/reference_host_cpp.patch (its SYCL includes and
namespace alias dropped, the engine's drop-in header included in place of
src/run_bfs_gpu.hpp), compiles against the reference's include/json.hpp,
src/run_bfs_cpu.hpp and src/graph_load.hpp and links against libbfsx.so.

The sources are copied into a temporary directory and patched there; the
test is skipped where /root/reference is absent (the GPU box).  The GPU run
of the binaries that oracle/Makefile's `dropin` target builds the same way is
in tests/test_dropin_gpu.py."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bfs-gpu-sycl_amd")
REF = "/root/reference"
PATCH = os.path.join(PKG, "host", "reference_host_cpp.patch")

pytestmark = pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "src", "host.cpp")),
                                reason="reference sources absent")


def patched_tree(tmp_path):
    src = tmp_path / "src"
    src.mkdir()
    for f in ("host.cpp", "run_bfs_cpu.hpp", "graph_load.hpp"):
        shutil.copy(os.path.join(REF, "src", f), src / f)
    p = subprocess.run(["patch", "-p1", "-d", str(tmp_path)], stdin=open(PATCH),
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stdout + p.stderr
    return src


def test_patch_removes_every_sycl_name(tmp_path):
    host = (patched_tree(tmp_path) / "host.cpp").read_text()
    assert "sycl" not in host.lower()
    assert "#include <bfsx_dropin.hpp>" in host
    # the caller's code below the includes is untouched
    orig = open(os.path.join(REF, "src", "host.cpp")).read()
    tail = orig[orig.index("int main("):]
    assert tail in host


@pytest.mark.parametrize("strategy", [(), ("-DUSE_GLOBAL_LOAD_BALANCE=1",),
                                      ("-DUSE_STRIDED_LOCAL_LOAD_BALANCE=1",)])
def test_reference_host_compiles_and_links(tmp_path, strategy):
    if not os.path.exists(os.path.join(PKG, "libbfsx.so")):
        pytest.skip("libbfsx.so not built")
    src = patched_tree(tmp_path)
    exe = tmp_path / "bfs_1.gpu"
    cmd = ["g++", "-O1", "-std=c++17", "-DNUM_GPU=1", "-DVERBOSE=0", *strategy,
           "-I", os.path.join(REF, "include"), "-I", str(src), "-I", os.path.join(PKG, "host"),
           "-I", os.path.join(ROOT, "include"), str(src / "host.cpp"), "-L", PKG, "-lbfsx",
           f"-Wl,-rpath,{PKG}", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-4000:]
    # the engine's entry points resolve from libbfsx.so
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True,
                        text=True, check=True).stdout
    for sym in ("bfsx_create", "bfsx_upload_csc", "bfsx_run", "bfsx_destroy"):
        assert sym in nm
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert os.path.join(PKG, "libbfsx.so") in ldd


def test_verbose_build_compiles(tmp_path):
    """VERBOSE=1 makes host.cpp print the per-offset and DeviceInfo tables
    (printDeviceInfo, from the drop-in header)."""
    if not os.path.exists(os.path.join(PKG, "libbfsx.so")):
        pytest.skip("libbfsx.so not built")
    src = patched_tree(tmp_path)
    p = subprocess.run(["g++", "-fsyntax-only", "-std=c++17", "-DNUM_GPU=1", "-DVERBOSE=1",
                        "-I", os.path.join(REF, "include"), "-I", str(src), "-I",
                        os.path.join(PKG, "host"), "-I", os.path.join(ROOT, "include"),
                        str(src / "host.cpp")], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-4000:]
