"""The C ABI's host code under AddressSanitizer (CPU, no GPU).

``csrc/pof_abi.hip`` (the entry points, set-up, transitions, initial
trajectories, batch bookkeeping) is compiled with host-side ASan
(``-Xarch_host`` before ``-fsanitize=address``: the device code is not
instrumented) and linked with the other objects of the in-tree build and the
driver ``tests/asan/host_abi_check.cpp``, which calls every host-only entry
point over all field kinds, orders 1..4 and grids of 2..300 points, plus the
argument checks.  ASan aborts on any out-of-bounds access or use after free,
and reports leaks at exit.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "parallel-in-time-ode-filters_amd" / "csrc"
OBJS = ROOT / "parallel-in-time-ode-filters_amd" / "_build"


@pytest.mark.slow
def test_host_abi_under_address_sanitizer(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    others = [OBJS / f"{s}.o" for s in ("pof_inst_d1.hip", "pof_inst_d2.hip", "pof_inst_d3.hip", "pof_inst_d4.hip",
                                        "pof_registry.hip", "pof_large.hip")]
    if not Path(hipcc).exists() or not all(o.exists() for o in others):
        pytest.skip("hipcc or the in-tree build objects are missing (run __graft_entry__.build())")
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    common = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    abi_o = tmp_path / "pof_abi_asan.o"
    r = subprocess.run([hipcc, *common, *san, "-c", str(CSRC / "pof_abi.hip"), "-o", str(abi_o)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # the driver is plain host C++, and the link names no HIP source: ROCm's clang++ with its ASan runtime
    clangxx = Path(hipcc).resolve().parent.parent / "lib" / "llvm" / "bin" / "clang++"
    if not clangxx.exists():
        clangxx = Path("/opt/rocm/llvm/bin/clang++")
    drv_o = tmp_path / "host_abi_check.o"
    r = subprocess.run([str(clangxx), "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-c",
                        str(ROOT / "tests" / "asan" / "host_abi_check.cpp"), "-o", str(drv_o)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = tmp_path / "host_abi_check"
    r = subprocess.run([str(clangxx), "-fsanitize=address", str(drv_o), str(abi_o), *map(str, others),
                        "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", HIP_VISIBLE_DEVICES="")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "0 failures" in r.stdout
