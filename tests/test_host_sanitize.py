"""The C-ABI's host code (plan construction, tile layout, split and offset
arithmetic of tvbk_api.cpp) built for the CPU with AddressSanitizer and
UndefinedBehaviorSanitizer and the ring-index bounds-check mode
(TVBK_BOUNDS_CHECK), linked with a stand-in HIP runtime whose kernel
launchers check every address the kernels would form from the plan
(tests/hostcheck/host_hip.cpp), and driven through plans shaped like the
benchmark configurations and the ABI's edge cases
(tests/hostcheck/check_plans.cpp).  The planted variant shortens every
allocation by one element and must report violations: the check is live."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tvb-root_amd", "csrc", "tvbk_api.cpp"),
       os.path.join(ROOT, "tests", "hostcheck", "host_hip.cpp"),
       os.path.join(ROOT, "tests", "hostcheck", "check_plans.cpp")]
HIP_INCLUDE = "/opt/rocm/include"


def _build(tmp_path, extra=()):
    exe = str(tmp_path / "check_plans")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-Wno-unused-result", "-DTVBK_BOUNDS_CHECK", "-D__HIP_PLATFORM_AMD__",
           "-I", HIP_INCLUDE, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tvb-root_amd", "csrc"),
           *extra, *SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _run(exe):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([exe], capture_output=True, text=True, env=env, timeout=900)


pytestmark = pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(HIP_INCLUDE),
                                reason="needs g++ and the HIP headers")


def test_plan_layout_under_sanitizers(tmp_path):
    r = _run(_build(tmp_path))
    assert r.returncode == 0, r.stderr[-3000:]
    assert "0 bounds violations, 0 failures" in r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_planted_overrun_is_reported(tmp_path):
    r = _run(_build(tmp_path, ["-DTVBK_HOSTCHECK_PLANT"]))
    assert r.returncode != 0
    assert "bounds:" in r.stderr
