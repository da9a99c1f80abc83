"""Build recipe of the HIP library ``libtvbk.so`` (gfx950).

Compiles ``tvb-root_amd/csrc/*.hip`` and ``*.cpp`` with hipcc into
``tvb-root_amd/tvb_hip/libtvbk.so`` in-tree, so the built library travels with
the source tree.  Run ``python -m tvb_hip._build`` (with ``tvb-root_amd`` on
``sys.path``) or call :func:`build_library`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)                 # tvb-root_amd/
REPO_ROOT = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO_ROOT, "include")
LIB_PATH = os.path.join(HERE, "libtvbk.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TVBK_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    # reproduce the reference's separate roundings: no fused multiply-add
    "-ffp-contract=off",
    # no packed FP32 instructions on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def build_library(verbose=False):
    """Compile every HIP/C++ source of the library; returns the .so path."""
    cmd = [HIPCC] + FLAGS + ["-I", INCLUDE, "-I", CSRC, "-o", LIB_PATH + ".tmp"] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
