"""Build recipe of the HIP library ``libtvbk.so`` (gfx950).

Compiles ``tvb-root_amd/csrc/*.hip`` and ``*.cpp`` with hipcc (in parallel) into
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
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # reproduce the reference's separate roundings: no fused multiply-add
    "-ffp-contract=off",
    # no packed FP32 instructions on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

# tvbk_inst.hip is compiled once per (precision, model/coupling combination)
# and once for the dense kernels; every other source once.
INST_SOURCE = "tvbk_inst.hip"
INST_VARIANTS = ([("%s_c%d" % (("f32", "f64", "mixed")[r], c), ["-DTVBK_INST_R=%d" % r, "-DTVBK_INST_COMBO=%d" % c])
                  for r in (0, 1, 2) for c in range(7)] + [("dense", ["-DTVBK_INST_DENSE"])])


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")) and f != INST_SOURCE)


def compile_jobs(objdir):
    jobs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        jobs.append([HIPCC] + FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", src, "-o", obj])
    for tag, defs in INST_VARIANTS:
        obj = os.path.join(objdir, "inst_%s.o" % tag)
        jobs.append([HIPCC] + FLAGS + defs + ["-I", INCLUDE, "-I", CSRC, "-c",
                                              os.path.join(CSRC, INST_SOURCE), "-o", obj])
    return jobs


def build_library(verbose=False, jobs=None):
    """Compile every HIP/C++ source of the library (in parallel) and link
    ``libtvbk.so``; returns its path."""
    import concurrent.futures as cf
    import shutil
    import tempfile
    if jobs is None:
        jobs = max(1, min(int(os.environ.get("MAX_JOBS", "8")), os.cpu_count() or 1, 16))
    objdir = tempfile.mkdtemp(prefix="tvbk_build_")
    try:
        cmds = compile_jobs(objdir)

        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        with cf.ThreadPoolExecutor(jobs) as ex:
            results = list(ex.map(run, cmds))
        for cmd, res in zip(cmds, results):
            if res.returncode != 0:
                raise RuntimeError("hipcc failed: %s\n%s%s" % (" ".join(cmd), res.stdout, res.stderr))
        objs = [c[c.index("-o") + 1] for c in cmds]
        link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB_PATH + ".tmp"] + objs
        if verbose:
            print(" ".join(link), flush=True)
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("link failed:\n" + res.stdout + res.stderr)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
