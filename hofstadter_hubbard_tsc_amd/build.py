"""Build libhhdmrg.so in-tree with hipcc for gfx950.

Usage: python -m hofstadter_hubbard_tsc_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhhdmrg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["hh_core.cpp", "hh_core2.cpp", "hh_runtime.cpp", "hh_measure.cpp", "hh_checkpoint.cpp", "hh_plans.cpp",
           "hh_files.cpp", "hh_davidson.cpp", "hh_batch.cpp",
           "hh_tune.cpp",
           "hh_engine.hip", "hh_precond.hip", "hh_heev.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(HERE, "..", "include"), "-I" + CSRC]
# packed FP32 instructions are kept out of every device object
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
        else:
            cmd = [HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        cmds.append(cmd)
        objs.append(obj)
    # one compiler process per source, at most `jobs` at a time
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for _ in pool.map(_run, cmds):
            pass
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", "--offload-arch=gfx950", "-o", LIB + ".tmp"] + objs + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    _run(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
