"""Build libpeec_hip.so in-tree for gfx950 with hipcc (no cmake/ninja).

Each translation unit is compiled to an object under solver_amd/_build/ and
linked into solver_amd/libpeec_hip.so.  integrals.hip is compiled with
-ffp-contract=off so the FP64 closed forms are evaluated exactly as written,
fft.hip so that every launch shape of a pass rounds the same way.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpeec_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["api.cpp", "integrals.hip", "fft.hip", "fft_lanes.hip", "operator.hip", "trsv.hip", "mfsolve.hip", "lu.hip", "gmres.hip", "bgmres.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-I", CSRC]
# integrals.hip: the closed forms evaluate exactly as written; fft.hip:
# every launch shape of the FFT passes rounds identically
PER_FILE = {"integrals.hip": ["-ffp-contract=off"], "fft.hip": ["-ffp-contract=off"], "fft_lanes.hip": ["-ffp-contract=off"]}


def build(verbose=False, jobs=None):
    """Compile every translation unit (in parallel, `jobs` at a time; at
    most 8 and the host's CPU count) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC, *COMMON, *PER_FILE.get(src, []), *lang, "-c", path, "-o", obj])
        objs.append(obj)
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    running, failed = [], None
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append((cmd, subprocess.Popen(cmd)))
        while len([p for _, p in running if p.poll() is None]) >= jobs:
            running[[p.poll() is None for _, p in running].index(True)][1].wait()
    for cmd, p in running:
        if p.wait() != 0 and failed is None:
            failed = cmd
    if failed is not None:
        raise subprocess.CalledProcessError(1, failed)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
