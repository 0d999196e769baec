"""Build recipe of the HIP library (gfx950).

Compiles every source under platoanalyze_amd/csrc into
platoanalyze_amd/libplatoanalyze_amd.so with hipcc.  Packed FP32 instructions
are disabled for all device code.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libplatoanalyze_amd.so"
ARCH = os.environ.get("PA_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def build(verbose=True, jobs=8):
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs, procs = [], []
    for src in sources():
        obj = objdir / (src.name + ".o")
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(str(obj))
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    tmp = str(LIB) + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def _wait(procs):
    while procs:
        p, src = procs.pop(0)
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {src}")


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
