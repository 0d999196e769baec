"""Build recipe of the C++ host-layer test program (tests/cpp/reference_cases)
against include/platoanalyze_amd.hpp and the built library."""
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
BIN = HERE / "reference_cases"


def build(verbose=True):
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{ROOT / 'include'}",
           str(HERE / "reference_cases.cpp"), "-o", str(BIN) + ".tmp", f"-L{ROOT / 'platoanalyze_amd'}",
           "-lplatoanalyze_amd", "-Wl,-rpath,$ORIGIN/../../platoanalyze_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    Path(str(BIN) + ".tmp").replace(BIN)
    return BIN


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
