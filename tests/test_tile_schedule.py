"""Host-side check of the cell-tile assembly schedules (CPU only).

Compiles tests/host/tile_schedule_check.cpp (the library's host code with the
HIP memory calls mapped to host memory) and runs it on several grids, odd and
degenerate shapes included: the tile tables must reproduce a direct per-cell
assembly of the elastic Jacobian and residual (1e-13 relative) and satisfy the
invariants the tile kernels rely on.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "host" / "tile_schedule_check.cpp"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not Path("/opt/rocm/include/hip/hip_runtime.h").exists():
        pytest.skip("g++ or the HIP headers are not available")
    exe = tmp_path_factory.mktemp("tiles") / "tile_schedule_check"
    cmd = [gxx, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(SRC), "-o", str(exe),
           "-lpthread", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("dims", [(20, 10, 10), (3, 5, 7), (1, 1, 1), (7, 2, 9), (5, 1, 3), (9, 6, 5), (2, 4, 4)])
def test_tile_schedule_reproduces_assembly(checker, dims):
    r = subprocess.run([str(checker), *map(str, dims)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
