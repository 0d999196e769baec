"""The C++ host layer's Exodus II / netCDF classic reader
(Plato::amd::readExodusData in include/platoanalyze_amd.hpp) on damaged
files: truncations at many points, a variable naming a dimension that does
not exist, a side set whose element and side arrays differ in length, and
absurd counts.  Each must end in the reader's own error (status 3 of the
test program), never in a crash; an AddressSanitizer build of the same
program, when it builds here, must report no invalid access."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from platoanalyze_amd import exodus

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
BIN = HERE / "cpp" / "reference_cases"


@pytest.fixture(scope="module")
def programs(tmp_path_factory):
    if not BIN.exists():
        import importlib.util
        spec = importlib.util.spec_from_file_location("pa_cpp_build", HERE / "cpp" / "build.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build(verbose=False)
    progs = [BIN]
    # host-only AddressSanitizer build (g++ on the test program; the HIP
    # library it links is not instrumented)
    asan = tmp_path_factory.mktemp("asan") / "reference_cases_asan"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address", "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}",
           str(HERE / "cpp" / "reference_cases.cpp"), "-o", str(asan), f"-L{ROOT / 'platoanalyze_amd'}",
           "-lplatoanalyze_amd", f"-Wl,-rpath,{ROOT / 'platoanalyze_amd'}"]
    if shutil.which("g++") and subprocess.run(cmd, capture_output=True).returncode == 0:
        progs.append(asan)
    return progs


def run(prog, path):
    env = {"ASAN_OPTIONS": "detect_leaks=0", "PATH": "/usr/bin:/bin"}
    return subprocess.run([str(prog), "--exodus", str(path)], capture_output=True, text=True, timeout=60, env=env)


@pytest.fixture(scope="module")
def valid(tmp_path_factory):
    coords, conn = oracle.box_mesh(3, 2, 2, 2.0, 1.0, 1.0)
    X = coords.reshape(-1, 3)
    sides = {"x+": oracle.side_set(coords, conn, 0, 2.0)}
    nodes = {"x-": np.nonzero(X[:, 0] == 0.0)[0].astype(np.int32)}
    path = tmp_path_factory.mktemp("exo") / "box.exo"
    exodus.write(path, coords, conn, sides, nodes)
    return path, path.read_bytes()


def test_valid_file_reads(programs, valid):
    path, _ = valid
    for prog in programs:
        r = run(prog, path)
        assert r.returncode == 0, r.stdout + r.stderr
        assert '"cells": 72' in r.stdout and '"side_sets": 1' in r.stdout and '"node_sets": 1' in r.stdout


def test_truncated_files_fail_cleanly(programs, valid, tmp_path):
    path, data = valid
    cuts = sorted(set(list(range(0, 600, 7)) + [len(data) // 2, len(data) - 8, len(data) - 1]))
    for prog in programs:
        for cut in cuts:
            bad = tmp_path / f"cut{cut}.exo"
            bad.write_bytes(data[:cut])
            r = run(prog, bad)
            assert r.returncode in (0, 3), (prog.name, cut, r.returncode, r.stdout, r.stderr[-2000:])
            assert "AddressSanitizer" not in r.stderr, (cut, r.stderr[-2000:])


def test_inconsistent_headers_fail_cleanly(programs, valid, tmp_path):
    path, data = valid
    cases = {}
    # variable "connect1": name record, then ndims, then the dimension ids
    i = data.find(b"connect1")
    j = i + (len(b"connect1") + 3) // 4 * 4
    ndims = int.from_bytes(data[j:j + 4], "big")
    assert ndims == 2
    cases["bad_dim_id"] = data[:j + 4] + (100000).to_bytes(4, "big") + data[j + 8:]
    # side set arrays of different lengths: shrink the side dimension of side_ss1
    k = data.find(b"num_side_ss1")
    kk = k + (len(b"num_side_ss1") + 3) // 4 * 4
    cases["huge_side_count"] = data[:kk] + (0x7fffffff).to_bytes(4, "big") + data[kk + 4:]
    # num_side_sets absurdly large
    m = data.find(b"num_side_sets")
    mm = m + (len(b"num_side_sets") + 3) // 4 * 4
    cases["huge_set_count"] = data[:mm] + (0x7ffffff0).to_bytes(4, "big") + data[mm + 4:]
    for prog in programs:
        for name, blob in cases.items():
            bad = tmp_path / f"{name}.exo"
            bad.write_bytes(blob)
            r = run(prog, bad)
            assert r.returncode == 3, (prog.name, name, r.returncode, r.stdout, r.stderr[-2000:])
            assert "AddressSanitizer" not in r.stderr
