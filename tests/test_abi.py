"""The C-ABI library loads and exports every entry point the header declares
(no device calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "platoanalyze_amd" / "libplatoanalyze_amd.so"))
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_device_work():
    import platoanalyze_amd as pa
    # invalid grid: fails on the host with a message, no exception across the ABI
    st = pa.lib.pa_box_mesh(0, 1, 1, 1.0, 1.0, 1.0, None, None)
    assert st == 1
    assert "positive" in pa.lib.pa_last_error().decode()


def test_mesh_validation_errors_without_device_work():
    """pa_mesh_create validates its input on the host before any device
    allocation: out-of-range connectivity and empty meshes are errors."""
    import platoanalyze_amd as pa
    h = ctypes.c_void_p()
    coords = np.zeros(12)
    bad = np.array([0, 1, 2, 4], np.int32)  # node 4 of a 4-node mesh
    st = pa.lib.pa_mesh_create(coords.ctypes.data_as(ctypes.c_void_p), 4, bad.ctypes.data_as(ctypes.c_void_p), 1,
                               ctypes.byref(h))
    assert st != 0 and "out of range" in pa.lib.pa_last_error().decode()
    neg = np.array([0, 1, -1, 3], np.int32)
    st = pa.lib.pa_mesh_create(coords.ctypes.data_as(ctypes.c_void_p), 4, neg.ctypes.data_as(ctypes.c_void_p), 1,
                               ctypes.byref(h))
    assert st != 0 and "out of range" in pa.lib.pa_last_error().decode()
    st = pa.lib.pa_mesh_create(coords.ctypes.data_as(ctypes.c_void_p), 4, bad.ctypes.data_as(ctypes.c_void_p), 0,
                               ctypes.byref(h))
    assert st != 0 and "empty" in pa.lib.pa_last_error().decode()


def test_box_mesh_matches_oracle():
    import oracle
    import platoanalyze_amd as pa
    for dims in [(2, 2, 2), (5, 3, 4)]:
        c1, k1 = pa.box_mesh(*dims, 2.0, 1.0, 0.5)
        c2, k2 = oracle.box_mesh(*dims, 2.0, 1.0, 0.5)
        assert np.array_equal(k1, k2)
        assert np.array_equal(c1, c2)
