"""Exodus II reader/writer (platoanalyze_amd.exodus), CPU only: a box mesh
with its side and node sets survives a write/read round trip; the TETRA side
numbering maps each side to the face opposite the node the Exodus II
convention names; malformed inputs raise."""
import numpy as np
import pytest

import oracle
from platoanalyze_amd import exodus


def box_sets(coords, conn):
    X = coords.reshape(-1, 3)
    sides, nodes = {}, {}
    for axis, name in enumerate("xyz"):
        L = X[:, axis].max()
        for sign, val in (("-", 0.0), ("+", L)):
            sides[name + sign] = oracle.side_set(coords, conn, axis, val)
            nodes[name + sign] = np.nonzero(np.abs(X[:, axis] - val) < 1e-12)[0].astype(np.int32)
    return sides, nodes


def test_round_trip_box(tmp_path):
    coords, conn = oracle.box_mesh(4, 3, 2, 2.0, 1.0, 1.0)
    sides, nodes = box_sets(coords, conn)
    path = tmp_path / "box.exo"
    exodus.write(path, coords, conn, sides, nodes)
    m = exodus.read(path)
    assert np.array_equal(m.coords, coords) and np.array_equal(m.conn, conn)
    assert m.block_names == ["body"]
    assert set(m.side_sets) == set(sides) and set(m.node_sets) == set(nodes)
    for name, (cells, ln) in sides.items():
        c2, l2 = m.side_sets[name]
        assert np.array_equal(c2, cells)
        # same faces (node order inside a face follows the Exodus side numbering)
        f1 = np.sort(conn.reshape(-1, 4)[cells[:, None], ln.reshape(-1, 3)], axis=1)
        f2 = np.sort(conn.reshape(-1, 4)[c2[:, None], l2.reshape(-1, 3)], axis=1)
        assert np.array_equal(f1, f2)
    for name, n in nodes.items():
        assert np.array_equal(m.node_sets[name], n)


def test_tetra_side_numbering():
    # side s holds three local nodes and omits one: side 1 omits node 3 (1-based),
    # side 2 omits node 1, side 3 omits node 2, side 4 omits node 4
    omitted = [set(range(4)) - set(f.tolist()) for f in exodus.TET_SIDES]
    assert omitted == [{2}, {0}, {1}, {3}]
    for s, f in enumerate(exodus.TET_SIDES):
        assert exodus.side_of_face(f[::-1]) == s
    with pytest.raises(ValueError):
        exodus.side_of_face([0, 0, 1])


def test_unnamed_sets_and_errors(tmp_path):
    from scipy.io import netcdf_file
    coords, conn = oracle.box_mesh(1, 1, 1)
    path = tmp_path / "bare.exo"
    with netcdf_file(str(path), "w", version=2) as nc:
        nc.createDimension("num_dim", 3)
        nc.createDimension("num_nodes", coords.size // 3)
        nc.createDimension("num_el_blk", 1)
        nc.createDimension("num_el_in_blk1", conn.size // 4)
        nc.createDimension("num_nod_per_el1", 4)
        nc.createDimension("num_node_sets", 1)
        nc.createDimension("num_nod_ns1", 2)
        nc.createVariable("coord", "d", ("num_dim", "num_nodes"))[:] = coords.reshape(-1, 3).T
        v = nc.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        v.elem_type = b"TETRA"
        v[:] = conn.reshape(-1, 4) + 1
        p = nc.createVariable("ns_prop1", "i", ("num_node_sets",))
        p[:] = [7]
        nc.createVariable("node_ns1", "i", ("num_nod_ns1",))[:] = [1, 8]
    m = exodus.read(path)
    assert np.array_equal(m.coords, coords) and np.array_equal(m.conn, conn)
    assert list(m.node_sets) == ["7"] and np.array_equal(m.node_sets["7"], [0, 7])
    bad = tmp_path / "hex.exo"
    with netcdf_file(str(bad), "w", version=2) as nc:
        nc.createDimension("num_dim", 3)
        nc.createDimension("num_nodes", 8)
        nc.createDimension("num_el_blk", 1)
        nc.createDimension("num_el_in_blk1", 1)
        nc.createDimension("num_nod_per_el1", 8)
        for ax in "xyz":
            nc.createVariable(f"coord{ax}", "d", ("num_nodes",))[:] = np.zeros(8)
        v = nc.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        v.elem_type = b"HEX8"
        v[:] = np.arange(1, 9).reshape(1, 8)
    with pytest.raises(ValueError, match="TETRA4"):
        exodus.read(bad)
