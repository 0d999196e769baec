"""Exodus II mesh files (netCDF classic and 64-bit-offset formats) for the
Tet4 path: the reference reads its meshes through Omega_h
(`Omega_h::read_mesh_file`, src/OmegaHMesh.cpp:13-27), then uses the
coordinates, the element-to-node connectivity, the side sets (for natural
BCs) and the node sets (for essential BCs).  This module reads the same
entities from an Exodus II file into host arrays, and writes them (used by the
tests and to hand synthetic meshes to other tools).

Conventions of the Exodus II format (SEACAS):
* element-to-node arrays `connect<b>` per element block, 1-based, blocks in
  file order give the global element numbering;
* side sets as (element, side) pairs `elem_ss<i>`, `side_ss<i>`, 1-based, with
  the TETRA side numbering side 1 = nodes (1, 2, 4), side 2 = (2, 3, 4),
  side 3 = (1, 4, 3), side 4 = (1, 3, 2);
* node sets `node_ns<i>`, 1-based;
* names in `eb_names`, `ss_names`, `ns_names` (a set without a name is named by
  its id from `ss_prop1` / `ns_prop1`).

Only TETRA4 blocks are on this path; other element types raise.  NetCDF-4
(HDF5) Exodus files are not readable here (no HDF5 library in this image):
convert them to the classic format first (e.g. `nccopy -k classic`).
"""
from dataclasses import dataclass, field

import numpy as np

# local node ordinals (0-based) of TETRA sides 1..4
TET_SIDES = np.array([[0, 1, 3], [1, 2, 3], [0, 3, 2], [0, 2, 1]], np.int32)
_TET_TYPES = {"TETRA", "TETRA4", "TET4", "TET"}


@dataclass
class ExodusMesh:
    coords: np.ndarray                      # N*3 float64, node-major
    conn: np.ndarray                        # C*4 int32, 0-based
    side_sets: dict = field(default_factory=dict)   # name -> (cells int32 F, local nodes int32 F*3)
    node_sets: dict = field(default_factory=dict)   # name -> nodes int32
    block_names: list = field(default_factory=list)

    @property
    def num_nodes(self):
        return self.coords.size // 3

    @property
    def num_cells(self):
        return self.conn.size // 4


def _names(nc, var, count, ids):
    out = []
    raw = nc.variables[var][:] if var in nc.variables else None
    for i in range(count):
        name = ""
        if raw is not None:
            name = b"".join(bytes(c) for c in np.asarray(raw[i]).ravel()).split(b"\0")[0].decode().strip()
        out.append(name if name else str(int(ids[i]) if ids is not None else i + 1))
    return out


def read(path):
    """ExodusMesh of a Tet4 Exodus II file (classic or 64-bit-offset netCDF)."""
    from scipy.io import netcdf_file

    with netcdf_file(str(path), "r", mmap=False) as nc:
        dims = nc.dimensions
        dim = int(dims.get("num_dim", 3))
        if dim != 3:
            raise ValueError(f"{path}: {dim}-dimensional mesh; this path is three-dimensional")
        N = int(dims["num_nodes"])
        if "coord" in nc.variables:
            xyz = np.asarray(nc.variables["coord"][:], np.float64).reshape(3, N)
        else:
            xyz = np.stack([np.asarray(nc.variables[v][:], np.float64) for v in ("coordx", "coordy", "coordz")])
        coords = np.ascontiguousarray(xyz.T).reshape(-1)
        nblk = int(dims.get("num_el_blk", 0))
        conns = []
        blk_ids = np.asarray(nc.variables["eb_prop1"][:]) if "eb_prop1" in nc.variables else None
        for b in range(1, nblk + 1):
            var = nc.variables.get(f"connect{b}")
            if var is None:
                continue
            etype = getattr(var, "elem_type", b"TETRA4")
            etype = (etype.decode() if isinstance(etype, bytes) else str(etype)).upper()
            c = np.asarray(var[:], np.int64)
            if etype not in _TET_TYPES or c.shape[1] != 4:
                raise ValueError(f"{path}: element block {b} is {etype} with {c.shape[1]} nodes; only TETRA4 "
                                 f"is on this path")
            conns.append(c - 1)
        conn = np.concatenate(conns).astype(np.int32).reshape(-1) if conns else np.zeros(0, np.int32)
        C = conn.size // 4
        if conn.size and (conn.min() < 0 or conn.max() >= N):
            raise ValueError(f"{path}: connectivity out of range")
        mesh = ExodusMesh(coords, conn, block_names=_names(nc, "eb_names", nblk, blk_ids))
        nss = int(dims.get("num_side_sets", 0))
        ss_ids = np.asarray(nc.variables["ss_prop1"][:]) if "ss_prop1" in nc.variables else None
        for i, name in enumerate(_names(nc, "ss_names", nss, ss_ids), start=1):
            if f"elem_ss{i}" not in nc.variables:
                continue
            el = np.asarray(nc.variables[f"elem_ss{i}"][:], np.int64) - 1
            sd = np.asarray(nc.variables[f"side_ss{i}"][:], np.int64) - 1
            if el.size and (el.min() < 0 or el.max() >= C or sd.min() < 0 or sd.max() > 3):
                raise ValueError(f"{path}: side set '{name}' out of range")
            mesh.side_sets[name] = (el.astype(np.int32), TET_SIDES[sd].reshape(-1).copy())
        nns = int(dims.get("num_node_sets", 0))
        ns_ids = np.asarray(nc.variables["ns_prop1"][:]) if "ns_prop1" in nc.variables else None
        for i, name in enumerate(_names(nc, "ns_names", nns, ns_ids), start=1):
            if f"node_ns{i}" not in nc.variables:
                continue
            nodes = np.asarray(nc.variables[f"node_ns{i}"][:], np.int64) - 1
            if nodes.size and (nodes.min() < 0 or nodes.max() >= N):
                raise ValueError(f"{path}: node set '{name}' out of range")
            mesh.node_sets[name] = nodes.astype(np.int32)
    return mesh


def side_of_face(local_nodes):
    """Exodus TETRA side (0-based) whose nodes are the given three local nodes."""
    key = sorted(int(a) for a in local_nodes)
    for s, face in enumerate(TET_SIDES):
        if sorted(face.tolist()) == key:
            return s
    raise ValueError(f"{local_nodes} is not a face of a tetrahedron")


def write(path, coords, conn, side_sets=None, node_sets=None, title="platoanalyze_amd mesh"):
    """Write a one-block TETRA4 Exodus II file (64-bit-offset netCDF).
    side_sets: name -> (cells, local nodes F*3); node_sets: name -> nodes."""
    from scipy.io import netcdf_file

    coords = np.asarray(coords, np.float64).reshape(-1, 3)
    conn = np.asarray(conn, np.int64).reshape(-1, 4)
    side_sets = side_sets or {}
    node_sets = node_sets or {}
    len_name = 33
    with netcdf_file(str(path), "w", version=2) as nc:
        nc.title = title.encode()
        nc.api_version = np.float32(8.0)
        nc.version = np.float32(8.0)
        nc.floating_point_word_size = np.int32(8)
        nc.createDimension("len_string", len_name)
        nc.createDimension("len_name", len_name)
        nc.createDimension("num_dim", 3)
        nc.createDimension("num_nodes", coords.shape[0])
        nc.createDimension("num_elem", conn.shape[0])
        nc.createDimension("num_el_blk", 1)
        nc.createDimension("num_el_in_blk1", conn.shape[0])
        nc.createDimension("num_nod_per_el1", 4)

        def names_var(var, dim, names):
            v = nc.createVariable(var, "c", (dim, "len_name"))
            arr = np.zeros((len(names), len_name), "S1")
            for i, n in enumerate(names):
                for j, ch in enumerate(n.encode()[:len_name - 1]):
                    arr[i, j] = bytes([ch])
            v[:] = arr

        def ids_var(var, dim, n):
            v = nc.createVariable(var, "i", (dim,))
            v.name = b"ID"
            v[:] = np.arange(1, n + 1, dtype=np.int32)

        for k, ax in enumerate("xyz"):
            nc.createVariable(f"coord{ax}", "d", ("num_nodes",))[:] = coords[:, k]
        names_var("coor_names", "num_dim", ["x", "y", "z"])
        names_var("eb_names", "num_el_blk", ["body"])
        ids_var("eb_prop1", "num_el_blk", 1)
        cv = nc.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        cv.elem_type = b"TETRA4"
        cv[:] = (conn + 1).astype(np.int32)
        if side_sets:
            nc.createDimension("num_side_sets", len(side_sets))
            names_var("ss_names", "num_side_sets", list(side_sets))
            ids_var("ss_prop1", "num_side_sets", len(side_sets))
            for i, (name, (cells, lnodes)) in enumerate(side_sets.items(), start=1):
                cells = np.asarray(cells, np.int64)
                lnodes = np.asarray(lnodes, np.int64).reshape(-1, 3)
                sides = np.array([side_of_face(f) for f in lnodes], np.int64)
                nc.createDimension(f"num_side_ss{i}", cells.size)
                nc.createVariable(f"elem_ss{i}", "i", (f"num_side_ss{i}",))[:] = (cells + 1).astype(np.int32)
                nc.createVariable(f"side_ss{i}", "i", (f"num_side_ss{i}",))[:] = (sides + 1).astype(np.int32)
        if node_sets:
            nc.createDimension("num_node_sets", len(node_sets))
            names_var("ns_names", "num_node_sets", list(node_sets))
            ids_var("ns_prop1", "num_node_sets", len(node_sets))
            for i, (name, nodes) in enumerate(node_sets.items(), start=1):
                nodes = np.asarray(nodes, np.int64)
                nc.createDimension(f"num_nod_ns{i}", nodes.size)
                nc.createVariable(f"node_ns{i}", "i", (f"num_nod_ns{i}",))[:] = (nodes + 1).astype(np.int32)
