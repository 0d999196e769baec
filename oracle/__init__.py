"""CPU FP64 oracle (test infrastructure only).

This package is the checker for the HIP library: the tests, the smoke check
and the CPU-baseline leg of bench.py load it.  The product library never
imports, links or calls anything here.

`build()` compiles pa_oracle.cpp (and the J2 restatement) with g++/OpenMP into
oracle/liboracle.so; `lib()` loads it with ctypes.  The numpy helpers below
wrap the C entry points.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
SOURCES = [HERE / "pa_oracle.cpp", HERE / "pa_oracle_j2.cpp"]

_lib = None


def build(verbose=False):
    srcs = [str(s) for s in SOURCES if s.exists()]
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
           "-o", str(LIB) + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(str(LIB) + ".tmp", LIB)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _declare(_lib)
    return _lib


class Problem(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_int),
        ("num_cells", ctypes.c_int),
        ("coords", ctypes.c_void_p),
        ("conn", ctypes.c_void_p),
        ("youngs", ctypes.c_double),
        ("poisson", ctypes.c_double),
        ("conductivity", ctypes.c_double),
        ("penalty_exponent", ctypes.c_double),
        ("penalty_min", ctypes.c_double),
        ("num_faces", ctypes.c_int),
        ("face_cells", ctypes.c_void_p),
        ("face_lnodes", ctypes.c_void_p),
        ("flux", ctypes.c_double * 3),
        ("has_body", ctypes.c_int),
        ("body_dof", ctypes.c_int),
        ("body_value", ctypes.c_double),
    ]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER(Problem)
    L.orc_set_threads.argtypes = [i]
    L.orc_set_threads.restype = i
    L.orc_box_mesh.argtypes = [i, i, i, d, d, d, vp, vp]
    L.orc_graph_nnz.argtypes = [i, i, vp]
    L.orc_graph_nnz.restype = ctypes.c_long
    L.orc_graph_fill.argtypes = [i, i, vp, vp, vp]
    L.orc_elastic_cell_chain.argtypes = [P, vp, vp, vp, vp, vp]
    for name in ["elastic_value", "thermal_value", "elastic_gradient_u", "thermal_gradient_u",
                 "elastic_gradient_u_T", "thermal_gradient_u_T", "elastic_gradient_z", "thermal_gradient_z",
                 "elastic_energy_gradient_u", "thermal_energy_gradient_u", "elastic_energy_gradient_z",
                 "thermal_energy_gradient_z"]:
        f = getattr(L, "orc_" + name)
        f.argtypes = [P, vp, vp, vp]
        f.restype = None
    for name in ["elastic_energy_value", "thermal_energy_value"]:
        f = getattr(L, "orc_" + name)
        f.argtypes = [P, vp, vp]
        f.restype = d
    L.orc_apply_block_constraints.argtypes = [i, i, vp, vp, vp, vp, i, vp, vp, d]
    L.orc_block_spmv.argtypes = [i, i, i, vp, vp, vp, vp, vp]


def set_threads(n):
    return lib().orc_set_threads(int(n))


# ---------------------------------------------------------------------------
# mesh helpers
# ---------------------------------------------------------------------------
def box_mesh(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    N = (nx + 1) * (ny + 1) * (nz + 1)
    C = 6 * nx * ny * nz
    coords = np.zeros(N * 3, np.float64)
    conn = np.zeros(C * 4, np.int32)
    lib().orc_box_mesh(nx, ny, nz, lx, ly, lz, _p(coords), _p(conn))
    return coords, conn


def graph(num_nodes, conn):
    C = conn.size // 4
    nnz = lib().orc_graph_nnz(num_nodes, C, _p(conn))
    rowmap = np.zeros(num_nodes + 1, np.int32)
    colidx = np.zeros(nnz, np.int32)
    lib().orc_graph_fill(num_nodes, C, _p(conn), _p(rowmap), _p(colidx))
    return rowmap, colidx


def side_set(coords, conn, axis, value, tol=1e-12):
    """Faces of cells whose three nodes lie on the plane x[axis] == value.

    Returns (cells, local node ordinals F x 3) in cell order, the layout
    Omega_h side sets give to SurfaceLoadIntegral."""
    X = coords.reshape(-1, 3)
    cn = conn.reshape(-1, 4)
    on = np.abs(X[cn][:, :, axis] - value) < tol  # C x 4
    cells, lnodes = [], []
    faces = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    for k, f in enumerate(faces):
        hit = on[:, f[0]] & on[:, f[1]] & on[:, f[2]]
        idx = np.nonzero(hit)[0]
        cells.append(idx)
        lnodes.append(np.tile(np.array(f, np.int32), (idx.size, 1)))
    cells = np.concatenate(cells).astype(np.int32)
    lnodes = np.concatenate(lnodes).astype(np.int32)
    order = np.argsort(cells, kind="stable")
    return cells[order].copy(), lnodes[order].reshape(-1).copy()


class Model:
    """Holds the arrays a Problem struct points to (keeps them alive)."""

    def __init__(self, coords, conn, *, youngs=1.0, poisson=0.3, conductivity=1.0, penalty_exponent=3.0,
                 penalty_min=1e-9, face_cells=None, face_lnodes=None, flux=(0.0, 0.0, 0.0), body=None):
        self.coords = np.ascontiguousarray(coords, np.float64)
        self.conn = np.ascontiguousarray(conn, np.int32)
        self.N = self.coords.size // 3
        self.C = self.conn.size // 4
        self.face_cells = np.ascontiguousarray(face_cells if face_cells is not None else np.zeros(0), np.int32)
        self.face_lnodes = np.ascontiguousarray(face_lnodes if face_lnodes is not None else np.zeros(0), np.int32)
        P = Problem()
        P.num_nodes, P.num_cells = self.N, self.C
        P.coords, P.conn = _p(self.coords).value, _p(self.conn).value
        P.youngs, P.poisson, P.conductivity = youngs, poisson, conductivity
        P.penalty_exponent, P.penalty_min = penalty_exponent, penalty_min
        P.num_faces = self.face_cells.size
        P.face_cells = _p(self.face_cells).value if self.face_cells.size else None
        P.face_lnodes = _p(self.face_lnodes).value if self.face_lnodes.size else None
        for k in range(3):
            P.flux[k] = flux[k]
        if body is not None:
            P.has_body, P.body_dof, P.body_value = 1, int(body[0]), float(body[1])
        self.P = P
        self.rowmap, self.colidx = graph(self.N, self.conn)
        self.nnzb = self.colidx.size

    def _vec(self, name, u, z, n):
        out = np.zeros(n, np.float64)
        getattr(lib(), "orc_" + name)(ctypes.byref(self.P), _p(np.ascontiguousarray(u, np.float64)),
                                      _p(np.ascontiguousarray(z, np.float64)), _p(out))
        return out

    # VectorFunction
    def value(self, physics, u, z):
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_value", u, z, self.N * dpn)

    def gradient_u(self, physics, u, z):
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_gradient_u", u, z, self.nnzb * dpn * dpn)

    def gradient_u_T(self, physics, u, z):
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_gradient_u_T", u, z, self.nnzb * dpn * dpn)

    def gradient_z(self, physics, u, z):
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_gradient_z", u, z, self.nnzb * dpn)

    # PhysicsScalarFunction (internal energy)
    def energy_value(self, physics, u, z):
        f = getattr(lib(), f"orc_{physics}_energy_value")
        return f(ctypes.byref(self.P), _p(np.ascontiguousarray(u, np.float64)), _p(np.ascontiguousarray(z, np.float64)))

    def energy_gradient_u(self, physics, u, z):
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_energy_gradient_u", u, z, self.N * dpn)

    def energy_gradient_z(self, physics, u, z):
        return self._vec(physics + "_energy_gradient_z", u, z, self.N)

    def cell_chain(self, u):
        C = self.C
        g = np.zeros(C * 12)
        e = np.zeros(C * 6)
        s = np.zeros(C * 6)
        r = np.zeros(C * 12)
        lib().orc_elastic_cell_chain(ctypes.byref(self.P), _p(np.ascontiguousarray(u, np.float64)), _p(g), _p(e),
                                     _p(s), _p(r))
        return g.reshape(C, 4, 3), e.reshape(C, 6), s.reshape(C, 6), r.reshape(C, 12)


def apply_block_constraints(dpn, rowmap, colidx, values, rhs, bc_dofs, bc_vals, scale=1.0):
    N = rowmap.size - 1
    bc_dofs = np.ascontiguousarray(bc_dofs, np.int32)
    bc_vals = np.ascontiguousarray(bc_vals, np.float64)
    lib().orc_apply_block_constraints(dpn, N, _p(rowmap), _p(colidx), _p(values), _p(rhs), bc_dofs.size,
                                      _p(bc_dofs), _p(bc_vals), scale)


def block_spmv(dpn_r, dpn_c, rowmap, colidx, values, x):
    N = rowmap.size - 1
    y = np.zeros(N * dpn_r)
    lib().orc_block_spmv(dpn_r, dpn_c, N, _p(rowmap), _p(colidx), _p(values), _p(np.ascontiguousarray(x)), _p(y))
    return y


def to_scipy(dpn_r, dpn_c, rowmap, colidx, values):
    """Expand a block CRS matrix into scipy CSR (test helper)."""
    import scipy.sparse as sp
    N = rowmap.size - 1
    nnzb = colidx.size
    rows = np.repeat(np.arange(N), np.diff(rowmap))
    bi = np.repeat(rows[:, None] * dpn_r, dpn_r * dpn_c, axis=1) + np.repeat(np.arange(dpn_r), dpn_c)[None, :]
    bj = np.repeat(colidx[:, None] * dpn_c, dpn_r * dpn_c, axis=1) + np.tile(np.arange(dpn_c), dpn_r)[None, :]
    ncols = (colidx.max() + 1) * dpn_c if nnzb else 0
    return sp.csr_matrix((values.reshape(nnzb, -1).ravel(), (bi.ravel(), bj.ravel())), shape=(N * dpn_r, N * dpn_c))
