"""CPU FP64 oracle (test infrastructure only).

This package is the checker for the HIP library: the tests, the smoke check
and the CPU-baseline leg of bench.py load it.  The product library never
imports, links or calls anything here.

`build()` compiles pa_oracle.cpp, the J2 restatement and the optimized CPU
variant of the elastic assembly (pa_cpu_fast.cpp, a CPU baseline) with
g++/OpenMP into oracle/liboracle.so; `lib()` loads it with ctypes.  The numpy
helpers below wrap the C entry points.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
SOURCES = [HERE / "pa_oracle.cpp", HERE / "pa_oracle_j2.cpp", HERE / "pa_cpu_fast.cpp"]

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
        ("body_cells", ctypes.c_void_p),
        ("body_cells_dof", ctypes.c_int),
    ]


class J2Problem(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_int),
        ("num_cells", ctypes.c_int),
        ("coords", ctypes.c_void_p),
        ("conn", ctypes.c_void_p),
        ("youngs", ctypes.c_double),
        ("poisson", ctypes.c_double),
        ("hardening_isotropic", ctypes.c_double),
        ("hardening_kinematic", ctypes.c_double),
        ("initial_yield_stress", ctypes.c_double),
        ("elastic_penalty_exponent", ctypes.c_double),
        ("elastic_min_ersatz", ctypes.c_double),
        ("plastic_penalty_exponent", ctypes.c_double),
        ("plastic_min_ersatz", ctypes.c_double),
        ("global_penalty_exponent", ctypes.c_double),
        ("global_min_ersatz", ctypes.c_double),
        ("pressure_scaling", ctypes.c_double),
        ("projection_penalty", ctypes.c_int),
        ("projection_penalty_exponent", ctypes.c_double),
        ("projection_min_ersatz", ctypes.c_double),
        ("work_penalty_exponent", ctypes.c_double),
        ("work_min_ersatz", ctypes.c_double),
        ("formulation", ctypes.c_int),
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
                 "thermal_energy_gradient_z", "elastic_gradient_x", "thermal_gradient_x",
                 "elastic_energy_gradient_x", "thermal_energy_gradient_x"]:
        f = getattr(L, "orc_" + name)
        f.argtypes = [P, vp, vp, vp]
        f.restype = None
    for name in ["elastic_energy_value", "thermal_energy_value"]:
        f = getattr(L, "orc_" + name)
        f.argtypes = [P, vp, vp]
        f.restype = d
    L.orc_apply_block_constraints.argtypes = [i, i, vp, vp, vp, vp, i, vp, vp, d]
    L.orc_block_spmv.argtypes = [i, i, i, vp, vp, vp, vp, vp]
    J = ctypes.POINTER(J2Problem)
    L.orc_j2_update_local_state.argtypes = [J, vp, vp, vp, vp]
    L.orc_j2_local_value.argtypes = [J, vp, vp, vp, vp, vp]
    L.orc_j2_local_gradient.argtypes = [J, i, vp, vp, vp, vp, vp]
    L.orc_j2_global_value_ws.argtypes = [J, vp, vp, vp, vp, vp]
    L.orc_j2_global_gradient_ws.argtypes = [J, i, vp, vp, vp, vp, vp]
    L.orc_batched_inverse.argtypes = [i, i, vp, vp]
    L.orc_j2_condensed_system.argtypes = [J, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_j2_global_value.argtypes = [J, vp, vp, vp, vp, vp]
    L.orc_j2_projection.argtypes = [J, vp, vp, vp]
    L.orc_j2_plastic_work.argtypes = [J, vp, vp, vp, vp, vp]
    L.orc_j2_plastic_work.restype = d
    L.orc_j2_plastic_work_partials.argtypes = [J, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_j2_cell_geometry.argtypes = [J, vp, vp]


def set_threads(n):
    return lib().orc_set_threads(int(n))


class FastElastic:
    """Optimized CPU elastic value + gradient_u (pa_cpu_fast.cpp): closed-form
    element stiffness, precomputed block ordinals, coloured atomic-free
    OpenMP assembly.  A CPU baseline, checked against the restatement."""

    def __init__(self, coords, conn, face_cells=None, face_lnodes=None, traction=(0.0, 0.0, 0.0)):
        L = lib()
        vp, i = ctypes.c_void_p, ctypes.c_int
        L.cpu_fast_create.argtypes = [i, i, vp, vp, i, vp, vp, vp]
        L.cpu_fast_create.restype = vp
        L.cpu_fast_destroy.argtypes = [vp]
        L.cpu_fast_nnzb.argtypes = [vp]
        L.cpu_fast_nnzb.restype = ctypes.c_longlong
        L.cpu_fast_colors.argtypes = [vp]
        L.cpu_fast_colors.restype = i
        L.cpu_fast_elastic.argtypes = [vp] + [ctypes.c_double] * 4 + [vp] * 4
        self._keep = [np.ascontiguousarray(coords, np.float64), np.ascontiguousarray(conn, np.int32)]
        fc = np.ascontiguousarray(face_cells if face_cells is not None else np.zeros(0), np.int32)
        fl = np.ascontiguousarray(face_lnodes if face_lnodes is not None else np.zeros(0), np.int32)
        t = np.ascontiguousarray(traction, np.float64)
        self.N, self.C = self._keep[0].size // 3, self._keep[1].size // 4
        self.h = L.cpu_fast_create(self.N, self.C, _p(self._keep[0]), _p(self._keep[1]), fc.size, _p(fc), _p(fl),
                                   _p(t))
        self.nnzb = L.cpu_fast_nnzb(self.h)
        self.colors = L.cpu_fast_colors(self.h)
        self.R = np.zeros(self.N * 3)
        self.values = np.zeros(self.nnzb * 9)

    def value_gradient_u(self, youngs, poisson, penalty_exponent, penalty_min, u, z):
        u, z = np.ascontiguousarray(u, np.float64), np.ascontiguousarray(z, np.float64)
        lib().cpu_fast_elastic(self.h, youngs, poisson, penalty_exponent, penalty_min, _p(u), _p(z), _p(self.R),
                               _p(self.values))
        return self.R, self.values

    def __del__(self):
        try:
            lib().cpu_fast_destroy(self.h)
        except Exception:
            pass


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
                 penalty_min=1e-9, face_cells=None, face_lnodes=None, flux=(0.0, 0.0, 0.0), body=None,
                 body_cells=None):
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
        if body_cells is not None:  # (dof, the load function's value at each cell's centroid)
            self.body_cells = np.ascontiguousarray(body_cells[1], np.float64)
            assert self.body_cells.size == self.C
            P.body_cells, P.body_cells_dof = _p(self.body_cells).value, int(body_cells[0])
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

    def gradient_x(self, physics, u, z):
        """(dR/dx)^T on the graph, 3 x dpn blocks: block (a, b)[i][j] = dR_{b,j}/dx_{a,i}."""
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_gradient_x", u, z, self.nnzb * 3 * dpn)

    # PhysicsScalarFunction (internal energy)
    def energy_value(self, physics, u, z):
        f = getattr(lib(), f"orc_{physics}_energy_value")
        return f(ctypes.byref(self.P), _p(np.ascontiguousarray(u, np.float64)), _p(np.ascontiguousarray(z, np.float64)))

    def energy_gradient_u(self, physics, u, z):
        dpn = 3 if physics == "elastic" else 1
        return self._vec(physics + "_energy_gradient_u", u, z, self.N * dpn)

    def energy_gradient_z(self, physics, u, z):
        return self._vec(physics + "_energy_gradient_z", u, z, self.N)

    def energy_gradient_x(self, physics, u, z):
        return self._vec(physics + "_energy_gradient_x", u, z, self.N * 3)

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


# ---------------------------------------------------------------------------
# J2 plasticity (stabilized mixed u-p, 4 dofs per node, 14 local states/cell)
# ---------------------------------------------------------------------------
def gather_nodal(conn, field, width):
    """Cell workset C x (4*width) of a nodal field (node*width + k)."""
    f = np.asarray(field, np.float64).reshape(-1, width)
    return np.ascontiguousarray(f[conn.reshape(-1, 4)].reshape(-1, 4 * width))


def batched_inverse(A):
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[-1]
    out = np.zeros_like(A)
    lib().orc_batched_inverse(A.size // (n * n), n, _p(A), _p(out))
    return out


class J2Model:
    """Holds a J2Problem and the arrays it points to.

    Parameter names follow the reference's input lists:
    J2 Plasticity (hardening moduli, initial yield stress, elastic/plastic
    properties penalty exponent and minimum ersatz), the global residual's
    Elliptic / Penalty Function (exponent, minimum value), Pressure Scaling,
    the projection's penalty (None = NoPenalty) and the plastic work
    criterion's penalty.

    formulation "mixed" is the reference's stabilized u-p formulation (4 dof
    per node: ux, uy, uz, p); "displacement" is a pure-displacement J2 (3 dof
    per node, sigma = Dev(mu)(eps - eps_p) + kappa tr(eps - eps_p) I, no
    stabilization or projection), which has no reference counterpart.  Cell
    worksets keep the node*4 + dim layout in both (the pressure slot is zero
    in the displacement formulation); nodal vectors and matrix blocks have
    `dpn` = 4 or 3 components."""

    def __init__(self, coords, conn, *, youngs, poisson, hardening_isotropic, hardening_kinematic,
                 initial_yield_stress, elastic_penalty_exponent=1.0, elastic_min_ersatz=1e-9,
                 plastic_penalty_exponent=0.5, plastic_min_ersatz=1e-4, global_penalty_exponent=1.0,
                 global_min_ersatz=1e-9, pressure_scaling=1.0, projection_penalty=None, work_penalty=(3.0, 1e-9),
                 formulation="mixed"):
        self.coords = np.ascontiguousarray(coords, np.float64)
        self.conn = np.ascontiguousarray(conn, np.int32)
        self.N = self.coords.size // 3
        self.C = self.conn.size // 4
        P = J2Problem()
        P.num_nodes, P.num_cells = self.N, self.C
        P.coords, P.conn = _p(self.coords).value, _p(self.conn).value
        P.youngs, P.poisson = youngs, poisson
        P.hardening_isotropic, P.hardening_kinematic = hardening_isotropic, hardening_kinematic
        P.initial_yield_stress = initial_yield_stress
        P.elastic_penalty_exponent, P.elastic_min_ersatz = elastic_penalty_exponent, elastic_min_ersatz
        P.plastic_penalty_exponent, P.plastic_min_ersatz = plastic_penalty_exponent, plastic_min_ersatz
        P.global_penalty_exponent, P.global_min_ersatz = global_penalty_exponent, global_min_ersatz
        P.pressure_scaling = pressure_scaling
        if projection_penalty is None:
            P.projection_penalty = 0
        else:
            P.projection_penalty = 1
            P.projection_penalty_exponent, P.projection_min_ersatz = projection_penalty
        P.work_penalty_exponent, P.work_min_ersatz = work_penalty
        P.formulation = {"mixed": 0, "displacement": 1}[formulation]
        self.dpn = 4 if P.formulation == 0 else 3
        self.P = P
        self.rowmap, self.colidx = graph(self.N, self.conn)
        self.nnzb = self.colidx.size

    @staticmethod
    def _a(x):
        return _p(np.ascontiguousarray(x, np.float64))

    def update_local_state(self, u, cp, z):
        c = np.zeros(self.C * 14)
        u, cp, z = [np.ascontiguousarray(x, np.float64) for x in (u, cp, z)]
        lib().orc_j2_update_local_state(ctypes.byref(self.P), _p(u), _p(cp), _p(z), _p(c))
        return c

    def local_value(self, u, c, cp, z):
        H = np.zeros(self.C * 14)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, c, cp, z)]
        lib().orc_j2_local_value(ctypes.byref(self.P), *[_p(a) for a in args], _p(H))
        return H

    def local_gradient(self, which, u, c, cp, z):
        """which: 'c', 'u', 'z', 'cp' -> C x 14 x W."""
        k = {"c": 0, "u": 1, "z": 2, "cp": 3}[which]
        W = {0: 14, 1: 16, 2: 4, 3: 14}[k]
        out = np.zeros(self.C * 14 * W)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, c, cp, z)]
        lib().orc_j2_local_gradient(ctypes.byref(self.P), k, *[_p(a) for a in args], _p(out))
        return out.reshape(self.C, 14, W)

    def global_value_ws(self, u, c, pi_ws, z):
        out = np.zeros(self.C * 16)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, c, pi_ws, z)]
        lib().orc_j2_global_value_ws(ctypes.byref(self.P), *[_p(a) for a in args], _p(out))
        return out.reshape(self.C, 16)

    def global_gradient_ws(self, which, u, c, pi_ws, z):
        """which: 'u', 'c', 'z', 'pi' -> C x 16 x W."""
        k = {"u": 0, "c": 1, "z": 2, "pi": 3}[which]
        W = {0: 16, 1: 14, 2: 4, 3: 12}[k]
        out = np.zeros(self.C * 16 * W)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, c, pi_ws, z)]
        lib().orc_j2_global_gradient_ws(ctypes.byref(self.P), k, *[_p(a) for a in args], _p(out))
        return out.reshape(self.C, 16, W)

    def global_value(self, u, c, pi, z):
        R = np.zeros(self.N * self.dpn)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, c, pi, z)]
        lib().orc_j2_global_value(ctypes.byref(self.P), *[_p(a) for a in args], _p(R))
        return R

    def condensed_system(self, u, c, cp, pi, z):
        """(block-CRS values nnzb x dpn^2, residual dpn N) of one Newton linearization."""
        vals = np.zeros(self.nnzb * self.dpn * self.dpn)
        R = np.zeros(self.N * self.dpn)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, c, cp, pi, z)]
        lib().orc_j2_condensed_system(ctypes.byref(self.P), *[_p(a) for a in args], _p(self.rowmap),
                                      _p(self.colidx), _p(vals), _p(R))
        return vals, R

    def projection(self, u, z):
        pi = np.zeros(self.N * 3)
        u, z = np.ascontiguousarray(u, np.float64), np.ascontiguousarray(z, np.float64)
        lib().orc_j2_projection(ctypes.byref(self.P), _p(u), _p(z), _p(pi))
        return pi

    def plastic_work(self, u, up, c, cp, z):
        args = [np.ascontiguousarray(x, np.float64) for x in (u, up, c, cp, z)]
        return lib().orc_j2_plastic_work(ctypes.byref(self.P), *[_p(a) for a in args])

    def plastic_work_partials(self, u, up, c, cp, z):
        gu, gup = np.zeros(self.N * self.dpn), np.zeros(self.N * self.dpn)
        gc, gcp = np.zeros(self.C * 14), np.zeros(self.C * 14)
        gz = np.zeros(self.N)
        args = [np.ascontiguousarray(x, np.float64) for x in (u, up, c, cp, z)]
        lib().orc_j2_plastic_work_partials(ctypes.byref(self.P), *[_p(a) for a in args], _p(gu), _p(gup), _p(gc),
                                           _p(gcp), _p(gz))
        return gu, gup, gc, gcp, gz

    def solve(self, z, bc_dofs, bc_values, num_steps=1, max_newton=10, stop_tol=1e-8, current_norm_tol=1e-8):
        """Pseudo-time loop with the condensed Newton solve.

        src/PlasticityProblem.hpp:784-868 (steps, load factor t_k = (k+1)/K,
        Dirichlet increment on the first Newton iteration, projection update
        after each step) and src/NewtonRaphsonSolver.hpp:555-610 (elastic
        trial update, assemble -R_t and K_t, constraints, absolute-measure
        stopping test, u += du, local update).  Linear solves: scipy direct.
        Returns per-step lists of u (dpn N), c (C*14), pi (N*3; zero in the
        displacement formulation, which has no projection) and the Newton
        iteration counts."""
        import scipy.sparse.linalg as spla
        z = np.ascontiguousarray(z, np.float64)
        bc_dofs = np.ascontiguousarray(bc_dofs, np.int32)
        bc_values = np.ascontiguousarray(bc_values, np.float64)
        us, cs, pis, iters = [], [], [], []
        D = self.dpn
        up = np.zeros(self.N * D)
        cp = np.zeros(self.C * 14)
        pi = np.zeros(self.N * 3)
        g_prev = np.zeros_like(bc_values)
        for k in range(num_steps):
            t = (1.0 / num_steps) * (k + 1)
            g = bc_values * t
            dg = g - g_prev if k > 0 else g.copy()
            g_prev = g
            u = up.copy()
            c = self.update_local_state(u, cp, z)
            it = 0
            ref = np.finfo(float).max
            measure = np.finfo(float).max
            while True:
                vals, R = self.condensed_system(u, c, cp, pi, z)
                rhs = -R
                apply_block_constraints(D, self.rowmap, self.colidx, vals, rhs, bc_dofs,
                                        dg if it == 0 else np.zeros_like(dg))
                norm = float(np.linalg.norm(rhs))
                if it == 0:
                    ref = norm
                else:
                    measure = abs(ref - norm)
                    ref = norm
                if measure < stop_tol or norm < current_norm_tol or it >= max_newton:
                    break
                K = to_scipy(D, D, self.rowmap, self.colidx, vals)
                du = spla.spsolve(K.tocsc(), rhs)
                u = u + du
                c = self.update_local_state(u, cp, z)
                it += 1
            us.append(u)
            cs.append(c)
            iters.append(it)
            pis.append(pi)
            if k + 1 < num_steps and D == 4:
                pi = self.projection(u, z)
            up, cp = u, c
        return us, cs, pis, iters

    # -- criterion and its path-dependent adjoint ---------------------------
    def cell_geometry(self):
        g = np.zeros(self.C * 12)
        v = np.zeros(self.C)
        lib().orc_j2_cell_geometry(ctypes.byref(self.P), _p(g), _p(v))
        return g.reshape(self.C, 4, 3), v

    def _projection_weight(self, z):
        """Cell weight w(rho) of the projection and its derivative dw/drho."""
        rho = np.asarray(z)[self.conn.reshape(-1, 4)].mean(axis=1)
        if not self.P.projection_penalty:
            return np.ones(self.C), np.zeros(self.C)
        p, emin = self.P.projection_penalty_exponent, self.P.projection_min_ersatz
        return emin + (1 - emin) * rho ** p, (1 - emin) * p * rho ** (p - 1)

    def criterion_value(self, z, us, cs, multiplier=1.0):
        """Plastic work summed over the steps (src/PlasticityProblem.hpp:925-956)."""
        total = 0.0
        for k in range(len(us)):
            up = us[k - 1] if k else np.zeros(self.N * self.dpn)
            cp = cs[k - 1] if k else np.zeros(self.C * 14)
            total += self.plastic_work(us[k], up, cs[k], cp, z)
        return multiplier * total

    def criterion_gradient_z(self, z, us, cs, pis, bc_dofs, multiplier=1.0):
        """dF/dz by the backward recursion of src/PathDependentAdjointSolver.hpp
        (:277-351 local RHS, :394-469 global RHS, :737-770 projection adjoint,
        :772-798 transposed condensed tangent solve, :824-880 local adjoint,
        :517-551 control partials) driven as in src/PlasticityProblem.hpp:1035-1066,
        plus the criterion's own control partial summed over steps.

        The projection control partial is evaluated with the step's own
        pressure, as the reference's forward-state bookkeeping passes it
        (src/PlasticityProblem.hpp:1095-1112); it vanishes for NoPenalty."""
        import scipy.sparse.linalg as spla
        z = np.ascontiguousarray(z, np.float64)
        K, N, C, D = len(us), self.N, self.C, self.dpn
        mixed = D == 4
        cn = self.conn.reshape(-1, 4)
        grads, vols = self.cell_geometry()
        w, dw = self._projection_weight(z)
        ps = self.P.pressure_scaling
        rowsum = np.zeros(N)
        np.add.at(rowsum, cn.ravel(), np.repeat(0.25 * vols * w, 4))
        zeroC, zeroN = np.zeros(C * 14), np.zeros(N * D)

        def assemble(ws, width):  # cell workset C x (4*width) -> nodal
            out = np.zeros(N * width)
            idx = (cn[:, :, None] * width + np.arange(width)).reshape(C, -1)
            np.add.at(out, idx.ravel(), ws.reshape(C, -1).ravel())
            return out

        def to_nodal_layout(ws16):  # C x 16 (node*4 + dim) -> C x 4*D
            return ws16.reshape(C, 4, 4)[:, :, :D].reshape(C, 4 * D)

        def to_cell_layout(wsD):  # C x 4*D -> C x 16, pressure slot zero
            out = np.zeros((C, 4, 4))
            out[:, :, :D] = wsD.reshape(C, 4, D)
            return out.reshape(C, 16)

        def partials(k):
            up = us[k - 1] if k else zeroN
            cp = cs[k - 1] if k else zeroC
            return [multiplier * a for a in self.plastic_work_partials(us[k], up, cs[k], cp, z)]

        dfdz = np.zeros(N)
        lam1 = mu1 = gam1 = None
        for k in range(K - 1, -1, -1):
            u, c, pi = us[k], cs[k], pis[k]
            cp = cs[k - 1] if k else zeroC
            pi_ws = gather_nodal(self.conn, pi, 3)
            invJ = batched_inverse(self.local_gradient("c", u, c, cp, z))
            dHdu = self.local_gradient("u", u, c, cp, z)
            dHdz = self.local_gradient("z", u, c, cp, z)
            dRdc = self.global_gradient_ws("c", u, c, pi_ws, z)
            dRdz = self.global_gradient_ws("z", u, c, pi_ws, z)
            dRdpi = self.global_gradient_ws("pi", u, c, pi_ws, z)
            gu, gup, gc, gcp, gz = partials(k)
            gc = gc.reshape(C, 14)
            b_local = gc.copy()
            rhs = gu.copy()
            if k < K - 1:
                gu1, gup1, gc1, gcp1, gz1 = partials(k + 1)
                dHdcp1 = self.local_gradient("cp", us[k + 1], cs[k + 1], c, z)
                b_local += gcp1.reshape(C, 14) + np.einsum("eij,ei->ej", dHdcp1, mu1)
                rhs += gup1
                if mixed:
                    # DpDn^T gamma_{k+1}: dQ_{a,d}/dp_b = -1/4 V w ps dN_b/dx_d
                    gw = gather_nodal(self.conn, gam1, 3).reshape(C, 4, 3).sum(axis=1)  # sum_a gamma_{a,d}
                    dp = -0.25 * (vols * w * ps)[:, None] * np.einsum("ebd,ed->eb", grads, gw)
                    pr = np.zeros(N)
                    np.add.at(pr, cn.ravel(), dp.ravel())
                    rhs[3::4] += pr
            t = np.einsum("eji,ej->ei", invJ, b_local)
            rhs -= assemble(to_nodal_layout(np.einsum("eji,ej->ei", dHdu, t)), D)
            rhs = -rhs
            vals, _ = self.condensed_system(u, c, cp, pi, z)
            # transposed assembly (assemble_jacobian_transpose_pod): block (r, c)
            # of K^T is block (c, r) of K transposed; the graph is symmetric
            rmap, cidx = self.rowmap, self.colidx
            rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(rmap))
            pos = np.searchsorted(rows * N + cidx, cidx.astype(np.int64) * N + rows)
            bvals = np.ascontiguousarray(vals.reshape(-1, D, D)[pos].transpose(0, 2, 1)).ravel()
            apply_block_constraints(D, rmap, cidx, bvals, rhs, bc_dofs, np.zeros(len(bc_dofs)))
            lam = spla.spsolve(to_scipy(D, D, rmap, cidx, bvals).tocsc(), rhs)
            lam_ws = to_cell_layout(gather_nodal(self.conn, lam, D))
            b2 = gc + np.einsum("eji,ej->ei", dRdc, lam_ws)
            if k < K - 1:
                b2 += gcp1.reshape(C, 14) + np.einsum("eij,ei->ej", dHdcp1, mu1)
            mu = -np.einsum("eji,ej->ei", invJ, b2)
            if k == 0 or not mixed:
                gam = np.zeros(N * 3)
            else:
                r = -assemble(np.einsum("eji,ej->ei", dRdpi, lam_ws), 3)
                gam = (r.reshape(N, 3) / rowsum[:, None]).ravel()
            dfdz += assemble(np.einsum("eji,ej->ei", dRdz, lam_ws), 1)
            dfdz += assemble(np.einsum("eji,ej->ei", dHdz, mu), 1)
            if mixed and self.P.projection_penalty and k > 0:
                # dQ_{a,d}/dz_b = 1/4 * 1/4 V dw (Pi_gp - ps grad p_k)_d
                pk = u[3::4][cn]
                gp = ps * np.einsum("eb,ebd->ed", pk, grads)
                pig = pi_ws.reshape(C, 4, 3).mean(axis=1)
                gam_ws = gather_nodal(self.conn, gam, 3).reshape(C, 4, 3).sum(axis=1)
                q = 0.25 * 0.25 * vols * dw * np.einsum("ed,ed->e", pig - gp, gam_ws)
                dfdz += assemble(np.repeat(q[:, None], 4, axis=1), 1)
            dfdz += gz
            lam1, mu1, gam1 = lam, mu, gam
        return dfdz
