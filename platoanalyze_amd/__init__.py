"""Python host mirror of the MI355X Tet4 assembly library.

The compute path is libplatoanalyze_amd.so (hand-written HIP kernels for
gfx950 behind the C ABI in include/platoanalyze_amd.h).  This module loads it
with ctypes and mirrors the reference's interface for this path:

  * ``VectorFunction``  ~ Plato::Elliptic::VectorFunction<PhysicsT>
    (src/elliptic/VectorFunction.hpp:115-710): value, gradient_u,
    gradient_u_T, gradient_z, gradient_x;
  * ``PhysicsScalarFunction`` ~ Plato::Elliptic::PhysicsScalarFunction
    (src/elliptic/PhysicsScalarFunction.hpp:225-494): value, gradient_u,
    gradient_z, gradient_x;
  * ``apply_block_constraints`` ~ Plato::applyBlockConstraints
    (src/ApplyConstraints.hpp:40-112).

Arrays live in device memory (``DeviceArray``).  There is no CPU fallback:
importing without the built library raises.  Errors from the library are
raised as ``PlatoError`` (the reference throws std::runtime_error through
ANALYZE_THROWERR, src/AnalyzeMacros.hpp:29-33).
"""
import ctypes
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libplatoanalyze_amd.so"

FP64, FP32 = 0, 1
MECHANICS, THERMAL = 0, 1
INTERNAL_ELASTIC_ENERGY, INTERNAL_THERMAL_ENERGY = 0, 1


PRECOND_NONE, PRECOND_JACOBI, PRECOND_BLOCK_JACOBI = 0, 1, 2


class PlatoError(RuntimeError):
    pass


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("sweeps", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("converged", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


PRECOND_MULTIGRID = 3


SMOOTHERS = {"default": 0, "jacobi": 1, "chebyshev": 2}


class MultigridOptions(ctypes.Structure):
    _fields_ = [("max_levels", ctypes.c_int), ("pre_sweeps", ctypes.c_int), ("post_sweeps", ctypes.c_int),
                ("omega", ctypes.c_double), ("max_coarse_nodes", ctypes.c_int), ("symmetric", ctypes.c_int),
                ("smoother", ctypes.c_int), ("coarse_sweeps", ctypes.c_int), ("coarse_max_nodes", ctypes.c_int)]


class ProblemOptions(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("precond", ctypes.c_int), ("tol", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("replace_every", ctypes.c_int), ("check_every", ctypes.c_int)]


class J2Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "youngs", "poisson", "hardening_isotropic", "hardening_kinematic", "initial_yield_stress",
        "elastic_penalty_exponent", "elastic_min_ersatz", "plastic_penalty_exponent", "plastic_min_ersatz",
        "global_penalty_exponent", "global_min_ersatz", "pressure_scaling")] + [
        ("projection_penalty", ctypes.c_int), ("projection_penalty_exponent", ctypes.c_double),
        ("projection_min_ersatz", ctypes.c_double), ("work_penalty_exponent", ctypes.c_double),
        ("work_min_ersatz", ctypes.c_double), ("formulation", ctypes.c_int)]


J2_MIXED, J2_DISPLACEMENT = 0, 1


class PlasticityOptions(ctypes.Structure):
    _fields_ = [("num_steps", ctypes.c_int), ("max_newton", ctypes.c_int), ("newton_stop_tol", ctypes.c_double),
                ("newton_current_norm_tol", ctypes.c_double), ("linear_max_iter", ctypes.c_int),
                ("linear_tol", ctypes.c_double), ("krylov", ctypes.c_int), ("restart", ctypes.c_int),
                ("checkpoint_stride", ctypes.c_int)]


KRYLOV_GMRES, KRYLOV_BICGSTAB = 0, 1


WRT_U, WRT_C, WRT_Z, WRT_C_PREV, WRT_PI = 0, 1, 2, 3, 4


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    ll, f = ctypes.c_longlong, ctypes.c_float
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "pa_last_error": ([], ctypes.c_char_p),
        "pa_version": ([], ctypes.c_char_p),
        "pa_device_count": ([ip], i),
        "pa_malloc": ([ctypes.POINTER(vp), sz], i),
        "pa_free": ([vp], i),
        "pa_memcpy_h2d": ([vp, vp, sz], i),
        "pa_memcpy_d2h": ([vp, vp, sz], i),
        "pa_memcpy_d2d": ([vp, vp, sz, vp], i),
        "pa_memset": ([vp, i, sz, vp], i),
        "pa_copy_kernel": ([vp, vp, sz, vp], i),
        "pa_stream_create": ([ctypes.POINTER(vp)], i),
        "pa_stream_destroy": ([vp], i),
        "pa_stream_synchronize": ([vp], i),
        "pa_event_create": ([ctypes.POINTER(vp)], i),
        "pa_event_destroy": ([vp], i),
        "pa_event_record": ([vp, vp], i),
        "pa_event_synchronize": ([vp], i),
        "pa_event_elapsed_ms": ([ctypes.POINTER(ctypes.c_float), vp, vp], i),
        "pa_box_mesh": ([i, i, i, d, d, d, vp, vp], i),
        "pa_mesh_create": ([vp, i, vp, i, ctypes.POINTER(vp)], i),
        "pa_mesh_destroy": ([vp], i),
        "pa_mesh_sizes": ([vp, ip, ip, ctypes.POINTER(ctypes.c_int64)], i),
        "pa_mesh_graph": ([vp, vp, vp], i),
        "pa_mesh_assembly_info": ([vp, ip, ip, ip], i),
        "pa_mesh_grid_info": ([vp, ip], i),
        "pa_mesh_graph_dev": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp)], i),
        "pa_mesh_add_side_set": ([vp, ctypes.c_char_p, i, vp, vp], i),
        "pa_mesh_add_node_set": ([vp, ctypes.c_char_p, i, vp], i),
        "pa_mesh_node_set": ([vp, ctypes.c_char_p, ip, vp], i),
        "pa_vector_function_create": ([vp, i, vp, vp, ctypes.POINTER(vp)], i),
        "pa_vector_function_destroy": ([vp], i),
        "pa_vector_function_dofs_per_node": ([vp, ip], i),
        "pa_vector_function_add_uniform_natural_bc": ([vp, ctypes.c_char_p, vp], i),
        "pa_vector_function_add_body_load": ([vp, i, d], i),
        "pa_vector_function_add_body_load_function": ([vp, i, ctypes.c_char_p], i),
        "pa_expression_eval": ([ctypes.c_char_p, i, vp, vp, vp, d, vp], i),
        "pa_vector_function_value": ([vp, vp, vp, vp, vp], i),
        "pa_vector_function_gradient_u": ([vp, vp, vp, vp, i, vp], i),
        "pa_vector_function_gradient_u_T": ([vp, vp, vp, vp, i, vp], i),
        "pa_vector_function_value_gradient_u": ([vp, vp, vp, vp, vp, i, vp], i),
        "pa_vector_function_gradient_z": ([vp, vp, vp, vp, vp], i),
        "pa_vector_function_gradient_x": ([vp, vp, vp, vp, vp], i),
        "pa_scalar_function_create": ([vp, i, vp, vp, ctypes.POINTER(vp)], i),
        "pa_scalar_function_destroy": ([vp], i),
        "pa_scalar_function_value": ([vp, vp, vp, vp, vp], i),
        "pa_scalar_function_value_noalloc": ([vp, vp, vp, vp, vp], i),
        "pa_scalar_function_gradient_u": ([vp, vp, vp, vp, vp], i),
        "pa_scalar_function_gradient_z": ([vp, vp, vp, vp, vp], i),
        "pa_scalar_function_gradient_x": ([vp, vp, vp, vp, vp], i),
        "pa_problem_criterion_gradient_x": ([vp, vp, vp, vp, vp], i),
        "pa_apply_block_constraints": ([vp, i, vp, i, vp, i, vp, vp, d, vp], i),
        "pa_block_spmv": ([vp, i, i, vp, i, vp, d, vp, vp], i),
        "pa_solver_create": ([vp, i, i, ctypes.POINTER(vp)], i),
        "pa_solver_destroy": ([vp], i),
        "pa_solver_pcg": ([vp, vp, i, vp, vp, d, i, i, ctypes.POINTER(SolveStats), vp], i),
        "pa_problem_create": ([vp, vp, i, vp, vp, ctypes.POINTER(ProblemOptions), ctypes.POINTER(vp)], i),
        "pa_problem_destroy": ([vp], i),
        "pa_problem_set_options": ([vp, ctypes.POINTER(ProblemOptions)], i),
        "pa_problem_solution": ([vp, vp, vp, ctypes.POINTER(SolveStats), vp], i),
        "pa_problem_criterion_value": ([vp, vp, vp, vp, vp], i),
        "pa_problem_criterion_gradient": ([vp, vp, vp, vp, vp], i),
        "pa_problem_set_multigrid": ([vp, i, i, i, ctypes.POINTER(MultigridOptions)], i),
        "pa_solver_set_multigrid": ([vp, vp], i),
        "pa_multigrid_create": ([vp, i, i, i, i, ctypes.POINTER(MultigridOptions), ctypes.POINTER(vp)], i),
        "pa_multigrid_destroy": ([vp], i),
        "pa_multigrid_levels": ([vp, ip, ip], i),
        "pa_multigrid_setup": ([vp, vp, i, i, vp, vp], i),
        "pa_multigrid_apply": ([vp, vp, vp, vp], i),
        "pa_multigrid_level_operator": ([vp, i, vp, ip, ip, ip], i),
        "pa_j2_params_default": ([ctypes.POINTER(J2Params)], None),
        "pa_plasticity_create": ([vp, ctypes.POINTER(J2Params), ctypes.POINTER(vp)], i),
        "pa_plasticity_destroy": ([vp], i),
        "pa_j2_update_local_state": ([vp, vp, vp, vp, vp, vp], i),
        "pa_j2_local_value": ([vp, vp, vp, vp, vp, vp, vp], i),
        "pa_j2_local_gradient": ([vp, i, vp, vp, vp, vp, vp, vp], i),
        "pa_j2_global_value": ([vp, vp, vp, vp, vp, vp, vp], i),
        "pa_j2_global_gradient": ([vp, i, vp, vp, vp, vp, vp, vp], i),
        "pa_j2_condensed_system": ([vp, vp, vp, vp, vp, vp, vp, vp, i, vp], i),
        "pa_j2_projection": ([vp, vp, vp, vp, vp], i),
        "pa_j2_plastic_work": ([vp, vp, vp, vp, vp, vp, vp, vp], i),
        "pa_plasticity_options_default": ([ctypes.POINTER(PlasticityOptions)], None),
        "pa_plasticity_problem_create": ([vp, i, vp, vp, ctypes.POINTER(PlasticityOptions), ctypes.POINTER(vp)], i),
        "pa_plasticity_problem_destroy": ([vp], i),
        "pa_plasticity_problem_set_multigrid": ([vp, i, i, i, ctypes.POINTER(MultigridOptions)], i),
        "pa_plasticity_problem_solution": ([vp, vp, ip, vp], i),
        "pa_plasticity_problem_state": ([vp, i, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)], i),
        "pa_plasticity_problem_stats": ([vp, i, ip, ip, ctypes.POINTER(d)], i),
        "pa_plasticity_problem_state_bytes": ([vp, ctypes.POINTER(sz)], i),
        "pa_plasticity_problem_criterion_value": ([vp, vp, d, ctypes.POINTER(d), vp], i),
        "pa_plasticity_problem_criterion_gradient": ([vp, vp, d, vp, vp], i),
        # warm-start network kernels
        "pa_cnn_conv3d": ([i, i, i, vp, i, i, vp, vp, i, vp, i, i, vp, i, i, vp], i),
        "pa_cnn_pack_weights": ([vp, i, i, i, vp, vp], i),
        "pa_cnn_conv3d_wgrad_parts": ([i, i, i, vp, i, i, vp, i, i, ll, vp, vp], i),
        "pa_cnn_sum_parts": ([ll, i, vp, vp, vp], i),
        "pa_cnn_bias_parts": ([ll, vp, i, i, ll, vp, vp], i),
        "pa_cnn_pool": ([i, i, i, vp, i, i, vp, i, vp], i),
        "pa_cnn_pool_bwd": ([i, i, i, vp, i, i, vp, i, i, vp], i),
        "pa_cnn_upsample": ([i, i, i, vp, i, i, vp, i, vp], i),
        "pa_cnn_upsample_bwd": ([i, i, i, vp, i, i, vp, i, vp], i),
        "pa_cnn_relu_bwd": ([ll, i, vp, i, vp, i, vp], i),
        "pa_cnn_mse": ([ll, i, i, vp, i, vp, vp, i, f, vp, i, vp, vp], i),
        "pa_cnn_adam": ([ll, f, f, f, f, vp, vp, vp, vp, vp], i),
        "pa_cnn_energy_loss": ([i, i, i, i, i, vp, i, vp, vp, i, f, f, vp, i, vp, vp], i),
        "pa_cnn_to_bf16": ([ll, i, vp, i, vp, i, vp], i),
        "pa_cnn_add_bf16": ([ll, i, vp, i, vp, i, vp], i),
        "pa_cnn_nodes_to_voxels": ([i, i, i, i, vp, d, d, vp, i, i, vp], i),
        "pa_cnn_voxels_to_nodes": ([i, i, i, i, vp, i, d, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


lib = _load()


def _check(status):
    if status != 0:
        raise PlatoError(lib.pa_last_error().decode())


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def version():
    return lib.pa_version().decode()


def device_count():
    n = ctypes.c_int(0)
    _check(lib.pa_device_count(ctypes.byref(n)))
    return n.value


# ----------------------------------------------------------------- device
class DeviceArray:
    """A flat device allocation with a numpy dtype."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib.pa_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self._ptr = p

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype).ravel()
        d = cls(a.size, a.dtype)
        if a.size:
            _check(lib.pa_memcpy_h2d(d._ptr, _np_ptr(a), a.nbytes))
        return d

    @property
    def ptr(self):
        return self._ptr

    def numpy(self):
        out = np.empty(self.size, self.dtype)
        if self.size:
            _check(lib.pa_memcpy_d2h(_np_ptr(out), self._ptr, self.nbytes))
        return out

    def zero(self, stream=None):
        _check(lib.pa_memset(self._ptr, 0, self.nbytes, stream))

    def copy_from(self, other, stream=None):
        assert other.nbytes == self.nbytes
        _check(lib.pa_memcpy_d2d(self._ptr, other._ptr, self.nbytes, stream))

    def copy_kernel_from(self, other, stream=None):
        assert other.nbytes == self.nbytes
        _check(lib.pa_copy_kernel(self._ptr, other._ptr, self.nbytes, _sh(stream)))

    def free(self):
        if self._ptr is not None and self._ptr.value:
            lib.pa_free(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _check(lib.pa_stream_create(ctypes.byref(s)))
        self.handle = s

    def synchronize(self):
        _check(lib.pa_stream_synchronize(self.handle))

    def __del__(self):
        try:
            lib.pa_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(lib.pa_event_create(ctypes.byref(e)))
        self.handle = e

    def record(self, stream=None):
        _check(lib.pa_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def synchronize(self):
        _check(lib.pa_event_synchronize(self.handle))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _check(lib.pa_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle))
        return ms.value

    def __del__(self):
        try:
            lib.pa_event_destroy(self.handle)
        except Exception:
            pass


def _sh(stream):
    return stream.handle if isinstance(stream, Stream) else stream


def synchronize():
    _check(lib.pa_stream_synchronize(None))


# ------------------------------------------------------------------- mesh
def box_mesh(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    """Structured box split into 6 tets per hex (host arrays coords, conn)."""
    N = (nx + 1) * (ny + 1) * (nz + 1)
    C = 6 * nx * ny * nz
    coords = np.empty(N * 3, np.float64)
    conn = np.empty(C * 4, np.int32)
    _check(lib.pa_box_mesh(nx, ny, nz, lx, ly, lz, _np_ptr(coords), _np_ptr(conn)))
    return coords, conn


def box_side_set(coords, conn, axis, value, tol=1e-12):
    """Faces of cells with all three face nodes on the plane x[axis] == value,
    as (cells, local node ordinals F x 3), cell-ascending."""
    X = coords.reshape(-1, 3)
    cn = conn.reshape(-1, 4)
    on = np.abs(X[cn][:, :, axis] - value) < tol
    cells, lnodes = [], []
    for f in [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]:
        idx = np.nonzero(on[:, f[0]] & on[:, f[1]] & on[:, f[2]])[0]
        cells.append(idx)
        lnodes.append(np.tile(np.array(f, np.int32), (idx.size, 1)))
    cells = np.concatenate(cells).astype(np.int32)
    lnodes = np.concatenate(lnodes).astype(np.int32)
    order = np.argsort(cells, kind="stable")
    return cells[order].copy(), lnodes[order].reshape(-1).copy()



def evaluate_function(function, x=None, y=None, z=None, t=0.0):
    """A 'Function' string at points (x, y, z) and time t, by the library's
    evaluator (host only, no GPU)."""
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64).ravel() for a in (x, y, z)]
    n = max([a.size for a in arrs if a is not None], default=1)
    out = np.empty(n, np.float64)
    ptr = [None if a is None else _np_ptr(a) for a in arrs]
    _check(lib.pa_expression_eval(function.encode(), n, ptr[0], ptr[1], ptr[2], float(t), _np_ptr(out)))
    return out

class Mesh:
    """Mesh + block-CRS graph on the device (Omega_h mesh + NodeNodeGraph)."""

    def __init__(self, coords, conn):
        coords = np.ascontiguousarray(coords, np.float64).ravel()
        conn = np.ascontiguousarray(conn, np.int32).ravel()
        self.coords_host, self.conn_host = coords, conn
        h = ctypes.c_void_p()
        _check(lib.pa_mesh_create(_np_ptr(coords), coords.size // 3, _np_ptr(conn), conn.size // 4, ctypes.byref(h)))
        self.handle = h
        N, C, nnzb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(lib.pa_mesh_sizes(h, ctypes.byref(N), ctypes.byref(C), ctypes.byref(nnzb)))
        self.num_nodes, self.num_cells, self.num_blocks = N.value, C.value, nnzb.value

    @classmethod
    def box(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, side_sets=True):
        coords, conn = box_mesh(nx, ny, nz, lx, ly, lz)
        m = cls(coords, conn)
        m.box_dims = (nx, ny, nz, lx, ly, lz)
        if side_sets:
            for axis, name in enumerate("xyz"):
                L = (lx, ly, lz)[axis]
                for sign, val in (("-", 0.0), ("+", L)):
                    cells, ln = box_side_set(coords, conn, axis, val)
                    m.add_side_set(name + sign, cells, ln)
                    X = coords.reshape(-1, 3)
                    m.add_node_set(name + sign, np.nonzero(np.abs(X[:, axis] - val) < 1e-12)[0])
        return m

    @classmethod
    def from_exodus(cls, path):
        """Mesh of a Tet4 Exodus II file with its side sets and node sets
        (platoanalyze_amd.exodus; the reference reads meshes through Omega_h,
        src/OmegaHMesh.cpp:13-27)."""
        from . import exodus
        em = exodus.read(path)
        m = cls(em.coords, em.conn)
        for name, (cells, ln) in em.side_sets.items():
            m.add_side_set(name, cells, ln)
        for name, nodes in em.node_sets.items():
            m.add_node_set(name, nodes)
        return m

    def assembly_info(self):
        """(cell tiles, distinct tile schedules, distinct row patterns) of the assembly kernels."""
        t, tp, rp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.pa_mesh_assembly_info(self.handle, ctypes.byref(t), ctypes.byref(tp), ctypes.byref(rp)))
        return t.value, tp.value, rp.value

    def grid_info(self):
        """Node grid (x, y, z counts) of the streaming grid kernels, or None
        when the mesh is not a Kuhn-split grid."""
        d = (ctypes.c_int * 3)()
        _check(lib.pa_mesh_grid_info(self.handle, d))
        return tuple(d) if d[0] else None

    def graph(self):
        rowmap = np.empty(self.num_nodes + 1, np.int32)
        colidx = np.empty(self.num_blocks, np.int32)
        _check(lib.pa_mesh_graph(self.handle, _np_ptr(rowmap), _np_ptr(colidx)))
        return rowmap, colidx

    def add_side_set(self, name, cells, local_nodes):
        cells = np.ascontiguousarray(cells, np.int32)
        local_nodes = np.ascontiguousarray(local_nodes, np.int32)
        _check(lib.pa_mesh_add_side_set(self.handle, name.encode(), cells.size, _np_ptr(cells), _np_ptr(local_nodes)))

    def add_node_set(self, name, nodes):
        nodes = np.ascontiguousarray(nodes, np.int32)
        _check(lib.pa_mesh_add_node_set(self.handle, name.encode(), nodes.size, _np_ptr(nodes)))

    def node_set(self, name):
        n = ctypes.c_int()
        _check(lib.pa_mesh_node_set(self.handle, name.encode(), ctypes.byref(n), None))
        out = np.empty(n.value, np.int32)
        _check(lib.pa_mesh_node_set(self.handle, name.encode(), ctypes.byref(n), _np_ptr(out)))
        return out

    def __del__(self):
        try:
            lib.pa_mesh_destroy(self.handle)
        except Exception:
            pass


class CrsMatrix:
    """Block CRS matrix on the mesh graph (src/alg/CrsMatrix.hpp:19-111)."""

    def __init__(self, mesh, bi, bj, values, precision):
        self.mesh, self.bi, self.bj, self.values, self.precision = mesh, bi, bj, values, precision
        self.num_rows = mesh.num_nodes * bi
        self.num_cols = mesh.num_nodes * bj
        self.is_block_matrix = bi * bj > 1

    def entries(self):
        return self.values.numpy()

    def spmv(self, x, y, beta=0.0, stream=None):
        _check(lib.pa_block_spmv(self.mesh.handle, self.bi, self.bj, self.values.ptr, self.precision, x.ptr, beta,
                                 y.ptr, _sh(stream)))
        return y


def _prec(p):
    if p in (FP64, "fp64", "f64", np.float64):
        return FP64, np.float64
    if p in (FP32, "fp32", "f32", np.float32):
        return FP32, np.float32
    raise PlatoError(f"unknown precision {p!r}")


def _dev(x, dtype=np.float64):
    return x if isinstance(x, DeviceArray) else DeviceArray.from_numpy(np.asarray(x, dtype), dtype)


class VectorFunction:
    """Elliptic::VectorFunction<Mechanics<3> | Thermal<3>> on one domain."""

    def __init__(self, mesh, physics, *, youngs=1.0, poisson=0.3, conductivity=1.0, penalty_exponent=3.0,
                 penalty_min=1e-9):
        self.mesh = mesh
        ph = {"Mechanics": MECHANICS, "Thermal": THERMAL, MECHANICS: MECHANICS, THERMAL: THERMAL}[physics]
        mat = np.array([youngs, poisson] if ph == MECHANICS else [conductivity, 0.0], np.float64)
        pen = np.array([penalty_exponent, penalty_min], np.float64)
        h = ctypes.c_void_p()
        _check(lib.pa_vector_function_create(mesh.handle, ph, _np_ptr(mat), _np_ptr(pen), ctypes.byref(h)))
        self.handle = h
        self.physics = ph
        self.dpn = 3 if ph == MECHANICS else 1

    def size(self):
        return self.mesh.num_nodes * self.dpn

    def add_uniform_natural_bc(self, side_set, values):
        v = np.zeros(3, np.float64)
        v[: len(values)] = values
        _check(lib.pa_vector_function_add_uniform_natural_bc(self.handle, side_set.encode(), _np_ptr(v)))

    def add_body_load(self, dof, value):
        _check(lib.pa_vector_function_add_body_load(self.handle, dof, value))

    def add_body_load_function(self, dof, function):
        """Body load 'Function' of x, y, z on any dof, evaluated at the cell
        centroids (src/BodyLoads.hpp:54-122)."""
        _check(lib.pa_vector_function_add_body_load_function(self.handle, dof, function.encode()))

    def value(self, u, z, out=None, stream=None):
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.size())
        _check(lib.pa_vector_function_value(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def gradient_u(self, u, z, precision=FP64, out=None, stream=None):
        p, dt = _prec(precision)
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_blocks * self.dpn * self.dpn, dt)
        _check(lib.pa_vector_function_gradient_u(self.handle, u.ptr, z.ptr, out.ptr, p, _sh(stream)))
        return CrsMatrix(self.mesh, self.dpn, self.dpn, out, p)

    def gradient_u_T(self, u, z, precision=FP64, out=None, stream=None):
        p, dt = _prec(precision)
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_blocks * self.dpn * self.dpn, dt)
        _check(lib.pa_vector_function_gradient_u_T(self.handle, u.ptr, z.ptr, out.ptr, p, _sh(stream)))
        return CrsMatrix(self.mesh, self.dpn, self.dpn, out, p)

    def value_gradient_u(self, u, z, precision=FP64, r_out=None, out=None, stream=None):
        p, dt = _prec(precision)
        u, z = _dev(u), _dev(z)
        r_out = r_out if r_out is not None else DeviceArray(self.size())
        out = out if out is not None else DeviceArray(self.mesh.num_blocks * self.dpn * self.dpn, dt)
        _check(lib.pa_vector_function_value_gradient_u(self.handle, u.ptr, z.ptr, r_out.ptr, out.ptr, p,
                                                       _sh(stream)))
        return r_out, CrsMatrix(self.mesh, self.dpn, self.dpn, out, p)

    def gradient_z(self, u, z, out=None, stream=None):
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_blocks * self.dpn)
        _check(lib.pa_vector_function_gradient_z(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return CrsMatrix(self.mesh, 1, self.dpn, out, FP64)

    def gradient_x(self, u, z, out=None, stream=None):
        """(dR/dx)^T on the node graph, 3 x dpn blocks: block (a, b)[i][j] = dR_{b,j}/dx_{a,i}
        (src/elliptic/VectorFunction.hpp:337-420)."""
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_blocks * 3 * self.dpn)
        _check(lib.pa_vector_function_gradient_x(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return CrsMatrix(self.mesh, 3, self.dpn, out, FP64)

    def __del__(self):
        try:
            lib.pa_vector_function_destroy(self.handle)
        except Exception:
            pass


class PhysicsScalarFunction:
    """Elliptic::PhysicsScalarFunction with an internal-energy criterion."""

    def __init__(self, mesh, function_type, *, youngs=1.0, poisson=0.3, conductivity=1.0, penalty_exponent=3.0,
                 penalty_min=1e-9):
        self.mesh = mesh
        t = {"Internal Elastic Energy": INTERNAL_ELASTIC_ENERGY,
             "Internal Thermal Energy": INTERNAL_THERMAL_ENERGY}[function_type]
        mat = np.array([youngs, poisson] if t == INTERNAL_ELASTIC_ENERGY else [conductivity, 0.0], np.float64)
        pen = np.array([penalty_exponent, penalty_min], np.float64)
        h = ctypes.c_void_p()
        _check(lib.pa_scalar_function_create(mesh.handle, t, _np_ptr(mat), _np_ptr(pen), ctypes.byref(h)))
        self.handle = h
        self.dpn = 3 if t == INTERNAL_ELASTIC_ENERGY else 1

    def value(self, u, z, out=None, stream=None):
        u, z = _dev(u), _dev(z)
        dev_out = out if out is not None else DeviceArray(1)
        _check(lib.pa_scalar_function_value(self.handle, u.ptr, z.ptr, dev_out.ptr, _sh(stream)))
        if out is not None:
            return out
        return float(dev_out.numpy()[0])

    def gradient_u(self, u, z, out=None, stream=None):
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_nodes * self.dpn)
        _check(lib.pa_scalar_function_gradient_u(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def gradient_z(self, u, z, out=None, stream=None):
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_nodes)
        _check(lib.pa_scalar_function_gradient_z(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def gradient_x(self, u, z, out=None, stream=None):
        """df/dx, N*3 doubles (src/elliptic/PhysicsScalarFunction.hpp:293-355)."""
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.mesh.num_nodes * 3)
        _check(lib.pa_scalar_function_gradient_x(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def __del__(self):
        try:
            lib.pa_scalar_function_destroy(self.handle)
        except Exception:
            pass


def apply_block_constraints(matrix, rhs, bc_dofs, bc_values, scale=1.0, stream=None):
    """Symmetric Dirichlet elimination on `matrix` (CrsMatrix) and `rhs`."""
    dofs = _dev(np.asarray(bc_dofs, np.int32), np.int32) if not isinstance(bc_dofs, DeviceArray) else bc_dofs
    vals = _dev(bc_values)
    _check(lib.pa_apply_block_constraints(matrix.mesh.handle, matrix.bi, matrix.values.ptr, matrix.precision, rhs.ptr,
                                          dofs.size, dofs.ptr, vals.ptr, scale, _sh(stream)))


class Solver:
    """PCG on block-CRS matrices of a mesh graph (FP64 vectors/reductions)."""

    def __init__(self, mesh, dpn, precond=PRECOND_BLOCK_JACOBI):
        h = ctypes.c_void_p()
        _check(lib.pa_solver_create(mesh.handle, dpn, precond, ctypes.byref(h)))
        self.handle, self.mesh, self.dpn = h, mesh, dpn

    def set_multigrid(self, mg):
        """Precondition with a Multigrid (set up by the caller; kept alive here)."""
        _check(lib.pa_solver_set_multigrid(self.handle, mg.handle))
        self._mg = mg

    def solve(self, matrix, b, x, tol=1e-8, max_iter=10000, check_every=10, stream=None):
        st = SolveStats()
        _check(lib.pa_solver_pcg(self.handle, matrix.values.ptr, matrix.precision, b.ptr, x.ptr, tol, max_iter,
                                 check_every, ctypes.byref(st), _sh(stream)))
        return st.as_dict()

    def __del__(self):
        try:
            lib.pa_solver_destroy(self.handle)
        except Exception:
            pass


class Problem:
    """Elliptic::Problem: solution, criterionValue, criterionGradient
    (src/elliptic/Problem.hpp:237-450), self-adjoint criterion."""

    def __init__(self, vector_function, criterion, bc_dofs, bc_values=None, *, mode=FP64,
                 precond=PRECOND_BLOCK_JACOBI, tol=1e-8, max_iter=20000, replace_every=200, check_every=10):
        self.vf, self.sf = vector_function, criterion
        dofs = np.ascontiguousarray(bc_dofs, np.int32)
        vals = np.zeros(dofs.size) if bc_values is None else np.ascontiguousarray(bc_values, np.float64)
        self.options = ProblemOptions(_prec(mode)[0], precond, tol, max_iter, replace_every, check_every)
        h = ctypes.c_void_p()
        _check(lib.pa_problem_create(self.vf.handle, self.sf.handle, dofs.size, _np_ptr(dofs), _np_ptr(vals),
                                     ctypes.byref(self.options), ctypes.byref(h)))
        self.handle = h
        self.last_stats = None

    def set_options(self, **kw):
        for k, v in kw.items():
            setattr(self.options, k, _prec(v)[0] if k == "mode" else v)
        _check(lib.pa_problem_set_options(self.handle, ctypes.byref(self.options)))

    def set_multigrid(self, nx, ny, nz, *, max_levels=0, pre_sweeps=2, post_sweeps=2, omega=0.6,
                      max_coarse_nodes=1000, smoother="default", coarse_sweeps=0, coarse_max_nodes=0):
        """Geometric multigrid preconditioner on the nx x ny x nz box mesh."""
        opts = MultigridOptions(max_levels, pre_sweeps, post_sweeps, omega, max_coarse_nodes, 0, SMOOTHERS[smoother],
                                coarse_sweeps, coarse_max_nodes)
        _check(lib.pa_problem_set_multigrid(self.handle, nx, ny, nz, ctypes.byref(opts)))
        self.options.precond = PRECOND_MULTIGRID

    def solution(self, z, out=None, stream=None):
        z = _dev(z)
        out = out if out is not None else DeviceArray(self.vf.size())
        st = SolveStats()
        _check(lib.pa_problem_solution(self.handle, z.ptr, out.ptr, ctypes.byref(st), _sh(stream)))
        self.last_stats = st.as_dict()
        return out

    def criterion_value(self, z, u, out=None, stream=None):
        z, u = _dev(z), _dev(u)
        dev = out if out is not None else DeviceArray(1)
        _check(lib.pa_problem_criterion_value(self.handle, z.ptr, u.ptr, dev.ptr, _sh(stream)))
        return dev if out is not None else float(dev.numpy()[0])

    def criterion_gradient(self, z, u, out=None, stream=None):
        z, u = _dev(z), _dev(u)
        out = out if out is not None else DeviceArray(self.vf.mesh.num_nodes)
        _check(lib.pa_problem_criterion_gradient(self.handle, z.ptr, u.ptr, out.ptr, _sh(stream)))
        return out

    def criterion_gradient_x(self, z, u, out=None, stream=None):
        """criterionGradientX of a self-adjoint problem: -df/dx, N*3 doubles
        (src/elliptic/Problem.hpp:492-542)."""
        z, u = _dev(z), _dev(u)
        out = out if out is not None else DeviceArray(self.vf.mesh.num_nodes * 3)
        _check(lib.pa_problem_criterion_gradient_x(self.handle, z.ptr, u.ptr, out.ptr, _sh(stream)))
        return out

    def __del__(self):
        try:
            lib.pa_problem_destroy(self.handle)
        except Exception:
            pass


class Multigrid:
    """Geometric multigrid preconditioner on a box mesh (FP32 V-cycles)."""

    def __init__(self, mesh, nx, ny, nz, dpn, *, max_levels=0, pre_sweeps=2, post_sweeps=2, omega=0.6,
                 max_coarse_nodes=1000, symmetric=False, smoother="default", coarse_sweeps=0, coarse_max_nodes=0):
        opts = MultigridOptions(max_levels, pre_sweeps, post_sweeps, omega, max_coarse_nodes, int(symmetric),
                                SMOOTHERS[smoother], coarse_sweeps, coarse_max_nodes)
        h = ctypes.c_void_p()
        _check(lib.pa_multigrid_create(mesh.handle, nx, ny, nz, dpn, ctypes.byref(opts), ctypes.byref(h)))
        self.handle, self.mesh, self.dpn = h, mesh, dpn

    def levels(self):
        L, c = ctypes.c_int(), ctypes.c_int()
        _check(lib.pa_multigrid_levels(self.handle, ctypes.byref(L), ctypes.byref(c)))
        return L.value, c.value

    def setup(self, matrix, bc_dofs=None, stream=None):
        dofs = DeviceArray.from_numpy(np.asarray(bc_dofs if bc_dofs is not None else [], np.int32), np.int32)
        _check(lib.pa_multigrid_setup(self.handle, matrix.values.ptr, matrix.precision, dofs.size, dofs.ptr,
                                      _sh(stream)))
        self._dofs = dofs

    def apply(self, r, z, stream=None):
        _check(lib.pa_multigrid_apply(self.handle, r.ptr, z.ptr, _sh(stream)))
        return z

    def level_operator(self, level):
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.pa_multigrid_level_operator(self.handle, level, None, ctypes.byref(nx), ctypes.byref(ny),
                                               ctypes.byref(nz)))
        n = (nx.value + 1) * (ny.value + 1) * (nz.value + 1)
        out = np.empty(n * 27 * self.dpn * self.dpn, np.float32)
        _check(lib.pa_multigrid_level_operator(self.handle, level, _np_ptr(out), None, None, None))
        return (nx.value, ny.value, nz.value), out.reshape(n, 27, self.dpn, self.dpn)

    def __del__(self):
        try:
            lib.pa_multigrid_destroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# J2 plasticity
# ---------------------------------------------------------------------------
def j2_params(**kw):
    """pa_j2_params with the reference's defaults, overridden by keywords.
    formulation: "mixed" (the reference's stabilized u-p form, 4 dofs per node)
    or "displacement" (pure-displacement J2, 3 dofs per node)."""
    p = J2Params()
    lib.pa_j2_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k == "formulation" and isinstance(v, str):
            if v not in ("mixed", "displacement"):
                raise PlatoError(f"unknown J2 formulation {v!r}")
            v = J2_MIXED if v == "mixed" else J2_DISPLACEMENT
        if not hasattr(p, k):
            raise PlatoError(f"unknown J2 parameter {k!r}")
        setattr(p, k, v)
    return p


class Plasticity:
    """Stabilized mixed u-p J2 plasticity on one mesh: the reference's
    GlobalVectorFunctionInc (src/GlobalVectorFunctionInc.hpp) and
    LocalVectorFunctionInc (src/LocalVectorFunctionInc.hpp) for
    Physics 'Plasticity', plus the projection and the plastic-work criterion.

    u: 4 dofs per node (ux, uy, uz, p), or 3 (ux, uy, uz) with
    formulation="displacement"; c: 14 local states per cell; pi: 3 per node.
    Worksets come back as (C, rows, W) numpy arrays with the local dof
    node*4 + d in either formulation."""

    def __init__(self, mesh, **params):
        self.mesh = mesh
        self.params = j2_params(**params)
        h = ctypes.c_void_p()
        _check(lib.pa_plasticity_create(mesh.handle, ctypes.byref(self.params), ctypes.byref(h)))
        self.handle = h
        self.N, self.C = mesh.num_nodes, mesh.num_cells
        self.dpn = 4 if self.params.formulation == J2_MIXED else 3

    # LocalVectorFunctionInc
    def update_local_state(self, u, c_prev, z, out=None, stream=None):
        u, cp, z = _dev(u), _dev(c_prev), _dev(z)
        out = out if out is not None else DeviceArray(self.C * 14)
        _check(lib.pa_j2_update_local_state(self.handle, u.ptr, cp.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def local_value(self, u, c, c_prev, z, out=None, stream=None):
        u, c, cp, z = _dev(u), _dev(c), _dev(c_prev), _dev(z)
        out = out if out is not None else DeviceArray(self.C * 14)
        _check(lib.pa_j2_local_value(self.handle, u.ptr, c.ptr, cp.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def local_gradient(self, wrt, u, c, c_prev, z, stream=None):
        W = {WRT_C: 14, WRT_U: 16, WRT_Z: 4, WRT_C_PREV: 14}[wrt]
        u, c, cp, z = _dev(u), _dev(c), _dev(c_prev), _dev(z)
        out = DeviceArray(self.C * 14 * W)
        _check(lib.pa_j2_local_gradient(self.handle, wrt, u.ptr, c.ptr, cp.ptr, z.ptr, out.ptr, _sh(stream)))
        return out.numpy().reshape(self.C, 14, W)

    # GlobalVectorFunctionInc
    def global_value(self, u, c, pi, z, out=None, stream=None):
        u, c, pi, z = _dev(u), _dev(c), _dev(pi), _dev(z)
        out = out if out is not None else DeviceArray(self.N * self.dpn)
        _check(lib.pa_j2_global_value(self.handle, u.ptr, c.ptr, pi.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def global_gradient(self, wrt, u, c, pi, z, stream=None):
        W = {WRT_U: 16, WRT_C: 14, WRT_Z: 4, WRT_PI: 12}[wrt]
        u, c, pi, z = _dev(u), _dev(c), _dev(pi), _dev(z)
        out = DeviceArray(self.C * 16 * W)
        _check(lib.pa_j2_global_gradient(self.handle, wrt, u.ptr, c.ptr, pi.ptr, z.ptr, out.ptr, _sh(stream)))
        return out.numpy().reshape(self.C, 16, W)

    def condensed_system(self, u, c, c_prev, pi, z, transpose=False, values=None, R=None, stream=None):
        """(CrsMatrix dpn x dpn of K_t or K_t^T, DeviceArray R_t or None)."""
        u, c, cp, pi, z = _dev(u), _dev(c), _dev(c_prev), _dev(pi), _dev(z)
        D = self.dpn
        values = values if values is not None else DeviceArray(self.mesh.num_blocks * D * D)
        R = None if transpose else (R if R is not None else DeviceArray(self.N * D))
        _check(lib.pa_j2_condensed_system(self.handle, u.ptr, c.ptr, cp.ptr, pi.ptr, z.ptr, values.ptr,
                                          R.ptr if R is not None else None, 1 if transpose else 0, _sh(stream)))
        return CrsMatrix(self.mesh, D, D, values, FP64), R

    def projection(self, u, z, out=None, stream=None):
        u, z = _dev(u), _dev(z)
        out = out if out is not None else DeviceArray(self.N * 3)
        _check(lib.pa_j2_projection(self.handle, u.ptr, z.ptr, out.ptr, _sh(stream)))
        return out

    def plastic_work(self, u, u_prev, c, c_prev, z, stream=None):
        u, up, c, cp, z = _dev(u), _dev(u_prev), _dev(c), _dev(c_prev), _dev(z)
        out = DeviceArray(1)
        _check(lib.pa_j2_plastic_work(self.handle, u.ptr, up.ptr, c.ptr, cp.ptr, z.ptr, out.ptr, _sh(stream)))
        return float(out.numpy()[0])

    def __del__(self):
        try:
            lib.pa_plasticity_destroy(self.handle)
        except Exception:
            pass


class PlasticityProblem:
    """PlasticityProblem (src/PlasticityProblem.hpp): pseudo-time steps with the
    condensed Newton solve, the plastic-work criterion and its path-dependent
    adjoint gradient.  Dirichlet values are the end values of the load ramp."""

    def __init__(self, plasticity, bc_dofs, bc_values=None, *, num_steps=20, max_newton=10, newton_stop_tol=1e-8,
                 newton_current_norm_tol=1e-8, linear_max_iter=20000, linear_tol=1e-12, krylov=KRYLOV_GMRES,
                 restart=50, checkpoint_stride=1):
        self.pl = plasticity
        dofs = np.ascontiguousarray(bc_dofs, np.int32)
        vals = np.zeros(dofs.size) if bc_values is None else np.ascontiguousarray(bc_values, np.float64)
        self.options = PlasticityOptions(num_steps, max_newton, newton_stop_tol, newton_current_norm_tol,
                                         linear_max_iter, linear_tol, krylov, restart, checkpoint_stride)
        h = ctypes.c_void_p()
        _check(lib.pa_plasticity_problem_create(plasticity.handle, dofs.size, _np_ptr(dofs), _np_ptr(vals),
                                                ctypes.byref(self.options), ctypes.byref(h)))
        self.handle = h
        self.num_steps = num_steps

    def set_multigrid(self, nx, ny, nz, *, max_levels=0, pre_sweeps=2, post_sweeps=2, omega=0.6,
                      max_coarse_nodes=1000, smoother="default", coarse_sweeps=0, coarse_max_nodes=0):
        """Geometric multigrid (dpn x dpn nodal blocks) as the Krylov preconditioner."""
        opts = MultigridOptions(max_levels, pre_sweeps, post_sweeps, omega, max_coarse_nodes, 0, SMOOTHERS[smoother],
                                coarse_sweeps, coarse_max_nodes)
        _check(lib.pa_plasticity_problem_set_multigrid(self.handle, nx, ny, nz, ctypes.byref(opts)))

    def state_bytes(self):
        """Device bytes held for the step states (u, c, pi)."""
        b = ctypes.c_size_t()
        _check(lib.pa_plasticity_problem_state_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def solution(self, z, stream=None):
        z = _dev(z)
        conv = ctypes.c_int()
        _check(lib.pa_plasticity_problem_solution(self.handle, z.ptr, ctypes.byref(conv), _sh(stream)))
        return bool(conv.value)

    def state(self, step):
        """(u, c, pi) of a step as numpy arrays."""
        pu, pc, pp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib.pa_plasticity_problem_state(self.handle, step, ctypes.byref(pu), ctypes.byref(pc),
                                               ctypes.byref(pp)))
        out = []
        for ptr, n in ((pu, self.pl.N * self.pl.dpn), (pc, self.pl.C * 14), (pp, self.pl.N * 3)):
            a = np.empty(n)
            _check(lib.pa_memcpy_d2h(_np_ptr(a), ptr, a.nbytes))
            out.append(a)
        return tuple(out)

    def stats(self, step):
        n, l, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib.pa_plasticity_problem_stats(self.handle, step, ctypes.byref(n), ctypes.byref(l),
                                               ctypes.byref(r)))
        return {"newton_iterations": n.value, "linear_iterations": l.value, "residual_norm": r.value}

    def criterion_value(self, z, multiplier=1.0, stream=None):
        z = _dev(z)
        v = ctypes.c_double()
        _check(lib.pa_plasticity_problem_criterion_value(self.handle, z.ptr, multiplier, ctypes.byref(v),
                                                         _sh(stream)))
        return v.value

    def criterion_gradient(self, z, multiplier=1.0, out=None, stream=None):
        z = _dev(z)
        out = out if out is not None else DeviceArray(self.pl.N)
        _check(lib.pa_plasticity_problem_criterion_gradient(self.handle, z.ptr, multiplier, out.ptr, _sh(stream)))
        return out

    def __del__(self):
        try:
            lib.pa_plasticity_problem_destroy(self.handle)
        except Exception:
            pass
