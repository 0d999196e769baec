"""Configuration 3 at full size: steady thermal on the 192x192x192 hex grid x6
Tet4 (42,467,328 cells, 7,189,057 nodes, 106,944,193 blocks), k = 1, SIMP
p = 3, z ~ U[0.1, 1] (seed 3), body source 1: the assembly the benchmark times.

* Oracle parity on three 4-hex x-slabs of the grid (both x boundaries and the
  middle), each its own mesh with the full grid's coordinates, densities and
  temperatures: the full grid's rows of the slab's complete node layers equal
  the oracle's (value and gradient_u FP64 at 1e-12, FP32 at 2e-6).  The
  reference computes the same residual in src/elliptic/ThermostaticResidual.hpp:128-223.
* Size-independent properties of the full operator: K bitwise equal to K^T;
  R(T) - R(0) = K T at 1e-11; constant temperatures in the null space of K
  (row sums zero); the FP32 values equal the FP64 ones to single rounding.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DIMS = (192, 192, 192)
TOL64, TOL32 = 1e-12, 2e-6


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def full():
    import platoanalyze_amd as pa
    from platoanalyze_amd.workloads import density
    assert pa.device_count() >= 1
    print("case: configuration 3, thermal 192^3, mesh + assembly", flush=True)
    mesh = pa.Mesh.box(*DIMS, 1.0, 1.0, 1.0, side_sets=False)
    N, C, nnzb = mesh.num_nodes, mesh.num_cells, mesh.num_blocks
    assert (C, N, nnzb) == (42467328, 7189057, 106944193)
    assert mesh.grid_info() == (193, 193, 193)
    z = density(DIMS, ("uniform", 0.1, 3))
    T = np.random.default_rng(13).standard_normal(N)
    vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_body_load(0, 1.0)
    zd, Td = pa.DeviceArray.from_numpy(z), pa.DeviceArray.from_numpy(T)
    R = vf.value(Td, zd)
    K = vf.gradient_u(Td, zd)
    K32 = vf.gradient_u(Td, zd, precision="fp32")
    pa.synchronize()
    return dict(pa=pa, mesh=mesh, vf=vf, z=z, T=T, zd=zd, Td=Td, R=R, K=K, K32=K32, Kh=K.entries(),
                K32h=K32.entries(), graph=mesh.graph())


@pytest.mark.parametrize("i0", [0, 94, 188])
def test_config3_slab_parity(full, i0):
    """Rows of node layers of a 4-hex slab starting at hex layer i0 that the
    slab holds completely, against the oracle on the slab mesh."""
    print(f"case: configuration 3 slab at x layer {i0} vs oracle", flush=True)
    mesh = full["mesh"]
    nx, ny, nz = DIMS
    h = 1.0 / nx
    _, conn = oracle.box_mesh(4, ny, nz, 4 * h, 1.0, 1.0)
    S = (ny + 1) * (nz + 1)
    sl = slice(i0 * S, (i0 + 5) * S)
    X = mesh.coords_host.reshape(-1, 3)[sl]  # the full grid's own coordinates
    z, T = full["z"][sl], full["T"][sl]
    om = oracle.Model(X.ravel(), conn, conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9, body=(0, 1.0))
    # complete layers: all five at the x boundaries' own faces, else the three inner ones
    lo = 0 if i0 == 0 else 1
    hi = 5 if i0 + 4 == nx else 4
    rows_s = np.arange(lo * S, hi * S)
    rows_f = rows_s + i0 * S
    Rs = om.value("thermal", T, z)
    Rf = full["R"].numpy()
    assert rel(Rf[rows_f], Rs[rows_s]) < TOL64
    Ks = om.gradient_u("thermal", T, z)
    rm_f, ci_f = full["graph"]
    a_f, b_f = rm_f[rows_f[0]], rm_f[rows_f[-1] + 1]
    a_s, b_s = om.rowmap[rows_s[0]], om.rowmap[rows_s[-1] + 1]
    assert b_f - a_f == b_s - a_s
    assert np.array_equal(om.colidx[a_s:b_s] + i0 * S, ci_f[a_f:b_f])
    assert rel(full["Kh"][a_f:b_f], Ks[a_s:b_s]) < TOL64
    assert rel(full["K32h"][a_f:b_f], Ks[a_s:b_s]) < TOL32


def test_config3_symmetry_and_fp32(full):
    print("case: configuration 3 K == K^T bitwise, FP32 vs FP64", flush=True)
    import scipy.sparse as sp
    rowmap, colidx = full["graph"]
    vals = full["Kh"]
    N = full["mesh"].num_nodes
    A = sp.csr_matrix((vals, colidx, rowmap), shape=(N, N))
    AT = A.T.tocsr()
    AT.sort_indices()
    assert np.array_equal(AT.indptr, rowmap) and np.array_equal(AT.indices, colidx)
    assert np.array_equal(AT.data, vals)
    del A, AT
    assert rel(full["K32h"], vals) < TOL32


def test_config3_linearity_and_null_space(full):
    print("case: configuration 3 R(T) - R(0) = K T, K 1 = 0", flush=True)
    pa, vf, K = full["pa"], full["vf"], full["K"]
    N = full["mesh"].num_nodes
    R0 = vf.value(np.zeros(N), full["zd"]).numpy()
    KT = pa.DeviceArray(N)
    K.spmv(full["Td"], KT)
    assert rel(full["R"].numpy() - R0, KT.numpy()) < 1e-11
    ones = pa.DeviceArray.from_numpy(np.ones(N))
    K1 = pa.DeviceArray(N)
    K.spmv(ones, K1)
    assert np.max(np.abs(K1.numpy())) < 1e-12 * np.max(np.abs(full["Kh"]))
    # the body source integrates to -f * (mean-density weighted volume)
    Rb = R0.sum()
    assert Rb < 0 and abs(Rb + np.mean(full["z"])) < 0.05
