"""GPU parity of the elliptic assembly path (through the C ABI) against the
oracle and the reference's golden vectors.

Tolerances: FP64 mode 1e-12 relative (max-norm error / max-norm of the
oracle value); FP32 values 2e-6 relative (single rounding of FP64-accumulated
entries).  Graph (rowmap/colidx) bit-exact.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "derivative_tests.json").read_text())["tests"]
TOL64 = 1e-12
TOL32 = 2e-6


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def gold(test, key):
    return np.asarray(GOLD[test]["gold"][key]["values"], float)


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


# ------------------------------------------------------------ goldens on GPU
def test_goldens_through_abi(pa):
    print("case: DerivativeTests goldens, 2x2x2 box", flush=True)
    mesh = pa.Mesh.box(2, 2, 2)
    N = mesh.num_nodes
    u = 1e-4 * np.arange(1, 3 * N + 1)
    T = 0.1 * np.arange(1, N + 1)
    z = np.ones(N)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1e6, poisson=0.3, penalty_exponent=1.0, penalty_min=0.0)
    assert rel(vf.value(u, z).numpy(), gold("ElastostaticResidual3D", "residual")) < 1e-13
    assert rel(vf.gradient_u(u, z).entries()[:36], gold("ElastostaticResidual3D", "jacobian_head")) < 1e-15
    assert rel(vf.gradient_z(u, z).entries()[:12], gold("ElastostaticResidual3D", "gradient_z_head")) < 2e-14
    sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1e6, poisson=0.3, penalty_exponent=1.0,
                                  penalty_min=0.0)
    assert abs(sf.value(u, z) - 48.15) / 48.15 < 1e-13
    assert rel(sf.gradient_u(u, z).numpy()[:21], gold("InternalElasticEnergy3D", "gradient_u")) < 1e-13
    assert rel(sf.gradient_z(u, z).numpy(), gold("InternalElasticEnergy3D", "gradient_z")) < 1e-13
    tf = pa.VectorFunction(mesh, "Thermal", conductivity=100.0, penalty_exponent=1.0, penalty_min=0.0)
    assert rel(tf.value(T, z).numpy(), gold("ThermostaticResidual3D", "residual")) < 1e-13
    assert rel(tf.gradient_u(T, z).entries()[:22], gold("ThermostaticResidual3D", "jacobian_head")) < 1e-15
    assert rel(tf.gradient_z(T, z).entries()[:22], gold("ThermostaticResidual3D", "gradient_z_head")) < 1e-14
    ts = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=100.0, penalty_exponent=1.0,
                                  penalty_min=0.0)
    assert abs(ts.value(T, z) - float(gold("InternalThermalEnergy3D", "value"))) < 1e-11
    assert rel(ts.gradient_u(T, z).numpy(), gold("InternalThermalEnergy3D", "gradient_u")) < 1e-13
    assert rel(ts.gradient_z(T, z).numpy(), gold("InternalThermalEnergy3D", "gradient_z")) < 1e-13


# ------------------------------------------------- config-1 parity vs oracle
def _cfg(pa, dims, L, seed, physics):
    nx, ny, nz = dims
    coords, conn = pa.box_mesh(nx, ny, nz, *L)
    mesh = pa.Mesh.box(nx, ny, nz, *L)
    N = mesh.num_nodes
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.3, 1.0, N)
    fc, fl = oracle.side_set(coords, conn, 0, L[0])
    if physics == "elastic":
        u = rng.standard_normal(3 * N) * 1e-2
        vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
        om = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9,
                          face_cells=fc, face_lnodes=fl, flux=(0.0, 0.0, -1.0))
    else:
        u = rng.standard_normal(N)
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.5,))
        vf.add_body_load(0, 1.0)
        om = oracle.Model(coords, conn, conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9, face_cells=fc,
                          face_lnodes=fl, flux=(0.5, 0.0, 0.0), body=(0, 1.0))
    return mesh, vf, om, u, z


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_config1_parity(pa, physics):
    print(f"case: 20x10x10 box, {physics}, seed 1", flush=True)
    mesh, vf, om, u, z = _cfg(pa, (20, 10, 10), (2.0, 1.0, 1.0), 1, physics)
    rowmap, colidx = mesh.graph()
    assert np.array_equal(rowmap, om.rowmap) and np.array_equal(colidx, om.colidx)
    R = vf.value(u, z).numpy()
    Rref = om.value(physics, u, z)
    assert rel(R, Rref) < TOL64
    Kref = om.gradient_u(physics, u, z)
    assert rel(vf.gradient_u(u, z).entries(), Kref) < TOL64
    assert rel(vf.gradient_u(u, z, precision="fp32").entries(), Kref) < TOL32
    assert rel(vf.gradient_u_T(u, z).entries(), om.gradient_u_T(physics, u, z)) < TOL64
    R2, K2 = vf.value_gradient_u(u, z)
    assert rel(R2.numpy(), Rref) < TOL64 and rel(K2.entries(), Kref) < TOL64
    assert rel(vf.gradient_z(u, z).entries(), om.gradient_z(physics, u, z)) < TOL64
    kind = "Internal Elastic Energy" if physics == "elastic" else "Internal Thermal Energy"
    mat = dict(youngs=1.0, poisson=0.3) if physics == "elastic" else dict(conductivity=1.0)
    sf = pa.PhysicsScalarFunction(mesh, kind, penalty_exponent=3.0, penalty_min=1e-9, **mat)
    assert abs(sf.value(u, z) - om.energy_value(physics, u, z)) < TOL64 * abs(om.energy_value(physics, u, z))
    assert rel(sf.gradient_u(u, z).numpy(), om.energy_gradient_u(physics, u, z)) < TOL64
    assert rel(sf.gradient_z(u, z).numpy(), om.energy_gradient_z(physics, u, z)) < TOL64


def test_deterministic_repeat(pa):
    mesh, vf, om, u, z = _cfg(pa, (8, 4, 4), (2.0, 1.0, 1.0), 3, "elastic")
    a = vf.gradient_u(u, z).entries()
    b = vf.gradient_u(u, z).entries()
    assert np.array_equal(a, b)


def test_constraints_and_spmv_vs_oracle(pa):
    print("case: constraints + spmv, 6x3x3 elastic", flush=True)
    mesh, vf, om, u, z = _cfg(pa, (6, 3, 3), (2.0, 1.0, 1.0), 5, "elastic")
    K = vf.gradient_u(u, z)
    rhs = pa.DeviceArray.from_numpy(np.arange(3 * mesh.num_nodes, dtype=float))
    dofs = np.concatenate([3 * mesh.node_set("x-")[:, None] + np.arange(3)]).ravel().astype(np.int32)
    g = np.linspace(-0.1, 0.2, dofs.size)
    pa.apply_block_constraints(K, rhs, dofs, g)
    vals = om.gradient_u("elastic", u, z)
    r0 = np.arange(3 * mesh.num_nodes, dtype=float)
    oracle.apply_block_constraints(3, om.rowmap, om.colidx, vals, r0, dofs, g)
    assert rel(K.entries(), vals) < TOL64
    assert rel(rhs.numpy(), r0) < TOL64
    x = np.random.default_rng(2).standard_normal(3 * mesh.num_nodes)
    y = pa.DeviceArray(3 * mesh.num_nodes)
    K.spmv(pa.DeviceArray.from_numpy(x), y)
    assert rel(y.numpy(), oracle.block_spmv(3, 3, om.rowmap, om.colidx, vals, x)) < 1e-14
    G = vf.gradient_z(u, z)
    lam = np.random.default_rng(3).standard_normal(3 * mesh.num_nodes)
    y1 = pa.DeviceArray(mesh.num_nodes)
    G.spmv(pa.DeviceArray.from_numpy(lam), y1)
    assert rel(y1.numpy(), oracle.block_spmv(1, 3, om.rowmap, om.colidx, om.gradient_z("elastic", u, z), lam)) < 1e-13


def test_constraints_fixed_order(pa):
    """Both x faces constrained (hundreds of constrained dofs next to shared
    free rows): the result matches the oracle and is the same bit for bit for
    a repeat and for the constraint list in another order."""
    print("case: constraints, 16x8x8 elastic, both x faces, permuted list", flush=True)
    mesh, vf, om, u, z = _cfg(pa, (16, 8, 8), (2.0, 1.0, 1.0), 7, "elastic")
    xm, xp = mesh.node_set("x-"), mesh.node_set("x+")
    dofs = np.concatenate([3 * xm[:, None] + np.arange(3), 3 * xp[:, None] + np.arange(3)]).ravel().astype(np.int32)
    g = np.random.default_rng(8).uniform(-0.2, 0.2, dofs.size)
    r_init = np.random.default_rng(9).standard_normal(3 * mesh.num_nodes)
    runs = []
    for perm in (np.arange(dofs.size), np.arange(dofs.size), np.random.default_rng(10).permutation(dofs.size)):
        K = vf.gradient_u(u, z)
        rhs = pa.DeviceArray.from_numpy(r_init)
        pa.apply_block_constraints(K, rhs, dofs[perm], g[perm])
        runs.append((K.entries(), rhs.numpy()))
    vals = om.gradient_u("elastic", u, z)
    r0 = r_init.copy()
    oracle.apply_block_constraints(3, om.rowmap, om.colidx, vals, r0, dofs, g)
    assert rel(runs[0][0], vals) < TOL64
    assert rel(runs[0][1], r0) < TOL64
    for k, r in runs[1:]:
        assert np.array_equal(k, runs[0][0]) and np.array_equal(r, runs[0][1])


# --------------------------------------- full-size properties (config 2 grid)
def test_config2_full_size_properties(pa):
    print("case: 128x64x64 elastic, size-independent properties", flush=True)
    nx, ny, nz = 128, 64, 64
    mesh = pa.Mesh.box(nx, ny, nz, 2.0, 1.0, 1.0)
    N = mesh.num_nodes
    assert (mesh.num_cells, N, mesh.num_blocks) == (3145728, 545025, 8008449)
    rng = np.random.default_rng(2)
    z = rng.uniform(0.0, 1.0, N)
    u = rng.standard_normal(3 * N)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
    R, K = vf.value_gradient_u(u, z)
    KT = vf.gradient_u_T(u, z)
    # symmetric operator: K^T assembled separately equals K bit for bit
    assert np.array_equal(K.entries(), KT.entries())
    # linearity: R(u) - R(0) == K u
    R0 = vf.value(np.zeros(3 * N), z).numpy()
    Ku = pa.DeviceArray(3 * N)
    K.spmv(pa.DeviceArray.from_numpy(u), Ku)
    assert rel(R.numpy() - R0, Ku.numpy()) < 1e-11
    # traction resultant: sum of R(0) = -t * area = (0, 0, 1)
    assert np.allclose(R0.reshape(-1, 3).sum(0), [0, 0, 1.0], atol=1e-10)
    # rigid translations are in the null space
    for d in range(3):
        t = np.zeros(3 * N)
        t[d::3] = 1.0
        y = pa.DeviceArray(3 * N)
        K.spmv(pa.DeviceArray.from_numpy(t), y)
        assert np.max(np.abs(y.numpy())) < 1e-12 * np.max(np.abs(K.entries()))
    # FP32 assembly agrees with FP64 to single rounding
    K32 = vf.gradient_u(u, z, precision="fp32")
    assert rel(K32.entries(), K.entries()) < TOL32
    # energy = 1/2 u.K u
    sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1.0, poisson=0.3, penalty_exponent=3.0,
                                  penalty_min=1e-9)
    e = sf.value(u, z)
    assert abs(e - 0.5 * u @ Ku.numpy()) < 1e-11 * abs(e)


# ------------------------------------------------- gradient_x (shape derivatives)
def test_gradient_x_goldens_through_abi(pa):
    print("case: gradient_x goldens, 2x2x2 box", flush=True)
    mesh = pa.Mesh.box(2, 2, 2)
    N = mesh.num_nodes
    u = 1e-4 * np.arange(1, 3 * N + 1)
    T = 0.1 * np.arange(1, N + 1)
    z = np.ones(N)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1e6, poisson=0.3, penalty_exponent=1.0, penalty_min=0.0)
    Gx = vf.gradient_x(u, z)
    assert (Gx.bi, Gx.bj) == (3, 3)
    assert rel(Gx.entries()[:36], gold("ElastostaticResidual3D", "gradient_x_head")) < 1e-13
    sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1e6, poisson=0.3, penalty_exponent=1.0,
                                  penalty_min=0.0)
    gx = sf.gradient_x(u, z).numpy()
    ref = gold("InternalElasticEnergy3D", "gradient_x")
    nz = ref != 0.0
    assert rel(gx[:ref.size][nz], ref[nz]) < 1e-13
    assert np.max(np.abs(gx[:ref.size][~nz])) < 1e-13
    tf = pa.VectorFunction(mesh, "Thermal", conductivity=100.0, penalty_exponent=1.0, penalty_min=0.0)
    assert rel(tf.gradient_x(T, z).entries()[:12], gold("ThermostaticResidual3D", "gradient_x_head")) < 1e-13
    ts = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=100.0, penalty_exponent=1.0,
                                  penalty_min=0.0)
    assert rel(ts.gradient_x(T, z).numpy()[:36], gold("InternalThermalEnergy3D", "gradient_x")) < 1e-13


def _skewed_mesh(pa, dims, L, seed):
    coords, conn = pa.box_mesh(*dims, *L)
    X = coords.reshape(-1, 3).copy()
    rng = np.random.default_rng(seed)
    h = min(L[0] / dims[0], L[1] / dims[1], L[2] / dims[2])
    X += rng.uniform(-0.12 * h, 0.12 * h, X.shape)
    X[np.isclose(coords.reshape(-1, 3)[:, 0], L[0]), 0] = L[0]
    coords = X.ravel()
    fc, fl = oracle.side_set(coords, conn, 0, L[0])
    mesh = pa.Mesh(coords, conn)
    mesh.add_side_set("x+", fc, fl)
    return coords, conn, fc, fl, mesh, rng


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_gradient_x_parity_skewed_mesh(pa, physics):
    """gradient_x of the vector function (natural BC and body source included),
    of the criterion, and the problem's criterionGradientX against the oracle
    (config FAD) on a 12x6x6 box with every node moved off the grid."""
    print(f"case: gradient_x, skewed 12x6x6 box, {physics}, seed 3", flush=True)
    coords, conn, fc, fl, mesh, rng = _skewed_mesh(pa, (12, 6, 6), (2.0, 1.0, 1.0), 3)
    N = mesh.num_nodes
    z = rng.uniform(0.3, 1.0, N)
    if physics == "elastic":
        st = rng.standard_normal(3 * N) * 1e-2
        vf = pa.VectorFunction(mesh, "Mechanics", youngs=2.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-3)
        vf.add_uniform_natural_bc("x+", (0.1, 0.2, -1.0))
        om = oracle.Model(coords, conn, youngs=2.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-3,
                          face_cells=fc, face_lnodes=fl, flux=(0.1, 0.2, -1.0))
        sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=2.0, poisson=0.3,
                                      penalty_exponent=3.0, penalty_min=1e-3)
    else:
        st = rng.standard_normal(N)
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.5, penalty_exponent=3.0, penalty_min=1e-3)
        vf.add_uniform_natural_bc("x+", (0.7,))
        vf.add_body_load(0, 1.0)
        om = oracle.Model(coords, conn, conductivity=1.5, penalty_exponent=3.0, penalty_min=1e-3, face_cells=fc,
                          face_lnodes=fl, flux=(0.7, 0.0, 0.0), body=(0, 1.0))
        sf = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=1.5, penalty_exponent=3.0,
                                      penalty_min=1e-3)
    assert rel(vf.gradient_x(st, z).entries(), om.gradient_x(physics, st, z)) < TOL64
    gref = om.energy_gradient_x(physics, st, z)
    assert rel(sf.gradient_x(st, z).numpy(), gref) < TOL64
    # criterionGradientX, self-adjoint: -df/dx exactly
    dofs = np.arange(3 if physics == "elastic" else 1)
    P = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-10, max_iter=10)
    assert np.array_equal(P.criterion_gradient_x(z, st).numpy(), -sf.gradient_x(st, z).numpy())
    # (dR/dx)^T lambda through the 3 x dpn block SpMV against scipy
    Gx = vf.gradient_x(st, z)
    lam = pa.DeviceArray.from_numpy(rng.standard_normal(st.size))
    y = pa.DeviceArray(3 * N)
    Gx.spmv(lam, y)
    dpn = 3 if physics == "elastic" else 1
    Gs = oracle.to_scipy(3, dpn, om.rowmap, om.colidx, om.gradient_x(physics, st, z))
    assert rel(y.numpy(), Gs @ lam.numpy()) < TOL64


def test_gradient_x_config1_parity(pa):
    print("case: gradient_x, 20x10x10 box, elastic and thermal, seed 1", flush=True)
    for physics in ("elastic", "thermal"):
        mesh, vf, om, u, z = _cfg(pa, (20, 10, 10), (2.0, 1.0, 1.0), 1, physics)
        assert rel(vf.gradient_x(u, z).entries(), om.gradient_x(physics, u, z)) < TOL64
        kind = "Internal Elastic Energy" if physics == "elastic" else "Internal Thermal Energy"
        mat = dict(youngs=1.0, poisson=0.3) if physics == "elastic" else dict(conductivity=1.0)
        sf = pa.PhysicsScalarFunction(mesh, kind, penalty_exponent=3.0, penalty_min=1e-9, **mat)
        assert rel(sf.gradient_x(u, z).numpy(), om.energy_gradient_x(physics, u, z)) < TOL64


def test_natural_load_cache_follows_changes(pa):
    """value() on a tile mesh adds the natural BCs' nodal load, kept between
    calls: a load added later, or a side set redefined, must show up."""
    print("case: natural-BC load cache, 8x4x4 box, elastic", flush=True)
    dims, L = (8, 4, 4), (2.0, 1.0, 1.0)
    coords, conn = pa.box_mesh(*dims, *L)
    mesh = pa.Mesh.box(*dims, *L)
    assert mesh.assembly_info()[0] > 0  # tiles
    N = mesh.num_nodes
    rng = np.random.default_rng(5)
    z = rng.uniform(0.3, 1.0, N)
    u = rng.standard_normal(3 * N) * 1e-2
    fc, fl = oracle.side_set(coords, conn, 0, L[0])
    kw = dict(youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)

    def ref(flux, faces=True):
        m = oracle.Model(coords, conn, face_cells=fc if faces else None, face_lnodes=fl if faces else None,
                         flux=flux, **kw)
        return m.value("elastic", u, z)

    vf = pa.VectorFunction(mesh, "Mechanics", **kw)
    assert rel(vf.value(u, z).numpy(), ref((0.0, 0.0, 0.0), faces=False)) < TOL64
    vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
    assert rel(vf.value(u, z).numpy(), ref((0.0, 0.0, -1.0))) < TOL64
    assert rel(vf.value(u, z).numpy(), ref((0.0, 0.0, -1.0))) < TOL64  # cached
    vf.add_uniform_natural_bc("x+", (0.5, 0.0, 0.0))
    assert rel(vf.value(u, z).numpy(), ref((0.5, 0.0, -1.0))) < TOL64
    mesh.add_side_set("x+", np.zeros(0, np.int32), np.zeros(0, np.int32))  # redefined: no faces
    assert rel(vf.value(u, z).numpy(), ref((0.0, 0.0, 0.0), faces=False)) < TOL64
