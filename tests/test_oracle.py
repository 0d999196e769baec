"""Pins the CPU oracle before it is trusted as the checker.

* every golden vector of unit_tests/DerivativeTests.cpp (tests/golden/);
* the FAD Jacobian against an independent numpy B^T C B (np.linalg.inv of the
  cell Jacobian, dense 6x6 C) and against central differences of the residual;
* dR/dz and the criterion gradients against central differences;
* closed forms: patch test, uniform-strain energy, rigid-body null space,
  symmetry, graph size N + 2E of the Kuhn split;
* the Dirichlet elimination invariants.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = json.loads((Path(__file__).parent / "golden" / "derivative_tests.json").read_text())["tests"]


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def gold(test, key):
    return np.asarray(GOLD[test]["gold"][key]["values"], float)


@pytest.fixture(scope="module")
def box2():
    coords, conn = oracle.box_mesh(2, 2, 2)
    N = coords.size // 3
    u = 1e-4 * np.arange(1, 3 * N + 1)
    T = 0.1 * np.arange(1, N + 1)
    z = np.ones(N)
    el = oracle.Model(coords, conn, youngs=1e6, poisson=0.3, penalty_exponent=1.0, penalty_min=0.0)
    th = oracle.Model(coords, conn, conductivity=100.0, penalty_exponent=1.0, penalty_min=0.0)
    return dict(coords=coords, conn=conn, N=N, u=u, T=T, z=z, el=el, th=th)


# ------------------------------------------------------------------ goldens
def test_functor_chain_goldens(box2):
    g, e, s, r = box2["el"].cell_chain(box2["u"])
    assert rel(g[:6], gold("3D", "gradient")) < 1e-13
    assert rel(e[:6], gold("3D", "strain")) < 1e-13
    assert rel(s[:6], gold("3D", "stress")) < 1e-13
    assert rel(r[:2], gold("3D", "result")) < 1e-13


def test_elastostatic_goldens(box2):
    el, u, z = box2["el"], box2["u"], box2["z"]
    assert rel(el.value("elastic", u, z), gold("ElastostaticResidual3D", "residual")) < 1e-13
    assert rel(el.gradient_u("elastic", u, z)[:36], gold("ElastostaticResidual3D", "jacobian_head")) < 1e-15
    assert rel(el.gradient_z("elastic", u, z)[:12], gold("ElastostaticResidual3D", "gradient_z_head")) < 2e-14


def test_internal_elastic_energy_goldens(box2):
    el, u, z = box2["el"], box2["u"], box2["z"]
    assert abs(el.energy_value("elastic", u, z) - 48.15) / 48.15 < 1e-13
    assert rel(el.energy_gradient_u("elastic", u, z)[:21], gold("InternalElasticEnergy3D", "gradient_u")) < 1e-13
    assert rel(el.energy_gradient_z("elastic", u, z), gold("InternalElasticEnergy3D", "gradient_z")) < 1e-13


def test_thermostatic_goldens(box2):
    th, T, z = box2["th"], box2["T"], box2["z"]
    assert rel(th.value("thermal", T, z), gold("ThermostaticResidual3D", "residual")) < 1e-13
    assert rel(th.gradient_u("thermal", T, z)[:22], gold("ThermostaticResidual3D", "jacobian_head")) < 1e-15
    assert rel(th.gradient_z("thermal", T, z)[:22], gold("ThermostaticResidual3D", "gradient_z_head")) < 1e-14


def test_internal_thermal_energy_goldens(box2):
    th, T, z = box2["th"], box2["T"], box2["z"]
    v = th.energy_value("thermal", T, z)
    assert abs(v - float(gold("InternalThermalEnergy3D", "value"))) / v < 1e-13
    assert rel(th.energy_gradient_u("thermal", T, z), gold("InternalThermalEnergy3D", "gradient_u")) < 1e-13
    assert rel(th.energy_gradient_z("thermal", T, z), gold("InternalThermalEnergy3D", "gradient_z")) < 1e-13


def test_gradient_x_goldens(box2):
    el, th, u, T, z = box2["el"], box2["th"], box2["u"], box2["T"], box2["z"]
    assert rel(el.gradient_x("elastic", u, z)[:36], gold("ElastostaticResidual3D", "gradient_x_head")) < 1e-13
    assert rel(th.gradient_x("thermal", T, z)[:12], gold("ThermostaticResidual3D", "gradient_x_head")) < 1e-13
    ge = el.energy_gradient_x("elastic", u, z)
    ref = gold("InternalElasticEnergy3D", "gradient_x")
    nz = ref != 0.0  # the reference checks its zero entries as |value| < 1e-13
    assert rel(ge[:ref.size][nz], ref[nz]) < 1e-13
    assert np.max(np.abs(ge[:ref.size][~nz])) < 1e-13
    assert rel(th.energy_gradient_x("thermal", T, z)[:36], gold("InternalThermalEnergy3D", "gradient_x")) < 1e-13


def test_graph_rows_match_reference_structure(box2):
    rowmap, colidx = box2["el"].rowmap, box2["el"].colidx
    # DerivativeTests.cpp:1651-1656 (thermal Jacobian head) implies these rows
    assert list(colidx[rowmap[0]:rowmap[1]]) == [0, 1, 3, 4, 9, 10, 12, 13]
    assert list(colidx[rowmap[1]:rowmap[2]]) == [0, 1, 2, 4, 5, 10, 11, 13, 14]


@pytest.mark.parametrize("n", [(2, 2, 2), (20, 10, 10), (3, 4, 5)])
def test_graph_size_kuhn(n):
    nx, ny, nz = n
    coords, conn = oracle.box_mesh(nx, ny, nz)
    N = coords.size // 3
    # edges of the Kuhn split: axis edges, one diagonal per face of each
    # grid plane family, one body diagonal per hex
    axis = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    faces = nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz
    E = axis + faces + nx * ny * nz
    rowmap, colidx = oracle.graph(N, conn)
    assert colidx.size == N + 2 * E
    if n == (20, 10, 10):
        assert colidx.size == 33621
    for i in range(N):
        row = colidx[rowmap[i]:rowmap[i + 1]]
        assert np.all(np.diff(row) > 0) and i in row


# ------------------------------------------------ independent numpy checks
def numpy_stiffness(coords, conn, E, nu, z, p, emin):
    """Dense K from B^T C B with np.linalg.inv of the cell Jacobian."""
    X = coords.reshape(-1, 3)
    cn = conn.reshape(-1, 4)
    N = X.shape[0]
    c = E / ((1 + nu) * (1 - 2 * nu))
    Cm = np.zeros((6, 6))
    Cm[:3, :3] = c * nu
    np.fill_diagonal(Cm[:3, :3], c * (1 - nu))
    Cm[3, 3] = Cm[4, 4] = Cm[5, 5] = c * (1 - 2 * nu) / 2
    K = np.zeros((3 * N, 3 * N))
    for e in range(cn.shape[0]):
        xe = X[cn[e]]
        J = (xe[:3] - xe[3]).T
        Jinv = np.linalg.inv(J)
        g = np.vstack([Jinv, -Jinv.sum(0)])
        V = abs(np.linalg.det(J)) / 6
        B = np.zeros((6, 12))
        for a in range(4):
            B[0, 3 * a], B[1, 3 * a + 1], B[2, 3 * a + 2] = g[a]
            B[3, 3 * a + 1], B[3, 3 * a + 2] = g[a, 2], g[a, 1]
            B[4, 3 * a + 0], B[4, 3 * a + 2] = g[a, 2], g[a, 0]
            B[5, 3 * a + 0], B[5, 3 * a + 1] = g[a, 1], g[a, 0]
        rho = z[cn[e]].mean()
        s = emin + (1 - emin) * rho ** p
        Ke = s * V * B.T @ Cm @ B
        idx = (3 * cn[e][:, None] + np.arange(3)).ravel()
        K[np.ix_(idx, idx)] += Ke
    return K


@pytest.fixture(scope="module")
def small():
    coords, conn = oracle.box_mesh(3, 2, 2, 1.5, 1.0, 1.0)
    N = coords.size // 3
    rng = np.random.default_rng(7)
    z = rng.uniform(0.3, 1.0, N)
    u = rng.standard_normal(3 * N) * 1e-2
    T = rng.standard_normal(N)
    fc, fl = oracle.side_set(coords, conn, 0, 1.5)
    el = oracle.Model(coords, conn, youngs=2.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-3,
                      face_cells=fc, face_lnodes=fl, flux=(0.1, 0.2, -1.0))
    th = oracle.Model(coords, conn, conductivity=1.5, penalty_exponent=3.0, penalty_min=1e-3, face_cells=fc,
                      face_lnodes=fl, flux=(0.7, 0, 0), body=(0, 1.0))
    return dict(coords=coords, conn=conn, N=N, z=z, u=u, T=T, el=el, th=th)


def test_jacobian_equals_numpy_btcb(small):
    el = small["el"]
    Kd = numpy_stiffness(small["coords"], small["conn"], 2.0, 0.3, small["z"], 3.0, 1e-3)
    K = oracle.to_scipy(3, 3, el.rowmap, el.colidx, el.gradient_u("elastic", small["u"], small["z"])).toarray()
    assert rel(K, Kd) < 1e-13
    KT = oracle.to_scipy(3, 3, el.rowmap, el.colidx, el.gradient_u_T("elastic", small["u"], small["z"])).toarray()
    assert rel(KT, Kd.T) < 1e-13


def test_residual_is_linear_in_u(small):
    el, u, z = small["el"], small["u"], small["z"]
    K = oracle.to_scipy(3, 3, el.rowmap, el.colidx, el.gradient_u("elastic", u, z))
    R0 = el.value("elastic", np.zeros_like(u), z)
    assert rel(el.value("elastic", u, z) - R0, K @ u) < 1e-12
    # traction only: R(0) = -t A / 3 per face node; total = -t * area(x+ face) = -t * 1 * 1
    Rz = R0.reshape(-1, 3).sum(0)
    assert np.allclose(Rz, [-0.1, -0.2, 1.0], rtol=1e-13, atol=1e-15)


def test_rigid_translation_null_space(small):
    el = small["el"]
    K = oracle.to_scipy(3, 3, el.rowmap, el.colidx, el.gradient_u("elastic", small["u"], small["z"]))
    for d in range(3):
        t = np.zeros(3 * small["N"])
        t[d::3] = 1.0
        assert np.max(np.abs(K @ t)) < 1e-12 * np.max(np.abs(K.data))


def test_patch_test_interior_residual_zero():
    coords, conn = oracle.box_mesh(4, 4, 4)
    X = coords.reshape(-1, 3)
    A = np.array([[1e-3, 2e-3, -1e-3], [0.5e-3, -1e-3, 3e-3], [2e-3, 1e-3, 0.5e-3]])
    u = (X @ A.T).ravel()
    el = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=1.0, penalty_min=0.0)
    R = el.value("elastic", u, np.ones(X.shape[0])).reshape(-1, 3)
    interior = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    assert np.max(np.abs(R[interior])) < 1e-14


def test_uniform_strain_energy_closed_form(box2):
    # u = 1e-4 (k+1) is linear in the node index, hence a uniform strain field;
    # W = 1/2 sigma:eps * volume(=1) with the golden strain/stress
    eps = gold("3D", "strain")[:6]
    sig = gold("3D", "stress")[:6]
    assert abs(0.5 * eps @ sig - 48.15) < 1e-10
    assert abs(box2["el"].energy_value("elastic", box2["u"], box2["z"]) - 0.5 * eps @ sig) < 1e-11


def fd(f, x, h, idx):
    out = []
    for i in idx:
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        out.append((f(xp) - f(xm)) / (2 * h))
    return np.array(out)


def test_gradient_z_central_differences(small):
    el, th, u, T, z = small["el"], small["th"], small["u"], small["T"], small["z"]
    idx = [0, 5, 11, small["N"] - 1]
    for m, phys, st in ((el, "elastic", u), (th, "thermal", T)):
        dpn = 3 if phys == "elastic" else 1
        Gt = oracle.to_scipy(1, dpn, m.rowmap, m.colidx, m.gradient_z(phys, st, z)).toarray()  # N x N*dpn
        d = fd(lambda zz: m.value(phys, st, zz), z, 1e-6, idx)  # rows: dR/dz_i
        assert rel(Gt[idx], d) < 1e-7


def test_thermal_jacobian_central_differences(small):
    th, T, z = small["th"], small["T"], small["z"]
    K = oracle.to_scipy(1, 1, th.rowmap, th.colidx, th.gradient_u("thermal", T, z)).toarray()
    d = fd(lambda tt: th.value("thermal", tt, z), T, 1e-5, [0, 7, 20])
    assert rel(K[:, [0, 7, 20]].T, d) < 1e-8


def test_energy_gradients_central_differences(small):
    el, th, u, T, z = small["el"], small["th"], small["u"], small["T"], small["z"]
    for m, phys, st in ((el, "elastic", u), (th, "thermal", T)):
        gu = m.energy_gradient_u(phys, st, z)
        gz = m.energy_gradient_z(phys, st, z)
        iu = [0, 4, 13]
        assert rel(gu[iu], fd(lambda x: m.energy_value(phys, x, z), st, 1e-6, iu)) < 1e-7
        iz = [1, 6, 17]
        assert rel(gz[iz], fd(lambda x: m.energy_value(phys, st, x), z, 1e-6, iz)) < 1e-7


def _small_models(coords, conn, fc, fl):
    el = oracle.Model(coords, conn, youngs=2.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-3,
                      face_cells=fc, face_lnodes=fl, flux=(0.1, 0.2, -1.0))
    th = oracle.Model(coords, conn, conductivity=1.5, penalty_exponent=3.0, penalty_min=1e-3, face_cells=fc,
                      face_lnodes=fl, flux=(0.7, 0, 0), body=(0, 1.0))
    return el, th


@pytest.fixture(scope="module")
def skewed():
    # the small box with interior and boundary nodes moved off the grid, so no
    # cell is a right-angled Kuhn tet
    coords, conn = oracle.box_mesh(3, 2, 2, 1.5, 1.0, 1.0)
    X = coords.reshape(-1, 3).copy()
    rng = np.random.default_rng(5)
    X += rng.uniform(-0.06, 0.06, X.shape)
    X[np.isclose(coords.reshape(-1, 3)[:, 0], 1.5), 0] = 1.5  # x+ face stays planar
    coords = X.ravel()
    fc, fl = oracle.side_set(coords, conn, 0, 1.5)
    N = X.shape[0]
    z = rng.uniform(0.3, 1.0, N)
    u = rng.standard_normal(3 * N) * 1e-2
    T = rng.standard_normal(N)
    return dict(coords=coords, conn=conn, fc=fc, fl=fl, N=N, z=z, u=u, T=T)


def test_gradient_x_central_differences(skewed):
    """Rows of (dR/dx)^T and df/dx against central differences of the oracle's
    value in the node coordinates, natural BC and body source included; the
    nodes cover a corner, the loaded x+ face and the interior."""
    S = skewed
    X = S["coords"]
    el, th = _small_models(X, S["conn"], S["fc"], S["fl"])
    X3 = X.reshape(-1, 3)
    loaded = int(np.flatnonzero(np.isclose(X3[:, 0], 1.5))[5])
    nodes = [0, loaded, 17]
    h = 1e-6
    for phys, st, dpn in (("elastic", S["u"], 3), ("thermal", S["T"], 1)):
        m = el if phys == "elastic" else th
        Gx = oracle.to_scipy(3, dpn, m.rowmap, m.colidx, m.gradient_x(phys, st, S["z"])).toarray()
        gx = m.energy_gradient_x(phys, st, S["z"])
        for a in nodes:
            for i in range(3):
                Xp, Xm = X.copy(), X.copy()
                Xp[3 * a + i] += h
                Xm[3 * a + i] -= h
                mp = _small_models(Xp, S["conn"], S["fc"], S["fl"])[0 if phys == "elastic" else 1]
                mm = _small_models(Xm, S["conn"], S["fc"], S["fl"])[0 if phys == "elastic" else 1]
                dR = (mp.value(phys, st, S["z"]) - mm.value(phys, st, S["z"])) / (2 * h)
                assert rel(Gx[3 * a + i], dR) < 1e-7, (phys, a, i)
                df = (mp.energy_value(phys, st, S["z"]) - mm.energy_value(phys, st, S["z"])) / (2 * h)
                assert abs(gx[3 * a + i] - df) < 1e-7 * np.max(np.abs(gx)), (phys, a, i)


def test_gradient_x_translation_invariance(skewed):
    """A rigid translation of the whole mesh changes neither R nor f: the
    coordinate derivatives summed over all nodes vanish per direction."""
    S = skewed
    el, th = _small_models(S["coords"], S["conn"], S["fc"], S["fl"])
    for m, phys, st, dpn in ((el, "elastic", S["u"], 3), (th, "thermal", S["T"], 1)):
        Gx = oracle.to_scipy(3, dpn, m.rowmap, m.colidx, m.gradient_x(phys, st, S["z"])).toarray()
        col = np.abs(Gx).max()
        for i in range(3):
            assert np.max(np.abs(Gx[i::3].sum(0))) < 1e-12 * col
        gx = m.energy_gradient_x(phys, st, S["z"]).reshape(-1, 3)
        assert np.max(np.abs(gx.sum(0))) < 1e-12 * np.abs(gx).max()


def test_elastic_energy_is_half_u_K_u(small):
    el, u, z = small["el"], small["u"], small["z"]
    K = oracle.to_scipy(3, 3, el.rowmap, el.colidx, el.gradient_u("elastic", u, z))
    assert abs(el.energy_value("elastic", u, z) - 0.5 * u @ (K @ u)) < 1e-12 * abs(u @ (K @ u))
    assert rel(el.energy_gradient_u("elastic", u, z), K @ u) < 1e-12


def test_block_constraints_invariants(small):
    el, u, z = small["el"], small["u"], small["z"]
    vals = el.gradient_u("elastic", u, z)
    K0 = oracle.to_scipy(3, 3, el.rowmap, el.colidx, vals.copy()).toarray()
    rhs = np.arange(3 * small["N"], dtype=float)
    dofs = np.array([0, 1, 2, 10, 31], np.int32)
    g = np.array([0.1, -0.2, 0.3, 0.0, 0.5])
    rhs0 = rhs.copy()
    oracle.apply_block_constraints(3, el.rowmap, el.colidx, vals, rhs, dofs, g)
    K = oracle.to_scipy(3, 3, el.rowmap, el.colidx, vals).toarray()
    for d in dofs:
        e = np.zeros(K.shape[0])
        e[d] = 1
        assert np.array_equal(K[d], e) and np.array_equal(K[:, d], e)
    free = np.setdiff1d(np.arange(K.shape[0]), dofs)
    gfull = np.zeros(K.shape[0])
    gfull[dofs] = g
    assert rel(rhs[free], rhs0[free] - (K0 @ gfull)[free]) < 1e-13
    assert np.array_equal(rhs[dofs], g)
