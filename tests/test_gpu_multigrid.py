"""GPU checks of the geometric multigrid preconditioner.

* the Galerkin operator of level 1 equals P^T A P computed with numpy from the
  assembled (constrained) fine matrix and the trilinear prolongation with the
  Dirichlet dofs masked out (FP32 storage: 1e-5 relative);
* MG-preconditioned CG solves the SIMP problem to the FP64 tolerance and
  agrees with the block-Jacobi solution; it needs far fewer iterations;
* the V-cycle is symmetric (x.M y == y.M x), as CG requires, with damped
  Jacobi and with Chebyshev smoothing;
* with a symmetric operator the fine level's half stencil (the diagonal and
  the blocks towards larger offsets, the others read transposed from the
  neighbour) gives the full stencil's V-cycle bit for bit.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def prolongation(dims, dpn, mask):
    nx, ny, nz = dims
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    fid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    cid = lambda i, j, k: (i * (cy + 1) + j) * (cz + 1) + k

    def w1(f):
        return [(f // 2, 1.0)] if f % 2 == 0 else [((f - 1) // 2, 0.5), ((f + 1) // 2, 0.5)]

    rows, cols, vals = [], [], []
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz + 1):
                for I, wi in w1(i):
                    for J, wj in w1(j):
                        for K, wk in w1(k):
                            for d in range(dpn):
                                f = fid(i, j, k) * dpn + d
                                if mask[f]:
                                    rows.append(f)
                                    cols.append(cid(I, J, K) * dpn + d)
                                    vals.append(wi * wj * wk)
    nf = (nx + 1) * (ny + 1) * (nz + 1) * dpn
    nc = (cx + 1) * (cy + 1) * (cz + 1) * dpn
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc)), (cx, cy, cz)


def stencil_to_dense(dims, st):
    nx, ny, nz = dims
    n = (nx + 1) * (ny + 1) * (nz + 1)
    D = st.shape[-1]
    A = np.zeros((n * D, n * D))
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz + 1):
                r = nid(i, j, k)
                for s in range(27):
                    dx, dy, dz = s // 9 - 1, (s // 3) % 3 - 1, s % 3 - 1
                    a, b, c = i + dx, j + dy, k + dz
                    if 0 <= a <= nx and 0 <= b <= ny and 0 <= c <= nz:
                        col = nid(a, b, c)
                        A[r * D:(r + 1) * D, col * D:(col + 1) * D] = st[r, s]
    return A


def elastic_setup(pa, dims, seed, lo=0.05):
    nx, ny, nz = dims
    mesh = pa.Mesh.box(nx, ny, nz, 2.0, 1.0, 1.0)
    z = np.random.default_rng(seed).uniform(lo, 1.0, mesh.num_nodes)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
    sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1.0, poisson=0.3, penalty_exponent=3.0,
                                  penalty_min=1e-9)
    dofs = (3 * mesh.node_set("x-")[:, None] + np.arange(3)).ravel().astype(np.int32)
    return mesh, vf, sf, z, dofs


def test_galerkin_level_operator(pa):
    print("case: Galerkin P^T A P on 4x2x2 elastic", flush=True)
    dims = (4, 2, 2)
    mesh, vf, sf, z, dofs = elastic_setup(pa, dims, 3)
    n = 3 * mesh.num_nodes
    K = vf.gradient_u(np.zeros(n), z)
    rhs = pa.DeviceArray(n)
    rhs.zero()
    pa.apply_block_constraints(K, rhs, dofs, np.zeros(dofs.size))
    mg = pa.Multigrid(mesh, *dims, 3)
    assert mg.levels()[0] == 2
    mg.setup(K, dofs)
    cdims, st = mg.level_operator(1)
    assert cdims == (2, 1, 1)
    mask = np.ones(n, bool)
    mask[dofs] = False
    P, _ = prolongation(dims, 3, mask)
    A = oracle.to_scipy(3, 3, *mesh.graph(), K.entries()).toarray()
    ref = P.T @ A @ P
    assert rel(stencil_to_dense(cdims, st), ref) < 1e-5


@pytest.mark.parametrize("symmetric", [False, True])
def test_galerkin_second_level(pa, symmetric):
    """Level 1 from the fine stencil (full or half layout) and level 2 from
    level 1's 27-point stencil against P^T A P with numpy, on 8x4x4."""
    print(f"case: Galerkin levels 1 and 2 on 8x4x4 elastic, symmetric={symmetric}", flush=True)
    dims = (8, 4, 4)
    mesh, vf, sf, z, dofs = elastic_setup(pa, dims, 6)
    n = 3 * mesh.num_nodes
    K = vf.gradient_u(np.zeros(n), z)
    rhs = pa.DeviceArray(n)
    rhs.zero()
    pa.apply_block_constraints(K, rhs, dofs, np.zeros(dofs.size))
    mg = pa.Multigrid(mesh, *dims, 3, symmetric=symmetric)
    assert mg.levels()[0] == 3
    mg.setup(K, dofs)
    mask = np.ones(n, bool)
    mask[dofs] = False
    P1, c1 = prolongation(dims, 3, mask)
    A = oracle.to_scipy(3, 3, *mesh.graph(), K.entries()).toarray()
    cdims, st1 = mg.level_operator(1)
    assert cdims == c1 == (4, 2, 2)
    A1 = stencil_to_dense(cdims, st1)
    assert rel(A1, P1.T @ A @ P1) < 1e-5
    P2, c2 = prolongation(cdims, 3, np.ones(A1.shape[0], bool))
    cdims2, st2 = mg.level_operator(2)
    assert cdims2 == c2 == (2, 1, 1)
    assert rel(stencil_to_dense(cdims2, st2), P2.T @ A1 @ P2) < 1e-5


@pytest.mark.parametrize("smoother,coarse_sweeps", [("jacobi", 0), ("chebyshev", 0), ("chebyshev", 6)])
def test_vcycle_is_symmetric(pa, smoother, coarse_sweeps):
    """x.M y == y.M x and M positive, with damped Jacobi and with the
    Chebyshev smoother (dampings from the power-iteration eigenvalue bound,
    post-smoothing in reverse order)."""
    print(f"case: V-cycle symmetry, {smoother}, coarse sweeps {coarse_sweeps}", flush=True)
    dims = (8, 4, 4)
    mesh, vf, sf, z, dofs = elastic_setup(pa, dims, 5)
    n = 3 * mesh.num_nodes
    K = vf.gradient_u(np.zeros(n), z, precision="fp32")
    rhs = pa.DeviceArray(n)
    rhs.zero()
    pa.apply_block_constraints(K, rhs, dofs, np.zeros(dofs.size))
    mg = pa.Multigrid(mesh, *dims, 3, symmetric=True, smoother=smoother, coarse_sweeps=coarse_sweeps)
    mg.setup(K, dofs)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    x[dofs] = 0
    y[dofs] = 0
    Mx = mg.apply(pa.DeviceArray.from_numpy(x), pa.DeviceArray(n)).numpy()
    My = mg.apply(pa.DeviceArray.from_numpy(y), pa.DeviceArray(n)).numpy()
    assert abs(y @ Mx - x @ My) <= 1e-4 * abs(y @ Mx)
    assert x @ Mx > 0 and y @ My > 0


def test_chebyshev_needs_equal_sweeps(pa):
    """A Chebyshev V-cycle is symmetric only with the post-smoother the
    adjoint of the pre-smoother: unequal sweep counts are refused."""
    mesh = pa.Mesh.box(4, 2, 2, 2.0, 1.0, 1.0)
    with pytest.raises(pa.PlatoError):
        pa.Multigrid(mesh, 4, 2, 2, 3, pre_sweeps=1, post_sweeps=2, smoother="chebyshev")
    pa.Multigrid(mesh, 4, 2, 2, 3, pre_sweeps=1, post_sweeps=2, smoother="jacobi")


@pytest.mark.parametrize("physics", ["Mechanics", "Thermal"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_half_stencil_vcycle_bitwise(pa, physics, precision):
    """The assembled Kuhn-grid Jacobians are symmetric bit for bit, and so
    are they after the Dirichlet elimination: the half fine stencil then
    reads the same blocks as the full one, transposed, and the V-cycles are
    equal bit for bit (on a second setup as well)."""
    dims = (6, 4, 4)
    print(f"case: half-stencil V-cycle, {physics} {precision}, {dims}", flush=True)
    mesh, vf, sf, z, dofs = elastic_setup(pa, dims, 9)
    dpn = 3
    if physics == "Thermal":
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
        dofs = mesh.node_set("x-").astype(np.int32)
        dpn = 1
    n = dpn * mesh.num_nodes
    rng = np.random.default_rng(4)
    for rep in range(2):
        K = vf.gradient_u(np.zeros(n), z * (1.0 + 0.1 * rep), precision=precision)
        rhs = pa.DeviceArray(n)
        rhs.zero()
        pa.apply_block_constraints(K, rhs, dofs, np.zeros(dofs.size))
        out = []
        for sym in (False, True):
            mg = pa.Multigrid(mesh, *dims, dpn, symmetric=sym)
            mg.setup(K, dofs)
            r = rng.standard_normal(n) if not sym else r
            out.append(mg.apply(pa.DeviceArray.from_numpy(r), pa.DeviceArray(n)).numpy())
        assert np.all(np.isfinite(out[0])) and np.any(out[0] != 0)
        assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
@pytest.mark.parametrize("coarse_sweeps", [0, 6])
def test_multigrid_pcg_matches_block_jacobi(pa, mode, coarse_sweeps):
    print(f"case: MG-PCG 16x8x8 elastic, {mode}, coarse sweeps {coarse_sweeps}", flush=True)
    dims = (16, 8, 8)
    mesh, vf, sf, z, dofs = elastic_setup(pa, dims, 7)
    P = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-10, max_iter=20000)
    u_bj = P.solution(z).numpy()
    it_bj = P.last_stats["iterations"]
    P2 = pa.Problem(vf, sf, dofs, mode=mode, tol=1e-10, max_iter=2000, replace_every=50)
    P2.set_multigrid(*dims, coarse_sweeps=coarse_sweeps)
    u_mg = P2.solution(z).numpy()
    st = P2.last_stats
    print("block-Jacobi", it_bj, "multigrid", st, flush=True)
    assert st["converged"] and st["rel_residual"] <= 1e-10, st
    assert st["iterations"] < it_bj / 4
    assert rel(u_mg, u_bj) < 1e-7
    g1 = P.criterion_gradient(z, pa.DeviceArray.from_numpy(u_bj)).numpy()
    g2 = P2.criterion_gradient(z, pa.DeviceArray.from_numpy(u_mg)).numpy()
    assert rel(g2, g1) < 1e-7


def test_multigrid_large_dense_coarsest(pa):
    """A coarsest level of 675 dofs, inverted by the stepwise Gauss-Jordan
    (one launch per pivot step): MG-PCG matches block-Jacobi PCG, and the
    exact solve on the 8x4x4 level needs no more iterations than the deeper
    hierarchy that smooths it."""
    print("case: MG-PCG 16x8x8 elastic, dense coarsest 8x4x4 (675 dofs)", flush=True)
    dims = (16, 8, 8)
    mesh, vf, sf, z, dofs = elastic_setup(pa, dims, 8)
    P = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-10, max_iter=20000)
    u_bj = P.solution(z).numpy()
    its = []
    for ml in (0, 2):
        P2 = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-10, max_iter=2000, check_every=1)
        P2.set_multigrid(*dims, max_levels=ml, max_coarse_nodes=2000)
        u_mg = P2.solution(z).numpy()
        st = P2.last_stats
        assert st["converged"] and st["rel_residual"] <= 1e-10, st
        assert rel(u_mg, u_bj) < 1e-7
        its.append(st["iterations"])
    print("iterations (deep, dense 8x4x4):", its, flush=True)
    assert its[1] <= its[0]
    mg = pa.Multigrid(mesh, *dims, 3, max_levels=2, max_coarse_nodes=2000)
    assert mg.levels() == (2, 675)


def test_multigrid_thermal(pa):
    print("case: MG-PCG 16x16x16 thermal", flush=True)
    dims = (16, 16, 16)
    mesh = pa.Mesh.box(*dims)
    z = np.random.default_rng(2).uniform(0.1, 1.0, mesh.num_nodes)
    vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_body_load(0, 1.0)
    sf = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=1.0, penalty_exponent=3.0,
                                  penalty_min=1e-9)
    dofs = mesh.node_set("z-").astype(np.int32)
    P = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-10, max_iter=20000)
    u1 = P.solution(z).numpy()
    P2 = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-10, max_iter=2000)
    P2.set_multigrid(*dims)
    u2 = P2.solution(z).numpy()
    assert P2.last_stats["converged"], P2.last_stats
    assert rel(u2, u1) < 1e-7


def j2_tangent(pa, dims):
    """Condensed J2 tangent (4x4 blocks) of the first elastic Newton step on a
    box clamped at x- with ux prescribed at x+, constrained."""
    mesh = pa.Mesh.box(*dims, 2.0, 1.0, 1.0)
    coords, conn = oracle.box_mesh(*dims, 2.0, 1.0, 1.0)
    om = oracle.J2Model(coords, conn, youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=0.0,
                        initial_yield_stress=1.0)
    N = om.N
    xm, xp = mesh.node_set("x-"), mesh.node_set("x+")
    dofs = np.concatenate([xm * 4, xm * 4 + 1, xm * 4 + 2, xp * 4]).astype(np.int32)
    g = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, 8e-4)])
    z = np.ones(N)
    c = om.update_local_state(np.zeros(4 * N), np.zeros(om.C * 14), z)
    vals, R = om.condensed_system(np.zeros(4 * N), c, np.zeros(om.C * 14), np.zeros(3 * N), z)
    rhs = -R
    oracle.apply_block_constraints(4, om.rowmap, om.colidx, vals, rhs, dofs, g)
    return mesh, om, vals, rhs, dofs


def test_multigrid_4x4_blocks_galerkin_and_vcycle(pa):
    print("case: 4x4-block multigrid on the condensed J2 tangent, 8x4x4", flush=True)
    dims = (8, 4, 4)
    mesh, om, vals, rhs, dofs = j2_tangent(pa, dims)
    K = pa.CrsMatrix(mesh, 4, 4, pa.DeviceArray.from_numpy(vals), pa.FP64)
    mg = pa.Multigrid(mesh, *dims, 4)
    mg.setup(K, dofs)
    cdims, st = mg.level_operator(1)
    mask = np.ones(4 * om.N, bool)
    mask[dofs] = False
    P = prolongation4(dims, mask)
    A = oracle.to_scipy(4, 4, om.rowmap, om.colidx, vals)
    ref = (P.T @ A @ P).toarray()
    got = stencil_to_dense(cdims, st)
    assert rel(got, ref) < 1e-5
    z = mg.apply(pa.DeviceArray.from_numpy(rhs), pa.DeviceArray(rhs.size)).numpy()
    assert np.all(np.isfinite(z))


def prolongation4(dims, mask):
    nx, ny, nz = dims
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    fid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    cid = lambda i, j, k: (i * (cy + 1) + j) * (cz + 1) + k

    def w1(f):
        return [(f // 2, 1.0)] if f % 2 == 0 else [((f - 1) // 2, 0.5), ((f + 1) // 2, 0.5)]

    rows, cols, vals = [], [], []
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz + 1):
                for I, wi in w1(i):
                    for J, wj in w1(j):
                        for K, wk in w1(k):
                            for d in range(4):
                                f = fid(i, j, k) * 4 + d
                                if mask[f]:
                                    rows.append(f)
                                    cols.append(cid(I, J, K) * 4 + d)
                                    vals.append(wi * wj * wk)
    nf = (nx + 1) * (ny + 1) * (nz + 1) * 4
    nc = (cx + 1) * (cy + 1) * (cz + 1) * 4
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))


def test_multigrid_preconditioned_plasticity_newton(pa):
    print("case: J2 10-step Newton with multigrid vs block Jacobi, 16x8x8", flush=True)
    dims = (16, 8, 8)
    mesh = pa.Mesh.box(*dims, 2.0, 1.0, 1.0)
    xm, xp = mesh.node_set("x-"), mesh.node_set("x+")
    dofs = np.concatenate([xm * 4, xm * 4 + 1, xm * 4 + 2, xp * 4])
    # ramp to 0.004 in 10 steps: plastic from step 4 on, within the reference Newton's range
    g = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, 0.004)])
    z = np.ones(mesh.num_nodes)
    sols = []
    for use_mg in (False, True):
        pl = pa.Plasticity(mesh, youngs=1e3, poisson=0.3, hardening_isotropic=10.0, initial_yield_stress=1.0)
        prob = pa.PlasticityProblem(pl, dofs, g, num_steps=10, max_newton=10, linear_tol=1e-10)
        if use_mg:
            prob.set_multigrid(*dims)
        ok = prob.solution(z)
        print("multigrid" if use_mg else "block jacobi", [prob.stats(k) for k in range(10)], flush=True)
        assert ok
        sols.append(prob.state(9)[0])
    assert rel(sols[1], sols[0]) < 1e-7
