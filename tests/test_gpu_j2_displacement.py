"""GPU parity of the pure-displacement J2 formulation (3 dofs per node; no
reference counterpart) against the oracle, whose displacement form is pinned
on the CPU by elastostatics in the elastic regime and by central differences
of every partial, the condensed tangent and the adjoint gradient.

* every workset entry point at a random state with plastic and elastic cells
  (cell worksets keep the node*4 + d layout, pressure slots zero): 1e-12
  relative to the largest entry;
* the condensed system (3x3 blocks) and its transpose, the plastic work;
* in the elastic regime the condensed tangent and residual equal the GPU
  elastostatic VectorFunction's (an independent GPU path);
* a 3-step plastic solve step by step (GMRES with block Jacobi, and BiCGStab
  with the 3x3-block multigrid), the criterion value and adjoint gradient,
  and the gradient against central differences of GPU forward solves.
"""
import numpy as np
import pytest

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


MAT = dict(youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=5.0, initial_yield_stress=1.0,
           elastic_penalty_exponent=3.0, elastic_min_ersatz=1e-9, plastic_penalty_exponent=2.5,
           plastic_min_ersatz=1e-9, global_penalty_exponent=3.0, global_min_ersatz=1e-9)


def both(pa, dims, lengths=(1.0, 1.0, 1.0), **params):
    mesh = pa.Mesh.box(*dims, *lengths)
    coords, conn = oracle.box_mesh(*dims, *lengths)
    pl = pa.Plasticity(mesh, formulation="displacement", **params)
    om = oracle.J2Model(coords, conn, formulation="displacement", **params)
    assert pl.dpn == 3 and om.dpn == 3
    return mesh, pl, om


@pytest.fixture(scope="module")
def state(pa):
    print("case: displacement J2 parity state on 4x2x2, E=1e3", flush=True)
    mesh, pl, om = both(pa, (4, 2, 2), **MAT)
    rng = np.random.default_rng(21)
    X = om.coords.reshape(-1, 3)
    u = rng.uniform(-1, 1, om.N * 3) * 2e-3
    u.reshape(-1, 3)[:] *= (X[:, 0:1] > 0.3 * X[:, 0].max())
    cp = np.zeros((om.C, 14))
    cp[:, 2:8] = rng.uniform(-1, 1, (om.C, 6)) * 1e-4
    cp[:, 8:] = rng.uniform(-1, 1, (om.C, 6)) * 0.05
    cp[:, 2:5] -= cp[:, 2:5].mean(axis=1, keepdims=True)
    cp[:, 8:11] -= cp[:, 8:11].mean(axis=1, keepdims=True)
    cp[:, 0] = rng.uniform(0, 1e-2, om.C)
    cp = cp.ravel()
    z = rng.uniform(0.5, 1.0, om.N)
    pi = rng.uniform(-1, 1, om.N * 3)  # read and ignored by this formulation
    c = om.update_local_state(u, cp, z)
    cc = c.reshape(-1, 14)
    assert (cc[:, 1] > 0).sum() > 10 and (cc[:, 1] == 0).sum() > 10
    return mesh, pl, om, u, c, cp, z, pi


def test_update_and_local_value(state):
    mesh, pl, om, u, c, cp, z, pi = state
    cg = pl.update_local_state(u, cp, z).numpy()
    assert rel(cg, c) < 1e-12
    assert np.array_equal(cg.reshape(-1, 14)[:, 1] > 0, c.reshape(-1, 14)[:, 1] > 0)
    c2 = c + np.random.default_rng(2).uniform(-1, 1, c.size) * 1e-4 * (np.abs(c) + 1e-3)
    c2.reshape(-1, 14)[:, 1] = c.reshape(-1, 14)[:, 1]
    assert rel(pl.local_value(u, c2, cp, z).numpy(), om.local_value(u, c2, cp, z)) < 1e-12


@pytest.mark.parametrize("wrt", ["c", "u", "z", "cp"])
def test_local_gradients(pa, state, wrt):
    mesh, pl, om, u, c, cp, z, pi = state
    code = {"c": pa.WRT_C, "u": pa.WRT_U, "z": pa.WRT_Z, "cp": pa.WRT_C_PREV}[wrt]
    got = pl.local_gradient(code, u, c, cp, z)
    assert rel(got, om.local_gradient(wrt, u, c, cp, z)) < 1e-12, wrt
    if wrt == "u":
        assert np.all(got[:, :, 3::4] == 0.0)


def test_global_value(state):
    mesh, pl, om, u, c, cp, z, pi = state
    R = pl.global_value(u, c, pi, z).numpy()
    assert R.size == 3 * om.N
    assert rel(R, om.global_value(u, c, pi, z)) < 1e-12


@pytest.mark.parametrize("wrt", ["u", "c", "z", "pi"])
def test_global_gradients(pa, state, wrt):
    mesh, pl, om, u, c, cp, z, pi = state
    code = {"u": pa.WRT_U, "c": pa.WRT_C, "z": pa.WRT_Z, "pi": pa.WRT_PI}[wrt]
    got = pl.global_gradient(code, u, c, pi, z)
    ref = om.global_gradient_ws(wrt, u, c, oracle.gather_nodal(om.conn, pi, 3), z)
    assert np.all(got[:, 3::4, :] == 0.0)  # no pressure rows
    if wrt == "pi":
        assert np.all(got == 0.0) and np.all(ref == 0.0)
    else:
        assert rel(got, ref) < 1e-12, wrt


def test_condensed_system_and_transpose(state):
    mesh, pl, om, u, c, cp, z, pi = state
    K, R = pl.condensed_system(u, c, cp, pi, z)
    vals, Rref = om.condensed_system(u, c, cp, pi, z)
    assert K.entries().size == 9 * om.nnzb
    assert rel(K.entries(), vals) < 1e-12
    assert rel(R.numpy(), Rref) < 1e-12
    KT, none = pl.condensed_system(u, c, cp, pi, z, transpose=True)
    assert none is None
    A = oracle.to_scipy(3, 3, om.rowmap, om.colidx, vals)
    AT = oracle.to_scipy(3, 3, om.rowmap, om.colidx, KT.entries())
    assert abs(AT - A.T).max() <= 1e-12 * abs(A).max()
    K2, R2 = pl.condensed_system(u, c, cp, pi, z)
    assert np.array_equal(K2.entries(), K.entries()) and np.array_equal(R2.numpy(), R.numpy())


def test_plastic_work(state):
    mesh, pl, om, u, c, cp, z, pi = state
    up = u * 0.5
    w_gpu = pl.plastic_work(u, up, c, cp, z)
    w_ref = om.plastic_work(u, up, c, cp, z)
    assert abs(w_gpu - w_ref) <= 1e-12 * abs(w_ref)


def test_elastic_regime_matches_gpu_elastostatics(pa):
    """No plastic strain: the condensed tangent and residual are the SIMP
    elastostatic Jacobian and residual of the same E, nu and penalty."""
    print("case: displacement J2 elastic regime vs GPU elastostatics, 5x3x2", flush=True)
    mesh = pa.Mesh.box(5, 3, 2, 2.0, 1.0, 1.0)
    pl = pa.Plasticity(mesh, formulation="displacement", youngs=7.0, poisson=0.27, hardening_isotropic=1.0,
                       hardening_kinematic=0.0, initial_yield_stress=1e6, global_penalty_exponent=3.0,
                       global_min_ersatz=1e-3)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=7.0, poisson=0.27, penalty_exponent=3.0, penalty_min=1e-3)
    rng = np.random.default_rng(3)
    N, C = mesh.num_nodes, mesh.num_cells
    u = rng.standard_normal(3 * N) * 1e-2
    z = rng.uniform(0.2, 1.0, N)
    c0 = np.zeros(C * 14)
    c = pl.update_local_state(u, c0, z).numpy()
    assert np.all(c == 0.0)
    K, R = pl.condensed_system(u, c, c0, np.zeros(3 * N), z)
    assert rel(K.entries(), vf.gradient_u(u, z).entries()) < 1e-12
    assert rel(R.numpy(), vf.value(u, z).numpy()) < 1e-12


def displacement_problem(pa):
    mesh, pl, om = both(pa, (4, 2, 2), (2.0, 1.0, 1.0), youngs=1e3, poisson=0.3, hardening_isotropic=10.0,
                        hardening_kinematic=5.0, initial_yield_stress=1.0, elastic_penalty_exponent=3.0,
                        plastic_penalty_exponent=2.5, plastic_min_ersatz=1e-9, global_penalty_exponent=3.0)
    X = om.coords.reshape(-1, 3)
    xm = np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]
    xp = np.nonzero(np.abs(X[:, 0] - 2.0) < 1e-12)[0]
    dofs = np.concatenate([(3 * xm[:, None] + np.arange(3)).ravel(), 3 * xp])
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, 0.004)])
    order = np.argsort(dofs)
    return mesh, pl, om, dofs[order].astype(np.int32), vals[order]


@pytest.mark.parametrize("solver", ["gmres", "bicgstab_multigrid"])
def test_multistep_solve_and_adjoint(pa, solver):
    print(f"case: displacement J2 3-step solve on 4x2x2, {solver}", flush=True)
    mesh, pl, om, dofs, vals = displacement_problem(pa)
    z = np.random.default_rng(5).uniform(0.6, 1.0, om.N)
    opts = dict(num_steps=3, max_newton=25, newton_stop_tol=1e-11, newton_current_norm_tol=1e-11)
    if solver != "gmres":
        opts["krylov"] = pa.KRYLOV_BICGSTAB

    def problem():
        pr = pa.PlasticityProblem(pl, dofs, vals, **opts)
        if solver != "gmres":
            pr.set_multigrid(4, 2, 2)
        return pr

    prob = problem()
    assert prob.solution(z)
    us, cs, pis, its = om.solve(z, dofs, vals, num_steps=3, max_newton=25, stop_tol=1e-11, current_norm_tol=1e-11)
    for k in range(3):
        u, cst, pi = prob.state(k)
        assert u.size == 3 * om.N
        assert rel(u, us[k]) < 1e-8, k
        assert np.all(pi == 0.0)
    assert (cs[-1].reshape(-1, 14)[:, 1] > 0).sum() > 0
    F0 = prob.criterion_value(z, -1.0)
    assert abs(F0 - om.criterion_value(z, us, cs, -1.0)) <= 1e-9 * abs(F0)
    grad = prob.criterion_gradient(z, -1.0).numpy()
    assert rel(grad, om.criterion_gradient_z(z, us, cs, pis, dofs, -1.0)) < 1e-7
    if solver != "gmres":
        return
    dz = np.random.default_rng(6).uniform(-1, 1, om.N)
    h = 1e-6

    def F(zz):
        pr = problem()
        pr.solution(zz)
        return pr.criterion_value(zz, -1.0)

    fd = (F(z + h * dz) - F(z - h * dz)) / (2 * h)
    assert abs(fd - grad @ dz) <= 1e-6 * abs(fd)
