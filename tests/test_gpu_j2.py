"""GPU parity of the J2 plasticity path against the (golden-pinned) oracle.

* every entry point (return map, local residual and its four workset
  Jacobians, global residual and its four workset Jacobians, the condensed
  Newton system and its transpose, projection, plastic work) at a random
  state with plastic and elastic cells: 1e-12 relative to the largest entry;
* the reference goldens through the GPU problem driver: the elastic
  mixed u-p Newton solve (ElasticSolution3D, 1e-4) and the 4-step
  plastic-work criterion value and adjoint gradient (CriterionTest_3D, 1e-4),
  plus agreement with the oracle's direct-solve path at 1e-8;
* a multi-step plastic solve against the oracle step by step;
* the adjoint gradient against central differences of GPU forward solves;
* determinism: two assemblies are bitwise identical.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "j2_tests.json").read_text())["tests"]
PARAM_KEYS = ["youngs", "poisson", "hardening_isotropic", "hardening_kinematic", "initial_yield_stress",
              "elastic_penalty_exponent", "elastic_min_ersatz", "plastic_penalty_exponent", "plastic_min_ersatz",
              "global_penalty_exponent", "global_min_ersatz", "pressure_scaling"]


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def both(pa, dims, lengths=(1.0, 1.0, 1.0), projection_penalty=None, work_penalty=(3.0, 1e-9), **params):
    mesh = pa.Mesh.box(*dims, *lengths)
    coords, conn = oracle.box_mesh(*dims, *lengths)
    kw = dict(params)
    gpu_kw = dict(kw)
    if projection_penalty is not None:
        gpu_kw.update(projection_penalty=1, projection_penalty_exponent=projection_penalty[0],
                      projection_min_ersatz=projection_penalty[1])
    gpu_kw.update(work_penalty_exponent=work_penalty[0], work_min_ersatz=work_penalty[1])
    pl = pa.Plasticity(mesh, **gpu_kw)
    om = oracle.J2Model(coords, conn, projection_penalty=projection_penalty, work_penalty=work_penalty, **kw)
    return mesh, pl, om


def random_state(om, seed, scale=2e-3):
    rng = np.random.default_rng(seed)
    X = om.coords.reshape(-1, 3)
    u = rng.uniform(-1, 1, om.N * 4) * scale
    u[3::4] *= 100.0
    u.reshape(-1, 4)[:, :3] *= (X[:, 0:1] > 0.3 * X[:, 0].max())
    cp = np.zeros((om.C, 14))
    cp[:, 2:8] = rng.uniform(-1, 1, (om.C, 6)) * 1e-4
    cp[:, 8:] = rng.uniform(-1, 1, (om.C, 6)) * 0.05
    cp[:, 2:5] -= cp[:, 2:5].mean(axis=1, keepdims=True)
    cp[:, 8:11] -= cp[:, 8:11].mean(axis=1, keepdims=True)
    cp[:, 0] = rng.uniform(0, 1e-2, om.C)
    z = rng.uniform(0.5, 1.0, om.N)
    pi = rng.uniform(-1, 1, om.N * 3)
    return u, cp.ravel(), z, pi


MAT = dict(youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=5.0, initial_yield_stress=1.0,
           elastic_penalty_exponent=3.0, elastic_min_ersatz=1e-9, plastic_penalty_exponent=2.5,
           plastic_min_ersatz=1e-9, global_penalty_exponent=3.0, global_min_ersatz=1e-9, pressure_scaling=2.0)


@pytest.fixture(scope="module")
def state(pa):
    print("case: J2 parity state on 4x2x2, E=1e3", flush=True)
    mesh, pl, om = both(pa, (4, 2, 2), projection_penalty=(3.0, 1e-3), **MAT)
    u, cp, z, pi = random_state(om, 11)
    c = om.update_local_state(u, cp, z)
    cc = c.reshape(-1, 14)
    assert (cc[:, 1] > 0).sum() > 10 and (cc[:, 1] == 0).sum() > 10
    return mesh, pl, om, u, c, cp, z, pi


def test_update_and_local_value(state):
    mesh, pl, om, u, c, cp, z, pi = state
    cg = pl.update_local_state(u, cp, z).numpy()
    assert rel(cg, c) < 1e-12
    assert np.array_equal(cg.reshape(-1, 14)[:, 1] > 0, c.reshape(-1, 14)[:, 1] > 0)
    Hg = pl.local_value(u, c, cp, z).numpy()
    Ho = om.local_value(u, c, cp, z)
    # at the return-map state H vanishes up to rounding: compare against the oracle at a perturbed c
    c2 = c + np.random.default_rng(2).uniform(-1, 1, c.size) * 1e-4 * (np.abs(c) + 1e-3)
    c2.reshape(-1, 14)[:, 1] = c.reshape(-1, 14)[:, 1]
    assert rel(pl.local_value(u, c2, cp, z).numpy(), om.local_value(u, c2, cp, z)) < 1e-12
    assert np.max(np.abs(Hg - Ho)) < 1e-12 * max(1.0, np.max(np.abs(c)))


@pytest.mark.parametrize("wrt", ["c", "u", "z", "cp"])
def test_local_gradients(pa, state, wrt):
    mesh, pl, om, u, c, cp, z, pi = state
    code = {"c": pa.WRT_C, "u": pa.WRT_U, "z": pa.WRT_Z, "cp": pa.WRT_C_PREV}[wrt]
    got = pl.local_gradient(code, u, c, cp, z)
    ref = om.local_gradient(wrt, u, c, cp, z)
    assert rel(got, ref) < 1e-12, wrt


def test_global_value(state):
    mesh, pl, om, u, c, cp, z, pi = state
    assert rel(pl.global_value(u, c, pi, z).numpy(), om.global_value(u, c, pi, z)) < 1e-12


@pytest.mark.parametrize("wrt", ["u", "c", "z", "pi"])
def test_global_gradients(pa, state, wrt):
    mesh, pl, om, u, c, cp, z, pi = state
    code = {"u": pa.WRT_U, "c": pa.WRT_C, "z": pa.WRT_Z, "pi": pa.WRT_PI}[wrt]
    got = pl.global_gradient(code, u, c, pi, z)
    ref = om.global_gradient_ws(wrt, u, c, oracle.gather_nodal(om.conn, pi, 3), z)
    assert rel(got, ref) < 1e-12, wrt


def test_condensed_system_and_transpose(state):
    mesh, pl, om, u, c, cp, z, pi = state
    K, R = pl.condensed_system(u, c, cp, pi, z)
    vals, Rref = om.condensed_system(u, c, cp, pi, z)
    assert rel(K.entries(), vals) < 1e-12
    assert rel(R.numpy(), Rref) < 1e-12
    KT, none = pl.condensed_system(u, c, cp, pi, z, transpose=True)
    assert none is None
    A = oracle.to_scipy(4, 4, om.rowmap, om.colidx, vals)
    AT = oracle.to_scipy(4, 4, om.rowmap, om.colidx, KT.entries())
    assert abs(AT - A.T).max() <= 1e-12 * abs(A).max()
    # determinism
    K2, R2 = pl.condensed_system(u, c, cp, pi, z)
    assert np.array_equal(K2.entries(), K.entries()) and np.array_equal(R2.numpy(), R.numpy())


def test_projection_and_plastic_work(state):
    mesh, pl, om, u, c, cp, z, pi = state
    assert rel(pl.projection(u, z).numpy(), om.projection(u, z)) < 1e-12
    up = u * 0.5
    w_gpu = pl.plastic_work(u, up, c, cp, z)
    w_ref = om.plastic_work(u, up, c, cp, z)
    assert abs(w_gpu - w_ref) <= 1e-12 * abs(w_ref)


def node_set(coords, axis, value):
    X = coords.reshape(-1, 3)
    return np.nonzero(np.abs(X[:, axis] - value) < 1e-12)[0]


def test_elastic_solution_golden(pa):
    print("case: ElasticSolution3D golden through the GPU Newton driver", flush=True)
    g = GOLD["ElastoPlasticity_ElasticSolution3D"]
    inp = g["inputs"]
    mesh, pl, om = both(pa, (1, 1, 1), **{k: inp[k] for k in PARAM_KEYS if k in inp})
    coords = om.coords
    xm, xp = node_set(coords, 0, 0.0), node_set(coords, 0, 1.0)
    dofs = np.concatenate([xm * 4, xm * 4 + 1, xm * 4 + 2, xp * 4 + 1]).astype(np.int32)
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, -0.001)])
    prob = pa.PlasticityProblem(pl, dofs, vals, num_steps=1, max_newton=inp["max_newton"])
    prob.solution(np.ones(om.N))
    u, c, _ = prob.state(0)
    gold = np.asarray(g["gold"])
    nz = gold != 0
    assert np.all(np.abs(u[nz] - gold[nz]) <= inp["tolerance"] * np.abs(gold[nz]))
    us, cs, pis, its = om.solve(np.ones(om.N), dofs, vals, num_steps=1, max_newton=inp["max_newton"])
    assert rel(u, us[0]) < 1e-8


def criterion_problem(pa, z_value=None):
    g = GOLD["ElastoPlasticity_CriterionTest_3D"]
    inp = g["inputs"]
    keys = {k: inp[k] for k in PARAM_KEYS if k in inp}
    mesh, pl, om = both(pa, (2, 2, 2), work_penalty=(inp["work_penalty_exponent"], inp["work_min_ersatz"]), **keys)
    c = om.coords
    ns = lambda ax, v: node_set(c, ax, v)
    dofs = np.concatenate([ns(0, 0) * 4, ns(1, 0) * 4 + 1, ns(1, 1) * 4 + 1, ns(2, 0) * 4 + 2,
                           ns(0, 1) * 4]).astype(np.int32)
    vals = np.zeros(dofs.size)
    vals[-ns(0, 1).size:] = 0.002
    return g, inp, pl, om, dofs, vals


def test_plastic_work_criterion_golden(pa):
    print("case: CriterionTest_3D golden (4 plastic steps) value + adjoint gradient", flush=True)
    g, inp, pl, om, dofs, vals = criterion_problem(pa)
    z = np.ones(om.N)
    prob = pa.PlasticityProblem(pl, dofs, vals, num_steps=inp["num_steps"], max_newton=inp["max_newton"])
    prob.solution(z)
    F = prob.criterion_value(z, inp["multiplier"])
    assert abs(F - g["gold_value"]) <= inp["tolerance"] * abs(g["gold_value"])
    dF = prob.criterion_gradient(z, inp["multiplier"]).numpy()
    gold = np.asarray(g["gold"])
    assert np.all(np.abs(dF - gold) <= inp["tolerance"] * np.abs(gold))
    us, cs, pis, its = om.solve(z, dofs, vals, num_steps=inp["num_steps"], max_newton=inp["max_newton"])
    for k in range(inp["num_steps"]):
        u, c, pi = prob.state(k)
        assert rel(u, us[k]) < 1e-8, k
        assert rel(c, cs[k]) < 1e-6, k
    ref = om.criterion_gradient_z(z, us, cs, pis, dofs, inp["multiplier"])
    assert rel(dF, ref) < 1e-8


def test_multistep_plastic_solve_and_adjoint_central_differences(pa):
    print("case: 3-step plastic solve on 4x2x2 with random density; adjoint vs FD", flush=True)
    mesh, pl, om = both(pa, (4, 2, 2), (2.0, 1.0, 1.0), youngs=1e3, poisson=0.3, hardening_isotropic=10.0,
                        hardening_kinematic=5.0, initial_yield_stress=1.0, elastic_penalty_exponent=3.0,
                        plastic_penalty_exponent=2.5, plastic_min_ersatz=1e-9, global_penalty_exponent=3.0,
                        pressure_scaling=1.0)
    c = om.coords
    xm, xp = node_set(c, 0, 0.0), node_set(c, 0, 2.0)
    dofs = np.concatenate([xm * 4, xm * 4 + 1, xm * 4 + 2, xp * 4]).astype(np.int32)
    # end displacement 0.003: plastic in steps 2 and 3 and inside the reference
    # Newton's convergence range (at 0.005 its full-step Newton stalls)
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, 0.003)])
    z = np.random.default_rng(5).uniform(0.6, 1.0, om.N)
    opts = dict(num_steps=3, max_newton=25, newton_stop_tol=1e-11, newton_current_norm_tol=1e-11)
    prob = pa.PlasticityProblem(pl, dofs, vals, **opts)
    assert prob.solution(z)
    us, cs, pis, its = om.solve(z, dofs, vals, num_steps=3, max_newton=25, stop_tol=1e-11, current_norm_tol=1e-11)
    for k in range(3):
        u, cst, pi = prob.state(k)
        assert rel(u, us[k]) < 1e-8, k
        assert rel(pi, pis[k]) < 1e-7, k
    assert (cs[-1].reshape(-1, 14)[:, 1] > 0).sum() > 0
    F0 = prob.criterion_value(z, -1.0)
    assert abs(F0 - om.criterion_value(z, us, cs, -1.0)) <= 1e-9 * abs(F0)
    grad = prob.criterion_gradient(z, -1.0).numpy()
    assert rel(grad, om.criterion_gradient_z(z, us, cs, pis, dofs, -1.0)) < 1e-7
    dz = np.random.default_rng(6).uniform(-1, 1, om.N)
    h = 1e-6

    def F(zz):
        pr = pa.PlasticityProblem(pl, dofs, vals, **opts)
        pr.solution(zz)
        return pr.criterion_value(zz, -1.0)

    fd = (F(z + h * dz) - F(z - h * dz)) / (2 * h)
    assert abs(fd - grad @ dz) <= 1e-6 * abs(fd)
