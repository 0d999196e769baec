"""CPU checks that pin the J2 plasticity oracle.

* the reference's own golden vectors (tests/golden/j2_tests.json): local
  residual, local-state update, mixed u-p global residual, the elastic
  Newton solve, the 4-step plastic-work criterion value and its adjoint
  gradient, and the batched inverse;
* every dual-number partial (local: c, u, z, c_prev; global: u, c, z, Pi;
  plastic work: u, u_prev, c, c_prev, z) against central differences;
* the closed-form radial return zeroes the local residual;
* the condensed tangent equals the derivative of the reduced residual
  R(u, c(u)) with c(u) the return map (implicit function theorem), and the
  condensed residual equals R there;
* the lumped projection reproduces the gradient of a linear pressure field;
* the adjoint gradient of the plastic work against central differences of
  the whole forward solve.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = json.loads((Path(__file__).parent / "golden" / "j2_tests.json").read_text())["tests"]


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def node_set(coords, axis, value):
    X = coords.reshape(-1, 3)
    return np.nonzero(np.abs(X[:, axis] - value) < 1e-12)[0]


def model(coords, conn, inputs, **extra):
    keys = ["youngs", "poisson", "hardening_isotropic", "hardening_kinematic", "initial_yield_stress",
            "elastic_penalty_exponent", "elastic_min_ersatz", "plastic_penalty_exponent", "plastic_min_ersatz",
            "global_penalty_exponent", "global_min_ersatz"]
    kw = {k: inputs[k] for k in keys if k in inputs}
    kw.setdefault("hardening_isotropic", 0.0)
    kw.setdefault("hardening_kinematic", 0.0)
    kw.setdefault("initial_yield_stress", 1.0)
    if "work_penalty_exponent" in inputs:
        kw["work_penalty"] = (inputs["work_penalty_exponent"], inputs["work_min_ersatz"])
    kw.update(extra)
    return oracle.J2Model(coords, conn, **kw)


def shear_state(coords, value):
    """+= value on node sets x+ (uy, uz), y+ (ux, uz), z+ (ux, uy), 4 dofs/node."""
    u = np.zeros(coords.size // 3 * 4)
    for axis, dofs in ((0, (1, 2)), (1, (0, 2)), (2, (0, 1))):
        for d in dofs:
            u[node_set(coords, axis, 1.0) * 4 + d] += value
    return u


# ---------------------------------------------------------------------------
# goldens
# ---------------------------------------------------------------------------
def test_local_residual_golden():
    g = GOLD["J2Plasticity_Evaluate3D"]
    inp = g["inputs"]
    coords, conn = oracle.box_mesh(1, 1, 1)
    m = model(coords, conn, inp)
    C, N = m.C, m.N
    H = m.local_value(shear_state(coords, inp["boundary_value"]), np.tile(inp["local_state"], C),
                      np.tile(inp["prev_local_state"], C), np.full(N, inp["control"])).reshape(C, 14)
    for e in range(C):
        assert rel(H[e], g["gold"]) < inp["tolerance"]


def test_local_update_golden():
    g = GOLD["J2Plasticity_UpdateLocalState3D"]
    inp = g["inputs"]
    coords, conn = oracle.box_mesh(1, 1, 1)
    m = model(coords, conn, inp)
    c = m.update_local_state(shear_state(coords, inp["boundary_value"]), np.tile(inp["prev_local_state"], m.C),
                             np.full(m.N, inp["control"])).reshape(m.C, 14)
    gold = np.asarray(g["gold"])
    assert gold[5] == pytest.approx(2.0 * 0.18231319)
    for e in range(m.C):
        assert np.all(np.abs(c[e] - gold) <= inp["tolerance"] * np.abs(gold))


def test_global_residual_golden():
    g = GOLD["ElastoPlasticity_Residual3D_Elastic"]
    inp = g["inputs"]
    coords, conn = oracle.box_mesh(1, 1, 1)
    m = model(coords, conn, inp)
    n = np.arange(m.N)
    u = (np.array([1e-7, 2e-7, 3e-7, 4e-7])[None, :] * n[:, None]).ravel()
    e, a, d = np.meshgrid(np.arange(m.C), np.arange(4), np.arange(3), indexing="ij")
    pi_ws = (4e-7 * (a + 1) * (d + 1) * (e + 1)).reshape(m.C, 12)
    R = m.global_value_ws(u, np.zeros(m.C * 14), pi_ws, np.ones(m.N))
    gold = np.asarray(g["gold"]).reshape(6, 16)
    assert np.all(np.abs(R - gold) <= inp["tolerance"] * np.abs(gold))


def test_elastic_solution_golden():
    g = GOLD["ElastoPlasticity_ElasticSolution3D"]
    inp = g["inputs"]
    coords, conn = oracle.box_mesh(1, 1, 1)
    m = model(coords, conn, inp)
    xm, xp = node_set(coords, 0, 0.0), node_set(coords, 0, 1.0)
    dofs = np.concatenate([xm * 4, xm * 4 + 1, xm * 4 + 2, xp * 4 + 1]).astype(np.int32)
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, -0.001)])
    us, cs, pis, its = m.solve(np.ones(m.N), dofs, vals, num_steps=1, max_newton=inp["max_newton"])
    gold = np.asarray(g["gold"])
    nz = gold != 0
    assert np.all(np.abs(us[0][nz] - gold[nz]) <= inp["tolerance"] * np.abs(gold[nz]))
    assert np.max(np.abs(us[0][~nz])) == 0.0
    assert np.all(cs[0].reshape(m.C, 14)[:, 1] == 0.0)  # elastic everywhere


def criterion_setup(z_value=1.0):
    g = GOLD["ElastoPlasticity_CriterionTest_3D"]
    inp = g["inputs"]
    coords, conn = oracle.box_mesh(2, 2, 2)
    m = model(coords, conn, inp)
    ns = lambda ax, v: node_set(coords, ax, v)
    dofs = np.concatenate([ns(0, 0) * 4, ns(1, 0) * 4 + 1, ns(1, 1) * 4 + 1, ns(2, 0) * 4 + 2,
                           ns(0, 1) * 4]).astype(np.int32)
    vals = np.zeros(dofs.size)
    vals[-ns(0, 1).size:] = 0.002
    return g, m, dofs, vals


def test_plastic_work_criterion_and_adjoint_gradient_golden():
    g, m, dofs, vals = criterion_setup()
    inp = g["inputs"]
    z = np.ones(m.N)
    us, cs, pis, its = m.solve(z, dofs, vals, num_steps=inp["num_steps"], max_newton=inp["max_newton"])
    assert (cs[-1].reshape(m.C, 14)[:, 1] > 0).sum() > 0  # the golden exercises the plastic branch
    F = m.criterion_value(z, us, cs, inp["multiplier"])
    assert abs(F - g["gold_value"]) <= inp["tolerance"] * abs(g["gold_value"])
    dF = m.criterion_gradient_z(z, us, cs, pis, dofs, inp["multiplier"])
    gold = np.asarray(g["gold"])
    assert np.all(np.abs(dF - gold) <= inp["tolerance"] * np.abs(gold))


def test_adjoint_gradient_central_differences():
    g, m, dofs, vals = criterion_setup()
    rng = np.random.default_rng(4)
    z = rng.uniform(0.7, 1.0, m.N)
    dz = rng.uniform(-1, 1, m.N)

    def F(zz):
        us, cs, pis, _ = m.solve(zz, dofs, vals, num_steps=4, max_newton=30, stop_tol=1e-14, current_norm_tol=1e-14)
        return m.criterion_value(zz, us, cs, -1.0), (us, cs, pis)

    f0, (us, cs, pis) = F(z)
    assert (cs[-1].reshape(m.C, 14)[:, 1] > 0).sum() > 0
    grad = m.criterion_gradient_z(z, us, cs, pis, dofs, -1.0)
    h = 1e-6
    fd = (F(z + h * dz)[0] - F(z - h * dz)[0]) / (2 * h)
    assert abs(fd - grad @ dz) <= 1e-6 * abs(fd)


def test_batched_inverse_golden_and_identity():
    g = GOLD["ElastoPlasticity_InverseMatrixWorkset"]
    A = np.stack([np.array([[-2.0, 1.5], [1.0, -0.5]]) / (1.0 + n) for n in range(3)])
    inv = oracle.batched_inverse(A)
    gold = np.asarray(g["gold"]).reshape(2, 2)
    for n in range(3):
        assert np.all(np.abs(inv[n] - (1.0 + n) * gold) <= 1e-6 * np.abs((1.0 + n) * gold))
    rng = np.random.default_rng(0)
    B = rng.standard_normal((5, 14, 14)) + 14 * np.eye(14)
    assert np.max(np.abs(B @ oracle.batched_inverse(B) - np.eye(14))) < 1e-13


# ---------------------------------------------------------------------------
# partial derivatives against central differences
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def plastic_state():
    coords, conn = oracle.box_mesh(2, 1, 1)
    m = oracle.J2Model(coords, conn, youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=5.0,
                       initial_yield_stress=1.0, elastic_penalty_exponent=3.0, elastic_min_ersatz=1e-9,
                       plastic_penalty_exponent=2.5, plastic_min_ersatz=1e-9, global_penalty_exponent=3.0,
                       pressure_scaling=2.0)
    rng = np.random.default_rng(1)
    X = coords.reshape(-1, 3)
    u = rng.uniform(-1, 1, m.N * 4) * 2e-3
    u[3::4] *= 100.0
    u.reshape(-1, 4)[:, :3] *= (X[:, 0:1] > 0.9)  # the first hex column stays elastic
    cp = np.zeros((m.C, 14))
    cp[:, 2:8] = rng.uniform(-1, 1, (m.C, 6)) * 1e-4
    cp[:, 8:] = rng.uniform(-1, 1, (m.C, 6)) * 0.05
    # plastic strain and backstress are deviatoric along any loading path
    cp[:, 2:5] -= cp[:, 2:5].mean(axis=1, keepdims=True)
    cp[:, 8:11] -= cp[:, 8:11].mean(axis=1, keepdims=True)
    cp[:, 0] = rng.uniform(0, 1e-2, m.C)
    cp = cp.ravel()
    z = rng.uniform(0.5, 1.0, m.N)
    c = m.update_local_state(u, cp, z)
    cc = c.reshape(m.C, 14)
    assert (cc[:, 1] > 1e-8).sum() >= 2 and (cc[:, 1] == 0).sum() >= 1  # both branches present
    pi = rng.uniform(-1, 1, m.N * 3)
    return m, u, c, cp, z, pi


def central(f, x, k, h):
    xp, xm = x.copy(), x.copy()
    xp[k] += h
    xm[k] -= h
    return (f(xp) - f(xm)) / (2 * h)


def test_local_partials_central_differences(plastic_state):
    m, u, c, cp, z, pi = plastic_state
    cn = m.conn.reshape(-1, 4)
    for which in ("c", "u", "z", "cp"):
        J = m.local_gradient(which, u, c, cp, z)
        for e in (0, 3, 7):
            for k in range(J.shape[2]):
                if which == "c":
                    x0, h = c, 1e-7 * max(1.0, abs(c[e * 14 + k]))
                    f = lambda x: m.local_value(u, x, cp, z)[e * 14:(e + 1) * 14]
                    if k == 1 and c[e * 14 + 1] <= 0.0:
                        # the branch switches at c1 = 0: stay on the elastic side
                        xm = x0.copy()
                        xm[e * 14 + 1] -= h
                        fd = (f(x0) - f(xm)) / h
                    else:
                        fd = central(f, x0, e * 14 + k, h)
                elif which == "cp":
                    h = 1e-7 * max(1.0, abs(cp[e * 14 + k]))
                    f = lambda x: m.local_value(u, c, x, z)[e * 14:(e + 1) * 14]
                    fd = central(f, cp, e * 14 + k, h)
                elif which == "u":
                    gi = cn[e, k // 4] * 4 + k % 4
                    h = 1e-7
                    f = lambda x: m.local_value(x, c, cp, z)[e * 14:(e + 1) * 14]
                    fd = central(f, u, gi, h)
                else:
                    # a node shared by several cells: the workset partial is per
                    # local node, so perturb only this cell's view of z
                    h = 1e-7
                    def f(dzk, e=e, k=k):
                        zz = z.copy()
                        zz[cn[e, k]] += dzk
                        return m.local_value(u, c, cp, zz)[e * 14:(e + 1) * 14]
                    fd = (f(h) - f(-h)) / (2 * h)
                scale = max(np.max(np.abs(J[e])), 1e-12)
                assert np.max(np.abs(J[e, :, k] - fd)) <= 1e-6 * scale, (which, e, k)


def test_global_partials_central_differences(plastic_state):
    m, u, c, cp, z, pi = plastic_state
    cn = m.conn.reshape(-1, 4)
    pi_ws = oracle.gather_nodal(m.conn, pi, 3)
    for which in ("u", "c", "z", "pi"):
        J = m.global_gradient_ws(which, u, c, pi_ws, z)
        for e in (0, 5, 11):
            scale = max(np.max(np.abs(J[e])), 1e-12)
            for k in range(J.shape[2]):
                if which == "u":
                    f = lambda x: m.global_value_ws(x, c, pi_ws, z)[e]
                    fd = central(f, u, cn[e, k // 4] * 4 + k % 4, 1e-7)
                elif which == "c":
                    f = lambda x: m.global_value_ws(u, x, pi_ws, z)[e]
                    fd = central(f, c, e * 14 + k, 1e-7)
                elif which == "pi":
                    f = lambda x: m.global_value_ws(u, c, x.reshape(m.C, 12), z)[e]
                    fd = central(f, pi_ws.ravel(), e * 12 + k, 1e-7)
                else:
                    def f(dzk, e=e, k=k):
                        zz = z.copy()
                        zz[cn[e, k]] += dzk
                        return m.global_value_ws(u, c, pi_ws, zz)[e]
                    fd = (f(1e-7) - f(-1e-7)) / 2e-7
                assert np.max(np.abs(J[e, :, k] - fd)) <= 1e-6 * scale, (which, e, k)


def test_radial_return_zeroes_local_residual(plastic_state):
    m, u, c, cp, z, pi = plastic_state
    H = m.local_value(u, c, cp, z).reshape(m.C, 14)
    scale = np.max(np.abs(c.reshape(m.C, 14)), axis=1)
    assert np.all(np.max(np.abs(H), axis=1) <= 1e-12 * np.maximum(scale, 1.0))


def test_condensed_tangent_is_reduced_derivative(plastic_state):
    """With c(u) the return map, d/du R(u, c(u)) = dR/du - dR/dc (dH/dc)^-1 dH/du."""
    m, u, c, cp, z, pi = plastic_state
    vals, Rt = m.condensed_system(u, c, cp, pi, z)
    R = m.global_value(u, c, pi, z)
    assert rel(Rt, R) < 1e-12  # H(u, c(u)) = 0, so the condensed residual is R
    K = oracle.to_scipy(4, 4, m.rowmap, m.colidx, vals)
    rng = np.random.default_rng(3)
    du = rng.uniform(-1, 1, u.size) * 1e-3
    du[3::4] *= 100.0

    def reduced(x):
        return m.global_value(x, m.update_local_state(x, cp, z), pi, z)

    h = 1e-4
    fd = (reduced(u + h * du) - reduced(u - h * du)) / (2 * h)
    c_plus = m.update_local_state(u + h * du, cp, z).reshape(m.C, 14)[:, 1] > 0
    c_minus = m.update_local_state(u - h * du, cp, z).reshape(m.C, 14)[:, 1] > 0
    assert np.array_equal(c_plus, c.reshape(m.C, 14)[:, 1] > 0) and np.array_equal(c_plus, c_minus)
    assert rel(K @ du, fd) < 1e-6


def test_projection_of_linear_pressure():
    coords, conn = oracle.box_mesh(3, 2, 2, 1.5, 1.0, 1.0)
    m = oracle.J2Model(coords, conn, youngs=1.0, poisson=0.3, hardening_isotropic=0.0, hardening_kinematic=0.0,
                       initial_yield_stress=1.0, pressure_scaling=3.0, projection_penalty=(3.0, 1e-3))
    X = coords.reshape(-1, 3)
    a = np.array([0.7, -1.3, 2.1])
    u = np.zeros(m.N * 4)
    u[3::4] = X @ a + 0.4
    z = np.random.default_rng(5).uniform(0.1, 1.0, m.N)
    pi = m.projection(u, z).reshape(-1, 3)
    assert np.max(np.abs(pi - 3.0 * a)) < 1e-12 * np.max(np.abs(a)) * 3 * 10


def test_plastic_work_partials_central_differences(plastic_state):
    m, u, c, cp, z, pi = plastic_state
    rng = np.random.default_rng(7)
    up = u * 0.5
    parts = m.plastic_work_partials(u, up, c, cp, z)
    args = [u, up, c, cp, z]
    for i, g in enumerate(parts):
        d = rng.uniform(-1, 1, g.size)
        h = 1e-6

        def W(s):
            a = [x.copy() for x in args]
            a[i] = a[i] + s * d
            return m.plastic_work(*a)

        fd = (W(h) - W(-h)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * max(abs(fd), 1e-12), i
