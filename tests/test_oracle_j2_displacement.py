"""CPU checks that pin the oracle's pure-displacement J2 formulation (3 dof per
node).  It has no reference counterpart; what pins it:

* in the elastic regime (no plastic strain) its global residual is the
  reference-pinned elastostatic residual (Mechanics, SIMP on the whole
  stiffness) with the same E, nu and penalty;
* every dual-number partial of its global residual and plastic work against
  central differences, and the pressure slots of the cell worksets are zero;
* the closed-form radial return zeroes the (shared) local residual;
* the condensed tangent is the derivative of the reduced residual;
* the adjoint gradient of the plastic work against central differences of the
  whole forward solve.
"""
import numpy as np
import pytest

import oracle


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def central(f, x, k, h):
    xp, xm = x.copy(), x.copy()
    xp[k] += h
    xm[k] -= h
    return (f(xp) - f(xm)) / (2 * h)


def test_elastic_regime_is_elastostatics():
    coords, conn = oracle.box_mesh(3, 2, 2, 2.0, 1.0, 1.0)
    m = oracle.J2Model(coords, conn, youngs=7.0, poisson=0.27, hardening_isotropic=1.0, hardening_kinematic=0.0,
                       initial_yield_stress=1.0, global_penalty_exponent=3.0, global_min_ersatz=1e-3,
                       formulation="displacement")
    el = oracle.Model(coords, conn, youngs=7.0, poisson=0.27, penalty_exponent=3.0, penalty_min=1e-3)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(m.N * 3) * 1e-2
    z = rng.uniform(0.2, 1.0, m.N)
    c = np.zeros(m.C * 14)
    R = m.global_value(u, c, np.zeros(m.N * 3), z)
    assert rel(R, el.value("elastic", u, z)) < 1e-13
    # and its Jacobian with no plastic flow is the elastic stiffness
    vals, Rt = m.condensed_system(u, c, c, np.zeros(m.N * 3), z)
    assert rel(vals, el.gradient_u("elastic", u, z)) < 1e-13
    assert rel(Rt, R) < 1e-13


@pytest.fixture(scope="module")
def plastic_state():
    coords, conn = oracle.box_mesh(2, 1, 1)
    m = oracle.J2Model(coords, conn, youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=5.0,
                       initial_yield_stress=1.0, elastic_penalty_exponent=3.0, elastic_min_ersatz=1e-9,
                       plastic_penalty_exponent=2.5, plastic_min_ersatz=1e-9, global_penalty_exponent=3.0,
                       formulation="displacement")
    rng = np.random.default_rng(1)
    X = coords.reshape(-1, 3)
    u = rng.uniform(-1, 1, m.N * 3) * 2e-3
    u.reshape(-1, 3)[:] *= (X[:, 0:1] > 0.9)  # the first hex column stays elastic
    cp = np.zeros((m.C, 14))
    cp[:, 2:8] = rng.uniform(-1, 1, (m.C, 6)) * 1e-4
    cp[:, 8:] = rng.uniform(-1, 1, (m.C, 6)) * 0.05
    cp[:, 2:5] -= cp[:, 2:5].mean(axis=1, keepdims=True)
    cp[:, 8:11] -= cp[:, 8:11].mean(axis=1, keepdims=True)
    cp[:, 0] = rng.uniform(0, 1e-2, m.C)
    cp = cp.ravel()
    z = rng.uniform(0.5, 1.0, m.N)
    c = m.update_local_state(u, cp, z)
    cc = c.reshape(m.C, 14)
    assert (cc[:, 1] > 1e-8).sum() >= 2 and (cc[:, 1] == 0).sum() >= 1  # both branches present
    return m, u, c, cp, z


def test_global_partials_central_differences(plastic_state):
    m, u, c, cp, z = plastic_state
    cn = m.conn.reshape(-1, 4)
    pi_ws = np.zeros((m.C, 12))
    for which in ("u", "c", "z"):
        J = m.global_gradient_ws(which, u, c, pi_ws, z)
        assert np.all(J[:, 3::4, :] == 0.0)  # the pressure rows are empty
        if which == "u":
            assert np.all(J[:, :, 3::4] == 0.0)
        for e in (0, 5, 11):
            scale = max(np.max(np.abs(J[e])), 1e-12)
            for k in range(J.shape[2]):
                if which == "u":
                    if k % 4 == 3:
                        continue
                    f = lambda x: m.global_value_ws(x, c, pi_ws, z)[e]
                    fd = central(f, u, cn[e, k // 4] * 3 + k % 4, 1e-7)
                elif which == "c":
                    f = lambda x: m.global_value_ws(u, x, pi_ws, z)[e]
                    fd = central(f, c, e * 14 + k, 1e-7)
                else:
                    def f(dzk, e=e, k=k):
                        zz = z.copy()
                        zz[cn[e, k]] += dzk
                        return m.global_value_ws(u, c, pi_ws, zz)[e]
                    fd = (f(1e-7) - f(-1e-7)) / 2e-7
                assert np.max(np.abs(J[e, :, k] - fd)) <= 1e-6 * scale, (which, e, k)


def test_local_u_partials_central_differences(plastic_state):
    m, u, c, cp, z = plastic_state
    cn = m.conn.reshape(-1, 4)
    J = m.local_gradient("u", u, c, cp, z)
    assert np.all(J[:, :, 3::4] == 0.0)
    for e in (0, 3, 7):
        scale = max(np.max(np.abs(J[e])), 1e-12)
        for k in range(16):
            if k % 4 == 3:
                continue
            f = lambda x: m.local_value(x, c, cp, z)[e * 14:(e + 1) * 14]
            fd = central(f, u, cn[e, k // 4] * 3 + k % 4, 1e-7)
            assert np.max(np.abs(J[e, :, k] - fd)) <= 1e-6 * scale, (e, k)


def test_radial_return_zeroes_local_residual(plastic_state):
    m, u, c, cp, z = plastic_state
    H = m.local_value(u, c, cp, z).reshape(m.C, 14)
    scale = np.max(np.abs(c.reshape(m.C, 14)), axis=1)
    assert np.all(np.max(np.abs(H), axis=1) <= 1e-12 * np.maximum(scale, 1.0))


def test_condensed_tangent_is_reduced_derivative(plastic_state):
    m, u, c, cp, z = plastic_state
    pi = np.zeros(m.N * 3)
    vals, Rt = m.condensed_system(u, c, cp, pi, z)
    R = m.global_value(u, c, pi, z)
    assert rel(Rt, R) < 1e-12
    K = oracle.to_scipy(3, 3, m.rowmap, m.colidx, vals)
    du = np.random.default_rng(3).uniform(-1, 1, u.size) * 1e-3

    def reduced(x):
        return m.global_value(x, m.update_local_state(x, cp, z), pi, z)

    h = 1e-4
    fd = (reduced(u + h * du) - reduced(u - h * du)) / (2 * h)
    c_plus = m.update_local_state(u + h * du, cp, z).reshape(m.C, 14)[:, 1] > 0
    c_minus = m.update_local_state(u - h * du, cp, z).reshape(m.C, 14)[:, 1] > 0
    assert np.array_equal(c_plus, c.reshape(m.C, 14)[:, 1] > 0) and np.array_equal(c_plus, c_minus)
    assert rel(K @ du, fd) < 1e-6


def test_plastic_work_partials_central_differences(plastic_state):
    m, u, c, cp, z = plastic_state
    rng = np.random.default_rng(7)
    up = u * 0.5
    parts = m.plastic_work_partials(u, up, c, cp, z)
    assert parts[0].size == m.N * 3 and parts[1].size == m.N * 3
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


def displacement_problem():
    coords, conn = oracle.box_mesh(3, 2, 2, 2.0, 1.0, 1.0)
    m = oracle.J2Model(coords, conn, youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=0.0,
                       initial_yield_stress=1.0, formulation="displacement")
    X = coords.reshape(-1, 3)
    left = np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]
    right = np.nonzero(np.abs(X[:, 0] - 2.0) < 1e-12)[0]
    dofs = np.concatenate([(3 * left[:, None] + np.arange(3)).ravel(), 3 * right])
    vals = np.concatenate([np.zeros(3 * left.size), np.full(right.size, 0.006)])
    order = np.argsort(dofs)
    return m, dofs[order].astype(np.int32), vals[order]


def test_forward_solve_converges_and_yields():
    m, dofs, vals = displacement_problem()
    z = np.full(m.N, 1.0)
    us, cs, pis, iters = m.solve(z, dofs, vals, num_steps=3, max_newton=20, stop_tol=1e-12, current_norm_tol=1e-12)
    assert all(it < 20 for it in iters)
    assert (cs[-1].reshape(m.C, 14)[:, 1] > 0).sum() > 0
    # the converged state satisfies R_t = 0 away from the constraints
    vals_, R = m.condensed_system(us[-1], cs[-1], cs[-2], np.zeros(m.N * 3), z)
    free = np.setdiff1d(np.arange(m.N * 3), dofs)
    assert np.max(np.abs(R[free])) < 1e-9 * np.max(np.abs(R[dofs]))


def test_adjoint_gradient_central_differences():
    m, dofs, vals = displacement_problem()
    rng = np.random.default_rng(4)
    z = rng.uniform(0.7, 1.0, m.N)
    dz = rng.uniform(-1, 1, m.N)

    def F(zz):
        us, cs, pis, _ = m.solve(zz, dofs, vals, num_steps=3, max_newton=30, stop_tol=1e-14, current_norm_tol=1e-14)
        return m.criterion_value(zz, us, cs, -1.0), (us, cs, pis)

    f0, (us, cs, pis) = F(z)
    assert (cs[-1].reshape(m.C, 14)[:, 1] > 0).sum() > 0
    grad = m.criterion_gradient_z(z, us, cs, pis, dofs, -1.0)
    h = 1e-6
    fd = (F(z + h * dz)[0] - F(z - h * dz)[0]) / (2 * h)
    assert abs(fd - grad @ dz) <= 1e-6 * abs(fd)
