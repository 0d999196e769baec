"""Configuration 4 at full size: J2 plasticity on the 64x32x32 hex grid x6
Tet4 (393,216 cells, 70,785 nodes), E = 1e3, nu = 0.3, yield 1, isotropic
hardening 10, x- clamped, ux ramped on x+ to 0.008 over 10 steps; both the
reference's stabilized mixed u-p form (4 dofs per node) and the
pure-displacement form (3 dofs per node).

The GPU problem driver runs the first six of the ten load steps (a ramp to
0.0048 in six steps is the same ramp), with the benchmark's solver
(FGMRES(200) to 1e-8, block multigrid).  At the states of steps 2 and 6 (the
plastic zone spreading from the clamped corners) every kernel the Newton step
and the adjoint use is checked against the oracle at 1e-12: the return map
(update_local_state), the Schur-condensed tangent K_t, residual R_t and the
transposed tangent (src/NewtonRaphsonSolver.hpp:204-330), and the plastic
work increment (src/PlasticWorkCriterion.hpp:163-265).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DIMS, LENGTHS = (64, 32, 32), (2.0, 1.0, 1.0)
PARAMS = dict(youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=0.0, initial_yield_stress=1.0)
END, STEPS_OF_10, RUN = 0.008, 10, 6


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", params=["mixed", "displacement"])
def run(request):
    import platoanalyze_amd as pa
    assert pa.device_count() >= 1
    form = request.param
    print(f"case: configuration 4, J2 {form}, {RUN} of {STEPS_OF_10} load steps on the GPU", flush=True)
    mesh = pa.Mesh.box(*DIMS, *LENGTHS)
    assert (mesh.num_cells, mesh.num_nodes) == (393216, 70785)
    pl = pa.Plasticity(mesh, formulation=form, **PARAMS)
    D = pl.dpn
    xm, xp = mesh.node_set("x-"), mesh.node_set("x+")
    dofs = np.concatenate([xm * D, xm * D + 1, xm * D + 2, xp * D])
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, END * RUN / STEPS_OF_10)])
    order = np.argsort(dofs)
    dofs, vals = dofs[order].astype(np.int32), vals[order]
    N = mesh.num_nodes
    z = np.ones(N)
    prob = pa.PlasticityProblem(pl, dofs, vals, num_steps=RUN, max_newton=10, linear_tol=1e-8, restart=200)
    prob.set_multigrid(*DIMS, pre_sweeps=4, post_sweeps=4)
    assert prob.solution(z)
    coords, conn = oracle.box_mesh(*DIMS, *LENGTHS)
    om = oracle.J2Model(coords, conn, formulation=form, **PARAMS)
    states = [prob.state(k) for k in range(RUN)]
    return pa, pl, om, z, states


@pytest.mark.parametrize("k", [1, RUN - 1])
def test_config4_kernels_at_solved_states(run, k):
    pa, pl, om, z, states = run
    print(f"case: configuration 4 kernels at load step {k + 1}", flush=True)
    u, c, pi = states[k]
    up, cp, _ = states[k - 1]
    plastic = int((c.reshape(-1, 14)[:, 1] > 0).sum())
    if k == RUN - 1:
        assert plastic > 1000, plastic  # the plastic zone has spread
    # the return map from the previous step's state
    cg = pl.update_local_state(u, cp, z).numpy()
    co = om.update_local_state(u, cp, z)
    assert rel(cg, co) < 1e-12
    assert np.array_equal(cg.reshape(-1, 14)[:, 1] > 0, co.reshape(-1, 14)[:, 1] > 0)
    # the condensed Newton system and its transpose
    K, R = pl.condensed_system(u, c, cp, pi, z)
    vals, Rref = om.condensed_system(u, c, cp, pi, z)
    assert rel(K.entries(), vals) < 1e-12
    assert rel(R.numpy(), Rref) < 1e-12
    KT, _ = pl.condensed_system(u, c, cp, pi, z, transpose=True)
    D = pl.dpn
    A = oracle.to_scipy(D, D, om.rowmap, om.colidx, vals)
    AT = oracle.to_scipy(D, D, om.rowmap, om.colidx, KT.entries())
    assert abs(AT - A.T).max() <= 1e-12 * abs(A).max()
    # the plastic work increment
    w_gpu = pl.plastic_work(u, up, c, cp, z)
    w_ref = om.plastic_work(u, up, c, cp, z)
    assert abs(w_gpu - w_ref) <= 1e-12 * max(abs(w_ref), 1e-300)
