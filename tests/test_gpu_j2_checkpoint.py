"""Memory-bounded checkpointing of the J2 path-dependent adjoint.

With checkpoint stride m the problem keeps only steps m-1, 2m-1, ... and the
last, and the criterion value and the backward recursion recompute the other
steps segment by segment.  Every reduction of the forward solve is
fixed-order, so the recomputed states, the criterion value and the gradient
equal the keep-every-step path bit for bit.  Both J2 formulations, strides
2, 3 (= ceil(sqrt(K))) and K (only the last step kept).

Also the adjoint's linear-solve status: a gradient whose adjoint solves stop
short of the tolerance is reported as an error, not returned silently.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

PARAMS = dict(youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=5.0, initial_yield_stress=1.0,
              elastic_penalty_exponent=3.0, plastic_penalty_exponent=2.5, plastic_min_ersatz=1e-9,
              global_penalty_exponent=3.0, pressure_scaling=1.0)
K = 7


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def setup(pa, form):
    mesh = pa.Mesh.box(6, 3, 3, 2.0, 1.0, 1.0)
    pl = pa.Plasticity(mesh, formulation=form, **PARAMS)
    D = pl.dpn
    xm, xp = mesh.node_set("x-"), mesh.node_set("x+")
    dofs = np.concatenate([xm * D, xm * D + 1, xm * D + 2, xp * D])
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, 0.006)])
    order = np.argsort(dofs)
    z = np.random.default_rng(3).uniform(0.6, 1.0, mesh.num_nodes)
    return pl, dofs[order].astype(np.int32), vals[order], z


def run(pa, pl, dofs, vals, z, stride, **kw):
    opts = dict(num_steps=K, max_newton=25, newton_stop_tol=1e-11, newton_current_norm_tol=1e-11, linear_tol=1e-12)
    opts.update(kw)
    prob = pa.PlasticityProblem(pl, dofs, vals, checkpoint_stride=stride, **opts)
    conv = prob.solution(z)
    return prob, conv


@pytest.mark.parametrize("form", ["mixed", "displacement"])
def test_checkpointed_gradient_is_bitwise_equal(pa, form):
    print(f"case: J2 checkpointing, {form}, 6x3x3, {K} steps, strides 1/2/3/{K}", flush=True)
    pl, dofs, vals, z = setup(pa, form)
    ref, conv = run(pa, pl, dofs, vals, z, 1)
    assert conv
    F = ref.criterion_value(z, -1.0)
    g = ref.criterion_gradient(z, -1.0).numpy()
    states = [ref.state(k) for k in range(K)]
    assert (states[-1][1].reshape(-1, 14)[:, 1] > 0).sum() > 0  # plastic cells
    full_bytes = ref.state_bytes()
    for m in (2, 3, K):
        prob, conv = run(pa, pl, dofs, vals, z, m)
        assert conv
        kept = [k for k in range(K) if k % m == m - 1 or k == K - 1]
        nslots = len(kept) + m - 1
        assert prob.state_bytes() * K == full_bytes * nslots
        for k in range(K):
            if k in kept:
                u, c, pi = prob.state(k)
                assert np.array_equal(u, states[k][0]) and np.array_equal(c, states[k][1]), (m, k)
                assert np.array_equal(pi, states[k][2]), (m, k)
            else:
                with pytest.raises(pa.PlatoError, match="not kept"):
                    prob.state(k)
        assert prob.criterion_value(z, -1.0) == F, m
        assert np.array_equal(prob.criterion_gradient(z, -1.0).numpy(), g), m
        # the value again after the backward sweep (segments recomputed once more)
        assert prob.criterion_value(z, -1.0) == F, m


def test_repeat_solves_are_bitwise_equal(pa):
    """Two problems solved from scratch, and a checkpointed third, on a grid
    whose constraints span several wavefronts: states, criterion value and
    gradient agree bit for bit (no order-dependent sums on the path)."""
    print("case: J2 repeat solves, displacement, 16x8x8, 4 steps, strides 1/1/2", flush=True)
    mesh = pa.Mesh.box(16, 8, 8, 2.0, 1.0, 1.0)
    pl = pa.Plasticity(mesh, formulation="displacement", **PARAMS)
    xm, xp = mesh.node_set("x-"), mesh.node_set("x+")
    dofs = np.concatenate([xm * 3, xm * 3 + 1, xm * 3 + 2, xp * 3])
    vals = np.concatenate([np.zeros(3 * xm.size), np.full(xp.size, 0.012)])
    order = np.argsort(dofs)
    dofs, vals = dofs[order].astype(np.int32), vals[order]
    z = np.random.default_rng(4).uniform(0.6, 1.0, mesh.num_nodes)
    out = []
    for m in (1, 1, 2):
        prob = pa.PlasticityProblem(pl, dofs, vals, num_steps=4, checkpoint_stride=m, max_newton=25,
                                    newton_stop_tol=1e-11, newton_current_norm_tol=1e-11, linear_tol=1e-12)
        assert prob.solution(z)
        out.append((prob.state(3)[0], prob.criterion_value(z, -1.0), prob.criterion_gradient(z, -1.0).numpy()))
    assert (out[0][0] != 0).any()
    for u, F, g in out[1:]:
        assert np.array_equal(u, out[0][0]) and F == out[0][1] and np.array_equal(g, out[0][2])


def test_gradient_matches_oracle_with_checkpoints(pa):
    print("case: J2 checkpointed adjoint vs oracle, mixed, stride 3", flush=True)
    pl, dofs, vals, z = setup(pa, "mixed")
    prob, conv = run(pa, pl, dofs, vals, z, 3)
    assert conv
    coords, conn = oracle.box_mesh(6, 3, 3, 2.0, 1.0, 1.0)
    om = oracle.J2Model(coords, conn, **PARAMS)
    us, cs, pis, its = om.solve(z, dofs, vals, num_steps=K, max_newton=25, stop_tol=1e-11, current_norm_tol=1e-11)
    g = prob.criterion_gradient(z, -1.0).numpy()
    ref = om.criterion_gradient_z(z, us, cs, pis, dofs, -1.0)
    assert np.max(np.abs(g - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_unconverged_adjoint_is_an_error(pa):
    print("case: adjoint linear solves capped at 2 iterations", flush=True)
    pl, dofs, vals, z = setup(pa, "mixed")
    prob, conv = run(pa, pl, dofs, vals, z, 1, linear_max_iter=2)
    with pytest.raises(pa.PlatoError, match="adjoint linear solve"):
        prob.criterion_gradient(z, -1.0)
