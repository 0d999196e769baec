"""A mesh read from an Exodus II file (Mesh.from_exodus) drives the GPU path
like the in-memory box: the elastic value, Jacobian and a compliance solve
agree with the box mesh's (the side-set faces come back in the Exodus side
order, so the traction term agrees to rounding, 1e-14)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_exodus_mesh_matches_box(tmp_path):
    import platoanalyze_amd as pa
    from platoanalyze_amd import exodus
    print("case: Exodus round trip 8x4x4 box, elastic value/gradient/solve", flush=True)
    dims, L = (8, 4, 4), (2.0, 1.0, 1.0)
    box = pa.Mesh.box(*dims, *L)
    coords, conn = pa.box_mesh(*dims, *L)
    X = coords.reshape(-1, 3)
    sides = {"x+": oracle.side_set(coords, conn, 0, 2.0)}
    nodes = {"x-": np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]}
    path = tmp_path / "box.exo"
    exodus.write(path, coords, conn, sides, nodes)
    mesh = pa.Mesh.from_exodus(path)
    assert mesh.num_nodes == box.num_nodes and np.array_equal(mesh.node_set("x-"), box.node_set("x-"))
    out = []
    for m in (box, mesh):
        vf = pa.VectorFunction(m, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
        sf = pa.PhysicsScalarFunction(m, "Internal Elastic Energy", youngs=1.0, poisson=0.3, penalty_exponent=3.0,
                                      penalty_min=1e-9)
        rng = np.random.default_rng(2)
        u = rng.standard_normal(3 * m.num_nodes) * 1e-2
        z = rng.uniform(0.3, 1.0, m.num_nodes)
        dofs = (3 * m.node_set("x-")[:, None] + np.arange(3)).ravel().astype(np.int32)
        P = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-11, max_iter=20000)
        us = P.solution(z).numpy()
        out.append((vf.value(u, z).numpy(), vf.gradient_u(u, z).entries(), us, P.criterion_value(z, us)))
    (R0, K0, u0, F0), (R1, K1, u1, F1) = out
    assert np.array_equal(K0, K1)
    assert rel(R1, R0) < 1e-14
    assert rel(u1, u0) < 1e-9 and abs(F1 - F0) <= 1e-9 * abs(F0)
