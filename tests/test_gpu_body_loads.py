"""Body loads given by a Function of x, y, z (BodyLoad::get,
src/BodyLoads.hpp:54-122, the function evaluated at the cell centroids) on
the GPU against the oracle's per-cell body load.

Meshes: a skewed general mesh (row kernels), a Kuhn box grid (grid kernels)
and the same box with the grid kernels off (tile kernels).  Physics: elastic
loads on each dof (a quadratic function of x, y, z and a constant, e.g.
gravity) and a thermal source.  Checked: value, value_gradient_u's residual,
gradient_z and gradient_x at 1e-12 relative (max norm), the problem's solution
against a scipy direct solve and its fused criterion gradient against the
reference composite dfdz + gradient_z^T lambda (lambda = -u) at 1e-12.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
QUAD = "1 + x^2 - 0.5*y*z"


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def centroid_values(pa, function, coords, conn):
    X = np.asarray(coords, float).reshape(-1, 3)[np.asarray(conn).reshape(-1, 4)].mean(axis=1)
    return pa.evaluate_function(function, X[:, 0], X[:, 1], X[:, 2])


def make_mesh(pa, kind):
    if kind == "skewed":
        coords, conn = pa.box_mesh(5, 4, 3, 1.5, 1.0, 1.0)
        X = coords.reshape(-1, 3).copy()
        X += np.random.default_rng(11).uniform(-0.04, 0.04, X.shape)
        coords = X.ravel()
        return coords, conn, pa.Mesh(coords, conn)
    coords, conn = pa.box_mesh(6, 5, 4, 1.5, 1.0, 1.0)
    old = os.environ.get("PA_GRID")
    os.environ["PA_GRID"] = "1" if kind == "grid" else "0"
    try:
        mesh = pa.Mesh(coords, conn)
    finally:
        if old is None:
            del os.environ["PA_GRID"]
        else:
            os.environ["PA_GRID"] = old
    assert (mesh.grid_info() is not None) == (kind == "grid")
    return coords, conn, mesh


@pytest.mark.parametrize("kind", ["skewed", "grid", "tiles"])
@pytest.mark.parametrize("load", [("elastic", 0, QUAD), ("elastic", 1, QUAD), ("elastic", 2, "-9.81"),
                                  ("thermal", 0, "2*x + y^2 - 0.5*z")])
def test_function_body_load_vs_oracle(pa, kind, load):
    physics, dof, function = load
    print(f"case: body load '{function}' on dof {dof}, {physics}, {kind} mesh", flush=True)
    coords, conn, mesh = make_mesh(pa, kind)
    N = mesh.num_nodes
    rng = np.random.default_rng(7 + dof)
    z = rng.uniform(0.2, 1.0, N)
    f = centroid_values(pa, function, coords, conn)
    if physics == "elastic":
        vf = pa.VectorFunction(mesh, "Mechanics", youngs=2.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-3)
        om = oracle.Model(coords, conn, youngs=2.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-3,
                          body_cells=(dof, f))
        u = rng.standard_normal(3 * N) * 1e-2
    else:
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.5, penalty_exponent=3.0, penalty_min=1e-3)
        om = oracle.Model(coords, conn, conductivity=1.5, penalty_exponent=3.0, penalty_min=1e-3,
                          body_cells=(dof, f))
        u = rng.standard_normal(N)
    vf.add_body_load_function(dof, function)
    R = om.value(physics, u, z)
    assert rel(vf.value(u, z).numpy(), R) < TOL
    R2, _ = vf.value_gradient_u(u, z)
    assert rel(R2.numpy(), R) < TOL
    assert rel(vf.gradient_z(u, z).entries(), om.gradient_z(physics, u, z)) < TOL
    assert rel(vf.gradient_x(u, z).entries(), om.gradient_x(physics, u, z)) < TOL


def test_function_source_problem(pa):
    """Thermal problem with a source given by a Function: the solution against
    a scipy direct solve of the oracle's system and the fused criterion
    gradient against the reference's composite formula."""
    print("case: thermal problem with a function source, 8x6x5 box", flush=True)
    dims, L = (8, 6, 5), (1.0, 1.0, 1.0)
    coords, conn = pa.box_mesh(*dims, *L)
    mesh = pa.Mesh.box(*dims, *L)
    N = mesh.num_nodes
    z = np.random.default_rng(3).uniform(0.3, 1.0, N)
    function = "a = 2; a + sin(pi*x)*y"
    f = centroid_values(pa, function, coords, conn)
    vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_body_load_function(0, function)
    sf = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=1.0, penalty_exponent=3.0,
                                  penalty_min=1e-9)
    om = oracle.Model(coords, conn, conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9, body_cells=(0, f))
    dofs = mesh.node_set("z-").astype(np.int32)
    P = pa.Problem(vf, sf, dofs, mode="fp64", tol=1e-12, max_iter=20000)
    u = P.solution(z).numpy()
    vals = om.gradient_u("thermal", np.zeros(N), z)
    rhs = -om.value("thermal", np.zeros(N), z)
    oracle.apply_block_constraints(1, om.rowmap, om.colidx, vals, rhs, dofs, np.zeros(dofs.size))
    ud = spla.spsolve(oracle.to_scipy(1, 1, om.rowmap, om.colidx, vals).tocsc(), rhs)
    assert rel(u, ud) < 1e-9
    g = P.criterion_gradient(z, u).numpy()
    comp = om.energy_gradient_z("thermal", u, z) + oracle.block_spmv(1, 1, om.rowmap, om.colidx,
                                                                     om.gradient_z("thermal", u, z), -u)
    assert rel(g, comp) < TOL
