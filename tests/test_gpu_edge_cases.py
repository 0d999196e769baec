"""Elliptic entry points on meshes outside the structured-box fast paths,
against the oracle (FP64, 1e-12 relative):

* a fan of 76 tets around one centre node (a triangulated sphere joined to
  its centre): the centre lies in 76 cells and its row has 41 blocks, which
  takes the wide-row kernel variants (32 lanes, up to 64 blocks per row);
* one node used by no cell (its row is the diagonal alone, all zero);
* a side set with no faces carrying a load (no contribution), next to a
  loaded one on part of the sphere.
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

import oracle

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def fan_mesh():
    rng = np.random.default_rng(11)
    P = rng.standard_normal((40, 3))
    P /= np.linalg.norm(P, axis=1)[:, None]
    tri = ConvexHull(P).simplices
    coords = np.vstack([[0.0, 0.0, 0.0], P, [3.0, 3.0, 3.0]])  # last node: in no cell
    conn = np.hstack([np.zeros((len(tri), 1), int), tri + 1]).astype(np.int32)
    # loaded faces: the outer faces (local nodes 1, 2, 3) of the cells with z > 0 centroids
    cells = np.flatnonzero(coords[conn[:, 1:]].mean(axis=1)[:, 2] > 0.0).astype(np.int32)
    lnodes = np.tile(np.array([1, 2, 3], np.int32), (cells.size, 1))
    return coords.ravel(), conn.ravel(), cells, lnodes.ravel(), rng


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_fan_mesh_all_entry_points(pa, physics):
    print(f"case: 76-tet fan mesh with an unused node, {physics}", flush=True)
    coords, conn, fc, fl, rng = fan_mesh()
    mesh = pa.Mesh(coords, conn)
    mesh.add_side_set("loaded", fc, fl)
    mesh.add_side_set("none", np.zeros(0, np.int32), np.zeros(0, np.int32))
    N = mesh.num_nodes
    assert N == 42
    z = rng.uniform(0.3, 1.0, N)
    if physics == "elastic":
        st = rng.standard_normal(3 * N) * 1e-2
        vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("loaded", (0.3, -0.2, 1.0))
        vf.add_uniform_natural_bc("none", (5.0, 5.0, 5.0))
        om = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9,
                          face_cells=fc, face_lnodes=fl, flux=(0.3, -0.2, 1.0))
        sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1.0, poisson=0.3,
                                      penalty_exponent=3.0, penalty_min=1e-9)
        dpn = 3
    else:
        st = rng.standard_normal(N)
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=2.0, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("loaded", (0.4,))
        vf.add_uniform_natural_bc("none", (5.0,))
        vf.add_body_load(0, 1.5)
        om = oracle.Model(coords, conn, conductivity=2.0, penalty_exponent=3.0, penalty_min=1e-9, face_cells=fc,
                          face_lnodes=fl, flux=(0.4, 0.0, 0.0), body=(0, 1.5))
        sf = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=2.0, penalty_exponent=3.0,
                                      penalty_min=1e-9)
        dpn = 1
    rowmap, colidx = mesh.graph()
    assert np.array_equal(rowmap, om.rowmap) and np.array_equal(colidx, om.colidx)
    assert np.diff(rowmap).max() == 41 and np.diff(rowmap)[-1] == 1
    R = vf.value(st, z).numpy()
    assert rel(R, om.value(physics, st, z)) < TOL64
    assert np.all(R[-dpn:] == 0.0)
    K = vf.gradient_u(st, z).entries()
    assert rel(K, om.gradient_u(physics, st, z)) < TOL64
    assert np.all(K[-dpn * dpn:] == 0.0)
    assert rel(vf.gradient_u_T(st, z).entries(), om.gradient_u_T(physics, st, z)) < TOL64
    R2, K2 = vf.value_gradient_u(st, z)
    assert rel(R2.numpy(), R) < TOL64 and rel(K2.entries(), K) < TOL64
    assert rel(vf.gradient_z(st, z).entries(), om.gradient_z(physics, st, z)) < TOL64
    assert rel(vf.gradient_x(st, z).entries(), om.gradient_x(physics, st, z)) < TOL64
    e = om.energy_value(physics, st, z)
    assert abs(sf.value(st, z) - e) <= TOL64 * abs(e)
    assert rel(sf.gradient_u(st, z).numpy(), om.energy_gradient_u(physics, st, z)) < TOL64
    assert rel(sf.gradient_z(st, z).numpy(), om.energy_gradient_z(physics, st, z)) < TOL64
    assert rel(sf.gradient_x(st, z).numpy(), om.energy_gradient_x(physics, st, z)) < TOL64


def test_j2_fan_mesh_all_entry_points(pa):
    """J2 plasticity (mixed u-p) on the fan: a 41-block row takes the gather
    kernel's reduced rows-per-workgroup launch."""
    print("case: J2 on the 76-tet fan mesh (41-block rows)", flush=True)
    coords, conn, _, _, rng = fan_mesh()
    coords = coords[:-3]  # drop the unused node: the lumped projection needs every node in a cell
    mesh = pa.Mesh(coords, conn)
    mat = dict(youngs=1e3, poisson=0.3, hardening_isotropic=10.0, hardening_kinematic=5.0, initial_yield_stress=1.0,
               elastic_penalty_exponent=3.0, elastic_min_ersatz=1e-9, plastic_penalty_exponent=2.5,
               plastic_min_ersatz=1e-9, global_penalty_exponent=3.0, global_min_ersatz=1e-9, pressure_scaling=2.0)
    pl = pa.Plasticity(mesh, projection_penalty=1, projection_penalty_exponent=3.0, projection_min_ersatz=1e-3,
                       work_penalty_exponent=3.0, work_min_ersatz=1e-9, **mat)
    om = oracle.J2Model(coords, conn, projection_penalty=(3.0, 1e-3), work_penalty=(3.0, 1e-9), **mat)
    N, C = om.N, om.C
    u = rng.uniform(-1, 1, N * 4) * 3e-4  # about 30 of the 76 cells plastic
    u[3::4] *= 100.0
    cp = np.zeros((C, 14))
    cp[:, 2:8] = rng.uniform(-1, 1, (C, 6)) * 1e-4
    cp[:, 8:] = rng.uniform(-1, 1, (C, 6)) * 0.05
    cp[:, 2:5] -= cp[:, 2:5].mean(axis=1, keepdims=True)
    cp[:, 8:11] -= cp[:, 8:11].mean(axis=1, keepdims=True)
    cp = cp.ravel()
    z = rng.uniform(0.5, 1.0, N)
    pi = rng.uniform(-1, 1, N * 3)
    c = om.update_local_state(u, cp, z)
    nplastic = (c.reshape(-1, 14)[:, 1] > 0).sum()
    assert 5 < nplastic < C - 5
    assert rel(pl.update_local_state(u, cp, z).numpy(), c) < TOL64
    for wrt, code in (("c", pa.WRT_C), ("u", pa.WRT_U), ("z", pa.WRT_Z), ("cp", pa.WRT_C_PREV)):
        assert rel(pl.local_gradient(code, u, c, cp, z), om.local_gradient(wrt, u, c, cp, z)) < TOL64, wrt
    assert rel(pl.global_value(u, c, pi, z).numpy(), om.global_value(u, c, pi, z)) < TOL64
    K, R = pl.condensed_system(u, c, cp, pi, z)
    vals, Rref = om.condensed_system(u, c, cp, pi, z)
    assert rel(K.entries(), vals) < TOL64 and rel(R.numpy(), Rref) < TOL64
    KT, _ = pl.condensed_system(u, c, cp, pi, z, transpose=True)
    A = oracle.to_scipy(4, 4, om.rowmap, om.colidx, vals)
    AT = oracle.to_scipy(4, 4, om.rowmap, om.colidx, KT.entries())
    assert abs(AT - A.T).max() <= 1e-12 * abs(A).max()
    assert rel(pl.projection(u, z).numpy(), om.projection(u, z)) < TOL64
    w = om.plastic_work(u, 0.5 * u, c, cp, z)
    assert abs(pl.plastic_work(u, 0.5 * u, c, cp, z) - w) <= TOL64 * abs(w)
