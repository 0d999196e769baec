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


def test_cpp_layer_reads_exodus(tmp_path):
    """readExodusMesh of the C++ host layer (its own netCDF classic reader)
    gives the mesh the Python reader gives: same sets, and the compliance deck
    assembles and solves to the Python path's results."""
    import json
    import os
    import subprocess
    from pathlib import Path

    import platoanalyze_amd as pa
    from platoanalyze_amd import exodus
    print("case: C++ readExodusMesh, 6x3x3 box", flush=True)
    coords, conn = pa.box_mesh(6, 3, 3, 2.0, 1.0, 1.0)
    X = coords.reshape(-1, 3)
    sides = {"x+": oracle.side_set(coords, conn, 0, 2.0)}
    nodes = {"x-": np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]}
    path = tmp_path / "box.exo"
    exodus.write(path, coords, conn, sides, nodes)
    binary = Path(__file__).resolve().parent / "cpp" / "reference_cases"
    r = subprocess.run([str(binary), "ExodusMesh"], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, PA_EXODUS_MESH=str(path)))
    assert r.returncode == 0, r.stderr[-2000:]
    o = json.loads(r.stdout.splitlines()[0])["outputs"]
    assert o["num_nodes"] == X.shape[0] and o["num_cells"] == conn.size // 4
    assert np.array_equal(o["x_minus"], nodes["x-"])
    mesh = pa.Mesh.from_exodus(path)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
    u, z = np.asarray(o["u"]), np.asarray(o["z"])
    assert rel(o["residual"], vf.value(u, z).numpy()) < 1e-15
    assert np.array_equal(np.asarray(o["jacobian"]), vf.gradient_u(u, z).entries())
    sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1.0, poisson=0.3, penalty_exponent=3.0,
                                  penalty_min=1e-9)
    dofs = (3 * nodes["x-"][:, None] + np.arange(3)).ravel().astype(np.int32)
    P = pa.Problem(vf, sf, np.sort(dofs), mode="fp64", tol=1e-12, max_iter=20000)
    us = P.solution(z).numpy()
    assert rel(o["state"], us) < 1e-9
    assert abs(o["compliance"] - P.criterion_value(z, us)) <= 1e-9 * abs(o["compliance"])
