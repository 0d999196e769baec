"""Cell-tile assembly kernels (one workgroup per block of grid rows, each
incident cell evaluated once into LDS) against the row-pattern kernels and the
oracle.

The tile kernels keep the row kernels' entry order, diagonal chunking and
residual addition tree; the compiler's fused multiply-add choices differ
between the kernels, so the two paths agree to a few units in the last place
(bound 1e-14 relative, FP32 values 1e-7).  Against the oracle both hold the
FP64 parity bar (1e-12 relative, max norm).
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _box(pa, dims, L, tiles):
    # the streaming grid kernels are off here (they take over box meshes):
    # these tests compare the tile kernels with the row kernels
    old = {k: os.environ.get(k) for k in ("PA_TILES", "PA_GRID")}
    os.environ["PA_TILES"] = "1" if tiles else "0"
    os.environ["PA_GRID"] = "0"
    try:
        return pa.Mesh.box(*dims, *L)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _functions(pa, mesh, physics):
    if physics == "elastic":
        vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
        sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1.0, poisson=0.3, penalty_exponent=3.0,
                                      penalty_min=1e-9)
    else:
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.5,))
        vf.add_body_load(0, 1.0)
        sf = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=1.0, penalty_exponent=3.0,
                                      penalty_min=1e-9)
    return vf, sf


def _outputs(vf, sf, u, z):
    R, K = vf.value_gradient_u(u, z)
    return {
        "value": vf.value(u, z).numpy(),
        "gradient_u": vf.gradient_u(u, z).entries(),
        "gradient_u_fp32": vf.gradient_u(u, z, precision="fp32").entries(),
        "gradient_u_T": vf.gradient_u_T(u, z).entries(),
        "fused_R": R.numpy(),
        "fused_K": K.entries(),
        "energy_gradient_u": sf.gradient_u(u, z).numpy(),
    }


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("dims", [(20, 10, 10), (3, 5, 7), (1, 1, 1), (7, 2, 9), (2, 4, 4), (5, 1, 3)])
def test_tiles_match_row_patterns(pa, dims, physics):
    print(f"case: tiles vs row patterns, {dims} box, {physics}", flush=True)
    L = (2.0, 1.0, 1.0)
    mt, mr = _box(pa, dims, L, True), _box(pa, dims, L, False)
    nt, ntp, nrp = mt.assembly_info()
    assert nt > 0 and 0 < ntp <= nt
    assert mr.assembly_info()[0] == 0 and mr.assembly_info()[2] > 0
    rng = np.random.default_rng(sum(dims))
    N = mt.num_nodes
    dpn = 3 if physics == "elastic" else 1
    u = rng.standard_normal(dpn * N) * 1e-2
    z = rng.uniform(0.3, 1.0, N)
    a = _outputs(*_functions(pa, mt, physics), u, z)
    b = _outputs(*_functions(pa, mr, physics), u, z)
    for k in a:
        assert rel(a[k], b[k]) < (1e-7 if k.endswith("fp32") else 1e-14), k
    # the transposed assembly of the self-adjoint operator is the direct one
    assert np.array_equal(a["gradient_u"], a["gradient_u_T"])


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
def test_tiles_vs_oracle(pa, physics):
    print(f"case: tiles vs oracle, 9x6x5 box, {physics}", flush=True)
    dims, L = (9, 6, 5), (2.0, 1.0, 1.0)
    mesh = _box(pa, dims, L, True)
    assert mesh.assembly_info()[0] > 0
    coords, conn = pa.box_mesh(*dims, *L)
    fc, fl = oracle.side_set(coords, conn, 0, L[0])
    rng = np.random.default_rng(11)
    N = mesh.num_nodes
    z = rng.uniform(0.3, 1.0, N)
    if physics == "elastic":
        u = rng.standard_normal(3 * N) * 1e-2
        om = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9,
                          face_cells=fc, face_lnodes=fl, flux=(0.0, 0.0, -1.0))
    else:
        u = rng.standard_normal(N)
        om = oracle.Model(coords, conn, conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9, face_cells=fc,
                          face_lnodes=fl, flux=(0.5, 0.0, 0.0), body=(0, 1.0))
    vf, sf = _functions(pa, mesh, physics)
    Rref, Kref = om.value(physics, u, z), om.gradient_u(physics, u, z)
    assert rel(vf.value(u, z).numpy(), Rref) < TOL64
    assert rel(vf.gradient_u(u, z).entries(), Kref) < TOL64
    R2, K2 = vf.value_gradient_u(u, z)
    assert rel(R2.numpy(), Rref) < TOL64 and rel(K2.entries(), Kref) < TOL64
    assert rel(sf.gradient_u(u, z).numpy(), om.energy_gradient_u(physics, u, z)) < TOL64


def test_renumbered_mesh_keeps_row_kernels(pa):
    """A node numbering that is not a lexicographic grid gets no tiles and the
    row kernels still match the oracle."""
    print("case: renumbered 6x4x3 box, elastic", flush=True)
    coords, conn = pa.box_mesh(6, 4, 3, 2.0, 1.0, 1.0)
    N = coords.size // 3
    perm = np.random.default_rng(5).permutation(N)  # new id of each old node
    inv = np.argsort(perm)
    coords2 = coords.reshape(-1, 3)[inv].ravel()
    conn2 = perm[conn.reshape(-1)].astype(np.int32)
    mesh = pa.Mesh(coords2, conn2)
    assert mesh.assembly_info()[0] == 0
    rng = np.random.default_rng(6)
    z = rng.uniform(0.3, 1.0, N)
    u = rng.standard_normal(3 * N) * 1e-2
    om = oracle.Model(coords2, conn2, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    assert rel(vf.value(u, z).numpy(), om.value("elastic", u, z)) < TOL64
    assert rel(vf.gradient_u(u, z).entries(), om.gradient_u("elastic", u, z)) < TOL64
