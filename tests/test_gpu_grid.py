"""Grid kernels on Kuhn-split hex grids (the streaming value of both physics
and thermal gradient_u, the brick kernel of the elastic gradient_u) against
the oracle and against the cell-tile kernels.

The grid dimensions cover every branch of the kernels' work split: one and
several z blocks of 64/63 rows, the z block whose last lane also owns the top
row (65 and 128 node layers), several line blocks of 4/3 lines, several x
blocks, and degenerate one-hex directions; for the elastic bricks (2 x 3 x 4
rows) partial bricks in every direction.  Bars: FP64 1e-12 relative to the
oracle (max norm), FP32 values 2e-6; against the tile kernels (same
mathematics, different rounding order) 1e-13.
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 2e-6

DIMS = [(1, 1, 1), (2, 2, 2), (3, 5, 7), (20, 10, 10), (2, 3, 64), (2, 3, 127), (2, 3, 128), (3, 7, 63),
        (4, 9, 5), (40, 2, 2), (1, 13, 2), (5, 4, 3)]


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def block_transpose(rowmap, colidx, vals, d):
    """The block CRS matrix with every block (i, j) replaced by block (j, i)^T."""
    B = vals.reshape(-1, d, d)
    out = np.empty_like(B)
    for i in range(rowmap.size - 1):
        for e in range(rowmap[i], rowmap[i + 1]):
            j = colidx[e]
            f = rowmap[j] + np.searchsorted(colidx[rowmap[j]:rowmap[j + 1]], i)
            out[e] = B[f].T
    return out.ravel()


def _mesh(pa, coords, conn, grid=True):
    old = os.environ.get("PA_GRID")
    os.environ["PA_GRID"] = "1" if grid else "0"
    try:
        return pa.Mesh(coords, conn)
    finally:
        if old is None:
            del os.environ["PA_GRID"]
        else:
            os.environ["PA_GRID"] = old


def _setup(pa, mesh, coords, conn, physics, L0):
    cells, ln = pa.box_side_set(coords, conn, 0, L0)
    mesh.add_side_set("x+", cells, ln)
    fc, fl = oracle.side_set(coords, conn, 0, L0)
    if physics == "elastic":
        vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
        sf = pa.PhysicsScalarFunction(mesh, "Internal Elastic Energy", youngs=1.0, poisson=0.3,
                                      penalty_exponent=3.0, penalty_min=1e-9)
        om = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9,
                          face_cells=fc, face_lnodes=fl, flux=(0.0, 0.0, -1.0))
    else:
        vf = pa.VectorFunction(mesh, "Thermal", conductivity=2.0, penalty_exponent=3.0, penalty_min=1e-9)
        vf.add_uniform_natural_bc("x+", (0.5,))
        vf.add_body_load(0, 1.5)
        sf = pa.PhysicsScalarFunction(mesh, "Internal Thermal Energy", conductivity=2.0, penalty_exponent=3.0,
                                      penalty_min=1e-9)
        om = oracle.Model(coords, conn, conductivity=2.0, penalty_exponent=3.0, penalty_min=1e-9, face_cells=fc,
                          face_lnodes=fl, flux=(0.5, 0.0, 0.0), body=(0, 1.5))
    return vf, sf, om


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("dims", DIMS)
def test_grid_vs_oracle(pa, dims, physics):
    print(f"case: grid kernels vs oracle, {dims} box, {physics}", flush=True)
    L = (2.0, 1.0, 1.0)
    coords, conn = pa.box_mesh(*dims, *L)
    mesh = _mesh(pa, coords, conn)
    assert mesh.grid_info() == tuple(d + 1 for d in dims)
    vf, sf, om = _setup(pa, mesh, coords, conn, physics, L[0])
    rng = np.random.default_rng(sum(dims) + (7 if physics == "thermal" else 0))
    N = mesh.num_nodes
    dpn = 3 if physics == "elastic" else 1
    u = rng.standard_normal(dpn * N) * (1e-2 if dpn == 3 else 1.0)
    z = rng.uniform(0.1, 1.0, N)
    Rref = om.value(physics, u, z)
    assert rel(vf.value(u, z).numpy(), Rref) < TOL64
    assert rel(sf.gradient_u(u, z).numpy(), om.energy_gradient_u(physics, u, z)) < TOL64
    R2, K2 = vf.value_gradient_u(u, z)
    assert rel(R2.numpy(), Rref) < TOL64
    Kref = om.gradient_u(physics, u, z)
    K = vf.gradient_u(u, z).entries()
    assert rel(K, Kref) < TOL64
    assert rel(K2.entries(), Kref) < TOL64
    # FP32 values: the residual still comes from the FP64 blocks
    R3, K3 = vf.value_gradient_u(u, z, precision="fp32")
    assert rel(R3.numpy(), Rref) < TOL64
    assert rel(K3.entries(), Kref) < TOL32
    if physics == "elastic":
        # the one-pass value_gradient_u writes the same matrix as gradient_u
        assert np.array_equal(K2.entries(), K)
    assert rel(vf.gradient_u(u, z, precision="fp32").entries(), Kref) < TOL32
    assert rel(vf.gradient_u_T(u, z).entries(), om.gradient_u_T(physics, u, z)) < TOL64
    if physics == "elastic" and N < 20000:
        # block (m, n) is block (n, m) transposed, bit for bit
        rowmap, colidx = mesh.graph()
        assert np.array_equal(block_transpose(rowmap, colidx, K, 3), K)


@pytest.mark.parametrize("physics", ["elastic", "thermal"])
@pytest.mark.parametrize("dims", [(20, 10, 10), (2, 3, 64), (9, 6, 5)])
def test_grid_vs_tiles(pa, dims, physics):
    print(f"case: grid vs tile kernels, {dims} box, {physics}", flush=True)
    L = (2.0, 1.0, 1.0)
    coords, conn = pa.box_mesh(*dims, *L)
    mg, mt = _mesh(pa, coords, conn, True), _mesh(pa, coords, conn, False)
    assert mg.grid_info() is not None and mt.grid_info() is None and mt.assembly_info()[0] > 0
    rng = np.random.default_rng(3)
    N = mg.num_nodes
    dpn = 3 if physics == "elastic" else 1
    u = rng.standard_normal(dpn * N) * 1e-2
    z = rng.uniform(0.3, 1.0, N)
    a, b = _setup(pa, mg, coords, conn, physics, L[0])[0], _setup(pa, mt, coords, conn, physics, L[0])[0]
    assert rel(a.value(u, z).numpy(), b.value(u, z).numpy()) < 1e-13
    assert rel(a.gradient_u(u, z).entries(), b.gradient_u(u, z).entries()) < 1e-13
    ka = a.gradient_u(u, z, precision="fp32").entries()
    kb = b.gradient_u(u, z, precision="fp32").entries()
    assert rel(ka, kb) < 1e-7


def test_grid_deterministic(pa):
    print("case: grid kernels bitwise repeat, 12x9x70 thermal and elastic", flush=True)
    coords, conn = pa.box_mesh(12, 9, 70, 1.0, 1.0, 1.0)
    mesh = pa.Mesh(coords, conn)
    N = mesh.num_nodes
    rng = np.random.default_rng(4)
    T, z = rng.standard_normal(N), rng.uniform(0.1, 1.0, N)
    vf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    vf.add_body_load(0, 1.0)
    assert np.array_equal(vf.value(T, z).numpy(), vf.value(T, z).numpy())
    assert np.array_equal(vf.gradient_u(T, z).entries(), vf.gradient_u(T, z).entries())
    ef = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    u = rng.standard_normal(3 * N) * 1e-2
    assert np.array_equal(ef.gradient_u(u, z).entries(), ef.gradient_u(u, z).entries())
    assert np.array_equal(ef.gradient_u(u, z, precision="fp32").entries(),
                          ef.gradient_u(u, z, precision="fp32").entries())


def test_shuffled_cells_keep_grid(pa):
    """Cell order and vertex order within cells do not matter to the grid
    check or the result (the kernels take each tet's geometry in their own
    vertex order)."""
    print("case: shuffled cells and rotated vertices, 6x5x4 box", flush=True)
    coords, conn = pa.box_mesh(6, 5, 4, 2.0, 1.0, 1.0)
    rng = np.random.default_rng(8)
    cn = conn.reshape(-1, 4)[rng.permutation(conn.size // 4)]
    cn = np.stack([np.roll(c, r) for c, r in zip(cn, rng.integers(0, 4, cn.shape[0]))])
    conn2 = cn.astype(np.int32).ravel()
    mesh = pa.Mesh(coords, conn2)
    assert mesh.grid_info() == (7, 6, 5)
    N = mesh.num_nodes
    u, z = rng.standard_normal(3 * N) * 1e-2, rng.uniform(0.3, 1.0, N)
    om = oracle.Model(coords, conn2, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    vf = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)
    assert rel(vf.value(u, z).numpy(), om.value("elastic", u, z)) < TOL64
    assert rel(vf.gradient_u(u, z).entries(), om.gradient_u("elastic", u, z)) < TOL64
    T = rng.standard_normal(N)
    ot = oracle.Model(coords, conn2, conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    tf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    assert rel(tf.gradient_u(T, z).entries(), ot.gradient_u("thermal", T, z)) < TOL64


def test_other_diagonal_is_not_grid(pa):
    """A hex split around another diagonal is not a Kuhn grid: the tile path
    keeps the mesh, and results still match the oracle."""
    print("case: one hex split around another diagonal, 3x3x3 box", flush=True)
    coords, conn = pa.box_mesh(3, 3, 3, 1.0, 1.0, 1.0)
    X = coords.reshape(-1, 3)
    cn = conn.reshape(-1, 4).copy()
    # hex 13 (the centre): mirror its six tets in x (corner v -> v ^ 4), a
    # split around the v100 -> v011 diagonal
    h = 13
    ids = cn[6 * h:6 * h + 6]
    nodes = np.unique(ids)
    base = nodes.min()
    S0, S1 = 16, 4
    def corner(n):
        d = n - base
        return (d // S0) << 2 | ((d % S0) // S1) << 1 | (d % S1)
    def node(v):
        return base + (v >> 2) * S0 + ((v >> 1) & 1) * S1 + (v & 1)
    cn[6 * h:6 * h + 6] = np.vectorize(lambda n: node(corner(n) ^ 4))(ids)
    conn2 = cn.astype(np.int32).ravel()
    mesh = pa.Mesh(coords, conn2)
    assert mesh.grid_info() is None
    rng = np.random.default_rng(9)
    N = mesh.num_nodes
    T, z = rng.standard_normal(N), rng.uniform(0.3, 1.0, N)
    ot = oracle.Model(coords, conn2, conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    tf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=3.0, penalty_min=1e-9)
    assert rel(tf.value(T, z).numpy(), ot.value("thermal", T, z)) < TOL64
    assert rel(tf.gradient_u(T, z).entries(), ot.gradient_u("thermal", T, z)) < TOL64


@pytest.mark.parametrize("p", [1.0, 2.0, 2.5])
def test_grid_penalty_exponents(pa, p):
    print(f"case: grid kernels with SIMP exponent {p}, 5x4x66", flush=True)
    coords, conn = pa.box_mesh(5, 4, 66, 1.0, 1.0, 2.0)
    mesh = pa.Mesh(coords, conn)
    N = mesh.num_nodes
    rng = np.random.default_rng(10)
    T, z = rng.standard_normal(N), rng.uniform(0.1, 1.0, N)
    u = rng.standard_normal(3 * N) * 1e-2
    ot = oracle.Model(coords, conn, conductivity=1.0, penalty_exponent=p, penalty_min=1e-3, body=(0, 1.0))
    tf = pa.VectorFunction(mesh, "Thermal", conductivity=1.0, penalty_exponent=p, penalty_min=1e-3)
    tf.add_body_load(0, 1.0)
    assert rel(tf.value(T, z).numpy(), ot.value("thermal", T, z)) < TOL64
    assert rel(tf.gradient_u(T, z).entries(), ot.gradient_u("thermal", T, z)) < TOL64
    oe = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=p, penalty_min=1e-3)
    ef = pa.VectorFunction(mesh, "Mechanics", youngs=1.0, poisson=0.3, penalty_exponent=p, penalty_min=1e-3)
    assert rel(ef.value(u, z).numpy(), oe.value("elastic", u, z)) < TOL64
    assert rel(ef.gradient_u(u, z).entries(), oe.gradient_u("elastic", u, z)) < TOL64
