"""Configuration 5 at full size: SIMP elastostatics on the 96x48x48 hex grid x6
Tet4 (1,327,104 cells, 232,897 nodes, 698,691 dofs) with GRF + tanh-projection
(beta 8) densities, and the learned warm start on that grid.

* Assembly (value, gradient_u FP64 and FP32) against the oracle on the full
  grid, field seed 100: 1e-12 / 2e-6.
* The network's forward pass at full size with the committed weights, layer
  by layer against float32 torch on the network's own stored bf16 operands
  (one bf16 ulp per layer, FP32 accumulation for the output layer).
* PCG (block Jacobi, FP64) from the learned start reaches ||r|| <= 1e-8 ||b||
  and its solution equals the one from zero to 1e-6.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DIMS, LENGTHS = (96, 48, 48), (2.0, 1.0, 1.0)
WEIGHTS = Path(__file__).resolve().parents[1] / "platoanalyze_amd" / "data" / "warmstart_96x48x48.npz"
ELASTIC = dict(youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def test_config5_assembly_vs_oracle(pa):
    from platoanalyze_amd.workloads import grf_density
    print("case: configuration 5 assembly, 96x48x48, GRF seed 100, vs oracle", flush=True)
    coords, conn = pa.box_mesh(*DIMS, *LENGTHS)
    mesh = pa.Mesh.box(*DIMS, *LENGTHS)
    assert (mesh.num_cells, mesh.num_nodes) == (1327104, 232897)
    z = grf_density(DIMS, 4.0, 8.0, 0.5, 100)
    u = np.random.default_rng(100).standard_normal(3 * mesh.num_nodes) * 1e-3
    vf = pa.VectorFunction(mesh, "Mechanics", **ELASTIC)
    vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
    fc, fl = oracle.side_set(coords, conn, 0, LENGTHS[0])
    om = oracle.Model(coords, conn, face_cells=fc, face_lnodes=fl, flux=(0.0, 0.0, -1.0), **ELASTIC)
    assert rel(vf.value(u, z).numpy(), om.value("elastic", u, z)) < 1e-12
    Kref = om.gradient_u("elastic", u, z)
    assert rel(vf.gradient_u(u, z).entries(), Kref) < 1e-12
    assert rel(vf.gradient_u(u, z, precision="fp32").entries(), Kref) < 2e-6


@pytest.fixture(scope="module")
def learned(pa):
    from platoanalyze_amd import warmstart as ws
    assert WEIGHTS.exists()
    net, scale = ws.load(str(WEIGHTS), DIMS)
    return ws, net, scale


def test_config5_network_forward_layers(pa, learned):
    torch = pytest.importorskip("torch")
    ws, net, scale = learned
    print("case: configuration 5 U-Net forward at 96x48x48 vs float32 torch, layer by layer", flush=True)
    feat = ws.Featurizer(DIMS)
    z = ws.densities(DIMS, [101], workers=1)[0]
    feat.input(pa.DeviceArray.from_numpy(z), net.x0)
    net.forward()
    flat = net.state()
    D, H, W = DIMS

    def t(view):
        v = ws.bf16_to_f32(view.numpy())
        d, h, w = view.grid
        return torch.tensor(v.reshape(d, h, w, -1)).permute(3, 0, 1, 2).unsqueeze(0).contiguous()

    def from_torch(x):
        return x.squeeze(0).permute(1, 2, 3, 0).reshape(-1, x.shape[1]).detach().numpy()

    def weights(name, cin, cout):
        a, b = net.offsets[name]
        w = ws.bf16_to_f32(ws.f32_to_bf16(flat[a:b]))
        wt = torch.tensor(w.reshape(3, 3, 3, cin, cout)).permute(4, 3, 0, 1, 2).contiguous()
        return wt, torch.tensor(flat[b:b + cout])

    shapes = {name: (cin, cout) for name, cin, cout, _ in ws.LAYERS}
    fwd = [("enc0a", net.x0, net.a0, 1), ("enc0b", net.a0, net.e0, 1), ("enc1a", net.p1, net.a1, 1),
           ("enc1b", net.a1, net.e1, 1), ("enc2a", net.p2, net.a2, 1), ("enc2b", net.a2, net.e2, 1),
           ("enc3a", net.p3, net.a3, 1), ("enc3b", net.a3, net.e3, 1), ("dec2", net.cat2, net.d2, 1),
           ("dec1", net.cat1, net.d1, 1), ("dec0", net.cat0, net.d0, 1)]
    with torch.no_grad():
        for name, x, y, relu in fwd:
            w, b = weights(name, *shapes[name])
            ref = torch.nn.functional.conv3d(t(x), w, b, padding=1)
            ref = from_torch(torch.relu(ref) if relu else ref).astype(np.float64)
            got = ws.bf16_to_f32(y.numpy()).astype(np.float64)
            bad = np.abs(got - ref) > 2.0 ** -7 * np.abs(ref) + 1e-6 * np.max(np.abs(ref))
            assert not bad.any(), (name, int(bad.sum()))
        w, b = weights("out", *shapes["out"])
        ref = from_torch(torch.nn.functional.conv3d(t(net.d0), w, b, padding=1))
        assert rel(net.out.numpy(), ref) < 1e-5
        pooled = from_torch(torch.nn.functional.avg_pool3d(t(net.e0), 2)).astype(np.float64)
        got = ws.bf16_to_f32(net.p1.numpy()).astype(np.float64)
        assert not (np.abs(got - pooled) > 2.0 ** -7 * np.abs(pooled) + 1e-6 * np.max(np.abs(pooled))).any()


def test_config5_pcg_from_learned_start(pa, learned):
    ws, net, scale = learned
    print("case: configuration 5 PCG to 1e-8 from the learned start, field seed 102", flush=True)
    box = ws.ElasticBox(DIMS, LENGTHS)
    feat = ws.Featurizer(DIMS)
    rec = ws.evaluate(net, scale, box, feat, [102], tol=1e-8, multigrid=False)[0]
    assert rec["bj_zero"]["rel_residual"] <= 1e-8
    assert rec["bj_warm"]["rel_residual"] <= 1e-8
    assert rec["solutions_rel_diff"] < 1e-6
