"""Learned warm start of the SIMP elastostatic solve (no counterpart in the
reference): a small 3D U-Net maps the voxelised density field of a box mesh
to its displacement field, which starts the preconditioned conjugate
gradients in place of zero.  It is trained on solutions the library's own
solver generates.  Every operation runs on the hand-written BF16 kernels of
csrc/pa_cnn.hip (FP32 accumulation, FP32 master weights, Adam).

Network (voxel grid D x H x W = the hex grid, divisible by 8):
  level 0: conv 16->16, conv 16->16    (input channels: density, x, y, z, 12 zero)
  level 1: pool, conv 16->32, conv 32->32
  level 2: pool, conv 32->64, conv 64->64
  level 3: pool, conv 64->64, conv 64->64
  decoder: up + concat(level 2) conv 128->64, up + concat(level 1) conv 96->32,
           up + concat(level 0) conv 48->16, conv 16->16 (outputs 0..2: u / scale)
3x3x3 convolutions with bias and ReLU (not on the last), 2x2x2 average
pooling, nearest upsampling.  The prediction maps back to nodes by averaging
the adjacent voxels.  In the solve it is scaled by the energy-optimal factor
alpha = (b . x) / (x . K x), so the start is never worse than zero in the
energy norm.
"""
import ctypes
import time

import numpy as np

from . import DeviceArray, PlatoError, _check, _sh, apply_block_constraints, lib

BF16 = np.uint16
LAYERS = [  # name, cin, cout, relu
    ("enc0a", 16, 16, 1), ("enc0b", 16, 16, 1),
    ("enc1a", 16, 32, 1), ("enc1b", 32, 32, 1),
    ("enc2a", 32, 64, 1), ("enc2b", 64, 64, 1),
    ("enc3a", 64, 64, 1), ("enc3b", 64, 64, 1),
    ("dec2", 128, 64, 1), ("dec1", 96, 32, 1), ("dec0", 48, 16, 1),
    ("out", 16, 16, 0),
]
IN_CHANNELS = 16
OUT_CHANNELS = 3
WMIN = 1e-3        # weight floor of void voxels in the weighted L2 loss
LOSS_WMIN = 1.0     # weight floor of the training loss (weights 1 + rho^3: nearly uniform in log space)
L2_WEIGHT = 0.05    # weight of the L2 term in the training loss


def _off(arr, elems):
    """Device pointer elems elements into a DeviceArray."""
    return ctypes.c_void_p(arr.ptr.value + int(elems) * arr.dtype.itemsize)


class View:
    """Channels-last tensor view (array, channel offset, stride, channels) on a voxel grid."""

    def __init__(self, arr, grid, c, stride=None, offset=0):
        self.arr, self.grid, self.c = arr, grid, c
        self.stride = c if stride is None else stride
        self.offset = offset

    @property
    def V(self):
        D, H, W = self.grid
        return D * H * W

    @property
    def ptr(self):
        return _off(self.arr, self.offset)

    def sub(self, c0, c):
        return View(self.arr, self.grid, c, self.stride, self.offset + c0)

    @staticmethod
    def new(grid, c, dtype=BF16):
        D, H, W = grid
        return View(DeviceArray(D * H * W * c, dtype), grid, c)

    def numpy(self):
        a = self.arr.numpy().reshape(-1, self.stride)[:, self.offset:self.offset + self.c]
        return a


def bf16_to_f32(a):
    return (np.asarray(a, np.uint32) << 16).view(np.float32)


def f32_to_bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


# -------------------------------------------------------------- primitives
def conv(x, wfrag, bias, cout, y=None, relu=False, yf=None, accumulate=False, stream=None):
    D, H, W = x.grid
    _check(lib.pa_cnn_conv3d(D, H, W, x.ptr, x.stride, x.c, wfrag.ptr, bias.ptr if bias is not None else None,
                             cout, y.ptr if y is not None else None, y.stride if y is not None else 0,
                             1 if relu else 0, yf.ptr if yf is not None else None, yf.stride if yf is not None else 0,
                             1 if accumulate else 0, _sh(stream)))


def pack(w, cin, cout, transposed, out, stream=None):
    _check(lib.pa_cnn_pack_weights(w, cin, cout, 1 if transposed else 0, out.ptr, _sh(stream)))


class WarmStartNet:
    def __init__(self, dims, seed=0, chunk=2048):
        D, H, W = dims
        if D % 8 or H % 8 or W % 8 or (D * H * W // 512) % 16:
            raise PlatoError("WarmStartNet: grid dimensions must be divisible by 8 and the coarsest level's voxel "
                             "count by 16")
        self.dims = (D, H, W)
        self.grids = [(D >> l, H >> l, W >> l) for l in range(4)]
        self.chunk = chunk
        rng = np.random.default_rng(seed)
        # parameters: per layer W[27][cin][cout] then b[cout], one flat FP32 array
        self.offsets, n = {}, 0
        host = []
        for name, cin, cout, relu in LAYERS:
            std = np.sqrt(2.0 / (27 * cin)) * (0.1 if name == "out" else 1.0)
            w = rng.standard_normal(27 * cin * cout) * std
            self.offsets[name] = (n, n + w.size)
            host += [w, np.zeros(cout)]
            n += w.size + cout
        self.nparams = n
        self.params = DeviceArray.from_numpy(np.concatenate(host).astype(np.float32))
        self.grads = DeviceArray(n, np.float32)
        self.m1 = DeviceArray.from_numpy(np.zeros(n, np.float32))
        self.m2 = DeviceArray.from_numpy(np.zeros(n, np.float32))
        self.step_count = 0
        self.frag = {name: DeviceArray(27 * cin * cout, BF16) for name, cin, cout, _ in LAYERS}
        self.fragT = {name: DeviceArray(27 * cin * cout, BF16) for name, cin, cout, _ in LAYERS}
        self.layer = {name: (cin, cout, relu) for name, cin, cout, relu in LAYERS}
        g0, g1, g2, g3 = self.grids
        A = View.new
        # activations (the concatenation buffers hold the skip tensors at their tails)
        self.x0 = A(g0, IN_CHANNELS)
        self.a0, self.cat0 = A(g0, 16), A(g0, 48)
        self.p1, self.a1, self.cat1 = A(g1, 16), A(g1, 32), A(g1, 96)
        self.p2, self.a2, self.cat2 = A(g2, 32), A(g2, 64), A(g2, 128)
        self.p3, self.a3, self.e3 = A(g3, 64), A(g3, 64), A(g3, 64)
        self.d2, self.d1, self.d0 = A(g2, 64), A(g1, 32), A(g0, 16)
        self.out = View.new(g0, 16, np.float32)
        self.e0, self.e1, self.e2 = self.cat0.sub(32, 16), self.cat1.sub(64, 32), self.cat2.sub(64, 64)
        # gradients
        self.dout = A(g0, 16)
        self.dd0, self.dcat0 = A(g0, 16), A(g0, 48)
        self.dd1, self.dcat1 = A(g1, 32), A(g1, 96)
        self.dd2, self.dcat2 = A(g2, 64), A(g2, 128)
        self.de3, self.da3, self.dp3 = A(g3, 64), A(g3, 64), A(g3, 64)
        self.da2, self.dp2 = A(g2, 64), A(g2, 32)
        self.da1, self.dp1 = A(g1, 32), A(g1, 16)
        self.da0 = A(g0, 16)
        V0 = D * H * W
        nchunks = (V0 + chunk - 1) // chunk
        self.wparts = DeviceArray(nchunks * 27 * 128 * 64, np.float32)
        self.bparts = DeviceArray(nchunks * 128, np.float32)
        self.loss_parts = DeviceArray((V0 + 255) // 256, np.float32)
        self.pack_all()

    # ------------------------------------------------------------ weights
    def _w(self, name, arr=None):
        a = self.params if arr is None else arr
        return _off(a, self.offsets[name][0])

    def _b(self, name, arr=None):
        a = self.params if arr is None else arr
        return _off(a, self.offsets[name][1])

    def pack_all(self, stream=None):
        for name, (cin, cout, _) in self.layer.items():
            pack(self._w(name), cin, cout, False, self.frag[name], stream)
            pack(self._w(name), cin, cout, True, self.fragT[name], stream)

    def state(self):
        return self.params.numpy()

    def load_state(self, flat):
        flat = np.asarray(flat, np.float32)
        if flat.size != self.nparams:
            raise PlatoError(f"WarmStartNet: {flat.size} parameters given, {self.nparams} expected")
        self.params = DeviceArray.from_numpy(flat)
        self.pack_all()

    # ------------------------------------------------------------ forward
    def _conv(self, name, x, y=None, yf=None, stream=None):
        cin, cout, relu = self.layer[name]
        assert x.c == cin, (name, x.c, cin)
        conv(x, self.frag[name], _Bias(self._b(name)), cout, y=y, relu=relu, yf=yf, stream=stream)

    def forward(self, stream=None):
        """Input in self.x0 (bf16); prediction (FP32, 16 channels, first 3 used) in self.out."""
        g0, g1, g2, g3 = self.grids
        s = stream
        self._conv("enc0a", self.x0, self.a0, stream=s)
        self._conv("enc0b", self.a0, self.e0, stream=s)
        _pool(self.e0, self.p1, s)
        self._conv("enc1a", self.p1, self.a1, stream=s)
        self._conv("enc1b", self.a1, self.e1, stream=s)
        _pool(self.e1, self.p2, s)
        self._conv("enc2a", self.p2, self.a2, stream=s)
        self._conv("enc2b", self.a2, self.e2, stream=s)
        _pool(self.e2, self.p3, s)
        self._conv("enc3a", self.p3, self.a3, stream=s)
        self._conv("enc3b", self.a3, self.e3, stream=s)
        _upsample(self.e3, self.cat2.sub(0, 64), s)
        self._conv("dec2", self.cat2, self.d2, stream=s)
        _upsample(self.d2, self.cat1.sub(0, 64), s)
        self._conv("dec1", self.cat1, self.d1, stream=s)
        _upsample(self.d1, self.cat0.sub(0, 32), s)
        self._conv("dec0", self.cat0, self.d0, stream=s)
        self._conv("out", self.d0, yf=self.out, stream=s)

    # ----------------------------------------------------------- backward
    def _grads(self, name, x, dy, stream=None):
        """Weight and bias gradients of layer name (input view x, output gradient dy)."""
        cin, cout, _ = self.layer[name]
        D, H, W = x.grid
        V = D * H * W
        nch = (V + self.chunk - 1) // self.chunk
        _check(lib.pa_cnn_conv3d_wgrad_parts(D, H, W, x.ptr, x.stride, cin, dy.ptr, dy.stride, cout, self.chunk,
                                             self.wparts.ptr, _sh(stream)))
        _check(lib.pa_cnn_sum_parts(27 * cin * cout, nch, self.wparts.ptr, self._w(name, self.grads), _sh(stream)))
        _check(lib.pa_cnn_bias_parts(V, dy.ptr, dy.stride, cout, self.chunk, self.bparts.ptr, _sh(stream)))
        _check(lib.pa_cnn_sum_parts(cout, nch, self.bparts.ptr, self._b(name, self.grads), _sh(stream)))

    def _dconv(self, name, dy, dx, stream=None):
        """dx = data gradient of layer name for output gradient dy."""
        cin, cout, _ = self.layer[name]
        conv(dy, self.fragT[name], None, cin, y=dx, stream=stream)

    def backward(self, stream=None):
        """Gradients of the loss whose output gradient is in self.dout, into self.grads."""
        s = stream
        self._grads("out", self.d0, self.dout, s)
        self._dconv("out", self.dout, self.dd0, s)
        _relu_bwd(self.d0, self.dd0, s)
        self._grads("dec0", self.cat0, self.dd0, s)
        self._dconv("dec0", self.dd0, self.dcat0, s)
        _upsample_bwd(self.dcat0.sub(0, 32), self.dd1, s)
        _relu_bwd(self.d1, self.dd1, s)
        self._grads("dec1", self.cat1, self.dd1, s)
        self._dconv("dec1", self.dd1, self.dcat1, s)
        _upsample_bwd(self.dcat1.sub(0, 64), self.dd2, s)
        _relu_bwd(self.d2, self.dd2, s)
        self._grads("dec2", self.cat2, self.dd2, s)
        self._dconv("dec2", self.dd2, self.dcat2, s)
        _upsample_bwd(self.dcat2.sub(0, 64), self.de3, s)
        # level 3
        _relu_bwd(self.e3, self.de3, s)
        self._grads("enc3b", self.a3, self.de3, s)
        self._dconv("enc3b", self.de3, self.da3, s)
        _relu_bwd(self.a3, self.da3, s)
        self._grads("enc3a", self.p3, self.da3, s)
        self._dconv("enc3a", self.da3, self.dp3, s)
        de2 = self.dcat2.sub(64, 64)
        _pool_bwd(self.dp3, de2, True, s)
        # level 2
        _relu_bwd(self.e2, de2, s)
        self._grads("enc2b", self.a2, de2, s)
        self._dconv("enc2b", de2, self.da2, s)
        _relu_bwd(self.a2, self.da2, s)
        self._grads("enc2a", self.p2, self.da2, s)
        self._dconv("enc2a", self.da2, self.dp2, s)
        de1 = self.dcat1.sub(64, 32)
        _pool_bwd(self.dp2, de1, True, s)
        # level 1
        _relu_bwd(self.e1, de1, s)
        self._grads("enc1b", self.a1, de1, s)
        self._dconv("enc1b", de1, self.da1, s)
        _relu_bwd(self.a1, self.da1, s)
        self._grads("enc1a", self.p1, self.da1, s)
        self._dconv("enc1a", self.da1, self.dp1, s)
        de0 = self.dcat0.sub(32, 16)
        _pool_bwd(self.dp1, de0, True, s)
        # level 0
        _relu_bwd(self.e0, de0, s)
        self._grads("enc0b", self.a0, de0, s)
        self._dconv("enc0b", de0, self.da0, s)
        _relu_bwd(self.a0, self.da0, s)
        self._grads("enc0a", self.x0, self.da0, s)

    def loss_and_grad(self, target, stream=None, weighted=True):
        """Mean squared error of the first 3 output channels against target (FP32 V x 3),
        weighted by the voxel stiffness rho^3 + WMIN of the input's density channel (an
        energy-norm proxy); gradient into self.dout."""
        V = self.out.V
        _check(lib.pa_cnn_mse(V, OUT_CHANNELS, 16, self.out.ptr, 16, target.ptr,
                              self.x0.ptr if weighted else None, self.x0.stride, WMIN, self.dout.ptr, 16,
                              self.loss_parts.ptr, _sh(stream)))
        return float(self.loss_parts.numpy().astype(np.float64).sum() / (V * OUT_CHANNELS))

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step_count += 1
        k = self.step_count
        lr_t = lr * np.sqrt(1 - b2 ** k) / (1 - b1 ** k)
        _check(lib.pa_cnn_adam(self.nparams, float(lr_t), b1, b2, eps, self.grads.ptr, self.params.ptr,
                               self.m1.ptr, self.m2.ptr, _sh(stream)))
        self.pack_all(stream)

    def energy_loss_and_grad(self, target, l2=L2_WEIGHT, stream=None):
        """Stiffness-weighted discrete H1 seminorm (+ l2 times the L2 norm) of the error of the
        first 3 output channels against target: the training loss, an energy-norm proxy
        (the norm in which the solve's error is measured); gradient into self.dout."""
        D, H, W = self.dims
        _check(lib.pa_cnn_energy_loss(D, H, W, OUT_CHANNELS, 16, self.out.ptr, 16, target.ptr, self.x0.ptr,
                                      self.x0.stride, LOSS_WMIN, l2, self.dout.ptr, 16, self.loss_parts.ptr,
                                      _sh(stream)))
        return float(self.loss_parts.numpy().astype(np.float64).sum() / (self.out.V * OUT_CHANNELS))

    def train_step(self, x0, target, lr, stream=None):
        self.x0.arr.copy_from(x0.arr, stream)
        self.forward(stream)
        loss = self.energy_loss_and_grad(target, stream=stream)
        self.backward(stream)
        self.adam(lr, stream=stream)
        return loss


class _Bias:
    def __init__(self, ptr):
        self.ptr = ptr


def _pool(x, y, stream=None):
    D, H, W = x.grid
    _check(lib.pa_cnn_pool(D, H, W, x.ptr, x.stride, x.c, y.ptr, y.stride, _sh(stream)))


def _pool_bwd(dy, dx, accumulate, stream=None):
    D, H, W = dx.grid
    _check(lib.pa_cnn_pool_bwd(D, H, W, dy.ptr, dy.stride, dx.c, dx.ptr, dx.stride, 1 if accumulate else 0,
                               _sh(stream)))


def _upsample(x, y, stream=None):
    D, H, W = y.grid
    _check(lib.pa_cnn_upsample(D, H, W, x.ptr, x.stride, x.c, y.ptr, y.stride, _sh(stream)))


def _upsample_bwd(dy, dx, stream=None):
    D, H, W = dx.grid
    _check(lib.pa_cnn_upsample_bwd(D, H, W, dy.ptr, dy.stride, dx.c, dx.ptr, dx.stride, _sh(stream)))


def _relu_bwd(y, dy, stream=None):
    _check(lib.pa_cnn_relu_bwd(y.V, y.c, y.ptr, y.stride, dy.ptr, dy.stride, _sh(stream)))


# ------------------------------------------------------------ data and use
class Featurizer:
    """Network input (bf16 V x 16: voxel density, x, y, z in [0, 1], zeros) and
    targets (FP32 V x 3: voxel displacement / scale) from nodal fields of a
    box mesh with hex counts dims."""

    def __init__(self, dims):
        self.dims = dims
        D, H, W = dims
        self.V = D * H * W
        i, j, k = np.meshgrid((np.arange(D) + 0.5) / D, (np.arange(H) + 0.5) / H, (np.arange(W) + 0.5) / W,
                              indexing="ij")
        base = np.zeros((self.V, IN_CHANNELS), np.float32)
        base[:, 1], base[:, 2], base[:, 3] = i.ravel(), j.ravel(), k.ravel()
        self.base = DeviceArray.from_numpy(base)
        self.scratch = DeviceArray(self.V * IN_CHANNELS, np.float32)

    def input(self, z_dev, out=None, stream=None):
        D, H, W = self.dims
        out = out if out is not None else View.new(self.dims, IN_CHANNELS)
        self.scratch.copy_from(self.base, stream)
        _check(lib.pa_cnn_nodes_to_voxels(D, H, W, 1, z_dev.ptr, 1.0, 0.0, self.scratch.ptr, IN_CHANNELS, 0,
                                          _sh(stream)))
        _check(lib.pa_cnn_to_bf16(self.V, IN_CHANNELS, self.scratch.ptr, IN_CHANNELS, out.ptr, out.stride,
                                  _sh(stream)))
        return out

    def target(self, u_dev, scale, out=None, stream=None):
        """Log-magnitude target q = sign(u) log(1 + |u| / scale) of the voxel displacement: it
        spans the many decades between the load-carrying material and soft patches under the
        load, which dominate the solution's energy."""
        D, H, W = self.dims
        out = out if out is not None else View.new(self.dims, OUT_CHANNELS, np.float32)
        _check(lib.pa_cnn_nodes_to_voxels(D, H, W, 3, u_dev.ptr, 1.0 / scale, 0.0, out.ptr, OUT_CHANNELS, 0,
                                          _sh(stream)))
        t = out.arr.numpy()
        out.arr.copy_from(DeviceArray.from_numpy((np.sign(t) * np.log1p(np.abs(t))).astype(np.float32)))
        return out

    def nodes(self, out_view, scale, u_dev, stream=None):
        """Nodal displacement (FP64, 3 per node) of a prediction view (inverse of the target map)."""
        D, H, W = self.dims
        q = out_view.numpy()[:, :OUT_CHANNELS].astype(np.float64)
        vox = DeviceArray.from_numpy((np.sign(q) * np.expm1(np.minimum(np.abs(q), 40.0))).astype(np.float32))
        _check(lib.pa_cnn_voxels_to_nodes(D, H, W, 3, vox.ptr, OUT_CHANNELS, scale, u_dev.ptr, _sh(stream)))
        return u_dev


def linear_system(vf, z, bc_dofs):
    """(K, b) of the Newton step from u = 0: K = gradient_u(0, z), b = -value(0, z),
    symmetric Dirichlet elimination with zero values (src/elliptic/Problem.hpp:248-281)."""
    n = vf.size()
    u0 = DeviceArray.from_numpy(np.zeros(n))
    R, K = vf.value_gradient_u(u0, z)
    b = DeviceArray.from_numpy(-R.numpy())
    apply_block_constraints(K, b, bc_dofs, np.zeros(len(bc_dofs)))
    return K, b


def energy_scaled(K, b, x, bc_dofs):
    """alpha x with alpha = (b . x) / (x . K x) after zeroing x on the Dirichlet dofs."""
    xh = x.numpy()
    xh[np.asarray(bc_dofs)] = 0.0
    x = DeviceArray.from_numpy(xh)
    Kx = DeviceArray(xh.size)
    K.spmv(x, Kx)
    xKx = float(xh @ Kx.numpy())
    alpha = float(b.numpy() @ xh) / xKx if xKx > 0 else 0.0
    return DeviceArray.from_numpy(alpha * xh), alpha


# ------------------------------------------------------- training pipeline
ELASTIC = dict(youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9)


def _grf(args):
    from .workloads import grf_density
    dims, corr, beta, eta, seed = args
    return grf_density(dims, corr, beta, eta, seed).astype(np.float64)


def densities(dims, seeds, corr=4.0, beta=8.0, eta=0.5, workers=8):
    """Nodal GRF densities with tanh projection (the benchmark's field family), one per seed,
    generated by a pool of threads (numpy releases the interpreter lock in the smoothing)."""
    from concurrent.futures import ThreadPoolExecutor
    args = [(dims, corr, beta, eta, int(s)) for s in seeds]
    if workers <= 1 or len(args) < 4:
        return [_grf(a) for a in args]
    with ThreadPoolExecutor(workers) as pool:
        return list(pool.map(_grf, args))


class ElasticBox:
    """The benchmark's elastic problem on a box: clamped x-, traction (0, 0, -1) on x+."""

    def __init__(self, dims, lengths=(2.0, 1.0, 1.0)):
        import platoanalyze_amd as pa
        self.dims, self.lengths = dims, lengths
        self.mesh = pa.Mesh.box(*dims, *lengths)
        self.vf = pa.VectorFunction(self.mesh, "Mechanics", **ELASTIC)
        self.vf.add_uniform_natural_bc("x+", (0.0, 0.0, -1.0))
        self.sf = pa.PhysicsScalarFunction(self.mesh, "Internal Elastic Energy", **ELASTIC)
        self.bc = np.sort((3 * self.mesh.node_set("x-")[:, None] + np.arange(3)).ravel()).astype(np.int32)
        self.problem = pa.Problem(self.vf, self.sf, self.bc, tol=1e-8, max_iter=20000)
        self.problem.set_multigrid(*dims)

    def solve(self, z):
        """Displacement (device, FP64) of density z by the multigrid-preconditioned solve."""
        return self.problem.solution(z)


def sample_scale(z, u):
    """Stiffness-weighted RMS displacement sqrt(sum z^3 |u|^2 / (3 sum z^3)) of one solution:
    the displacement magnitude of the load-carrying material, by which each target is
    normalised (soft pockets of the SIMP field move orders of magnitude more)."""
    w = np.asarray(z, np.float64) ** 3 + WMIN
    uu = np.asarray(u, np.float64).reshape(-1, 3)
    return float(np.sqrt(np.sum(w[:, None] * uu * uu) / (3.0 * np.sum(w))))


def generate(box, feat, seeds, log=None, workers=8):
    """Training pairs of the given seeds: inputs (bf16, n x V x 16) and targets (FP32,
    n x V x 3: the log-magnitude map of displacement / the sample's stiffness-weighted scale)
    on the device.
    The overall magnitude is left to the solve (the energy-optimal factor of the start)."""
    V, dims = feat.V, box.dims
    n = len(seeds)
    X = DeviceArray(n * V * IN_CHANNELS, BF16)
    T = DeviceArray(n * V * OUT_CHANNELS, np.float32)
    t0 = time.perf_counter()
    zs = densities(dims, seeds, workers=workers)
    if log:
        log(f"warmstart: {n} densities in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    scales = []
    for i, z in enumerate(zs):
        zd = DeviceArray.from_numpy(z)
        u = box.solve(zd)
        feat.input(zd, View(_sub_array(X, i * V * IN_CHANNELS, V * IN_CHANNELS), dims, IN_CHANNELS))
        scales.append(sample_scale(z, u.numpy()))
        feat.target(u, scales[-1], View(_sub_array(T, i * V * OUT_CHANNELS, V * OUT_CHANNELS), dims, OUT_CHANNELS))
        if log and (i + 1) % 100 == 0:
            log(f"warmstart: solved {i + 1}/{n} ({time.perf_counter() - t0:.1f} s)")
    return X, T, float(np.median(scales))


class _SubArray:
    """A DeviceArray-like window (pointer, dtype, size) into another allocation."""

    def __init__(self, base, offset, size):
        self.base, self.dtype, self.size = base, base.dtype, size
        self.nbytes = size * base.dtype.itemsize
        self._ptr = _off(base, offset)

    @property
    def ptr(self):
        return self._ptr

    def numpy(self):
        out = np.empty(self.size, self.dtype)
        _check(lib.pa_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self._ptr, self.nbytes))
        return out

    def copy_from(self, other, stream=None):
        _check(lib.pa_memcpy_d2d(self._ptr, other.ptr, self.nbytes, stream))


def _sub_array(base, offset, size):
    return _SubArray(base, offset, size)


def train(net, X, T, n, epochs, lr=2e-3, log=None, seed=0, log_every=200):
    """Adam over the n stored samples, cosine learning-rate decay; returns the loss history."""
    V = net.out.V
    rng = np.random.default_rng(seed)
    total = epochs * n
    history = []
    k = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        for i in rng.permutation(n):
            lr_k = 0.5 * lr * (1 + np.cos(np.pi * k / total))
            x = View(_sub_array(X, int(i) * V * IN_CHANNELS, V * IN_CHANNELS), net.dims, IN_CHANNELS)
            t = View(_sub_array(T, int(i) * V * OUT_CHANNELS, V * OUT_CHANNELS), net.dims, OUT_CHANNELS)
            history.append(net.train_step(x, t, lr_k))
            k += 1
            if log and k % log_every == 0:
                log(f"warmstart: step {k}/{total} loss {np.mean(history[-log_every:]):.4e} "
                    f"({time.perf_counter() - t0:.1f} s)")
    return history


def predict(net, feat, z_dev, scale, u_out=None):
    """Nodal displacement prediction (device, FP64) for density z."""
    feat.input(z_dev, net.x0)
    net.forward()
    u_out = u_out if u_out is not None else DeviceArray(3 * (np.prod(np.array(net.dims) + 1)))
    return feat.nodes(net.out, scale, u_out)


def save(path, net, scale, meta=None):
    np.savez_compressed(path, params=net.state(), scale=np.float64(scale), dims=np.array(net.dims),
                        meta=np.array(repr(meta or {})))


def load(path, dims=None):
    d = np.load(path, allow_pickle=False)
    dims = tuple(int(a) for a in d["dims"]) if dims is None else dims
    net = WarmStartNet(dims)
    net.load_state(d["params"])
    return net, float(d["scale"])


def energy_error(K, x, xs):
    """||x - xs||_K / ||xs||_K."""
    e = DeviceArray.from_numpy(x - xs)
    Ke = DeviceArray(e.size)
    K.spmv(e, Ke)
    s = DeviceArray.from_numpy(xs)
    Ks = DeviceArray(s.size)
    K.spmv(s, Ks)
    return float(np.sqrt(max(e.numpy() @ Ke.numpy(), 0.0) / (xs @ Ks.numpy())))


def evaluate(net, scale, box, feat, seeds, tol=1e-8, max_iter=50000, log=None, multigrid=True,
             block_jacobi=True):
    """Per field of the given seeds: PCG iterations and seconds to ||r|| <= tol ||b|| from zero
    and from the network's warm start (inference, voxel-to-node map and energy scaling
    included), with block-Jacobi and/or multigrid preconditioning; plus the warm start's
    relative error against the converged solution (from zero)."""
    import platoanalyze_amd as pa
    mesh, dims = box.mesh, box.dims
    assert multigrid or block_jacobi
    if block_jacobi:
        bj = pa.Solver(mesh, 3, pa.PRECOND_BLOCK_JACOBI)
    if multigrid:
        mg = pa.Multigrid(mesh, *dims, 3)
        mgs = pa.Solver(mesh, 3, pa.PRECOND_MULTIGRID)
        mgs.set_multigrid(mg)
    n = 3 * mesh.num_nodes
    out = []
    for seed, z in zip(seeds, densities(dims, seeds)):
        zd = DeviceArray.from_numpy(z)
        K, b = linear_system(box.vf, zd, box.bc)
        rec = {"seed": int(seed)}
        pa.synchronize()
        t0 = time.perf_counter()
        xp = predict(net, feat, zd, scale)
        x0, alpha = energy_scaled(K, b, xp, box.bc)
        pa.synchronize()
        t_ws = time.perf_counter() - t0
        x0h = x0.numpy()
        xs, xw_sol = None, None
        if block_jacobi:
            t0 = time.perf_counter()
            x = DeviceArray.from_numpy(np.zeros(n))
            st = bj.solve(K, b, x, tol, max_iter)
            rec["bj_zero"] = {"iterations": st["iterations"], "s": time.perf_counter() - t0,
                              "rel_residual": st["rel_residual"]}
            xs = x.numpy()
            t0 = time.perf_counter()
            xw = DeviceArray.from_numpy(x0h)
            st = bj.solve(K, b, xw, tol, max_iter)
            rec["bj_warm"] = {"iterations": st["iterations"], "s": time.perf_counter() - t0 + t_ws,
                              "rel_residual": st["rel_residual"], "warm_start_s": t_ws}
            xw_sol = xw.numpy()
        if multigrid:
            t0 = time.perf_counter()
            mg.setup(K, box.bc)
            t_setup = time.perf_counter() - t0
            x = DeviceArray.from_numpy(np.zeros(n))
            t0 = time.perf_counter()
            st = mgs.solve(K, b, x, tol, max_iter)
            rec["mg_zero"] = {"iterations": st["iterations"], "s": time.perf_counter() - t0 + t_setup,
                              "rel_residual": st["rel_residual"]}
            if xs is None:
                xs = x.numpy()
            t0 = time.perf_counter()
            xw = DeviceArray.from_numpy(x0h)
            st = mgs.solve(K, b, xw, tol, max_iter)
            rec["mg_warm"] = {"iterations": st["iterations"], "s": time.perf_counter() - t0 + t_setup + t_ws,
                              "rel_residual": st["rel_residual"]}
            if xw_sol is None:
                xw_sol = xw.numpy()
        rec["warm_start_rel_error"] = float(np.linalg.norm(x0h - xs) / np.linalg.norm(xs))
        rec["warm_start_energy_error"] = energy_error(K, x0h, xs)
        rec["alpha"] = alpha
        rec["solutions_rel_diff"] = float(np.linalg.norm(xw_sol - xs) / np.linalg.norm(xs))
        out.append(rec)
        if log:
            log(f"warmstart eval seed {seed}: " + ", ".join(
                f"{k} {v['iterations']} it {v['s']:.3f} s" for k, v in rec.items() if isinstance(v, dict)) +
                f", warm-start error {rec['warm_start_rel_error']:.3f} (energy {rec['warm_start_energy_error']:.3f}), "
                f"alpha {rec['alpha']:.3g}")
    return out


def summarize(records):
    keys = [k for k in records[0] if isinstance(records[0][k], dict)]
    s = {k: {"iterations_mean": float(np.mean([r[k]["iterations"] for r in records])),
             "s_mean": float(np.mean([r[k]["s"] for r in records]))} for k in keys}
    s["warm_start_rel_error_mean"] = float(np.mean([r["warm_start_rel_error"] for r in records]))
    s["warm_start_energy_error_mean"] = float(np.mean([r["warm_start_energy_error"] for r in records]))
    s["fields"] = len(records)
    return s


def main(argv=None):
    import argparse
    import json
    import sys

    import platoanalyze_amd as pa
    ap = argparse.ArgumentParser(description="train / evaluate the warm-start network")
    ap.add_argument("--dims", type=int, nargs=3, default=[96, 48, 48])
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=8, help="fields evaluated after training")
    ap.add_argument("--out", default=None, help="weights file to write (.npz)")
    ap.add_argument("--weights", default=None, help="evaluate these weights instead of training")
    ap.add_argument("--no-multigrid", action="store_true")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    dims = tuple(args.dims)
    box = ElasticBox(dims, (2.0, 1.0, 1.0))
    feat = Featurizer(dims)
    if args.weights:
        net, scale = load(args.weights, dims)
        hist = []
    else:
        net = WarmStartNet(dims, seed=0)
        t0 = time.perf_counter()
        X, T, scale = generate(box, feat, list(range(10000, 10000 + args.samples)), log=log)
        t_gen = time.perf_counter() - t0
        log(f"warmstart: data {t_gen:.1f} s, scale {scale:.4e}")
        t0 = time.perf_counter()
        hist = train(net, X, T, args.samples, args.epochs, lr=args.lr, log=log)
        pa.synchronize()
        t_train = time.perf_counter() - t0
        log(f"warmstart: training {t_train:.1f} s, final loss {np.mean(hist[-100:]):.4e}")
        if args.out:
            save(args.out, net, scale, {"samples": args.samples, "epochs": args.epochs, "lr": args.lr,
                                        "data_s": t_gen, "train_s": t_train})
        del X, T
    recs = evaluate(net, scale, box, feat, list(range(1, 1 + args.eval)), log=log, multigrid=not args.no_multigrid)
    res = {"dims": dims, "scale": scale, "loss_first": float(np.mean(hist[:50])) if hist else None,
           "loss_last": float(np.mean(hist[-100:])) if hist else None, "summary": summarize(recs), "fields": recs}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
