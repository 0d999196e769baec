"""Synthetic inputs of the benchmark workloads (host side, numpy).

Box meshes come from the library (pa_box_mesh).  Densities:
  * uniform:  z ~ U[lo, 1] per node (numpy PCG64, fixed seed);
  * grf:      nodal white noise N(0,1) smoothed by a separable Gaussian of
              standard deviation `corr` cells, standardised, mapped to
              clamp(0.5 + G/6, 0, 1) and passed through the tanh Heaviside
              projection with (beta, eta) (the formula of the reference's
              HyperbolicTangentProjection, src/HyperbolicTangentProjection.hpp:77).
"""
import numpy as np

WORKLOADS = {
    # name: (physics, (nx, ny, nz), (lx, ly, lz), density spec)
    "elastic_20x10x10": ("Mechanics", (20, 10, 10), (2.0, 1.0, 1.0), ("uniform", 0.3, 1)),
    "elastic_128x64x64": ("Mechanics", (128, 64, 64), (2.0, 1.0, 1.0), ("grf", 4.0, 8.0, 0.5, 2)),
    "thermal_192x192x192": ("Thermal", (192, 192, 192), (1.0, 1.0, 1.0), ("uniform", 0.1, 3)),
    "thermal_96x96x96": ("Thermal", (96, 96, 96), (1.0, 1.0, 1.0), ("uniform", 0.1, 3)),
}


def tanh_projection(x, beta, eta):
    num = np.tanh(beta * eta) + np.tanh(beta * (x - eta))
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    return num / den


def _gauss_smooth_axis(a, sigma, axis):
    r = int(np.ceil(4 * sigma))
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    a = np.moveaxis(a, axis, -1)
    pad = np.pad(a, [(0, 0)] * (a.ndim - 1) + [(r, r)], mode="reflect")
    out = np.zeros_like(a)
    for i, w in enumerate(k):
        out += w * pad[..., i:i + a.shape[-1]]
    return np.moveaxis(out, -1, axis)


def grf_density(dims, corr, beta, eta, seed):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nx + 1, ny + 1, nz + 1))
    for ax in range(3):
        g = _gauss_smooth_axis(g, corr, ax)
    g = (g - g.mean()) / g.std()
    t = np.clip(0.5 + g / 6.0, 0.0, 1.0)
    return tanh_projection(t, beta, eta).ravel()  # node order ix, iy, iz (z fastest)


def density(dims, spec):
    nx, ny, nz = dims
    n = (nx + 1) * (ny + 1) * (nz + 1)
    if spec[0] == "uniform":
        lo, seed = spec[1], spec[2]
        return np.random.default_rng(seed).uniform(lo, 1.0, n)
    _, corr, beta, eta, seed = spec
    return grf_density(dims, corr, beta, eta, seed)


def algorithmic_bytes(C, N, nnzb, dpn, value_bytes, with_residual=True, with_jacobian=True):
    """Unique bytes one assembly pass must touch, each array once:
    connectivity 16 C, coords 24 N, density 8 N, graph 4 nnzb + 4 (N+1),
    state 8 dpn N and residual 8 dpn N (value pass), matrix values
    value_bytes dpn^2 nnzb (Jacobian pass)."""
    b = 16 * C + 24 * N + 8 * N
    if with_residual:
        b += 16 * dpn * N
    if with_jacobian:
        b += 4 * nnzb + 4 * (N + 1) + value_bytes * dpn * dpn * nnzb
    return b
