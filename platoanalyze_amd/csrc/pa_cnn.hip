// Kernels of the warm-start network: a small 3D U-Net on the voxelised density
// field that predicts the displacement field of the SIMP elastostatic solve
// (no counterpart in the reference; it supplies the initial guess of the
// linear solve).  BF16 activations and weights, FP32 accumulation, FP32
// master weights and Adam moments.
//
// Tensors are channels-last views: element (voxel v, channel c) of a view
// (ptr, stride) sits at ptr[v * stride + c]; a concatenation is one buffer
// whose producers write at channel offsets.  Voxels are ordered (d, h, w),
// w fastest.
//
// 3x3x3 convolutions (zero padding) run as implicit GEMMs on the matrix
// cores (v_mfma_f32_16x16x16_bf16): one wave computes 16 consecutive voxels
// x every 16-channel output block; the K loop runs over the 27 taps x 16-
// channel input chunks.  Weights are pre-packed into the B-operand fragment
// order, so a wave's weight loads are contiguous.  The data-gradient is the
// same kernel on the mirrored, transposed weights; the weight-gradient is a
// GEMM over voxels split into chunks, with per-chunk partials summed in a
// fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pa_internal.h"

namespace pa {
namespace cnn {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

void check() { PA_HIP(hipGetLastError()); }

__device__ inline float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// round to nearest even (NaN kept quiet)
__device__ inline unsigned short f2bf(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct Grid {
    int D, H, W;
};

// neighbour of voxel v through tap t (offsets (t/9-1, t/3%3-1, t%3-1) in (d, h, w)); -1 outside
__device__ inline long neighbour(const Grid& g, long v, int t) {
    const int w = (int)(v % g.W), h = (int)((v / g.W) % g.H), d = (int)(v / ((long)g.W * g.H));
    const int dd = d + t / 9 - 1, hh = h + (t / 3) % 3 - 1, ww = w + t % 3 - 1;
    if (dd < 0 || dd >= g.D || hh < 0 || hh >= g.H || ww < 0 || ww >= g.W) return -1;
    return ((long)dd * g.H + hh) * g.W + ww;
}

// ---------------------------------------------------------------------------
// convolution, forward (and data-gradient on mirrored/transposed weights)
// ---------------------------------------------------------------------------
// Y[v][co] = act(bias[co] + sum_t sum_ci X[nb(v, t)][ci] W[t][ci][co]).
// Wf: B fragments, index ((t * nck + ck) * ncb + cb) * 64 + lane, 4 bf16 each:
// K rows ck*16 + (lane/16)*4 + j, column cb*16 + lane%16.
// Lane layout of v_mfma_f32_16x16x16_bf16: A row lane%16, K (lane/16)*4 + j;
// B column lane%16, K (lane/16)*4 + j; D rows (lane/16)*4 + i, column lane%16.
template <int NCB>
__global__ __launch_bounds__(256) void k_conv_fwd(Grid g, long V, const unsigned short* __restrict__ X, int xs, int nck,
                                                  const bf16x4* __restrict__ Wf, const float* __restrict__ bias,
                                                  unsigned short* __restrict__ Y, int ys, int relu,
                                                  float* __restrict__ Yf, int yfs, int accumulate) {
    const int lane = threadIdx.x & 63;
    const long v0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (v0 >= V) return;
    const int m = lane & 15, kg = lane >> 4;
    f32x4 acc[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; ++t) {
        const long nb = neighbour(g, v0 + m, t);
        const unsigned short* xr = X + (nb < 0 ? 0 : nb) * (long)xs + kg * 4;
        const bf16x4* wt = Wf + (long)t * nck * NCB * 64 + lane;
        for (int ck = 0; ck < nck; ++ck) {
            bf16x4 a = bf16x4{0, 0, 0, 0};
            if (nb >= 0) a = *reinterpret_cast<const bf16x4*>(xr + ck * 16);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) acc[cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wt[(ck * NCB + cb) * 64], acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int co = cb * 16 + m;
        const float b = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long v = v0 + kg * 4 + i;
            float y = acc[cb][i] + b;
            if (accumulate) y += bf2f(Y[v * ys + co]);
            if (relu) y = fmaxf(y, 0.f);
            if (Y) Y[v * ys + co] = f2bf(y);
            if (Yf) Yf[v * yfs + co] = y;
        }
    }
}

// Weight gradient partials: P[chunk][t][ci][co] = sum over the chunk's voxels
// of X[nb(v, t)][ci] dY[v][co]; one wave per (t, ci block, co block, chunk).
// A operand (16 ci x 16 voxels): lane row ci = lane%16, voxels (lane/16)*4 + j;
// B operand (16 voxels x 16 co): lane column co = lane%16, voxels (lane/16)*4 + j.
__global__ __launch_bounds__(64) void k_conv_wgrad(Grid g, long V, const unsigned short* __restrict__ X, int xs,
                                                   int Cin, const unsigned short* __restrict__ dY, int dys, int Cout,
                                                   long chunk, float* __restrict__ P) {
    const int lane = threadIdx.x;
    const int nci = Cin / 16, nco = Cout / 16;
    long b = blockIdx.x;
    const int cob = (int)(b % nco);
    b /= nco;
    const int cib = (int)(b % nci);
    b /= nci;
    const int t = (int)(b % 27);
    const long ch = b / 27;
    const int m = lane & 15, kg = lane >> 4;
    const long vb = ch * chunk, ve = vb + chunk < V ? vb + chunk : V;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long v0 = vb; v0 < ve; v0 += 16) {
        bf16x4 a, bb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long v = v0 + kg * 4 + j;
            const long nb = neighbour(g, v, t);
            a[j] = nb >= 0 ? (short)X[nb * xs + cib * 16 + m] : (short)0;
            bb[j] = (short)dY[v * dys + cob * 16 + m];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
    }
    // D rows ci = (lane/16)*4 + i, column co = lane%16
    float* out = P + ((ch * 27 + t) * (long)Cin) * Cout;
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(long)(cib * 16 + kg * 4 + i) * Cout + cob * 16 + m] = acc[i];
}

// dst[k] = sum over nparts of P[p * n + k], fixed order
__global__ void k_sum_parts(long n, int nparts, const float* __restrict__ P, float* __restrict__ dst) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float s = 0.f;
    for (int p = 0; p < nparts; ++p) s += P[(long)p * n + k];
    dst[k] = s;
}

// bias gradient partials: P[block][c] = sum over the block's voxels of dY[v][c]
__global__ void k_bias_parts(long V, const unsigned short* __restrict__ dY, int dys, int C, long chunk,
                             float* __restrict__ P) {
    const int c = threadIdx.x;
    if (c >= C) return;
    const long vb = blockIdx.x * chunk, ve = vb + chunk < V ? vb + chunk : V;
    float s = 0.f;
    for (long v = vb; v < ve; ++v) s += bf2f(dY[v * dys + c]);
    P[(long)blockIdx.x * C + c] = s;
}

// Pack FP32 weights W[t][ci][co] into B fragments (bf16).  transposed: the
// data-gradient weights W'[t][co][ci] = W[26 - t][ci][co].
__global__ void k_pack_weights(const float* __restrict__ Wm, int Cin, int Cout, int transposed,
                               unsigned short* __restrict__ Wf) {
    const int K = transposed ? Cout : Cin, N = transposed ? Cin : Cout;
    const int nck = K / 16, ncb = N / 16;
    const long total = 27L * nck * ncb * 64;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int lane = (int)(q % 64);
    long r = q / 64;
    const int cb = (int)(r % ncb);
    r /= ncb;
    const int ck = (int)(r % nck);
    const int t = (int)(r / nck);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = ck * 16 + (lane >> 4) * 4 + j, n = cb * 16 + (lane & 15);
        const float w = transposed ? Wm[((long)(26 - t) * Cin + n) * Cout + k] : Wm[((long)t * Cin + k) * Cout + n];
        Wf[q * 4 + j] = f2bf(w);
    }
}

// ---------------------------------------------------------------------------
// elementwise and resampling kernels
// ---------------------------------------------------------------------------
// 2x2x2 average pooling: Y (D/2, H/2, W/2) from X (D, H, W), C channels
__global__ void k_pool(Grid gy, const unsigned short* __restrict__ X, int xs, int C, unsigned short* __restrict__ Y,
                       int ys) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long V = (long)gy.D * gy.H * gy.W;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long v = q / C;
    const int w = (int)(v % gy.W), h = (int)((v / gy.W) % gy.H), d = (int)(v / ((long)gy.W * gy.H));
    const int Hx = gy.H * 2, Wx = gy.W * 2;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const long u = ((long)(2 * d + (k >> 2)) * Hx + 2 * h + ((k >> 1) & 1)) * Wx + 2 * w + (k & 1);
        s += bf2f(X[u * xs + c]);
    }
    Y[v * ys + c] = f2bf(0.125f * s);
}

// pooling gradient: dX[u][c] (+)= dY[parent(u)][c] / 8
__global__ void k_pool_bwd(Grid gx, const unsigned short* __restrict__ dY, int dys, int C,
                           unsigned short* __restrict__ dX, int dxs, int accumulate) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long V = (long)gx.D * gx.H * gx.W;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long u = q / C;
    const int w = (int)(u % gx.W), h = (int)((u / gx.W) % gx.H), d = (int)(u / ((long)gx.W * gx.H));
    const long p = ((long)(d / 2) * (gx.H / 2) + h / 2) * (gx.W / 2) + w / 2;
    float g = 0.125f * bf2f(dY[p * dys + c]);
    if (accumulate) g += bf2f(dX[u * dxs + c]);
    dX[u * dxs + c] = f2bf(g);
}

// nearest 2x upsampling: Y[u][c] = X[parent(u)][c]
__global__ void k_upsample(Grid gy, const unsigned short* __restrict__ X, int xs, int C,
                           unsigned short* __restrict__ Y, int ys) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long V = (long)gy.D * gy.H * gy.W;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long u = q / C;
    const int w = (int)(u % gy.W), h = (int)((u / gy.W) % gy.H), d = (int)(u / ((long)gy.W * gy.H));
    const long p = ((long)(d / 2) * (gy.H / 2) + h / 2) * (gy.W / 2) + w / 2;
    Y[u * ys + c] = X[p * xs + c];
}

// upsampling gradient: dX[p][c] = sum of dY over the 8 children of p
__global__ void k_upsample_bwd(Grid gx, const unsigned short* __restrict__ dY, int dys, int C,
                               unsigned short* __restrict__ dX, int dxs) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long V = (long)gx.D * gx.H * gx.W;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long v = q / C;
    const int w = (int)(v % gx.W), h = (int)((v / gx.W) % gx.H), d = (int)(v / ((long)gx.W * gx.H));
    const int Hy = gx.H * 2, Wy = gx.W * 2;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const long u = ((long)(2 * d + (k >> 2)) * Hy + 2 * h + ((k >> 1) & 1)) * Wy + 2 * w + (k & 1);
        s += bf2f(dY[u * dys + c]);
    }
    dX[v * dxs + c] = f2bf(s);
}

// ReLU gradient in place: dY[v][c] = Y[v][c] > 0 ? dY[v][c] : 0
__global__ void k_relu_bwd(long V, int C, const unsigned short* __restrict__ Y, int ys, unsigned short* __restrict__ dY,
                           int dys) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long v = q / C;
    if (!(bf2f(Y[v * ys + c]) > 0.f)) dY[v * dys + c] = 0;
}

// Weighted mean squared error over V voxels x nc channels of O (FP32, stride
// os) against T (FP32, nc per voxel), weight w_v = rho_v^3 + wmin with rho_v
// channel 0 of the input X (the voxel density; the SIMP stiffness, so the loss
// approximates an energy norm), or 1 without X: per-block partial sums of
// w e^2, and the gradient dO[v][c] = 2 w (O - T) / (V nc) as bf16 (channels
// nc..C-1 zero).
__global__ void k_mse(long V, int nc, int C, const float* __restrict__ O, int os, const float* __restrict__ T,
                      const unsigned short* __restrict__ X, int xs, float wmin, unsigned short* __restrict__ dO,
                      int dos, float* __restrict__ part) {
    __shared__ float sh[4];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (v < V) {
        float w = 1.f;
        if (X) {
            const float r = fmaxf(bf2f(X[v * xs]), 0.f);
            w = r * r * r + wmin;
        }
        const float scale = 2.f * w / (float)(V * nc);
        for (int c = 0; c < C; ++c) {
            float g = 0.f;
            if (c < nc) {
                const float e = O[v * os + c] - T[v * nc + c];
                e2 += w * e * e;
                g = scale * e;
            }
            dO[v * dos + c] = f2bf(g);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Energy-norm proxy of the prediction error e = O - T (first nc channels) on
// the voxel grid: loss = sum_v w_v (l2 |e_v|^2 + sum over the +x, +y, +z
// neighbours u of v of |e_u - e_v|^2) / (V nc), w_v = rho_v^3 + wmin from the
// input's density channel (a stiffness-weighted discrete H1 seminorm plus a
// small L2 term).  Per-block partial sums of the loss, and its gradient with
// respect to O as bf16 (channels nc..C-1 zero); each voxel gathers the terms
// it appears in, so no atomics are needed.
__global__ void k_energy_loss(Grid g, int nc, int C, const float* __restrict__ O, int os, const float* __restrict__ T,
                              const unsigned short* __restrict__ X, int xs, float wmin, float l2,
                              unsigned short* __restrict__ dO, int dos, float* __restrict__ part) {
    __shared__ float sh[4];
    const long V = (long)g.D * g.H * g.W;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.f;
    if (v < V) {
        const int w = (int)(v % g.W), h = (int)((v / g.W) % g.H), d = (int)(v / ((long)g.W * g.H));
        const long step[3] = {(long)g.H * g.W, g.W, 1};
        const bool has_next[3] = {d + 1 < g.D, h + 1 < g.H, w + 1 < g.W};
        const bool has_prev[3] = {d > 0, h > 0, w > 0};
        auto weight = [&](long u) {
            const float r = fmaxf(bf2f(X[u * xs]), 0.f);
            return r * r * r + wmin;
        };
        const float wv = weight(v);
        float wp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) wp[k] = has_prev[k] ? weight(v - step[k]) : 0.f;
        const float scale = 2.f / (float)(V * nc);
        for (int c = 0; c < C; ++c) {
            float gsum = 0.f;
            if (c < nc) {
                const float ev = O[v * os + c] - T[v * nc + c];
                acc += wv * l2 * ev * ev;
                gsum += wv * l2 * ev;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (has_next[k]) {
                        const long u = v + step[k];
                        const float du = (O[u * os + c] - T[u * nc + c]) - ev;
                        acc += wv * du * du;
                        gsum -= wv * du;
                    }
                    if (has_prev[k]) {
                        const long u = v - step[k];
                        const float dd = ev - (O[u * os + c] - T[u * nc + c]);
                        gsum += wp[k] * dd;
                    }
                }
            }
            dO[v * dos + c] = f2bf(scale * gsum);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Adam with bias correction (lr_t = lr sqrt(1 - b2^k) / (1 - b1^k) given)
__global__ void k_adam(long n, float lr_t, float b1, float b2, float eps, const float* __restrict__ g,
                       float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk, b = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    w[k] -= lr_t * a / (sqrtf(b) + eps);
}

// FP32 <-> BF16 views
__global__ void k_to_bf16(long V, int C, const float* __restrict__ X, int xs, unsigned short* __restrict__ Y, int ys) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long v = q / C;
    Y[v * ys + c] = f2bf(X[v * xs + c]);
}

__global__ void k_add_bf16(long V, int C, const unsigned short* __restrict__ X, int xs, unsigned short* __restrict__ Y,
                           int ys) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= V * C) return;
    const int c = (int)(q % C);
    const long v = q / C;
    Y[v * ys + c] = f2bf(bf2f(Y[v * ys + c]) + bf2f(X[v * xs + c]));
}

// Node field (N x nc, FP64, on the box mesh's node grid: node = (i (ny+1) + j)
// (nz+1) + k for node indices (i, j, k) along (x, y, z)) to voxels: the mean of
// each hex's 8 nodes, scaled, clipped to [-clip, clip] when clip > 0, as FP32
// (voxel grid (D, H, W) = (nx, ny, nz)).
__global__ void k_nodes_to_voxels(Grid g, int nc, const double* __restrict__ F, double scale, double clip,
                                  float* __restrict__ out, int os, int oc) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long V = (long)g.D * g.H * g.W;
    if (v >= V) return;
    const int k = (int)(v % g.W), j = (int)((v / g.W) % g.H), i = (int)(v / ((long)g.W * g.H));
    for (int c = 0; c < nc; ++c) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long n = ((long)(i + (q >> 2)) * (g.H + 1) + j + ((q >> 1) & 1)) * (g.W + 1) + k + (q & 1);
            s += F[n * nc + c];
        }
        double y = 0.125 * s * scale;
        if (clip > 0.0) y = fmin(fmax(y, -clip), clip);
        out[v * os + oc + c] = (float)y;
    }
}

// Voxel field (FP32, nc channels at stride is) to nodes: each node takes the
// mean of its (up to 8) adjacent voxels, times scale (FP64 out, N x nc).
__global__ void k_voxels_to_nodes(Grid g, int nc, const float* __restrict__ X, int xs, double scale,
                                  double* __restrict__ F) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)(g.D + 1) * (g.H + 1) * (g.W + 1);
    if (n >= N) return;
    const int k = (int)(n % (g.W + 1)), j = (int)((n / (g.W + 1)) % (g.H + 1)), i = (int)(n / ((long)(g.W + 1) * (g.H + 1)));
    for (int c = 0; c < nc; ++c) {
        double s = 0.0;
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int ii = i - (q >> 2), jj = j - ((q >> 1) & 1), kk = k - (q & 1);
            if (ii < 0 || ii >= g.D || jj < 0 || jj >= g.H || kk < 0 || kk >= g.W) continue;
            s += (double)X[(((long)ii * g.H + jj) * g.W + kk) * xs + c];
            ++cnt;
        }
        F[n * nc + c] = scale * s / cnt;
    }
}

int nblocks(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace
}  // namespace cnn
}  // namespace pa

using pa::guarded;
using pa::cnn::Grid;

extern "C" {

int pa_cnn_conv3d(int D, int H, int W, const void* x, int x_stride, int cin, const void* wfrag, const float* bias,
                  int cout, void* y, int y_stride, int relu, float* y_f32, int yf_stride, int accumulate,
                  void* stream) {
    return guarded([&] {
        PA_REQUIRE(x && wfrag && (y || y_f32) && D > 0 && H > 0 && W > 0, PA_ERR_ARG, "pa_cnn_conv3d: bad argument");
        PA_REQUIRE(cin % 16 == 0 && cout % 16 == 0 && cout <= 128 && cin >= 16, PA_ERR_ARG,
                   "pa_cnn_conv3d: channel counts must be multiples of 16 (output at most 128)");
        PA_REQUIRE(x_stride % 4 == 0 && x_stride >= cin && (!y || y_stride >= cout) && (!y_f32 || yf_stride >= cout),
                   PA_ERR_ARG, "pa_cnn_conv3d: bad strides");
        PA_REQUIRE(!accumulate || y, PA_ERR_ARG, "pa_cnn_conv3d: accumulate needs the bf16 output");
        const long V = (long)D * H * W;
        PA_REQUIRE(V % 16 == 0, PA_ERR_ARG, "pa_cnn_conv3d: the voxel count must be a multiple of 16");
        const Grid g{D, H, W};
        const int grid = pa::cnn::nblocks(V / 16, 4);
        const hipStream_t st = (hipStream_t)stream;
        const auto* X = static_cast<const unsigned short*>(x);
        const auto* Wf = static_cast<const pa::cnn::bf16x4*>(wfrag);
        auto* Y = static_cast<unsigned short*>(y);
        const int nck = cin / 16;
#define PA_CONV_CASE(N)                                                                                              \
    case N:                                                                                                          \
        pa::cnn::k_conv_fwd<N><<<grid, 256, 0, st>>>(g, V, X, x_stride, nck, Wf, bias, Y, y_stride, relu, y_f32,    \
                                                     yf_stride, accumulate);                                         \
        break;
        switch (cout / 16) {
            PA_CONV_CASE(1)
            PA_CONV_CASE(2)
            PA_CONV_CASE(3)
            PA_CONV_CASE(4)
            PA_CONV_CASE(5)
            PA_CONV_CASE(6)
            PA_CONV_CASE(7)
            PA_CONV_CASE(8)
        }
#undef PA_CONV_CASE
        pa::cnn::check();
    });
}

int pa_cnn_conv3d_wgrad_parts(int D, int H, int W, const void* x, int x_stride, int cin, const void* dy, int dy_stride,
                              int cout, long long chunk, float* parts, void* stream) {
    return guarded([&] {
        const long V = (long)D * H * W;
        PA_REQUIRE(x && dy && parts && cin % 16 == 0 && cout % 16 == 0 && chunk > 0 && chunk % 16 == 0 && V % 16 == 0,
                   PA_ERR_ARG, "pa_cnn_conv3d_wgrad_parts: bad argument");
        const long nchunks = (V + chunk - 1) / chunk;
        const long blocks = nchunks * 27 * (cin / 16) * (cout / 16);
        pa::cnn::k_conv_wgrad<<<(unsigned)blocks, 64, 0, (hipStream_t)stream>>>(
            Grid{D, H, W}, V, static_cast<const unsigned short*>(x), x_stride, cin,
            static_cast<const unsigned short*>(dy), dy_stride, cout, (long)chunk, parts);
        pa::cnn::check();
    });
}

int pa_cnn_sum_parts(long long n, int nparts, const float* parts, float* dst, void* stream) {
    return guarded([&] {
        PA_REQUIRE(parts && dst && n >= 0 && nparts > 0, PA_ERR_ARG, "pa_cnn_sum_parts: bad argument");
        if (n == 0) return;
        pa::cnn::k_sum_parts<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(n, nparts, parts, dst);
        pa::cnn::check();
    });
}

int pa_cnn_bias_parts(long long V, const void* dy, int dy_stride, int c, long long chunk, float* parts, void* stream) {
    return guarded([&] {
        PA_REQUIRE(dy && parts && c > 0 && c <= 1024 && chunk > 0, PA_ERR_ARG, "pa_cnn_bias_parts: bad argument");
        const long nb = (long)((V + chunk - 1) / chunk);
        pa::cnn::k_bias_parts<<<(unsigned)nb, c, 0, (hipStream_t)stream>>>(
            (long)V, static_cast<const unsigned short*>(dy), dy_stride, c, (long)chunk, parts);
        pa::cnn::check();
    });
}

int pa_cnn_pack_weights(const float* w, int cin, int cout, int transposed, void* wfrag, void* stream) {
    return guarded([&] {
        PA_REQUIRE(w && wfrag && cin % 16 == 0 && cout % 16 == 0, PA_ERR_ARG, "pa_cnn_pack_weights: bad argument");
        const long total = 27L * (cin / 16) * (cout / 16) * 64;
        pa::cnn::k_pack_weights<<<pa::cnn::nblocks(total, 256), 256, 0, (hipStream_t)stream>>>(
            w, cin, cout, transposed, static_cast<unsigned short*>(wfrag));
        pa::cnn::check();
    });
}

int pa_cnn_pool(int D, int H, int W, const void* x, int xs, int c, void* y, int ys, void* stream) {
    return guarded([&] {
        PA_REQUIRE(x && y && D % 2 == 0 && H % 2 == 0 && W % 2 == 0, PA_ERR_ARG, "pa_cnn_pool: bad argument");
        const Grid gy{D / 2, H / 2, W / 2};
        const long n = (long)gy.D * gy.H * gy.W * c;
        pa::cnn::k_pool<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            gy, static_cast<const unsigned short*>(x), xs, c, static_cast<unsigned short*>(y), ys);
        pa::cnn::check();
    });
}

int pa_cnn_pool_bwd(int D, int H, int W, const void* dy, int dys, int c, void* dx, int dxs, int accumulate,
                    void* stream) {
    return guarded([&] {
        PA_REQUIRE(dy && dx && D % 2 == 0 && H % 2 == 0 && W % 2 == 0, PA_ERR_ARG, "pa_cnn_pool_bwd: bad argument");
        const long n = (long)D * H * W * c;
        pa::cnn::k_pool_bwd<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            Grid{D, H, W}, static_cast<const unsigned short*>(dy), dys, c, static_cast<unsigned short*>(dx), dxs,
            accumulate);
        pa::cnn::check();
    });
}

// (D, H, W) is the fine (output) grid
int pa_cnn_upsample(int D, int H, int W, const void* x, int xs, int c, void* y, int ys, void* stream) {
    return guarded([&] {
        PA_REQUIRE(x && y && D % 2 == 0 && H % 2 == 0 && W % 2 == 0, PA_ERR_ARG, "pa_cnn_upsample: bad argument");
        const long n = (long)D * H * W * c;
        pa::cnn::k_upsample<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            Grid{D, H, W}, static_cast<const unsigned short*>(x), xs, c, static_cast<unsigned short*>(y), ys);
        pa::cnn::check();
    });
}

// (D, H, W) is the coarse (output) grid
int pa_cnn_upsample_bwd(int D, int H, int W, const void* dy, int dys, int c, void* dx, int dxs, void* stream) {
    return guarded([&] {
        PA_REQUIRE(dy && dx, PA_ERR_ARG, "pa_cnn_upsample_bwd: bad argument");
        const long n = (long)D * H * W * c;
        pa::cnn::k_upsample_bwd<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            Grid{D, H, W}, static_cast<const unsigned short*>(dy), dys, c, static_cast<unsigned short*>(dx), dxs);
        pa::cnn::check();
    });
}

int pa_cnn_relu_bwd(long long V, int c, const void* y, int ys, void* dy, int dys, void* stream) {
    return guarded([&] {
        PA_REQUIRE(y && dy, PA_ERR_ARG, "pa_cnn_relu_bwd: bad argument");
        const long n = (long)V * c;
        pa::cnn::k_relu_bwd<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            (long)V, c, static_cast<const unsigned short*>(y), ys, static_cast<unsigned short*>(dy), dys);
        pa::cnn::check();
    });
}

// partial sums of the squared error into part[ceil(V / 256)]
int pa_cnn_mse(long long V, int nc, int c, const float* o, int os, const float* target, const void* x, int xs,
               float wmin, void* dout, int dos, float* part, void* stream) {
    return guarded([&] {
        PA_REQUIRE(o && target && dout && part && nc <= c, PA_ERR_ARG, "pa_cnn_mse: bad argument");
        pa::cnn::k_mse<<<pa::cnn::nblocks(V, 256), 256, 0, (hipStream_t)stream>>>(
            (long)V, nc, c, o, os, target, static_cast<const unsigned short*>(x), xs, wmin,
            static_cast<unsigned short*>(dout), dos, part);
        pa::cnn::check();
    });
}

int pa_cnn_energy_loss(int D, int H, int W, int nc, int c, const float* o, int os, const float* target, const void* x,
                       int xs, float wmin, float l2, void* dout, int dos, float* part, void* stream) {
    return guarded([&] {
        PA_REQUIRE(o && target && x && dout && part && nc <= c, PA_ERR_ARG, "pa_cnn_energy_loss: bad argument");
        const long V = (long)D * H * W;
        pa::cnn::k_energy_loss<<<pa::cnn::nblocks(V, 256), 256, 0, (hipStream_t)stream>>>(
            Grid{D, H, W}, nc, c, o, os, target, static_cast<const unsigned short*>(x), xs, wmin, l2,
            static_cast<unsigned short*>(dout), dos, part);
        pa::cnn::check();
    });
}

int pa_cnn_adam(long long n, float lr_t, float b1, float b2, float eps, const float* g, float* w, float* m1, float* m2,
                void* stream) {
    return guarded([&] {
        PA_REQUIRE(g && w && m1 && m2, PA_ERR_ARG, "pa_cnn_adam: bad argument");
        pa::cnn::k_adam<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>((long)n, lr_t, b1, b2, eps, g, w,
                                                                                      m1, m2);
        pa::cnn::check();
    });
}

int pa_cnn_to_bf16(long long V, int c, const float* x, int xs, void* y, int ys, void* stream) {
    return guarded([&] {
        PA_REQUIRE(x && y, PA_ERR_ARG, "pa_cnn_to_bf16: bad argument");
        const long n = (long)V * c;
        pa::cnn::k_to_bf16<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            (long)V, c, x, xs, static_cast<unsigned short*>(y), ys);
        pa::cnn::check();
    });
}

int pa_cnn_add_bf16(long long V, int c, const void* x, int xs, void* y, int ys, void* stream) {
    return guarded([&] {
        PA_REQUIRE(x && y, PA_ERR_ARG, "pa_cnn_add_bf16: bad argument");
        const long n = (long)V * c;
        pa::cnn::k_add_bf16<<<pa::cnn::nblocks(n, 256), 256, 0, (hipStream_t)stream>>>(
            (long)V, c, static_cast<const unsigned short*>(x), xs, static_cast<unsigned short*>(y), ys);
        pa::cnn::check();
    });
}

int pa_cnn_nodes_to_voxels(int D, int H, int W, int nc, const double* nodes, double scale, double clip, float* out,
                           int os, int oc, void* stream) {
    return guarded([&] {
        PA_REQUIRE(nodes && out && oc + nc <= os, PA_ERR_ARG, "pa_cnn_nodes_to_voxels: bad argument");
        const long V = (long)D * H * W;
        pa::cnn::k_nodes_to_voxels<<<pa::cnn::nblocks(V, 256), 256, 0, (hipStream_t)stream>>>(
            Grid{D, H, W}, nc, nodes, scale, clip, out, os, oc);
        pa::cnn::check();
    });
}

int pa_cnn_voxels_to_nodes(int D, int H, int W, int nc, const float* vox, int vs, double scale, double* nodes,
                           void* stream) {
    return guarded([&] {
        PA_REQUIRE(nodes && vox && vs >= nc, PA_ERR_ARG, "pa_cnn_voxels_to_nodes: bad argument");
        const long N = (long)(D + 1) * (H + 1) * (W + 1);
        pa::cnn::k_voxels_to_nodes<<<pa::cnn::nblocks(N, 256), 256, 0, (hipStream_t)stream>>>(Grid{D, H, W}, nc, vox,
                                                                                                vs, scale, nodes);
        pa::cnn::check();
    });
}

}  // extern "C"
