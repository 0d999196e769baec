// Block stencils of matrices on structured node grids (gfx950).
//
// A matrix whose graph is a neighbourhood graph of an (nx + 1) x (ny + 1) x
// (nz + 1) node grid (z fastest) is stored as a block stencil: the blocks of
// node n in the order of the offsets (dx, dy, dz) in {-1, 0, 1}^3 that the
// graph holds (stencil index sidx = (dx + 1) 9 + (dy + 1) 3 + dz + 1).  A
// symmetric matrix (block (m, n) the transpose of (n, m)) keeps only the
// diagonal and the offsets of larger stencil index; the block of offset d of
// smaller index is the neighbour's block of offset -d, transposed.  The
// kernels give each node G = 4 or 8 lanes, lane l taking the offsets l,
// l + G, ...: no CRS index is read, and every load of a lane issues at once.
#pragma once

#include <hip/hip_runtime.h>

#include "pa_internal.h"

namespace pa {

struct Grid {
    int nx, ny, nz;  // cells
    __host__ __device__ int nodes() const { return (nx + 1) * (ny + 1) * (nz + 1); }
    __host__ __device__ int id(int i, int j, int k) const { return (i * (ny + 1) + j) * (nz + 1) + k; }
    __host__ __device__ void ijk(int n, int& i, int& j, int& k) const {
        k = n % (nz + 1);
        j = (n / (nz + 1)) % (ny + 1);
        i = n / ((nz + 1) * (ny + 1));
    }
    __host__ __device__ bool in(int i, int j, int k) const {
        return i >= 0 && i <= nx && j >= 0 && j <= ny && k >= 0 && k <= nz;
    }
};

// The offset of lane slot s = l + G m is described by 16 bits of word
// lane[m][l / 2]: dx + 1, dy + 1, dz + 1 (2 bits each), read transposed (bit
// 6), present (bit 7), stored slot read (bits 8-12).
constexpr int kFsSlots = 4;  // offsets per lane
struct BlockStencil {
    int S, SS;                   // offsets in the graph; stored blocks per node
    int G;                       // lanes per node
    int half;                    // some offsets are read transposed
    unsigned lane[kFsSlots][4];  // per-lane offset descriptions
    unsigned ssig[6];            // stencil index of stored slot t, 5 bits each, six per word
    int diag;                    // stored slot of the diagonal block
    double rny, rnz;             // 1 / (ny + 1), 1 / (nz + 1): node coordinates by products
};

// The stencil of a matrix on grid g whose graph holds the offsets `offsets`
// (bit sidx); half when the matrix is symmetric and the offsets come in pairs.
// G: lanes per node (0: 4 for at most 16 offsets, else 8).
inline BlockStencil make_block_stencil(const Grid& g, unsigned offsets, bool symmetric, int G = 0) {
    BlockStencil f{};
    int lanes[27], stored[27], slot_of[27];
    f.S = 0;
    for (int q = 0; q < 27; ++q)
        if ((offsets >> q) & 1u) lanes[f.S++] = q;
    bool pairs = true;
    for (int q = 0; q < 27; ++q) pairs = pairs && (((offsets >> q) & 1u) == ((offsets >> (26 - q)) & 1u));
    const bool half = symmetric && pairs;
    f.SS = 0;
    for (int q = 0; q < 27; ++q) slot_of[q] = -1;
    for (int s = 0; s < f.S; ++s)
        if (!half || lanes[s] >= 13) {
            slot_of[lanes[s]] = f.SS;
            stored[f.SS++] = lanes[s];
        }
    PA_REQUIRE(slot_of[13] >= 0, PA_ERR_ARG, "block stencil: the graph has no diagonal entries");
    f.diag = slot_of[13];
    f.G = G > 0 ? G : (f.S <= 4 * kFsSlots ? 4 : 8);
    PA_REQUIRE((f.G == 4 || f.G == 8) && f.S <= f.G * kFsSlots, PA_ERR_ARG, "block stencil: too many offsets");
    for (int s = 0; s < f.S; ++s) {
        const int q = lanes[s];
        const bool tr = half && q < 13;
        const unsigned info = (unsigned)(q / 9) | (unsigned)((q / 3) % 3) << 2 | (unsigned)(q % 3) << 4 |
                              (tr ? 1u << 6 : 0u) | 1u << 7 | (unsigned)slot_of[tr ? 26 - q : q] << 8;
        const int m = s / f.G, l = s % f.G;
        f.lane[m][l / 2] |= info << ((l % 2) * 16);
    }
    for (int t = 0; t < f.SS; ++t) f.ssig[t / 6] |= (unsigned)stored[t] << ((t % 6) * 5);
    f.half = half;
    f.rny = 1.0 / (g.ny + 1);
    f.rnz = 1.0 / (g.nz + 1);
    return f;
}

// Whether the mesh graph is a neighbourhood graph of grid g; its offsets.
inline bool grid_graph_offsets(const pa_mesh* m, const Grid& g, unsigned* offsets) {
    if (g.nodes() != m->N) return false;
    unsigned o = 0;
    for (int r = 0; r < m->N; ++r) {
        int i, j, k;
        g.ijk(r, i, j, k);
        for (int e = m->h_rowmap[r]; e < m->h_rowmap[r + 1]; ++e) {
            int a, b, c;
            g.ijk(m->h_colidx[e], a, b, c);
            if (a < i - 1 || a > i + 1 || b < j - 1 || b > j + 1 || c < k - 1 || c > k + 1) return false;
            o |= 1u << ((a - i + 1) * 9 + (b - j + 1) * 3 + (c - k + 1));
        }
    }
    *offsets = o;
    return true;
}

namespace {

// The CRS matrix of the mesh graph (VT values) into the stencil (OT values):
// thread per (node, stored slot); a slot the node's row has no entry for
// gets a zero block.
template <int D, class VT, class OT>
__global__ void k_crs_to_stencil(Grid g, BlockStencil fs, const int* __restrict__ rowmap,
                                 const int* __restrict__ colidx, const VT* __restrict__ vals, OT* __restrict__ A) {
    constexpr int BS = D * D;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)g.nodes() * fs.SS) return;
    const int n = (int)(q / fs.SS), t = (int)(q % fs.SS);
    const int sig = (int)((fs.ssig[t / 6] >> ((t % 6) * 5)) & 31u);
    int i, j, k;
    g.ijk(n, i, j, k);
    const int ii = i + sig / 9 - 1, jj = j + (sig / 3) % 3 - 1, kk = k + sig % 3 - 1;
    int e = -1;
    if (g.in(ii, jj, kk)) {
        const int c = g.id(ii, jj, kk);
        for (int f = rowmap[n]; f < rowmap[n + 1]; ++f)
            if (colidx[f] == c) {
                e = f;
                break;
            }
    }
    OT* dst = A + q * BS;
#pragma unroll
    for (int r = 0; r < BS; ++r) dst[r] = e >= 0 ? (OT)vals[(long)e * BS + r] : (OT)0;
}

// v from another lane of the quad / half row (DPP row control CTRL)
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffull), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL, class T>
__device__ inline T dpp_any(T v) {
    if constexpr (sizeof(T) == 8) return dpp_f64<CTRL>(v);
    else return dpp_f32<CTRL>(v);
}

// Sum over the G lanes of a node (G = 4 or 8), every lane getting it, in a
// fixed order: l ^ 1, l ^ 2, then the other quad.
template <int G, class T>
__device__ inline T node_lanes_sum(T v) {
    v += dpp_any<0xb1>(v);                          // quad_perm [1, 0, 3, 2]
    v += dpp_any<0x4e>(v);                          // quad_perm [2, 3, 0, 1]
    if constexpr (G == 8) v += dpp_any<0x141>(v);  // row_half_mirror
    return v;
}

// Node n of grid g: (i, j, k) by products with FP64 reciprocals ((n + 1/2) / d
// rounded down is exact for any int n).
__device__ inline void stencil_ijk(const Grid& g, const BlockStencil& fs, int n, int& i, int& j, int& k) {
    const int jk = (int)(((double)n + 0.5) * fs.rnz);
    k = n - jk * (g.nz + 1);
    i = (int)(((double)jk + 0.5) * fs.rny);
    j = jk - i * (g.ny + 1);
}

// Lane l's m-th offset of node n = (i, j, k): ok (present and inside the
// grid), the neighbour c, the element offset of the block read (the node's own
// block for an offset that is not ok, so every lane loads from a valid
// address) and whether it is read transposed.
template <int G, bool SYM>
__device__ inline void stencil_slot(const Grid& g, const BlockStencil& fs, int m, int l, int n, int i, int j, int k,
                                    int BS, bool& ok, bool& tr, long& c, long& base) {
    const unsigned w0 = __builtin_amdgcn_readfirstlane(fs.lane[m][0]);
    const unsigned w1 = __builtin_amdgcn_readfirstlane(fs.lane[m][1]);
    unsigned w = (l & 2) ? w1 : w0;
    if constexpr (G == 8) {
        const unsigned w2 = __builtin_amdgcn_readfirstlane(fs.lane[m][2]);
        const unsigned w3 = __builtin_amdgcn_readfirstlane(fs.lane[m][3]);
        w = (l & 4) ? ((l & 2) ? w3 : w2) : w;
    }
    const unsigned info = w >> ((l & 1) * 16);
    const int ii = i + (int)(info & 3u) - 1, jj = j + (int)((info >> 2) & 3u) - 1, kk = k + (int)((info >> 4) & 3u) - 1;
    ok = ((info >> 7) & 1u) && ii >= 0 && ii <= g.nx && jj >= 0 && jj <= g.ny && kk >= 0 && kk <= g.nz;
    tr = SYM && ((info >> 6) & 1u);
    c = ok ? ((long)ii * (g.ny + 1) + jj) * (g.nz + 1) + kk : (long)n;
    const long src = ok ? (long)((info >> 8) & 31u) : (long)fs.diag;
    base = ((tr ? c : (long)n) * fs.SS + src) * BS;
}

}  // namespace
}  // namespace pa
