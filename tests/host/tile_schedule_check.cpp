// Host-side check of the cell-tile assembly schedules (no GPU needed).
//
// Builds the library's mesh tables with the HIP allocation and copy calls
// mapped to host memory, then walks the tile schedules the way the tile
// kernels do (cells into a scratch "LDS", work units summed over their entry
// slots, diagonal parts added in groups of four, mirrors stored transposed,
// blocks staged and copied run by run; residual rows summed over their
// pairs) and compares the Jacobian and residual with a direct per-cell
// assembly (isotropic B^T C B, SIMP p = 3).  Also checks the invariants the
// kernels rely on: every value written exactly once, padded slots pointing at
// the zero cell, diagonal groups aligned to four lanes, wave shapes uniform
// over aligned groups of 64 units.
//
// usage: tile_schedule_check nx ny nz   (exit 0 on success)
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

static hipError_t host_malloc(void** p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
static hipError_t host_free(void* p) {
    std::free(p);
    return hipSuccess;
}
static hipError_t host_memcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
#define hipMalloc(p, n) host_malloc((void**)(p), (n))
#define hipFree(p) host_free(p)
#define hipMemcpy(d, s, n, k) host_memcpy((d), (s), (n), (k))
#include "../../platoanalyze_amd/csrc/pa_host.cpp"

#include <cmath>
#include <cstdio>
#include <vector>

using namespace pa;

static int failures = 0;
#define CHECK(c, ...)                 \
    do {                              \
        if (!(c)) {                   \
            if (failures++ < 10) {    \
                std::printf(__VA_ARGS__); \
                std::printf("\n");    \
            }                         \
        }                             \
    } while (0)

static double gradients(const double x[4][3], double g[4][3]) {
    double j[3][3];
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) j[d][a] = x[a][d] - x[3][d];
    const double det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                       j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                       j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
    // rows of J^-1 (cofactors / det)
    g[0][0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) / det;
    g[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) / det;
    g[0][2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) / det;
    g[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) / det;
    g[1][1] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) / det;
    g[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) / det;
    g[2][0] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) / det;
    g[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) / det;
    g[2][2] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) / det;
    for (int d = 0; d < 3; ++d) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
    return std::fabs(det);
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
    const int N = (nx + 1) * (ny + 1) * (nz + 1), C = 6 * nx * ny * nz;
    std::vector<double> xyz(3L * N);
    std::vector<int> conn(4L * C);
    if (pa_box_mesh(nx, ny, nz, 2.0, 1.0, 1.0, xyz.data(), conn.data()) != PA_OK) return 3;
    pa_mesh* m = nullptr;
    if (pa_mesh_create(xyz.data(), N, conn.data(), C, &m) != PA_OK) {
        std::printf("mesh: %s\n", pa_last_error());
        return 3;
    }
    if (!m->has_tiles) {
        std::printf("no tiles\n");
        return 4;
    }
    const long nb = m->nnzb;
    std::vector<double> z(N), u(3L * N);
    for (int i = 0; i < N; ++i) z[i] = 0.3 + 0.7 * ((i * 7919) % 1000) / 1000.0;
    for (long i = 0; i < 3L * N; ++i) u[i] = 1e-2 * (((i * 104729) % 2001) / 1000.0 - 1.0);
    const double E = 1.0, nu = 0.3, cc = E / ((1 + nu) * (1 - 2 * nu)), lam = cc * nu, mu = 0.5 * cc * (1 - 2 * nu);
    auto slot_of = [&](int r, int col) {
        for (int e = m->h_rowmap[r]; e < m->h_rowmap[r + 1]; ++e)
            if (m->h_colidx[e] == col) return e;
        return -1;
    };
    struct Cell {
        double g[4][3], w, f[4][3];
    };
    auto eval = [&](const int nd[4], Cell& c) {
        double x[4][3];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) x[a][d] = xyz[3L * nd[a] + d];
        const double ad = gradients(x, c.g);
        const double rho = (((z[nd[0]] + z[nd[1]]) + z[nd[2]]) + z[nd[3]]) / 4.0;
        c.w = (1e-9 + (1 - 1e-9) * rho * rho * rho) * ad / 6.0;
        double eps[6] = {0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 4; ++a) {
            const double* ua = &u[3L * nd[a]];
            eps[0] += ua[0] * c.g[a][0];
            eps[1] += ua[1] * c.g[a][1];
            eps[2] += ua[2] * c.g[a][2];
            eps[3] += ua[2] * c.g[a][1] + ua[1] * c.g[a][2];
            eps[4] += ua[2] * c.g[a][0] + ua[0] * c.g[a][2];
            eps[5] += ua[1] * c.g[a][0] + ua[0] * c.g[a][1];
        }
        const double tr = eps[0] + eps[1] + eps[2];
        const double s[6] = {lam * tr + 2 * mu * eps[0], lam * tr + 2 * mu * eps[1], lam * tr + 2 * mu * eps[2],
                             mu * eps[3], mu * eps[4], mu * eps[5]};
        for (int a = 0; a < 4; ++a) {
            const double* ga = c.g[a];
            c.f[a][0] = c.w * (s[0] * ga[0] + s[5] * ga[1] + s[4] * ga[2]);
            c.f[a][1] = c.w * (s[5] * ga[0] + s[1] * ga[1] + s[3] * ga[2]);
            c.f[a][2] = c.w * (s[4] * ga[0] + s[3] * ga[1] + s[2] * ga[2]);
        }
    };
    // direct per-cell assembly
    std::vector<double> Kref(9L * nb, 0.0), Rref(3L * N, 0.0);
    for (int c = 0; c < C; ++c) {
        const int* nd = &conn[4L * c];
        Cell e;
        eval(nd, e);
        for (int a = 0; a < 4; ++a) {
            for (int i = 0; i < 3; ++i) Rref[3L * nd[a] + i] += e.f[a][i];
            for (int b = 0; b < 4; ++b) {
                const long o = 9L * slot_of(nd[a], nd[b]);
                const double dot = e.g[a][0] * e.g[b][0] + e.g[a][1] * e.g[b][1] + e.g[a][2] * e.g[b][2];
                for (int i = 0; i < 3; ++i)
                    for (int n = 0; n < 3; ++n)
                        Kref[o + 3 * i + n] +=
                            e.w * (lam * e.g[a][i] * e.g[b][n] + mu * e.g[a][n] * e.g[b][i] + (i == n ? mu * dot : 0.0));
            }
        }
    }
    // the tile schedules
    std::vector<double> K(9L * nb, 0.0), R(3L * N, 0.0);
    std::vector<char> kw(9L * nb, 0), rw(N, 0);
    const int UE = m->tile_unit_entries;
    for (int b = 0; b < m->num_tiles; ++b) {
        const int4 h0 = m->d_tile.p[3 * b], h1 = m->d_tile.p[3 * b + 1], h2 = m->d_tile.p[3 * b + 2];
        const int base = h0.x, nrows = h1.x, ncells = h1.y, nunits = h1.z, nruns = h1.w;
        CHECK(ncells + 1 <= kTileCellsMax && nrows <= 64 && nunits <= kTileUnitsMax, "tile %d limits", b);
        std::vector<double> lds((ncells + 1) * kTileStride, 0.0);
        std::vector<Cell> cells(ncells);
        for (int c = 0; c < ncells; ++c) {
            const int4 d = m->d_tp_cell.p[h0.z + c];
            const int nd[4] = {base + d.x, base + d.y, base + d.z, base + d.w};
            // the halo slots the residual kernels read the nodes through
            {
                const int d1 = m->grid_dims[1], d2 = m->grid_dims[2];
                const int i0 = base / (d1 * d2), j0 = (base / d2) % d1, k0 = base % d2;
                const unsigned hs = m->d_tp_hslot.p[h0.z + c];
                for (int n = 0; n < 4; ++n) {
                    const int slot = (int)((hs >> (8 * n)) & 0xffu);
                    CHECK(slot < kTileHalo, "tile %d cell %d: halo slot %d out of range", b, c, slot);
                    const int a = slot / ((kTileY + 2) * (kTileZ + 2)), bb = (slot / (kTileZ + 2)) % (kTileY + 2),
                              cc = slot % (kTileZ + 2);
                    const int i = i0 - 1 + a, j = j0 - 1 + bb, k = k0 - 1 + cc;
                    CHECK(i >= 0 && i < m->grid_dims[0] && j >= 0 && j < d1 && k >= 0 && k < d2,
                          "tile %d cell %d: halo node outside the mesh", b, c);
                    CHECK((i * d1 + j) * d2 + k == nd[n], "tile %d cell %d: halo slot names node %d, not %d", b, c,
                          (i * d1 + j) * d2 + k, nd[n]);
                }
            }
            eval(nd, cells[c]);
            const double sw = std::sqrt(cells[c].w);
            for (int n = 0; n < 4; ++n)
                for (int k = 0; k < 3; ++k) lds[c * kTileStride + 3 * n + k] = sw * cells[c].g[n][k];
        }
        std::vector<int> rstart(nrows), rstage(nrows);
        for (int r = 0; r < nrows; ++r) {
            const int4 ri = m->d_tp_row.p[h0.y + r];
            rstart[r] = m->h_rowmap[base + ri.x];
            rstage[r] = ri.w;
            double acc[3] = {0, 0, 0};
            for (int p = 0; p < ri.y; ++p) {
                const unsigned e = m->d_tp_pair.p[ri.z + p];
                for (int i = 0; i < 3; ++i) acc[i] += cells[e >> 2].f[e & 3][i];
            }
            CHECK(!rw[base + ri.x], "row %d summed twice", base + ri.x);
            rw[base + ri.x] = 1;
            for (int i = 0; i < 3; ++i) R[3L * (base + ri.x) + i] = acc[i];
        }
        std::vector<double> M(9L * nunits, 0.0), stage(9L * h2.y, std::nan(""));
        for (int it = 0; it < nunits; ++it) {
            const uint2 un = m->d_tp_unit.p[h0.w + it];
            const unsigned eo = un.y & 0x3fffffu, cnt = (un.y >> 22) & 15u, shape = un.y >> 30;
            const bool diag = (un.y >> 26) & 1u;
            CHECK(shape != 1 || (diag && cnt == 6), "shape 1 unit %d", it);
            CHECK(shape != 2 || (!diag && cnt == 6), "shape 2 unit %d", it);
            CHECK(shape != 3 || (!diag && cnt == 4), "shape 3 unit %d", it);
            CHECK((m->d_tp_unit.p[h0.w + (it & ~63)].y >> 30) == shape, "wave shape not uniform at unit %d", it);
            CHECK(!diag || ((un.y >> 27) & 3u) == (unsigned)(it & 3), "diagonal part misaligned at unit %d", it);
            const int ne = shape == 1 || shape == 2 ? 6 : shape == 3 ? 4 : UE;
            const uint4* e4 = reinterpret_cast<const uint4*>(m->d_tp_ent.p) + eo;
            const unsigned ew[8] = {e4[0].x, e4[0].y, e4[0].z, e4[0].w, e4[1].x, e4[1].y, e4[1].z, e4[1].w};
            for (int e = (int)cnt; e < kTileUnitEntries; ++e)
                CHECK((ew[e] & 0xffffu) == (unsigned)(ncells * kTileStride) && (ew[e] >> 16) == (ew[e] & 0xffffu),
                      "padded slot not the zero cell");
            for (int e = 0; e < ne; ++e) {
                const double* A = &lds[ew[e] & 0xffffu];
                const double* B = &lds[ew[e] >> 16];
                for (int i = 0; i < 3; ++i)
                    for (int n = 0; n < 3; ++n) M[9L * it + 3 * i + n] += A[i] * B[n];
            }
        }
        for (int it = 0; it < nunits; ++it) {
            const uint2 un = m->d_tp_unit.p[h0.w + it];
            const bool diag = (un.y >> 26) & 1u;
            if (diag && ((un.y >> 27) & 3u) != 0u) continue;
            double Mt[9], Kb[9];
            for (int q = 0; q < 9; ++q)
                Mt[q] = diag ? (M[9L * it + q] + M[9L * (it + 1) + q]) + (M[9L * (it + 2) + q] + M[9L * (it + 3) + q])
                             : M[9L * it + q];
            const double tr = (Mt[0] + Mt[4]) + Mt[8];
            for (int i = 0; i < 3; ++i)
                for (int n = 0; n < 3; ++n) Kb[3 * i + n] = lam * Mt[3 * i + n] + mu * Mt[3 * n + i] + (i == n ? mu * tr : 0.0);
            const int r = un.x & 0xffu, sl = (un.x >> 8) & 0xffu;
            for (int q = 0; q < 9; ++q) stage[9L * (rstage[r] + sl) + q] = Kb[q];
            if ((un.y >> 29) & 1u) {
                const int mr = (un.x >> 16) & 0xffu, ms = un.x >> 24;
                for (int i = 0; i < 3; ++i)
                    for (int n = 0; n < 3; ++n) stage[9L * (rstage[mr] + ms) + 3 * i + n] = Kb[3 * n + i];
            }
        }
        for (int j = 0; j < nruns; ++j) {
            const int4 run = m->d_tp_run.p[h2.x + j];
            for (int i = 0; i < 9 * run.z; ++i) {
                const long o = 9L * rstart[run.x] + i;
                CHECK(!kw[o], "value %ld written twice", o);
                CHECK(!std::isnan(stage[9L * run.y + i]), "value %ld not staged", o);
                kw[o] = 1;
                K[o] = stage[9L * run.y + i];
            }
        }
    }
    for (long i = 0; i < 9L * nb; ++i) CHECK(kw[i], "value %ld never written", i);
    for (int i = 0; i < N; ++i) CHECK(rw[i], "row %d never summed", i);
    double ek = 0, mk = 0, er = 0, mr = 0;
    for (long i = 0; i < 9L * nb; ++i) ek = std::max(ek, std::fabs(K[i] - Kref[i])), mk = std::max(mk, std::fabs(Kref[i]));
    for (long i = 0; i < 3L * N; ++i) er = std::max(er, std::fabs(R[i] - Rref[i])), mr = std::max(mr, std::fabs(Rref[i]));
    std::printf("%dx%dx%d: %d tiles, %d schedules; Jacobian rel err %.2e, residual rel err %.2e\n", nx, ny, nz,
                m->num_tiles, m->num_tile_patterns, ek / mk, er / mr);
    CHECK(ek / mk < 1e-13 && er / mr < 1e-13, "assembly mismatch");
    pa_mesh_destroy(m);
    return failures ? 1 : 0;
}
