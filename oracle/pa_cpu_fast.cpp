// Optimized CPU variant of the elastic value + gradient_u assembly (CPU
// baseline only, test infrastructure like the rest of oracle/): the same
// mathematics as the restatement in pa_oracle.cpp without the reference's
// structure costs.  The graph, a per-cell map of the 16 (a, b) block ordinals
// and a conflict-free cell schedule are built once; each call evaluates the
// isotropic element stiffness in closed form,
// K_ab = s V (lambda g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I), forms
// R_a = sum_b K_ab u_b plus the uniform traction term, and adds both into the
// outputs with OpenMP and no atomics.  Schedule: the cells in 2T contiguous
// chunks (T threads), the even chunks in one parallel phase and the odd ones
// in a second, valid when no node is touched by chunks two or more apart
// (true for cell orders with locality, such as the box meshes); otherwise a
// greedy colouring (no two cells of a colour share a node), one phase per
// colour.  It exists so that the
// GPU speed-up is also stated against a reasonable CPU implementation, not
// only against the reference-structured one.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" void orc_graph_fill(int N, int C, const int* conn, int* rowmap, int* colidx);
extern "C" long orc_graph_nnz(int N, int C, const int* conn);

namespace {

struct Fast {
    int N = 0, C = 0;
    std::vector<double> coords;
    std::vector<int> conn;
    std::vector<int> rowmap, colidx;
    std::vector<int> ent;        // C*16: block ordinal of (a, b)
    // phases: cells phase_cells[phase_off[k] .. phase_off[k+1]) in parallel
    // chunks chunk_off[...] of phase k (chunks of one phase share no node)
    std::vector<int> phase_off;   // per phase: first chunk
    std::vector<int> chunk_off;   // per chunk: first position in phase_cells
    std::vector<int> phase_cells;
    bool chunked = false;
    std::vector<double> traction_load;  // N*3, the state-independent boundary term
};

int ordinal(const Fast& F, int row, int col) {
    const int* b = F.colidx.data() + F.rowmap[row];
    const int* e = F.colidx.data() + F.rowmap[row + 1];
    return (int)(std::lower_bound(b, e, col) - F.colidx.data());
}

}  // namespace

extern "C" {

void* cpu_fast_create(int N, int C, const double* coords, const int* conn, int nfaces, const int* face_cells,
                      const int* face_lnodes, const double* traction) {
    auto* F = new Fast();
    F->N = N;
    F->C = C;
    F->coords.assign(coords, coords + (size_t)N * 3);
    F->conn.assign(conn, conn + (size_t)C * 4);
    F->rowmap.resize(N + 1);
    // graph: the same routine as the restatement (sorted columns, diagonal included)
    F->colidx.resize(orc_graph_nnz(N, C, conn));
    orc_graph_fill(N, C, conn, F->rowmap.data(), F->colidx.data());
    F->ent.resize((size_t)C * 16);
    for (int c = 0; c < C; ++c)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) F->ent[(size_t)c * 16 + a * 4 + b] = ordinal(*F, conn[c * 4 + a], conn[c * 4 + b]);
    // contiguous chunks: valid when every node's touching chunks are at most one apart
    int T = 1;
#ifdef _OPENMP
    T = omp_get_max_threads();
#endif
    const int nch = 2 * T;
    if (C >= 64 * nch) {
        std::vector<int> lo(N, nch), hi(N, -1);
        for (int c = 0; c < C; ++c) {
            const int k = (int)((long long)c * nch / C);
            for (int a = 0; a < 4; ++a) {
                lo[conn[c * 4 + a]] = std::min(lo[conn[c * 4 + a]], k);
                hi[conn[c * 4 + a]] = std::max(hi[conn[c * 4 + a]], k);
            }
        }
        bool ok = true;
        for (int n = 0; n < N && ok; ++n) ok = hi[n] < 0 || hi[n] - lo[n] <= 1;
        if (ok) {
            F->chunked = true;
            F->phase_off = {0, T, 2 * T};
            for (int ph = 0; ph < 2; ++ph)
                for (int k = ph; k < nch; k += 2) {
                    F->chunk_off.push_back((int)F->phase_cells.size());
                    const int c0 = (int)((long long)k * C / nch), c1 = (int)((long long)(k + 1) * C / nch);
                    for (int c = c0; c < c1; ++c) F->phase_cells.push_back(c);
                }
            F->chunk_off.push_back((int)F->phase_cells.size());
        }
    }
    if (!F->chunked) {
    // greedy colouring: colour of a cell = smallest colour unused by the cells sharing its nodes
    std::vector<unsigned long long> used(N, 0ULL);
    std::vector<int> color(C);
    int ncolors = 0;
    for (int c = 0; c < C; ++c) {
        unsigned long long m = 0;
        for (int a = 0; a < 4; ++a) m |= used[conn[c * 4 + a]];
        int k = 0;
        while (k < 64 && (m >> k & 1ULL)) ++k;
        color[c] = k;  // k < 64: a node touches far fewer than 64 cells on the meshes used here
        ncolors = std::max(ncolors, k + 1);
        for (int a = 0; a < 4; ++a) used[conn[c * 4 + a]] |= 1ULL << k;
    }
    // one chunk per colour, one phase per chunk
    F->chunk_off.assign(ncolors + 1, 0);
    for (int c = 0; c < C; ++c) ++F->chunk_off[color[c] + 1];
    for (int k = 0; k < ncolors; ++k) F->chunk_off[k + 1] += F->chunk_off[k];
    F->phase_cells.resize(C);
    std::vector<int> pos(F->chunk_off.begin(), F->chunk_off.end() - 1);
    for (int c = 0; c < C; ++c) F->phase_cells[pos[color[c]]++] = c;
    for (int k = 0; k <= ncolors; ++k) F->phase_off.push_back(k);
    }
    // uniform traction: -t A / 3 per face node (the elastic residual's boundary term)
    F->traction_load.assign((size_t)N * 3, 0.0);
    for (int f = 0; f < nfaces; ++f) {
        const int c = face_cells[f];
        int n[3];
        for (int k = 0; k < 3; ++k) n[k] = conn[c * 4 + face_lnodes[f * 3 + k]];
        double e1[3], e2[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = coords[(size_t)n[0] * 3 + d] - coords[(size_t)n[2] * 3 + d];
            e2[d] = coords[(size_t)n[1] * 3 + d] - coords[(size_t)n[2] * 3 + d];
        }
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) F->traction_load[(size_t)n[k] * 3 + d] -= traction[d] * area / 3.0;
    }
    return F;
}

void cpu_fast_destroy(void* h) { delete static_cast<Fast*>(h); }

long long cpu_fast_nnzb(void* h) { return (long long)static_cast<Fast*>(h)->colidx.size(); }

// parallel phases of the schedule (2 when chunked, else the number of colours)
int cpu_fast_colors(void* h) { return (int)static_cast<Fast*>(h)->phase_off.size() - 1; }

// R (N*3) and the 3x3 block values (nnzb*9) of the SIMP elastostatic operator.
void cpu_fast_elastic(void* h, double E, double nu, double p, double emin, const double* u, const double* z,
                      double* R, double* values) {
    const Fast& F = *static_cast<Fast*>(h);
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    std::memcpy(R, F.traction_load.data(), (size_t)F.N * 3 * sizeof(double));
    std::memset(values, 0, F.colidx.size() * 9 * sizeof(double));
    const int nphases = (int)F.phase_off.size() - 1;
    for (int ph = 0; ph < nphases; ++ph) {
        // chunked: one thread per chunk (cells inside a chunk share nodes);
        // colouring: the colour's independent cells split statically
        const int k0 = F.phase_off[ph], k1 = F.phase_off[ph + 1];
        const int nunits = F.chunked ? k1 - k0 : F.chunk_off[k1] - F.chunk_off[k0];
#pragma omp parallel for schedule(static, 1)
        for (int q = 0; q < nunits; ++q) {
            const int i0 = F.chunked ? F.chunk_off[k0 + q] : F.chunk_off[k0] + q;
            const int i1 = F.chunked ? F.chunk_off[k0 + q + 1] : i0 + 1;
            for (int i = i0; i < i1; ++i) {
            const int c = F.phase_cells[i];
            const int* nn = &F.conn[(size_t)c * 4];
            double x[4][3];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) x[a][d] = F.coords[(size_t)nn[a] * 3 + d];
            const double j11 = x[0][0] - x[3][0], j12 = x[1][0] - x[3][0], j13 = x[2][0] - x[3][0];
            const double j21 = x[0][1] - x[3][1], j22 = x[1][1] - x[3][1], j23 = x[2][1] - x[3][1];
            const double j31 = x[0][2] - x[3][2], j32 = x[1][2] - x[3][2], j33 = x[2][2] - x[3][2];
            const double det = j11 * (j22 * j33 - j23 * j32) - j12 * (j21 * j33 - j23 * j31) +
                               j13 * (j21 * j32 - j22 * j31);
            const double id = 1.0 / det;
            double g[4][3];
            g[0][0] = (j22 * j33 - j23 * j32) * id;
            g[0][1] = (j13 * j32 - j12 * j33) * id;
            g[0][2] = (j12 * j23 - j13 * j22) * id;
            g[1][0] = (j23 * j31 - j21 * j33) * id;
            g[1][1] = (j11 * j33 - j13 * j31) * id;
            g[1][2] = (j13 * j21 - j11 * j23) * id;
            g[2][0] = (j21 * j32 - j22 * j31) * id;
            g[2][1] = (j12 * j31 - j11 * j32) * id;
            g[2][2] = (j11 * j22 - j12 * j21) * id;
            for (int d = 0; d < 3; ++d) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
            const double rho = 0.25 * (z[nn[0]] + z[nn[1]] + z[nn[2]] + z[nn[3]]);
            const double s = emin + (1.0 - emin) * std::pow(rho, p);
            const double w = s * std::fabs(det) / 6.0;
            const double wl = w * lam, wm = w * mu;
            double ue[4][3];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) ue[a][d] = u[(size_t)nn[a] * 3 + d];
            for (int a = 0; a < 4; ++a) {
                double ra[3] = {0.0, 0.0, 0.0};
                for (int b = 0; b < 4; ++b) {
                    const double dot = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
                    double K[3][3];
                    for (int ii = 0; ii < 3; ++ii)
                        for (int jj = 0; jj < 3; ++jj)
                            K[ii][jj] = wl * g[a][ii] * g[b][jj] + wm * g[a][jj] * g[b][ii] + (ii == jj ? wm * dot : 0.0);
                    double* blk = values + (size_t)F.ent[(size_t)c * 16 + a * 4 + b] * 9;
                    for (int ii = 0; ii < 3; ++ii)
                        for (int jj = 0; jj < 3; ++jj) {
                            blk[ii * 3 + jj] += K[ii][jj];
                            ra[ii] += K[ii][jj] * ue[b][jj];
                        }
                }
                for (int d = 0; d < 3; ++d) R[(size_t)nn[a] * 3 + d] += ra[d];
            }
            }
        }
    }
}

}  // extern "C"
