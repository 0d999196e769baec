// Host side of the library: error state, device helpers, the synthetic box
// mesh, and the one-time construction of the assembly graph and the
// node -> cell incidence the node-ordered kernels read.
#include <algorithm>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "pa_internal.h"

namespace pa {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

// Split [0, n) over host threads.
template <class F>
static void parallel_for(long n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536 || nt == 1) {
        f(0L, n);
        return;
    }
    std::vector<std::thread> th;
    long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        long b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : th) t.join();
}

// Row patterns.  A row's schedule is fixed by its sequence of incident cells'
// (local index of the row node, column slots of the four nodes); rows with the
// same sequence (all interior rows of a structured mesh, for instance) share
// one table, so the per-row index data the kernels read is one pattern id
// plus the incident cell ids.  For each pattern:
//   a[p]           local index of the row node in pair p's cell;
//   entries        (p << 2 | b) grouped by column slot, pairs ascending;
//   items          one per off-diagonal slot, and the diagonal slot cut into
//                  chunks of kDiagChunk entries (slot | start << 8 | count << 20).
// Limits of the pattern kernels: <= 32 pairs, <= 32 slots, <= 32 items,
// <= 4 diagonal chunks per row; a mesh outside them keeps the generic kernels.
static void build_row_patterns(pa_mesh* m, const std::vector<int>& off, const std::vector<int2>& pairs) {
    constexpr int kMaxPairs = 32, kMaxSlots = 32, kMaxItems = 32, kDiagChunk = 8;
    const int N = m->N;
    m->has_patterns = false;
    if (m->max_cells_per_node > kMaxPairs || m->max_row_len > kMaxSlots) return;
    std::unordered_map<std::string, int> index;
    std::vector<int> row_pat(N);
    std::vector<std::string> sigs;
    std::string sig;
    constexpr int kRec = 5 + 16;  // a, four slots, four int32 node deltas
    std::vector<int> conn(4L * m->C);
    PA_HIP(hipMemcpy(conn.data(), m->d_conn.p, conn.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < N; ++i) {
        const int np = off[i + 1] - off[i];
        const int rl = m->h_rowmap[i + 1] - m->h_rowmap[i];
        sig.assign(2 + kRec * (size_t)np, '\0');
        sig[0] = (char)np;
        sig[1] = (char)rl;
        for (int p = 0; p < np; ++p) {
            const int2 pr = pairs[off[i] + p];
            char* r = &sig[2 + kRec * p];
            r[0] = (char)(pr.x & 3);
            for (int b = 0; b < 4; ++b) r[1 + b] = (char)(((unsigned)pr.y >> (8 * b)) & 0xff);
            for (int b = 0; b < 4; ++b) {
                const int d = conn[(long)(pr.x >> 2) * 4 + b] - i;
                std::memcpy(r + 5 + 4 * b, &d, 4);
            }
        }
        auto it = index.find(sig);
        if (it == index.end()) {
            it = index.emplace(sig, (int)sigs.size()).first;
            sigs.push_back(sig);
        }
        row_pat[i] = it->second;
    }
    std::vector<int4> info, pa_delta;
    std::vector<unsigned char> pa_a, pa_ent;
    std::vector<unsigned> pa_item;
    std::vector<uint4> pa_item2;
    for (const std::string& s : sigs) {
        const int np = (unsigned char)s[0], rl = (unsigned char)s[1];
        auto A = [&](int p) { return (int)(unsigned char)s[2 + kRec * p]; };
        auto S = [&](int p, int b) { return (int)(unsigned char)s[3 + kRec * p + b]; };
        const int diag = np ? S(0, A(0)) : 0;
        while (pa_ent.size() % 4) pa_ent.push_back(0);  // kernels read the entries as 32-bit words
        const int a_off = (int)pa_a.size(), ent_off = (int)pa_ent.size(), item_off = (int)pa_item.size();
        for (int p = 0; p < np; ++p) {
            pa_a.push_back((unsigned char)A(p));
            int d[4];
            std::memcpy(d, &s[2 + kRec * p + 5], 16);
            pa_delta.push_back(make_int4(d[0], d[1], d[2], d[3]));
        }
        int first_diag = -1;
        std::vector<unsigned char> slot_ent;
        auto push_item = [&](int slot, int first, int count) {
            // 16-byte item: slot | count << 8, then up to 8 entry bytes
            unsigned w[4] = {(unsigned)slot | (unsigned)count << 8, 0u, 0u, 0u};
            for (int e = 0; e < count; ++e) w[1 + e / 4] |= (unsigned)slot_ent[first + e] << (8 * (e % 4));
            pa_item2.push_back(make_uint4(w[0], w[1], w[2], w[3]));
            pa_item.push_back((unsigned)slot | (unsigned)count << 20);
        };
        bool fits = true;
        for (int slot = 0; slot < rl; ++slot) {
            slot_ent.clear();
            for (int p = 0; p < np; ++p)
                for (int b = 0; b < 4; ++b)
                    if (S(p, b) == slot) slot_ent.push_back((unsigned char)(p << 2 | b));
            for (unsigned char e : slot_ent) pa_ent.push_back(e);
            const int count = (int)slot_ent.size();
            if (slot == diag) {
                first_diag = (int)pa_item.size() - item_off;
                for (int c = 0; c < count; c += kDiagChunk) push_item(slot, c, std::min(kDiagChunk, count - c));
                if (count == 0) push_item(slot, 0, 0);
            } else {
                if (count > 8) fits = false;
                push_item(slot, 0, std::min(count, 8));
            }
        }
        const int nitems = (int)pa_item.size() - item_off;
        const int ndiag = (np + kDiagChunk - 1) / kDiagChunk;
        if (!fits || nitems > kMaxItems || ndiag > 4) return;
        info.push_back(make_int4(np, rl, nitems, diag));
        info.push_back(make_int4(a_off, ent_off, item_off, first_diag));
    }
    m->num_patterns = (int)sigs.size();
    std::vector<int> pcell(pairs.size());
    for (size_t p = 0; p < pairs.size(); ++p) pcell[p] = pairs[p].x >> 2;
    m->d_pcell.upload(pcell.data(), pcell.size());
    m->d_row_pat.upload(row_pat.data(), row_pat.size());
    m->d_pat_info.upload(info.data(), info.size());
    m->d_pat_a.upload(pa_a.data(), pa_a.size());
    m->d_pat_delta.upload(pa_delta.data(), pa_delta.size());
    m->d_pat_item2.upload(pa_item2.data(), pa_item2.size());
    m->d_pat_ent.upload(pa_ent.data(), pa_ent.size());
    m->d_pat_item.upload(pa_item.data(), pa_item.size());
    m->has_patterns = true;
}

}  // namespace pa

using pa::guarded;

extern "C" {

const char* pa_last_error(void) { return pa::g_last_error.c_str(); }
const char* pa_version(void) { return "platoanalyze_amd 0.1 (gfx950)"; }

int pa_device_count(int* count) {
    return guarded([&] { PA_HIP(hipGetDeviceCount(count)); });
}
int pa_malloc(void** p, size_t bytes) {
    return guarded([&] { PA_HIP(hipMalloc(p, bytes)); });
}
int pa_free(void* p) {
    return guarded([&] { PA_HIP(hipFree(p)); });
}
int pa_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return guarded([&] { PA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int pa_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return guarded([&] { PA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int pa_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] { PA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream)); });
}
int pa_memset(void* p, int value, size_t bytes, void* stream) {
    return guarded([&] { PA_HIP(hipMemsetAsync(p, value, bytes, (hipStream_t)stream)); });
}
int pa_stream_create(void** s) {
    return guarded([&] { PA_HIP(hipStreamCreate((hipStream_t*)s)); });
}
int pa_stream_destroy(void* s) {
    return guarded([&] { PA_HIP(hipStreamDestroy((hipStream_t)s)); });
}
int pa_stream_synchronize(void* s) {
    return guarded([&] { PA_HIP(hipStreamSynchronize((hipStream_t)s)); });
}
int pa_event_create(void** e) {
    return guarded([&] { PA_HIP(hipEventCreate((hipEvent_t*)e)); });
}
int pa_event_destroy(void* e) {
    return guarded([&] { PA_HIP(hipEventDestroy((hipEvent_t)e)); });
}
int pa_event_record(void* e, void* s) {
    return guarded([&] { PA_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); });
}
int pa_event_synchronize(void* e) {
    return guarded([&] { PA_HIP(hipEventSynchronize((hipEvent_t)e)); });
}
int pa_event_elapsed_ms(float* ms, void* a, void* b) {
    return guarded([&] { PA_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b)); });
}

int pa_box_mesh(int nx, int ny, int nz, double lx, double ly, double lz, double* coords, int* conn) {
    return guarded([&] {
        PA_REQUIRE(nx > 0 && ny > 0 && nz > 0, PA_ERR_ARG, "pa_box_mesh: grid sizes must be positive");
        PA_REQUIRE(coords && conn, PA_ERR_ARG, "pa_box_mesh: null output");
        PA_REQUIRE((long long)6 * nx * ny * nz < (1LL << 30), PA_ERR_LIMIT, "pa_box_mesh: more than 2^30 cells");
        const long NY = ny + 1, NZ = nz + 1;
        for (long ix = 0; ix <= nx; ++ix)
            for (long iy = 0; iy <= ny; ++iy)
                for (long iz = 0; iz <= nz; ++iz) {
                    long n = (ix * NY + iy) * NZ + iz;
                    coords[n * 3 + 0] = lx * (double)ix / nx;
                    coords[n * 3 + 1] = ly * (double)iy / ny;
                    coords[n * 3 + 2] = lz * (double)iz / nz;
                }
        // Six tets per hex sharing the (0,0,0)-(1,1,1) diagonal; vertices as
        // bit offsets (dx<<2 | dy<<1 | dz), tet vertex order fixed.
        static const int tet[6][4] = {{0, 6, 2, 7}, {0, 2, 3, 7}, {0, 3, 1, 7},
                                      {0, 1, 5, 7}, {0, 5, 4, 7}, {0, 4, 6, 7}};
        pa::parallel_for(nx, [&](long b, long e) {
            for (long ix = b; ix < e; ++ix)
                for (long iy = 0; iy < ny; ++iy)
                    for (long iz = 0; iz < nz; ++iz) {
                        long h = (ix * ny + iy) * nz + iz;
                        for (int t = 0; t < 6; ++t)
                            for (int a = 0; a < 4; ++a) {
                                int v = tet[t][a];
                                long n = ((ix + (v >> 2)) * NY + (iy + ((v >> 1) & 1))) * NZ + (iz + (v & 1));
                                conn[(h * 6 + t) * 4 + a] = (int)n;
                            }
                    }
        });
    });
}

int pa_mesh_create(const double* coords, int N, const int* conn, int C, pa_mesh** out) {
    return guarded([&] {
        PA_REQUIRE(out, PA_ERR_ARG, "pa_mesh_create: null output");
        PA_REQUIRE(N > 0 && C > 0 && coords && conn, PA_ERR_ARG, "pa_mesh_create: empty mesh");
        PA_REQUIRE(C < (1 << 30), PA_ERR_LIMIT, "pa_mesh_create: more than 2^30 cells");
        for (long i = 0; i < 4L * C; ++i)
            PA_REQUIRE(conn[i] >= 0 && conn[i] < N, PA_ERR_ARG, "pa_mesh_create: connectivity out of range");
        auto* m = new pa_mesh();
        try {
            m->N = N;
            m->C = C;
            // node -> cell incidence, cells ascending per node
            std::vector<int> off(N + 1, 0);
            for (long i = 0; i < 4L * C; ++i) off[conn[i] + 1]++;
            for (int i = 0; i < N; ++i) off[i + 1] += off[i];
            std::vector<int> fill(off.begin(), off.end() - 1);
            std::vector<int> pc(4L * C), pa_(4L * C);
            for (int c = 0; c < C; ++c)
                for (int a = 0; a < 4; ++a) {
                    int n = conn[(long)c * 4 + a];
                    int p = fill[n]++;
                    pc[p] = c;
                    pa_[p] = a;
                }
            int maxc = 0;
            for (int i = 0; i < N; ++i) maxc = std::max(maxc, off[i + 1] - off[i]);
            m->max_cells_per_node = maxc;
            // rows: sorted distinct nodes of the incident cells (diagonal included)
            std::vector<int> rowlen(N);
            pa::parallel_for(N, [&](long b, long e) {
                std::vector<int> buf;
                for (long i = b; i < e; ++i) {
                    buf.clear();
                    for (int p = off[i]; p < off[i + 1]; ++p)
                        for (int a = 0; a < 4; ++a) buf.push_back(conn[(long)pc[p] * 4 + a]);
                    if (buf.empty()) buf.push_back((int)i);
                    std::sort(buf.begin(), buf.end());
                    rowlen[i] = (int)(std::unique(buf.begin(), buf.end()) - buf.begin());
                }
            });
            m->h_rowmap.assign(N + 1, 0);
            long long acc = 0;
            int maxr = 0;
            for (int i = 0; i < N; ++i) {
                acc += rowlen[i];
                PA_REQUIRE(acc < (1LL << 31), PA_ERR_LIMIT, "pa_mesh_create: more than 2^31 graph blocks");
                m->h_rowmap[i + 1] = (int)acc;
                maxr = std::max(maxr, rowlen[i]);
            }
            m->nnzb = acc;
            m->max_row_len = maxr;
            PA_REQUIRE(maxr <= 255, PA_ERR_LIMIT, "pa_mesh_create: node degree above 254");
            m->h_colidx.assign(acc, 0);
            std::vector<int2> pairs(4L * C);
            pa::parallel_for(N, [&](long b, long e) {
                std::vector<int> buf;
                for (long i = b; i < e; ++i) {
                    buf.clear();
                    for (int p = off[i]; p < off[i + 1]; ++p)
                        for (int a = 0; a < 4; ++a) buf.push_back(conn[(long)pc[p] * 4 + a]);
                    if (buf.empty()) buf.push_back((int)i);
                    std::sort(buf.begin(), buf.end());
                    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
                    std::copy(buf.begin(), buf.end(), m->h_colidx.begin() + m->h_rowmap[i]);
                    for (int p = off[i]; p < off[i + 1]; ++p) {
                        unsigned slots = 0;
                        for (int a = 0; a < 4; ++a) {
                            int n = conn[(long)pc[p] * 4 + a];
                            unsigned s = (unsigned)(std::lower_bound(buf.begin(), buf.end(), n) - buf.begin());
                            slots |= s << (8 * a);
                        }
                        pairs[p] = make_int2((pc[p] << 2) | pa_[p], (int)slots);
                    }
                }
            });
            std::vector<double> c4(4L * N);
            for (long i = 0; i < N; ++i) {
                c4[i * 4 + 0] = coords[i * 3 + 0];
                c4[i * 4 + 1] = coords[i * 3 + 1];
                c4[i * 4 + 2] = coords[i * 3 + 2];
                c4[i * 4 + 3] = 0.0;
            }
            m->d_coords.upload(coords, 3L * N);
            m->d_coords4.upload(c4.data(), c4.size());
            m->d_conn.upload(conn, 4L * C);
            m->d_rowmap.upload(m->h_rowmap.data(), m->h_rowmap.size());
            m->d_colidx.upload(m->h_colidx.data(), m->h_colidx.size());
            m->d_n2c_off.upload(off.data(), off.size());
            m->d_pairs.upload(pairs.data(), pairs.size());
            pa::build_row_patterns(m, off, pairs);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int pa_mesh_destroy(pa_mesh* m) {
    return guarded([&] { delete m; });
}

int pa_mesh_sizes(const pa_mesh* m, int* N, int* C, int64_t* nnzb) {
    return guarded([&] {
        PA_REQUIRE(m, PA_ERR_ARG, "pa_mesh_sizes: null mesh");
        if (N) *N = m->N;
        if (C) *C = m->C;
        if (nnzb) *nnzb = m->nnzb;
    });
}

int pa_mesh_graph(const pa_mesh* m, int* rowmap, int* colidx) {
    return guarded([&] {
        PA_REQUIRE(m, PA_ERR_ARG, "pa_mesh_graph: null mesh");
        if (rowmap) std::memcpy(rowmap, m->h_rowmap.data(), m->h_rowmap.size() * sizeof(int));
        if (colidx) std::memcpy(colidx, m->h_colidx.data(), m->h_colidx.size() * sizeof(int));
    });
}

int pa_mesh_graph_dev(const pa_mesh* m, const int** rowmap, const int** colidx) {
    return guarded([&] {
        PA_REQUIRE(m, PA_ERR_ARG, "pa_mesh_graph_dev: null mesh");
        if (rowmap) *rowmap = m->d_rowmap.p;
        if (colidx) *colidx = m->d_colidx.p;
    });
}

int pa_mesh_add_side_set(pa_mesh* m, const char* name, int F, const int* cells, const int* lnodes) {
    return guarded([&] {
        PA_REQUIRE(m && name, PA_ERR_ARG, "pa_mesh_add_side_set: null argument");
        PA_REQUIRE(F >= 0 && (F == 0 || (cells && lnodes)), PA_ERR_ARG, "pa_mesh_add_side_set: bad faces");
        pa::SideSet& s = m->side_sets[name];
        s.cells.assign(cells, cells + F);
        s.lnodes.assign(lnodes, lnodes + 3L * F);
        for (int f = 0; f < F; ++f) {
            PA_REQUIRE(cells[f] >= 0 && cells[f] < m->C, PA_ERR_ARG, "pa_mesh_add_side_set: cell out of range");
            for (int k = 0; k < 3; ++k)
                PA_REQUIRE(lnodes[3 * f + k] >= 0 && lnodes[3 * f + k] < 4, PA_ERR_ARG,
                           "pa_mesh_add_side_set: local node out of range");
        }
        // boundary node -> (face, position) lists, nodes ascending, faces ascending
        std::vector<std::pair<int, int>> nf;
        std::vector<int> conn(4L * m->C);
        PA_HIP(hipMemcpy(conn.data(), m->d_conn.p, conn.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int f = 0; f < F; ++f)
            for (int k = 0; k < 3; ++k) nf.push_back({conn[(long)cells[f] * 4 + lnodes[3 * f + k]], f * 4 + k});
        std::sort(nf.begin(), nf.end());
        std::vector<int> bnodes, boff{0}, bface;
        for (size_t i = 0; i < nf.size(); ++i) {
            if (bnodes.empty() || bnodes.back() != nf[i].first) {
                if (!bnodes.empty()) boff.push_back((int)bface.size());
                bnodes.push_back(nf[i].first);
            }
            bface.push_back(nf[i].second);
        }
        boff.push_back((int)bface.size());
        if (bnodes.empty()) boff.assign(1, 0);
        s.num_bnodes = (int)bnodes.size();
        s.d_cells.upload(s.cells.data(), s.cells.size());
        s.d_lnodes.upload(s.lnodes.data(), s.lnodes.size());
        s.d_bnodes.upload(bnodes.data(), bnodes.size());
        s.d_bface_off.upload(boff.data(), boff.size());
        s.d_bface.upload(bface.data(), bface.size());
    });
}

int pa_mesh_add_node_set(pa_mesh* m, const char* name, int n, const int* nodes) {
    return guarded([&] {
        PA_REQUIRE(m && name && n >= 0 && (n == 0 || nodes), PA_ERR_ARG, "pa_mesh_add_node_set: bad argument");
        for (int i = 0; i < n; ++i)
            PA_REQUIRE(nodes[i] >= 0 && nodes[i] < m->N, PA_ERR_ARG, "pa_mesh_add_node_set: node out of range");
        m->node_sets[name].assign(nodes, nodes + n);
    });
}

int pa_mesh_node_set(const pa_mesh* m, const char* name, int* n, int* nodes) {
    return guarded([&] {
        PA_REQUIRE(m && name, PA_ERR_ARG, "pa_mesh_node_set: null argument");
        auto it = m->node_sets.find(name);
        PA_REQUIRE(it != m->node_sets.end(), PA_ERR_ARG, std::string("pa_mesh_node_set: no node set ") + name);
        if (n) *n = (int)it->second.size();
        if (nodes) std::copy(it->second.begin(), it->second.end(), nodes);
    });
}

}  // extern "C"
