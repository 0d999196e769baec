// Host side of the library: error state, device helpers, the synthetic box
// mesh, and the one-time construction of the assembly graph and the
// node -> cell incidence the node-ordered kernels read.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "pa_internal.h"

namespace pa {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

// Split [0, n) over host threads.
template <class F>
static void parallel_for(long n, F&& f, long serial_below = 65536) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < serial_below || nt == 1) {
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

// Node grid of a lexicographically numbered mesh (z fastest): the run of
// nodes sharing node 0's (x, y), and the run sharing its x.  A numbering
// without such runs gets no tiles; one with them but not a true grid still
// gets correct tiles (arbitrary row groups), only fewer repeated ones.
static bool detect_grid(const double* xyz, int N, int d[3]) {
    int nz = 1;
    while (nz < N && xyz[3L * nz] == xyz[0] && xyz[3L * nz + 1] == xyz[1]) ++nz;
    int nyz = nz;
    while (nyz < N && xyz[3L * nyz] == xyz[0]) ++nyz;
    if (nz < 2 || nyz % nz != 0 || nyz / nz < 2 || N % nyz != 0) return false;
    d[0] = N / nyz, d[1] = nyz / nz, d[2] = nz;
    return true;
}

// One tile's schedule, relative to its first row.  The byte image of the
// tables is the tile's signature: tiles that are translates of each other
// (same row offsets, same cell node offsets, same column slots) coincide.
//   rows   row - first row, number of incident cells, offset of its pairs,
//          staging offset (blocks) of its output
//   runs   rows with consecutive ids (adjacent CRS ranges): first row index,
//          staging offset, blocks
//   cells  node ids - first row of each distinct incident cell, sorted
//   hslots per cell, the halo slots of its 4 nodes (8 bits each): the tile's
//          rows span a kTileX x kTileY x kTileZ block of the node grid and
//          every cell node lies in the block grown by one node on each side
//          (kTileHalo nodes, numbered like the grid, z fastest)
//   pairs  per row and incident cell (incidence order): tile cell << 2 | a
//   units  (row, column slot) blocks: x = row | slot << 8 | mirror row << 16
//          | mirror slot << 24, y = entry offset (16-byte groups) | count << 22
//          | diagonal << 26 | part << 27 | mirror << 29 | wave shape << 30.
//          A diagonal block is cut into kTileDiagParts parts on adjacent
//          units; a block whose column node is also a row of the tile is
//          computed once (by the lower row) and also stored transposed
//          (its mirror); the wave shape (per aligned group of 64 units)
//          selects an unpadded evaluation.
//   ents   per unit, kTileUnitEntries slots: LDS offsets (doubles) of the
//          scaled gradients g_a and g_b, 16 bits each; unused slots point at
//          the tile's zero cell (index = number of cells)
struct TileSchedule {
    std::vector<int4> rows, cells, runs;
    std::vector<unsigned short> pairs;
    std::vector<uint2> cslots;  // per cell: LDS pair slot (row * 32 + k) of its 4 nodes, 0xffff if not a row
    std::vector<unsigned> hslots;  // per cell: halo slots of its 4 nodes, 8 bits each
    std::vector<uint2> units;
    std::vector<unsigned> ents;
    int max_count = 0;
    std::string sig;
};

static bool tile_schedule(const pa_mesh* m, const std::vector<int>& conn, const std::vector<int>& off,
                          const std::vector<int2>& pairs, const std::vector<int>& rows, const int d[3],
                          TileSchedule& t) {
    const int base = rows[0];
    std::vector<int> cells;
    for (int r : rows)
        for (int p = off[r]; p < off[r + 1]; ++p) cells.push_back(pairs[p].x >> 2);
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    if ((int)cells.size() + 1 > kTileCellsMax) return false;  // + the zero cell
    auto key = [&](int c) {
        const int* q = &conn[4L * c];
        return std::array<int, 4>{q[0] - base, q[1] - base, q[2] - base, q[3] - base};
    };
    std::sort(cells.begin(), cells.end(), [&](int a, int b) { return key(a) < key(b); });
    std::vector<std::pair<int, int>> tc_of;  // (cell id, tile cell index)
    const int i0 = base / (d[1] * d[2]), j0 = (base / d[2]) % d[1], k0 = base % d[2];
    for (int i = 0; i < (int)cells.size(); ++i) {
        const auto k = key(cells[i]);
        t.cells.push_back(make_int4(k[0], k[1], k[2], k[3]));
        tc_of.push_back({cells[i], i});
        unsigned hs = 0;
        for (int n = 0; n < 4; ++n) {
            const int node = conn[4L * cells[i] + n];
            const int a = node / (d[1] * d[2]) - i0 + 1, b = (node / d[2]) % d[1] - j0 + 1, c = node % d[2] - k0 + 1;
            if (a < 0 || a >= kTileX + 2 || b < 0 || b >= kTileY + 2 || c < 0 || c >= kTileZ + 2) return false;
            hs |= (unsigned)((a * (kTileY + 2) + b) * (kTileZ + 2) + c) << (8 * n);
        }
        t.hslots.push_back(hs);
    }
    std::sort(tc_of.begin(), tc_of.end());
    auto tc = [&](int cell) { return std::lower_bound(tc_of.begin(), tc_of.end(), std::make_pair(cell, -1))->second; };
    const unsigned zc = (unsigned)cells.size() * kTileStride;
    struct Unit {
        int r, slot, mr, ms, part;
        bool diag, mirror;
        std::vector<unsigned> ent;
    };
    std::vector<Unit> units;
    int stage = 0;  // staging offset (blocks) of the current row
    auto row_index = [&](int node) {
        auto it = std::lower_bound(rows.begin(), rows.end(), node);
        return (it != rows.end() && *it == node) ? (int)(it - rows.begin()) : -1;
    };
    for (int ri = 0; ri < (int)rows.size(); ++ri) {
        const int r = rows[ri];
        const int np = off[r + 1] - off[r], rl = m->h_rowmap[r + 1] - m->h_rowmap[r];
        if (np > 32 || rl > 255) return false;
        t.rows.push_back(make_int4(r - base, np, (int)t.pairs.size(), stage));
        // rows with consecutive ids have adjacent CRS ranges: one output run
        if (ri > 0 && r == rows[ri - 1] + 1) t.runs.back().z += rl;
        else t.runs.push_back(make_int4(ri, stage, rl, 0));
        stage += rl;
        for (int p = off[r]; p < off[r + 1]; ++p)
            t.pairs.push_back((unsigned short)(tc(pairs[p].x >> 2) << 2 | (pairs[p].x & 3)));
        const int diag = np ? (int)(((unsigned)pairs[off[r]].y >> (8 * (pairs[off[r]].x & 3))) & 0xff) : 0;
        for (int slot = 0; slot < rl; ++slot) {
            const int col = m->h_colidx[m->h_rowmap[r] + slot];
            const int cj = slot == diag ? -1 : row_index(col);
            if (cj >= 0 && col < r) continue;  // computed by row col's unit and mirrored
            std::vector<unsigned> ent;
            for (int p = off[r]; p < off[r + 1]; ++p)
                for (int b = 0; b < 4; ++b)
                    if ((int)(((unsigned)pairs[p].y >> (8 * b)) & 0xff) == slot) {
                        const unsigned c = (unsigned)tc(pairs[p].x >> 2), a = (unsigned)(pairs[p].x & 3);
                        ent.push_back((c * kTileStride + 3 * a) | (c * kTileStride + 3 * (unsigned)b) << 16);
                    }
            if (slot == diag) {
                const int per = ((int)ent.size() + kTileDiagParts - 1) / kTileDiagParts;
                for (int k = 0; k < kTileDiagParts; ++k) {
                    const int b0 = std::min((int)ent.size(), k * per), b1 = std::min((int)ent.size(), b0 + per);
                    units.push_back({ri, slot, 0, 0, k, true, false, std::vector<unsigned>(ent.begin() + b0, ent.begin() + b1)});
                }
            } else {
                int ms = 0;
                if (cj >= 0) {  // slot of row r in row col
                    const int* c0 = &m->h_colidx[m->h_rowmap[col]];
                    ms = (int)(std::lower_bound(c0, c0 + (m->h_rowmap[col + 1] - m->h_rowmap[col]), r) - c0);
                }
                units.push_back({ri, slot, std::max(cj, 0), ms, 0, false, cj >= 0, std::move(ent)});
            }
        }
    }
    // diagonal parts first (their groups of kTileDiagParts stay on adjacent
    // lanes of one pass), then the others by decreasing entry count, so
    // most waves hold units of one shape
    static_assert(kTileThreads % kTileDiagParts == 0 && kTileThreads % 64 == 0, "groups within one pass");
    std::stable_sort(units.begin(), units.end(), [](const Unit& a, const Unit& b) {
        if (a.diag != b.diag) return a.diag;
        if (a.diag) return false;
        return a.ent.size() > b.ent.size();
    });
    // pair slots of the residual kernel: pair k of row ri lives at ri * 32 + k
    {
        std::vector<unsigned short> cs(4 * cells.size(), 0xffff);
        for (int ri = 0; ri < (int)t.rows.size(); ++ri)
            for (int k = 0; k < t.rows[ri].y; ++k) {
                const unsigned pe = t.pairs[t.rows[ri].z + k];
                cs[4 * (pe >> 2) + (pe & 3)] = (unsigned short)(ri * 32 + k);
            }
        for (size_t c = 0; c < cells.size(); ++c)
            t.cslots.push_back(make_uint2((unsigned)cs[4 * c] | (unsigned)cs[4 * c + 1] << 16,
                                          (unsigned)cs[4 * c + 2] | (unsigned)cs[4 * c + 3] << 16));
    }
    for (const Unit& u : units) t.max_count = std::max(t.max_count, (int)u.ent.size());
    if (t.max_count > kTileUnitEntries) return false;
    // shape of each wave (aligned group of 64 units): 1 all diagonal parts of
    // 6 entries, 2 all off-diagonal with 6, 3 all off-diagonal with 4, else 0
    auto shape = [](const Unit& u) {
        if (u.diag) return u.ent.size() == 6 ? 1u : 0u;
        return u.ent.size() == 6 ? 2u : (u.ent.size() == 4 ? 3u : 0u);
    };
    for (size_t g = 0; g < units.size(); g += 64) {
        const size_t e = std::min(units.size(), g + 64);
        unsigned cls = shape(units[g]);
        for (size_t i = g; i < e; ++i)
            if (shape(units[i]) != cls) cls = 0;
        for (size_t i = g; i < e; ++i) {
            const Unit& u = units[i];
            const unsigned eoff = (unsigned)t.ents.size() / 4;
            if (eoff >= (1u << 22)) return false;
            t.ents.insert(t.ents.end(), u.ent.begin(), u.ent.end());
            for (size_t k = u.ent.size(); k < (size_t)kTileUnitEntries; ++k) t.ents.push_back(zc | zc << 16);
            const unsigned x = (unsigned)u.r | (unsigned)u.slot << 8 | (unsigned)u.mr << 16 | (unsigned)u.ms << 24;
            const unsigned y = eoff | (unsigned)u.ent.size() << 22 | (u.diag ? 1u : 0u) << 26 | (unsigned)u.part << 27 |
                               (u.mirror ? 1u : 0u) << 29 | cls << 30;
            t.units.push_back(make_uint2(x, y));
        }
    }
    if ((int)t.units.size() > kTileUnitsMax) return false;
    auto put = [&](const void* p, size_t n) { t.sig.append(static_cast<const char*>(p), n); };
    const int counts[6] = {(int)t.rows.size(), (int)t.cells.size(), (int)t.pairs.size(), (int)t.units.size(),
                           (int)t.ents.size(), (int)t.runs.size()};
    put(counts, sizeof(counts));
    put(t.rows.data(), t.rows.size() * sizeof(int4));
    put(t.runs.data(), t.runs.size() * sizeof(int4));
    put(t.cells.data(), t.cells.size() * sizeof(int4));
    put(t.hslots.data(), t.hslots.size() * sizeof(unsigned));
    put(t.pairs.data(), t.pairs.size() * sizeof(unsigned short));
    put(t.units.data(), t.units.size() * sizeof(uint2));
    put(t.ents.data(), t.ents.size() * sizeof(unsigned));
    return true;
}

static void build_tiles(pa_mesh* m, const double* xyz, const std::vector<int>& conn, const std::vector<int>& off,
                        const std::vector<int2>& pairs) {
    m->has_tiles = false;
    const char* env = std::getenv("PA_TILES");
    if (env && env[0] == '0') return;
    if (m->max_cells_per_node > 32) return;
    int d[3];
    if (!detect_grid(xyz, m->N, d)) return;
    const int tx = (d[0] + kTileX - 1) / kTileX, ty = (d[1] + kTileY - 1) / kTileY, tz = (d[2] + kTileZ - 1) / kTileZ;
    const long nt = (long)tx * ty * tz;
    auto first_row = [&](long t) {
        const int i0 = (int)(t / ((long)ty * tz)) * kTileX, j0 = (int)((t / tz) % ty) * kTileY,
                  k0 = (int)(t % tz) * kTileZ;
        return std::array<int, 3>{i0, j0, k0};
    };
    std::unordered_map<std::string, int> index;
    std::vector<int4> tiles(3 * nt), info, trow, tcell, trun;
    std::vector<uint2> tcslot;
    std::vector<unsigned> thslot;
    std::vector<unsigned short> tpair;
    std::vector<uint2> tunit;
    std::vector<unsigned> tent;
    int cmax = 0, rmax = 0, emax = 0, bmax = 0;
    // schedules are built in parallel batches and deduplicated in tile order
    constexpr long kBatch = 2048;
    std::vector<TileSchedule> sched(std::min(nt, kBatch));
    std::vector<char> ok(sched.size());
    for (long t0 = 0; t0 < nt; t0 += kBatch) {
        const long nb = std::min(kBatch, nt - t0);
        std::fill(ok.begin(), ok.end(), 1);
        parallel_for(nb, [&](long b, long e) {
            std::vector<int> rows;
            for (long q = b; q < e; ++q) {
                const auto f = first_row(t0 + q);
                rows.clear();
                for (int i = f[0]; i < std::min(f[0] + kTileX, d[0]); ++i)
                    for (int j = f[1]; j < std::min(f[1] + kTileY, d[1]); ++j)
                        for (int k = f[2]; k < std::min(f[2] + kTileZ, d[2]); ++k)
                            rows.push_back((i * d[1] + j) * d[2] + k);
                sched[q] = TileSchedule();
                ok[q] = tile_schedule(m, conn, off, pairs, rows, d, sched[q]);
            }
        }, 64);
        for (long q = 0; q < nb; ++q) {
            if (!ok[q]) return;
            TileSchedule& s = sched[q];
            auto it = index.find(s.sig);
            if (it == index.end()) {
                if ((int)index.size() >= kTilePatternsMax) return;
                it = index.emplace(s.sig, (int)index.size()).first;
                const unsigned eoff = (unsigned)tent.size() / 4;
                if (eoff + s.ents.size() / 4 >= (1u << 22)) return;  // entry offsets are 22-bit
                for (auto& w : s.units) w.y = (w.y & ~0x3fffffu) | ((w.y & 0x3fffffu) + eoff);
                const int nblk = s.runs.empty() ? 0 : s.runs.back().y + s.runs.back().z;
                info.push_back(make_int4(0, (int)trow.size(), (int)tcell.size(), (int)tunit.size()));
                info.push_back(make_int4((int)s.rows.size(), (int)s.cells.size(), (int)s.units.size(),
                                         (int)s.runs.size()));
                info.push_back(make_int4((int)trun.size(), nblk, 0, 0));
                trun.insert(trun.end(), s.runs.begin(), s.runs.end());
                bmax = std::max(bmax, nblk);
                const int poff = (int)tpair.size();
                for (auto r : s.rows) trow.push_back(make_int4(r.x, r.y, r.z + poff, r.w));
                tcell.insert(tcell.end(), s.cells.begin(), s.cells.end());
                tcslot.insert(tcslot.end(), s.cslots.begin(), s.cslots.end());
                thslot.insert(thslot.end(), s.hslots.begin(), s.hslots.end());
                tpair.insert(tpair.end(), s.pairs.begin(), s.pairs.end());
                tunit.insert(tunit.end(), s.units.begin(), s.units.end());
                tent.insert(tent.end(), s.ents.begin(), s.ents.end());
                cmax = std::max(cmax, (int)s.cells.size());
                rmax = std::max(rmax, (int)s.rows.size());
                emax = std::max(emax, s.max_count);
            }
            const auto f = first_row(t0 + q);
            for (int k = 0; k < 3; ++k) tiles[3 * (t0 + q) + k] = info[3 * it->second + k];
            tiles[3 * (t0 + q)].x = (f[0] * d[1] + f[1]) * d[2] + f[2];
        }
    }
    if (tent.size() >= (1u << 31) || rmax > 64) return;
    m->num_tiles = (int)nt;
    m->num_tile_patterns = (int)index.size();
    m->tile_cells_max = cmax;
    m->tile_rows_max = rmax;
    m->tile_unit_entries = emax <= 6 ? 6 : 8;
    m->tile_blocks_max = bmax;
    m->grid_dims[0] = d[0], m->grid_dims[1] = d[1], m->grid_dims[2] = d[2];
    m->d_tile.upload(tiles.data(), tiles.size());
    m->d_tp_row.upload(trow.data(), trow.size());
    m->d_tp_cell.upload(tcell.data(), tcell.size());
    m->d_tp_cslot.upload(tcslot.data(), tcslot.size());
    m->d_tp_hslot.upload(thslot.data(), thslot.size());
    m->d_tp_run.upload(trun.data(), trun.size());
    m->d_tp_pair.upload(tpair.data(), tpair.size());
    m->d_tp_unit.upload(tunit.data(), tunit.size());
    m->d_tp_ent.upload(tent.data(), tent.size());
    m->has_tiles = true;
}

// Kuhn grid check for the streaming grid kernels: the nodes form a
// lexicographic grid (z fastest) and the cells are exactly the six tets of
// every hex around its v0 -> v7 diagonal (in any order, any vertex order), so
// the graph is the 15-point stencil clipped at the boundary.
static void build_kuhn_grid(pa_mesh* m, const double* xyz, const std::vector<int>& conn) {
    m->has_grid = false;
    const char* env = std::getenv("PA_GRID");
    if (env && env[0] == '0') return;
    int d[3];
    if (!detect_grid(xyz, m->N, d) || d[0] < 2) return;
    const long nh = (long)(d[0] - 1) * (d[1] - 1) * (d[2] - 1);
    if ((long)m->C != 6 * nh) return;
    // (only the numbering matters: the kernels take every cell's geometry
    // from its coordinates)
    static const int kTets[6][4] = {{0, 2, 6, 7}, {0, 2, 3, 7}, {0, 1, 3, 7}, {0, 1, 5, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
    std::vector<unsigned char> mask(nh, 0);
    const long s1 = d[2], s0 = (long)d[1] * d[2];
    for (long c = 0; c < m->C; ++c) {
        int ii[4], jj[4], kk[4];
        for (int a = 0; a < 4; ++a) {
            const long n = conn[4 * c + a];
            ii[a] = (int)(n / s0), jj[a] = (int)((n / s1) % d[1]), kk[a] = (int)(n % s1);
        }
        const int i0 = *std::min_element(ii, ii + 4), j0 = *std::min_element(jj, jj + 4),
                  k0 = *std::min_element(kk, kk + 4);
        if (i0 >= d[0] - 1 || j0 >= d[1] - 1 || k0 >= d[2] - 1) return;
        int v[4];
        for (int a = 0; a < 4; ++a) {
            const int di = ii[a] - i0, dj = jj[a] - j0, dk = kk[a] - k0;
            if (di > 1 || dj > 1 || dk > 1) return;
            v[a] = di << 2 | dj << 1 | dk;
        }
        std::sort(v, v + 4);
        int t = 0;
        while (t < 6 && !std::equal(v, v + 4, kTets[t])) ++t;
        if (t == 6) return;
        const long h = ((long)i0 * (d[1] - 1) + j0) * (d[2] - 1) + k0;
        if (mask[h] & (1u << t)) return;  // the same tet twice
        mask[h] |= (unsigned char)(1u << t);
    }
    for (long h = 0; h < nh; ++h)
        if (mask[h] != 0x3f) return;
    // every row is the clipped 15-point stencil
    for (int n = 0; n < m->N; ++n) {
        const int i = n / (int)s0, j = (n / (int)s1) % d[1], k = n % (int)s1;
        int len = 0;
        static const int off[15][3] = {{-1, -1, -1}, {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 0}, {0, -1, -1},
                                       {0, -1, 0},   {0, 0, -1},  {0, 0, 0},   {0, 0, 1},  {0, 1, 0},
                                       {0, 1, 1},    {1, 0, 0},   {1, 0, 1},   {1, 1, 0},  {1, 1, 1}};
        for (const auto& o : off) {
            const int a = i + o[0], b = j + o[1], c = k + o[2];
            len += a >= 0 && a < d[0] && b >= 0 && b < d[1] && c >= 0 && c < d[2];
        }
        if (m->h_rowmap[n + 1] - m->h_rowmap[n] != len) return;
    }
    m->kgrid[0] = d[0], m->kgrid[1] = d[1], m->kgrid[2] = d[2];
    m->has_grid = true;
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
            {
                std::vector<int> cv(conn, conn + 4L * C);
                pa::build_tiles(m, coords, cv, off, pairs);
                pa::build_kuhn_grid(m, coords, cv);
            }
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

int pa_mesh_assembly_info(const pa_mesh* m, int* tiles, int* tile_patterns, int* row_patterns) {
    return guarded([&] {
        PA_REQUIRE(m, PA_ERR_ARG, "pa_mesh_assembly_info: null mesh");
        if (tiles) *tiles = m->has_tiles ? m->num_tiles : 0;
        if (tile_patterns) *tile_patterns = m->has_tiles ? m->num_tile_patterns : 0;
        if (row_patterns) *row_patterns = m->has_patterns ? m->num_patterns : 0;
    });
}

int pa_mesh_grid_info(const pa_mesh* m, int* dims) {
    return guarded([&] {
        PA_REQUIRE(m && dims, PA_ERR_ARG, "pa_mesh_grid_info: null argument");
        for (int d = 0; d < 3; ++d) dims[d] = m->has_grid ? m->kgrid[d] : 0;
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
        ++m->side_set_version;
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
