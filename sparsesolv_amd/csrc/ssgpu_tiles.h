// Tile layout of a triangular sweep for the tiled sweep kernel (host code,
// no device dependencies).
//
// A tile is the rows of a band of consecutive levels that fall in one
// interval of row index.  The rows of a band are walked in row-index order --
// ascending for a forward sweep, whose rows read smaller indices, descending
// for a backward one -- and cut into intervals whose shared memory stays
// under a cap, with at most kTileThreads * kTileExtPerLane external inputs and
// 65535 slots.  When a row joins, each of its inputs is already placed: it
// lies at a lower level, so in an earlier band or earlier in this band's walk,
// hence in the open tile or in an earlier one.  Taking tiles in (band,
// interval) order therefore never waits on a later tile.  Inside a tile, rows
// are packed level by level (the steps), in walk order; the inputs of a row
// get LDS slots: the tile's own rows first (in packed order), then the
// external rows in the order the packed entries first name them.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <unordered_map>
#include <vector>

namespace ssgpu {

constexpr int kTileThreads = 64;     // one wave per tile
constexpr int kTileExtPerLane = 16;  // external inputs a lane polls (a tile has at most 64 * 16)
constexpr int kTileMaxSteps = 64;    // levels per tile (one per lane of the step table)

// bytes of a tile's shared memory: value slots (in-tile rows, then
// externals), per-row right-hand side and diagonal, entry values and slots,
// per-row row index and entry offsets
inline size_t tile_lds_bytes(int rows, int ext, int64_t ents, size_t sT, size_t sF) {
    return (size_t)(rows + ext) * sT + 2 * (size_t)rows * sT + (size_t)ents * sF + (2 * (size_t)rows + 1) * 4 +
           (size_t)ents * 2;
}

struct TileLayout {
    std::vector<int> tile_rp{0};    // ntiles + 1: packed rows of each tile
    std::vector<int> tile_sp{0};    // ntiles + 1: steps of each tile in step_rp
    std::vector<int> step_rp;       // first packed row of each step
    std::vector<int> tile_xp{0};    // ntiles + 1: external inputs of each tile in ext_rows
    std::vector<int> ext_rows;      // rows a tile reads from earlier tiles
    std::vector<int> tpos;          // packed row -> level-order position
    std::vector<int> tent{0};       // packed row -> first packed entry
    std::vector<unsigned short> tslot;   // packed entry -> LDS slot of its input
    size_t lds[4] = {0, 0, 0, 0};   // largest tile's bytes: FP64/FP64, FP64/FP32, complex/complex, complex/complex64
};

// n rows; order: level-order position -> row; row i's off-diagonal inputs are
// col[off[i] .. off[i+1]) in summation order; levptr: level starts in the
// level order.  vsize: bytes of a value of the matrix (8 or 16).
inline TileLayout tile_layout(int n, const std::vector<int>& order, const std::vector<int>& off, const int* col,
                              const std::vector<int>& levptr, bool forward, int band, size_t lds_cap,
                              size_t vsize) {
    const int nlev = (int)levptr.size() - 1;
    band = std::max(1, std::min(band, kTileMaxSteps));
    std::vector<int> lev_of_row(n);
    for (int l = 0; l < nlev; ++l)
        for (int p = levptr[l]; p < levptr[l + 1]; ++p) lev_of_row[order[p]] = l;
    const int nb = (nlev + band - 1) / band;
    struct Part {   // one band's tiles, with band-local offsets
        std::vector<int> rp{0}, sp{0}, steps, xp{0}, ext, pos, ent{0};
        std::vector<unsigned short> slot;
        size_t lds[4] = {0, 0, 0, 0};
    };
    std::vector<Part> parts(nb);
    const size_t kMaxExt = (size_t)kTileThreads * kTileExtPerLane;
    auto len_of = [&](int i) { return off[i + 1] - off[i]; };
    auto close_tile = [&](Part& P, std::vector<int>& cur) {
        std::stable_sort(cur.begin(), cur.end(),
                         [&](int a, int b) { return lev_of_row[order[a]] < lev_of_row[order[b]]; });
        const int q0 = (int)P.pos.size();
        std::unordered_map<int, int> slot_of;
        slot_of.reserve(cur.size() * 4);
        for (size_t k = 0; k < cur.size(); ++k) slot_of[order[cur[k]]] = (int)k;
        const int R = (int)cur.size();
        int X = 0;
        int64_t E = 0;
        for (size_t k = 0; k < cur.size(); ++k) {
            const int p = cur[k], i = order[p];
            if (k == 0 || lev_of_row[i] != lev_of_row[order[cur[k - 1]]]) P.steps.push_back(q0 + (int)k);
            P.pos.push_back(p);
            for (int e = off[i]; e < off[i + 1]; ++e) {
                const int c = col[e];
                auto it = slot_of.find(c);
                int sl;
                if (it != slot_of.end()) {
                    sl = it->second;
                } else {
                    sl = R + X++;
                    slot_of[c] = sl;
                    P.ext.push_back(c);
                }
                P.slot.push_back((unsigned short)sl);
            }
            E += len_of(i);
            P.ent.push_back(P.ent.back() + len_of(i));
        }
        P.rp.push_back((int)P.pos.size());
        P.sp.push_back((int)P.steps.size());
        P.xp.push_back((int)P.ext.size());
        const size_t sz[4][2] = {{8, 8}, {8, 4}, {16, 16}, {16, 8}};
        for (int k = 0; k < 4; ++k) P.lds[k] = std::max(P.lds[k], tile_lds_bytes(R, X, E, sz[k][0], sz[k][1]));
        cur.clear();
    };
    std::atomic<int> next{0};
    auto work = [&]() {
        std::vector<int> cur;
        std::unordered_map<int, char> seen;   // the open tile's rows (1) and external inputs (2)
        while (true) {
            const int b = next++;
            if (b >= nb) break;
            Part& P = parts[b];
            const int l0 = b * band, l1 = std::min(nlev, l0 + band);
            std::vector<int> ps;
            ps.reserve(levptr[l1] - levptr[l0]);
            for (int p = levptr[l0]; p < levptr[l1]; ++p) ps.push_back(p);
            std::sort(ps.begin(), ps.end(), [&](int a, int b2) {
                return forward ? order[a] < order[b2] : order[a] > order[b2];
            });
            cur.clear();
            seen.clear();
            int R = 0, X = 0;
            int64_t E = 0;
            for (int p : ps) {
                const int i = order[p], len = len_of(i);
                int add = 0;   // inputs new to the open tile (a row's columns are distinct)
                for (int e = off[i]; e < off[i + 1]; ++e) add += seen.count(col[e]) ? 0 : 1;
                const bool fits = tile_lds_bytes(R + 1, X + add, E + len, vsize, vsize) <= lds_cap &&
                                  (size_t)(X + add) <= kMaxExt && R + 1 + X + add <= 65535;
                if (R > 0 && !fits) {
                    close_tile(P, cur);
                    seen.clear();
                    R = X = 0;
                    E = 0;
                }
                for (int e = off[i]; e < off[i + 1]; ++e)
                    if (seen.emplace(col[e], 2).second) ++X;
                seen[i] = 1;
                cur.push_back(p);
                ++R;
                E += len;
            }
            if (R > 0) close_tile(P, cur);
        }
    };
    {
        const int hw = (int)std::thread::hardware_concurrency();
        const int nt = std::max(1, std::min({hw > 0 ? hw : 1, 16, std::max(nb, 1)}));
        std::vector<std::thread> th;
        for (int k = 0; k < nt; ++k) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    TileLayout out;
    for (const Part& P : parts) {
        const int rbase = (int)out.tpos.size(), sbase = (int)out.step_rp.size(), xbase = (int)out.ext_rows.size();
        const int ebase = out.tent.back();
        for (size_t k = 1; k < P.rp.size(); ++k) {
            out.tile_rp.push_back(rbase + P.rp[k]);
            out.tile_sp.push_back(sbase + P.sp[k]);
            out.tile_xp.push_back(xbase + P.xp[k]);
        }
        for (int q : P.steps) out.step_rp.push_back(rbase + q);
        out.ext_rows.insert(out.ext_rows.end(), P.ext.begin(), P.ext.end());
        out.tpos.insert(out.tpos.end(), P.pos.begin(), P.pos.end());
        for (size_t k = 1; k < P.ent.size(); ++k) out.tent.push_back(ebase + P.ent[k]);
        out.tslot.insert(out.tslot.end(), P.slot.begin(), P.slot.end());
        for (int k = 0; k < 4; ++k) out.lds[k] = std::max(out.lds[k], P.lds[k]);
    }
    return out;
}

}  // namespace ssgpu
