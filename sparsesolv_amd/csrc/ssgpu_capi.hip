// Host side of the gfx950 solver library: contexts, device matrices with
// their structural analysis, the shifted-IC factorization driver and the
// three Krylov drivers (ICCG/ICCOCG, IC-MRTR, Eisenstat SGS-MRTR), exported
// through the C-ABI in include/sparsesolv_amd.h.
//
// The drivers follow the reference's control flow line by line
// (MatSolverICCG.cpp:178-328 / 477-626, MatSolverMRTR_ICver.cpp:183-373 /
// 529-722, MatSolverMRTR_SGSver.cpp:42-282 / 302-539): the same prologue
// (b = 0, early exit, normalizer), the same per-iteration bookkeeping
// (log, convergence, best iterate, divergence) and the same scalar
// recurrences, evaluated on the host in double / complex double.  Vector work
// runs on the device; the dot products and norms are the only operations whose
// rounding differs from the reference (fixed-tree sums instead of Eigen's
// packet order).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ssgpu_host.h"
#include "ssgpu_tiles.h"

namespace ssgpu {

int abmc_order(int n, const int* rp, const int* ci, int bs, std::vector<int>& perm);

thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }

// ------------------------------------------------------------------ memory
// device bytes held by the library in this process, and their high-water mark
static std::atomic<int64_t> g_resident{0}, g_peak{0};

void DevBuf::alloc(Context* c, size_t b) {
    release();
    ctx = c;
    if (b == 0) b = 8;
    SSGPU_HIP(hipSetDevice(c->device));
    SSGPU_HIP(hipMalloc(&p, b));
    bytes = b;
    c->resident += (int64_t)b;
    const int64_t now = g_resident += (int64_t)b;
    int64_t pk = g_peak.load();
    while (now > pk && !g_peak.compare_exchange_weak(pk, now)) {
    }
}
void DevBuf::release() {
    if (p) {
        (void)hipFree(p);
        if (ctx) ctx->resident -= (int64_t)bytes;
        g_resident -= (int64_t)bytes;
    }
    p = nullptr;
    bytes = 0;
}

// Device scalars of a context: reduction results in the first slots, and in
// the last slot the flag a sync-free sweep sets when it exceeds its spin
// limit.  Every scalar read-back copies the whole block, so a solve sees the
// flag at its next convergence check and stops there.
DevBuf& scal_buf(Context* c) {
    if (c->d_scal.bytes == 0) {
        c->d_scal.alloc(c, kScalSlots * sizeof(double));
        launch_fill<int>(c->stream, 2 * kScalSlots, c->d_scal.as<int>(), 0);
    }
    return c->d_scal;
}
// copy the scalar block to the pinned host copy and synchronize; a set sweep
// error flag is cleared and reported
const double* read_scalars(Context* c) {
    DevBuf& d = scal_buf(c);
    SSGPU_HIP(hipMemcpyAsync(c->h_scal, d.p, kScalSlots * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SSGPU_HIP(hipStreamSynchronize(c->stream));
    if (reinterpret_cast<const int*>(c->h_scal)[kErrInt] != 0) {
        launch_fill<int>(c->stream, 1, d.as<int>() + kErrInt, 0);
        SSGPU_HIP(hipStreamSynchronize(c->stream));
        throw Error{HIP_ERROR, "a sync-free sweep exceeded its spin limit (a dependency was never produced)"};
    }
    return c->h_scal;
}


// levels of the DAG "row i depends on the entries of its row with index on
// the solved side"; forward = entries with col < i, backward = col > i.
template <typename CI>
static std::vector<int> levels(int n, const std::vector<int>& rp, const CI& ci, bool forward) {
    std::vector<int> lev(n, 0);
    if (forward) {
        for (int i = 0; i < n; ++i) {
            int m = -1;
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] < i) m = std::max(m, lev[ci[k]]);
            lev[i] = m + 1;
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int m = -1;
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] > i) m = std::max(m, lev[ci[k]]);
            lev[i] = m + 1;
        }
    }
    return lev;
}

// Build a LevelCsr from a host CSR (rp, ci) plus per-entry source map.
template <typename CI, typename SRC>
static void build_level_csr(Context* c, int n, const std::vector<int>& rp, const CI& ci, const SRC& src,
                            bool forward, LevelCsr& out, std::vector<int>* pos_out,
                            std::vector<int>* order_out = nullptr) {
    const std::vector<int> lev = levels(n, rp, ci, forward);
    int nlev = 0;
    for (int i = 0; i < n; ++i) nlev = std::max(nlev, lev[i] + 1);
    std::vector<int> cnt(nlev + 1, 0);
    for (int i = 0; i < n; ++i) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
    out.levptr = cnt;
    std::vector<int> order(n);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < n; ++i) order[fill[lev[i]]++] = i;
    }
    std::vector<int> prp(n + 1, 0);
    hvec<int> pci(rp[n]), prow(n), map(rp[n]);
    if (pos_out) pos_out->assign(n, 0);
    for (int p = 0; p < n; ++p) prp[p + 1] = prp[p] + (rp[order[p] + 1] - rp[order[p]]);
    parallel_for(n, [&](int p) {
        const int i = order[p];
        prow[p] = i;
        if (pos_out) (*pos_out)[i] = p;
        int e = prp[p];
        for (int k = rp[i]; k < rp[i + 1]; ++k, ++e) {
            pci[e] = ci[k];
            map[e] = src[k];
        }
    });
    const int64_t e = prp[n];
    out.nlev = nlev;
    out.nnz = e;
    if (order_out) *order_out = order;
    upload(c, out.prp, prp);
    upload(c, out.pci, pci);
    upload(c, out.prow, prow);
    upload(c, out.map, map);
}

// Per original row: its off-diagonal entries (col, source in the factor,
// source in A) in summation order, and its diagonal's sources; flat arrays
// with row offsets (entries of row i at off[i] .. off[i+1]).
// mean rows per level from which a sweep runs one launch per level
constexpr int64_t kLevelSyncRows = 131072;
// most levels for which the one-launch-per-level sweep is timed by the tuner
constexpr int kTuneMaxLevels = 512;

struct SfRows {
    std::vector<int> off, df, da;
    hvec<int> col, sf, sa;   // written in full by the caller after alloc()
    explicit SfRows(int n) : off(n + 1, 0), df(n, -1), da(n, -1) {}
    void alloc(bool with_sa = true) {   // with_sa false: no matrix-value sources (sa stays empty)
        for (size_t i = 0; i + 1 < off.size(); ++i) off[i + 1] += off[i];
        col.resize(off.back());
        sf.resize(off.back());
        if (with_sa) sa.resize(off.back());
    }
    int len(int i) const { return off[i + 1] - off[i]; }
};

// with_f / with_a: keep the value maps from the factor / from the matrix
// (a structure filled from only one of them needs no map for the other)
static void build_sf(Context* c, int n, const std::vector<int>& order, const SfRows& rows, const LevelCsr& lv,
                     SfCsr& out, bool with_f, bool with_a) {
    int kmax = 0;
    for (int i = 0; i < n; ++i) kmax = std::max(kmax, rows.len(i));
    const int nsl = (n + 63) / 64;
    std::vector<int> soff(nsl + 1, 0), rlen(n);
    for (int sl = 0; sl < nsl; ++sl) {
        int w = 0;
        for (int p = 64 * sl; p < std::min(n, 64 * sl + 64); ++p) w = std::max(w, rows.len(order[p]));
        soff[sl + 1] = soff[sl] + 64 * w;
    }
    const int64_t nell = soff[nsl];
    // padding slots: column 0, no value source (-1); each slice is filled by
    // the thread that writes its rows
    hvec<int> ecol(std::max<int64_t>(nell, 1)), ef(with_f ? std::max<int64_t>(nell, 1) : 0),
        ea(with_a ? std::max<int64_t>(nell, 1) : 0);
    std::vector<int> df(n), da(n);
    if (nell == 0) {
        ecol[0] = 0;
        if (with_f) ef[0] = -1;
        if (with_a) ea[0] = -1;
    }
    parallel_for(nsl, [&](int sl) {
        for (int64_t e = soff[sl]; e < soff[sl + 1]; ++e) {
            ecol[e] = 0;
            if (with_f) ef[e] = -1;
            if (with_a) ea[e] = -1;
        }
        for (int p = 64 * sl; p < std::min(n, 64 * sl + 64); ++p) {
            const int i = order[p];
            const int a = rows.off[i], len = rows.len(i);
            rlen[p] = len;
            const int base = soff[sl] + (p % 64);
            for (int k = 0; k < len; ++k) {
                ecol[base + 64 * k] = rows.col[a + k];
                if (with_f) ef[base + 64 * k] = rows.sf[a + k];
                if (with_a) ea[base + 64 * k] = rows.sa[a + k];
            }
            df[p] = rows.df[i];
            da[p] = rows.da[i];
        }
    });
    out.kmax = kmax <= 64 ? std::max(kmax, 1) : 0;
    out.nell = nell;
    out.srow = lv.prow.as<int>();
    out.lv = &lv;
    // Wide levels (a multi-color order: a few levels of many rows) run one
    // launch per level: plain loads, no readiness polling.  Narrow levels (the
    // natural order: ~10^3 levels of ~10^3 rows) stay sync-free.
    out.by_level = lv.nlev > 0 && (int64_t)n >= kLevelSyncRows * (int64_t)lv.nlev;
    if (const char* e = std::getenv("SSGPU_SWEEP_MODE")) {
        if (!std::strcmp(e, "level")) out.by_level = true;
        if (!std::strcmp(e, "sf")) out.by_level = false;
    }
    out.h_soff = soff;
    upload(c, out.slice_off, soff);
    upload(c, out.rlen, rlen);
    upload(c, out.ecol, ecol);
    if (!out.by_level) {   // the inputs' level-order positions, for the position-published sweep
        hvec<int> pos(std::max(n, 1));
        parallel_for(n, [&](int p) { pos[order[p]] = p; });
        parallel_for(nsl, [&](int sl) {
            for (int64_t e = soff[sl]; e < soff[sl + 1]; ++e) ecol[e] = pos[ecol[e]];
        });
        if (nell == 0) ecol[0] = 0;
        upload(c, out.ecolp, ecol);
    }
    if (with_f) {
        upload(c, out.emap_f, ef);
        upload(c, out.dmap_f, df);
    }
    if (with_a) {
        upload(c, out.emap_a, ea);
        upload(c, out.dmap_a, da);
    }
}

// tiles of a sweep on the device (the layout is made by tile_layout)
// mean rows per level up to which tiles are built (see analyse)
constexpr int64_t kTileRowsPerLevel = 64;

static void build_tiles(Context* c, int n, const std::vector<int>& order, const SfRows& rows, const LevelCsr& lv,
                        const std::vector<int>& soff, bool forward, int band, size_t lds_cap, size_t vsize,
                        TileCsr& out) {
    const TileLayout t = tile_layout(n, order, rows.off, rows.col.data(), lv.levptr, forward, band, lds_cap, vsize);
    out.ntiles = (int)t.tile_rp.size() - 1;
    // each packed entry's place in the sliced-ELL arrays (entry k of position p
    // at soff[p / 64] + p % 64 + 64 k), so a tile's entries stage in flat batches
    std::vector<int> teidx(std::max<size_t>(t.tslot.size(), 1), 0);
    parallel_for((int)t.tpos.size(), [&](int q) {
        const int p = t.tpos[q];
        const int base = soff[p / kSfRows] + p % kSfRows;
        for (int e = t.tent[q]; e < t.tent[q + 1]; ++e) teidx[e] = base + kSfRows * (e - t.tent[q]);
    });
    upload(c, out.teidx, teidx);
    for (int k = 0; k < 4; ++k) out.lds[k] = t.lds[k];
    upload(c, out.tile_rp, t.tile_rp);
    upload(c, out.tile_sp, t.tile_sp);
    upload(c, out.step_rp, t.step_rp);
    upload(c, out.tile_xp, t.tile_xp);
    upload(c, out.ext_rows, t.ext_rows);
    upload(c, out.tpos, t.tpos);
    upload(c, out.tent, t.tent);
    upload(c, out.tslot, t.tslot);
}

// upper(A) (SprsTrans.cpp:387-394: entries col >= row, diagonal first) in
// backward level order and its sliced-ELL copy, for the SGS sweeps; built
// on first use
static void ensure_upper(Matrix& m) {
    if (m.us_built) return;
    Context* c = m.ctx;
    const int n = m.n;
    const auto& rp = m.h_rp;
    const auto& ci = m.h_ci;
    std::vector<int> urp(n + 1, 0), uci, usrc, order_s;
    m.upper_diag_ok = true;
    for (int i = 0; i < n; ++i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] >= i) { uci.push_back(ci[k]); usrc.push_back(k); }
        urp[i + 1] = (int)uci.size();
        if (urp[i + 1] == urp[i] || uci[urp[i]] != i) m.upper_diag_ok = false;
    }
    build_level_csr(c, n, urp, uci, usrc, false, m.Us, nullptr, &order_s);
    SfRows rs(n);
    for (int i = 0; i < n; ++i) rs.off[i + 1] = (urp[i + 1] - urp[i]) - ((urp[i + 1] > urp[i] && uci[urp[i]] == i) ? 1 : 0);
    rs.alloc();
    for (int i = 0; i < n; ++i) {
        int q = rs.off[i];
        for (int k = urp[i]; k < urp[i + 1]; ++k) {
            if (k == urp[i] && uci[k] == i) { rs.da[i] = usrc[k]; continue; }
            rs.col[q] = uci[k];
            rs.sf[q] = -1;
            rs.sa[q] = usrc[k];
            ++q;
        }
    }
    build_sf(c, n, order_s, rs, m.Us, m.sfUs, false, true);   // SGS values come from S A S
    m.us_built = true;
}

static void analyse(Matrix& m) {
    Context* c = m.ctx;
    const int n = m.n;
    const auto& rp = m.h_rp;
    const auto& ci = m.h_ci;
    // SpMV row blocks: <= kSpmvRows rows and <= kSpmvCap entries (or one long row)
    std::vector<int> blk{0};
    {
        int r = 0;
        while (r < n) {
            int r1 = r + 1;
            while (r1 < n && r1 - r < kSpmvRows && rp[r1 + 1] - rp[r] <= kSpmvCap) ++r1;
            blk.push_back(r1);
            r = r1;
        }
    }
    m.nblk = (int)blk.size() - 1;
    upload(c, m.blk, blk);
    if (const char* g = std::getenv("SSGPU_SWEEP_GRID")) m.sweep_grid = std::max(1, std::atoi(g));
    if (const char* g = std::getenv("SSGPU_SWEEP_IDLE")) m.sweep_idle = std::min(5, std::max(1, std::atoi(g)));
    if (const char* g = std::getenv("SSGPU_BLOCK_GRID")) m.block_grid = std::max(1, std::atoi(g));
    std::vector<int> diag(n, -1);
    parallel_for(n, [&](int i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) { diag[i] = k; break; }
    });
    upload(c, m.diag, diag);
    // lower(A): entries col <= row, order kept (SprsTrans.cpp:372-379)
    std::vector<int> lrp(n + 1, 0);
    parallel_for(n, [&](int i) {
        int cnt = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) cnt += ci[k] <= i;
        lrp[i + 1] = cnt;
    });
    for (int i = 0; i < n; ++i) lrp[i + 1] += lrp[i];
    hvec<int> lci(lrp[n]), lsrc(lrp[n]);
    parallel_for(n, [&](int i) {
        int q = lrp[i];
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] <= i) { lci[q] = ci[k]; lsrc[q] = k; ++q; }
    });
    m.lower_diag_ok = true;
    for (int i = 0; i < n; ++i)
        if (lrp[i + 1] == lrp[i] || lci[lrp[i + 1] - 1] != i) m.lower_diag_ok = false;
    m.lmax_row = 0;
    for (int i = 0; i < n; ++i) m.lmax_row = std::max(m.lmax_row, lrp[i + 1] - lrp[i]);
    std::vector<int> pos, order_l, order_u;
    build_level_csr(c, n, lrp, lci, lsrc, true, m.L, &pos, &order_l);
    upload(c, m.ppos, pos);
    // transpose(lower) (SprsTrans.cpp:312-319): row j holds (j, i) for lower entries (i, j),
    // ascending i; its source is the entry's position in the level-ordered L.
    {
        std::vector<int> urp(n + 1, 0);
        for (size_t k = 0; k < lci.size(); ++k) urp[lci[k] + 1]++;
        for (int i = 0; i < n; ++i) urp[i + 1] += urp[i];
        hvec<int> uci(lci.size()), usrc(lci.size());
        std::vector<int> fill(urp.begin(), urp.end() - 1);
        std::vector<int> lstart(n);   // start of original row i inside level-ordered L
        {
            std::vector<int> prp_h(n + 1, 0);
            // recompute level-ordered offsets on host from pos
            std::vector<int> len(n);
            for (int i = 0; i < n; ++i) len[i] = lrp[i + 1] - lrp[i];
            std::vector<int> order(n);
            for (int i = 0; i < n; ++i) order[pos[i]] = i;
            for (int p = 0; p < n; ++p) prp_h[p + 1] = prp_h[p] + len[order[p]];
            for (int i = 0; i < n; ++i) lstart[i] = prp_h[pos[i]];
        }
        for (int i = 0; i < n; ++i)
            for (int k = lrp[i]; k < lrp[i + 1]; ++k) {
                const int j = lci[k];
                const int q = fill[j]++;
                uci[q] = i;
                usrc[q] = lstart[i] + (k - lrp[i]);
            }
        build_level_csr(c, n, urp, uci, usrc, false, m.Uic, nullptr, &order_u);
        m.uic_built = true;
        // sync-free copies: lower(A) (off-diagonals, diagonal = last entry) and transpose(lower)
        SfRows rl(n), ru(n);
        // off-diagonal counts: row i of lower(A) and row i of its transpose,
        // less the diagonal when the row has it (the last entry of a lower row)
        parallel_for(n, [&](int i) {
            const int dg = (lrp[i + 1] > lrp[i] && lci[lrp[i + 1] - 1] == i) ? 1 : 0;
            rl.off[i + 1] = lrp[i + 1] - lrp[i] - dg;
            ru.off[i + 1] = urp[i + 1] - urp[i] - dg;
        });
        rl.alloc();
        ru.alloc(false);   // the transpose is filled from the factor only
        parallel_for(n, [&](int i) {   // rows of lower(A): each row's own range
            const int a = lrp[i], e = lrp[i + 1];
            int ql = rl.off[i];
            for (int k = a; k < e; ++k) {
                const int j = lci[k];
                const int f = lstart[i] + (k - a);
                if (j < i) {
                    rl.col[ql] = j; rl.sf[ql] = f; rl.sa[ql] = lsrc[k];
                    ++ql;
                } else if (j == i && k == e - 1) {
                    rl.df[i] = f; rl.da[i] = lsrc[k];
                    ru.df[i] = f;
                }
            }
        });
        // row j of the transpose: the transposed structure's row j (rows i of
        // lower(A) holding column j, ascending, each entry's place in the
        // level-ordered factor) without its diagonal
        parallel_for(n, [&](int j) {
            int q = ru.off[j];
            for (int e = urp[j]; e < urp[j + 1]; ++e)
                if (uci[e] != j) {
                    ru.col[q] = uci[e];
                    ru.sf[q] = usrc[e];
                    ++q;
                }
        });
        build_sf(c, n, order_l, rl, m.L, m.sfL, true, true);        // IC factor, or SGS lower(S A S)
        build_sf(c, n, order_u, ru, m.Uic, m.sfUic, true, false);   // the factor's transpose only
        // Tiles for the tiled sweep when levels are narrow (at most
        // kTileRowsPerLevel rows on average: chains, strips), where a 64-row
        // chunk of the sync-free sweep holds rows of several levels and pays a
        // global round trip per level; there tiles measured 15-21% faster
        // (a 200,000-row tridiagonal chain, an 8 x 50,000 grid).  On wide
        // levels (3-D grids) they lose, and are not built.  The tuner decides
        // whether they run.  SSGPU_TILE=1 builds them always, 0 never.
        const char* te = std::getenv("SSGPU_TILE");
        int band = 64;
        size_t cap = 64 << 10;
        if (const char* e = std::getenv("SSGPU_TILE_BAND")) band = std::max(1, std::min(kTileMaxSteps, std::atoi(e)));
        if (const char* e = std::getenv("SSGPU_TILE_LDS")) cap = (size_t)std::max(4096, std::atoi(e));
        const bool narrow = (int64_t)n <= kTileRowsPerLevel * (int64_t)std::max(1, m.L.nlev);
        const bool want = te ? std::strcmp(te, "0") != 0 : narrow;
        if (want && m.sfL.kmax > 0 && m.sfUic.kmax > 0 && m.L.nlev >= 2) {
            m.sfL.tile.reset(new TileCsr());
            m.sfUic.tile.reset(new TileCsr());
            build_tiles(c, n, order_l, rl, m.L, m.sfL.h_soff, true, band, cap, m.vsize, *m.sfL.tile);
            build_tiles(c, n, order_u, ru, m.Uic, m.sfUic.h_soff, false, band, cap, m.vsize, *m.sfUic.tile);
        }
    }
    // sync-free grid: 128 waves keep ~10^3-row levels fed with the least
    // polling; levels of thousands of rows want more waves in flight
    if (!std::getenv("SSGPU_SWEEP_GRID") && m.L.nlev > 0 && (int64_t)n >= 2048 * (int64_t)m.L.nlev)
        m.sweep_grid = 1024;
    // upper(A) for the SGS sweeps: built on first use (ensure_upper)
}

// ------------------------------------------------------------------ helpers
template <typename T> struct Host;
template <> struct Host<double> {
    static double zero() { return 0.0; }
    static double from(const double* p) { return p[0]; }
    static double div(double a, double b) { return a / b; }
};
template <> struct Host<zc> {
    static zc zero() { return zc{0.0, 0.0}; }
    static zc from(const double* p) { return zc{p[0], p[1]}; }
    static zc div(zc a, zc b) { return zdiv(a, b); }
};
inline zc operator*(double a, zc b) { return zc{a * b.re, a * b.im}; }
inline zc hdiv(double a, zc b) { return zdiv(zc{a, 0.0}, b); }
inline double hdiv(double a, double b) { return a / b; }

template <typename T> constexpr int kparts() { return sizeof(T) / sizeof(double); }

// Device work space of a solve carved from caller-supplied memory instead of
// the context's pools (the allocation-free entry point): vector slots of n
// values, the reduction partials, the scalar block (whose last slot is the
// sweep error flag) and the sweep tickets.
struct Workspace {
    char* vecs = nullptr;          // nslots vectors of vec_bytes each
    size_t vec_bytes = 0;
    int slot_of[16];               // work-vector index -> slot (-1: not planned for this solve)
    double* partial = nullptr;
    double* scal = nullptr;        // kScalSlots doubles
    int* tickets = nullptr;        // 16 ints
    char* xpos = nullptr;          // 2 vectors of vec_bytes: the position-published sweeps' polled arrays
};

struct Solver {
    Matrix& m;
    Context* c;
    hipStream_t s;
    const ssgpu_options& o;
    Workspace* ws = nullptr;   // set: every device buffer of the solve comes from it
    long spin_limit = kSweepSpinLimit;   // polling passes without progress before a sweep wave gives up
    std::vector<double> log;
    int iterations = 0;
    double final_res = 0.0;
    // c: the context whose stream and work space the solve uses
    Solver(Matrix& mm, Context* cc, const ssgpu_options& oo) : m(mm), c(cc), s(cc->stream), o(oo) {
        // SSGPU_SPIN_LIMIT lowers the sync-free sweeps' spin limit (tests of
        // the error path force the limit to 0)
        if (const char* e = std::getenv("SSGPU_SPIN_LIMIT")) spin_limit = std::max(0L, std::atol(e));
    }
    Solver(Matrix& mm, const ssgpu_options& oo) : Solver(mm, mm.ctx, oo) {}

    template <typename T> T* vec(size_t idx) {
        if (ws) {
            if (idx >= 16 || ws->slot_of[idx] < 0) throw Error{HIP_ERROR, "work vector outside the planned work space"};
            return reinterpret_cast<T*>(ws->vecs + ws->vec_bytes * (size_t)ws->slot_of[idx]);
        }
        while (c->work.size() <= idx) c->work.emplace_back(new DevBuf());
        DevBuf& b = *c->work[idx];
        const size_t need = sizeof(T) * (size_t)std::max(m.n, 1);
        if (b.bytes < need) b.alloc(c, need);
        return b.as<T>();
    }
    double* partial() {
        if (ws) return ws->partial;
        const size_t need = sizeof(double) * 8 * (size_t)std::max(m.nblk, kMaxReduceBlocks);
        if (c->partial.bytes < need) c->partial.alloc(c, need);
        return c->partial.as<double>();
    }
    double* scal() { return ws ? ws->scal : scal_buf(c).as<double>(); }
    // the scalar block to the context's pinned copy (synchronizes); a set sweep
    // error flag is cleared and reported
    const double* read_scal() {
        if (!ws) return read_scalars(c);
        SSGPU_HIP(hipMemcpyAsync(c->h_scal, ws->scal, kScalSlots * sizeof(double), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        if (reinterpret_cast<const int*>(c->h_scal)[kErrInt] != 0) {
            launch_fill<int>(s, 1, reinterpret_cast<int*>(ws->scal) + kErrInt, 0);
            SSGPU_HIP(hipStreamSynchronize(s));
            throw Error{HIP_ERROR, "a sync-free sweep exceeded its spin limit (a dependency was never produced)"};
        }
        return c->h_scal;
    }
    // finalize K doubles from nb partials and read them back (synchronizes)
    const double* fetch(int nb, int K) {
        launch_finalize(s, nb, K, partial(), scal());
        return read_scal();
    }
    template <typename T> double norm(const T* a) {
        launch_norm2<T>(s, m.n, a, partial());
        return std::sqrt(fetch(reduce_blocks(m.n), 1)[0]);
    }
    template <typename T> T dot(const T* a, const T* b) {
        launch_dot<T>(s, m.n, a, b, partial());
        return Host<T>::from(fetch(reduce_blocks(m.n), kparts<T>()));
    }
    template <typename T> void spmv(const T* v, const T* x, T* y) {
        launch_spmv<T>(s, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), v, x, y, nullptr, nullptr, m.spmv_stage_kb);
    }
    template <typename T> T spmv_dot(const T* v, const T* x, T* y) {   // y = A x, returns x.y
        launch_spmv<T>(s, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), v, x, y, x, partial(), m.spmv_stage_kb);
        return Host<T>::from(fetch(m.nblk, kparts<T>()));
    }

    // --- preconditioner sweeps ------------------------------------------------
    template <typename T, typename F>
    void sweep(const LevelCsr& S, int mode, const F* v, const T* rhs, const T* aux, const T* dd, T* x) {
        for (int l = 0; l < S.nlev; ++l)
            launch_sweep<T, F>(s, mode, S.levptr[l], S.levptr[l + 1], S.prp.as<int>(), S.pci.as<int>(), v,
                               S.prow.as<int>(), rhs, aux, dd, x);
    }
    // --- sync-free sweep plumbing ------------------------------------------------
    int* tickets() {
        if (ws) return ws->tickets;
        if (c->d_tickets.bytes == 0) c->d_tickets.alloc(c, 64);
        return c->d_tickets.as<int>();
    }
    // the array a sync-free sweep over S polls when it publishes at level
    // positions (slot k of two), or null when S sweeps another way
    template <typename T> T* pos_buf(const SfCsr& S, int k) {
        if (S.by_level || (S.use_tile && S.tile) || !S.ecolp.bytes) return nullptr;
        if (ws) return reinterpret_cast<T*>(ws->xpos + ws->vec_bytes * (size_t)k);
        DevBuf& b = c->d_xpos[k];
        const size_t need = sizeof(T) * (size_t)std::max(m.n, 1);
        if (b.bytes < need) b.alloc(c, need);
        return b.as<T>();
    }
    int* errflag() { return reinterpret_cast<int*>(scal()) + kErrInt; }
    bool sf_ok() const { return m.sfL.kmax > 0 && m.sfUic.kmax > 0; }
    template <typename T, typename F>
    void sf_sweep(SfCsr& S, int mode, bool reverse, int* ticket, const F* ev, const F* dv, const T* rhs,
                  const T* aux, const T* dd, T* x, T* xp = nullptr) {
        (void)reverse;   // backward structures are already in their own level order
        if (S.use_tile && S.tile && !S.by_level) {
            const int pair = (sizeof(T) == 16 ? 2 : 0) + (sizeof(F) < sizeof(T) ? 1 : 0);
            launch_sweep_tile<T, F>(s, mode, S.kmax, m.tile_grid, S.tile->lds[pair], S.tile->args(), ticket,
                                    errflag(), spin_limit, S.srow, S.slice_off.as<int>(), S.rlen.as<int>(), ev, dv,
                                    rhs, aux, dd, m.sweep_idle, x);
            return;
        }
        if (S.by_level)
            launch_sweep_levels<T, F>(s, mode, S.kmax, S.lv->nlev, S.lv->levptr.data(), S.srow,
                                      S.slice_off.as<int>(), S.rlen.as<int>(), S.ecol.as<int>(), ev, dv, rhs, aux,
                                      dd, x);
        else if (xp)   // xp: pos_buf of S, sentinel-filled by the caller
            launch_sweep_sf_pos<T, F>(s, mode, S.kmax, m.n, m.sweep_grid, ticket, errflag(), spin_limit, S.srow,
                                      S.slice_off.as<int>(), S.rlen.as<int>(), S.ecolp.as<int>(), ev, dv, rhs, aux,
                                      dd, m.sweep_idle, xp, x);
        else
            launch_sweep_sf<T, F>(s, mode, S.kmax, m.n, m.sweep_grid, ticket, errflag(), spin_limit, S.srow,
                                  S.slice_off.as<int>(), S.rlen.as<int>(), S.ecol.as<int>(), ev, dv, rhs, aux, dd,
                                  m.sweep_idle, x);
    }
    // values of a sliced-ELL copy from a source array (factor or matrix values)
    template <typename T> void sf_values(SfCsr& S, const T* src, bool from_factor, bool with_f32) {
        using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
        const int* em = from_factor ? S.emap_f.as<int>() : S.emap_a.as<int>();
        const int* dm = from_factor ? S.dmap_f.as<int>() : S.dmap_a.as<int>();
        const int64_t ne = std::max<int64_t>(S.nell, 1);
        if (S.eval.bytes < sizeof(T) * ne) S.eval.alloc(c, sizeof(T) * ne);
        if (S.dval.bytes < sizeof(T) * m.n) S.dval.alloc(c, sizeof(T) * m.n);
        launch_gather_pad<T, T>(s, S.nell, em, src, S.eval.as<T>());
        launch_gather_pad<T, T>(s, m.n, dm, src, S.dval.as<T>());
        if (with_f32) {
            if (S.eval32.bytes < sizeof(F) * ne) S.eval32.alloc(c, sizeof(F) * ne);
            if (S.dval32.bytes < sizeof(F) * m.n) S.dval32.alloc(c, sizeof(F) * m.n);
            launch_gather_pad<T, F>(s, S.nell, em, src, S.eval32.as<F>());
            launch_gather_pad<T, F>(s, m.n, dm, src, S.dval32.as<F>());
        }
    }

    // z = (L D L^T)^{-1} r (MatSolverBase.cpp:177-231); tmp is scratch
    template <typename T> void ic_apply(const T* r, T* z, T* tmp) {
        using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
        const bool f32 = o.factor_precision == SSGPU_FACTOR_F32;
        if (sf_ok()) {
            // a sync-free sweep polls its level-position array (xf, xb) and
            // stores by row; the others poll (tiles) or read their outputs
            T* xf = pos_buf<T>(m.sfL, 0);
            T* xb = pos_buf<T>(m.sfUic, 1);
            if (!m.sfL.by_level || !m.sfUic.by_level)
                launch_sf_prepare<T>(s, m.n, xf ? xf : tmp, xb ? xb : z, tickets());
            if (f32) {
                sf_sweep<T, F>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval32.as<F>(), m.sfL.dval32.as<F>(), r,
                               nullptr, nullptr, tmp, xf);
                sf_sweep<T, F>(m.sfUic, BWD_IC, true, tickets() + 1, m.sfUic.eval32.as<F>(), nullptr, nullptr, tmp,
                               m.diagD.as<T>(), z, xb);
            } else {
                sf_sweep<T, T>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval.as<T>(), m.sfL.dval.as<T>(), r,
                               nullptr, nullptr, tmp, xf);
                sf_sweep<T, T>(m.sfUic, BWD_IC, true, tickets() + 1, m.sfUic.eval.as<T>(), nullptr, nullptr, tmp,
                               m.diagD.as<T>(), z, xb);
            }
            return;
        }
        if (f32) {
            sweep<T, F>(m.L, FWD_DIV_LAST, m.Lv32.as<F>(), r, nullptr, nullptr, tmp);
            sweep<T, F>(m.Uic, BWD_IC, m.Uv32.as<F>(), nullptr, tmp, m.diagD.as<T>(), z);
        } else {
            sweep<T, T>(m.L, FWD_DIV_LAST, m.Lv.as<T>(), r, nullptr, nullptr, tmp);
            sweep<T, T>(m.Uic, BWD_IC, m.Uv.as<T>(), nullptr, tmp, m.diagD.as<T>(), z);
        }
    }
    // one IC sweep: forward (r -> t) or backward (t -> z, t given as `in`)
    template <typename T> void ic_sweep_one(bool forward, const T* in, T* out, T* aux) {
        using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
        const bool f32 = o.factor_precision == SSGPU_FACTOR_F32;
        (void)aux;
        if (sf_ok()) {
            T* xp = pos_buf<T>(forward ? m.sfL : m.sfUic, 0);
            if (!(forward ? m.sfL : m.sfUic).by_level) launch_sf_prepare<T>(s, m.n, xp ? xp : out, nullptr, tickets());
            if (forward) {
                if (f32) sf_sweep<T, F>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval32.as<F>(), m.sfL.dval32.as<F>(), in, nullptr, nullptr, out, xp);
                else sf_sweep<T, T>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval.as<T>(), m.sfL.dval.as<T>(), in, nullptr, nullptr, out, xp);
            } else {
                if (f32) sf_sweep<T, F>(m.sfUic, BWD_IC, true, tickets(), m.sfUic.eval32.as<F>(), nullptr, nullptr, in, m.diagD.as<T>(), out, xp);
                else sf_sweep<T, T>(m.sfUic, BWD_IC, true, tickets(), m.sfUic.eval.as<T>(), nullptr, nullptr, in, m.diagD.as<T>(), out, xp);
            }
            return;
        }
        if (f32) {
            if (forward) sweep<T, F>(m.L, FWD_DIV_LAST, m.Lv32.as<F>(), in, nullptr, nullptr, out);
            else sweep<T, F>(m.Uic, BWD_IC, m.Uv32.as<F>(), nullptr, in, m.diagD.as<T>(), out);
        } else {
            if (forward) sweep<T, T>(m.L, FWD_DIV_LAST, m.Lv.as<T>(), in, nullptr, nullptr, out);
            else sweep<T, T>(m.Uic, BWD_IC, m.Uv.as<T>(), nullptr, in, m.diagD.as<T>(), out);
        }
    }
    // plain sweeps on the current lower / upper values (SGS): values loaded by
    // sgs_values; with an FP32 preparation (o.factor_precision) the sliced-ELL
    // sweeps read the FP32 copies (rows longer than 64 entries keep FP64)
    template <typename T> void fr(const T* r, T* z) {
        using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
        if (m.sfL.kmax > 0) {
            T* xp = pos_buf<T>(m.sfL, 0);
            if (!m.sfL.by_level) launch_sf_prepare<T>(s, m.n, xp ? xp : z, nullptr, tickets());
            if (o.factor_precision == SSGPU_FACTOR_F32)
                sf_sweep<T, F>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval32.as<F>(), m.sfL.dval32.as<F>(), r,
                               nullptr, nullptr, z, xp);
            else
                sf_sweep<T, T>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval.as<T>(), m.sfL.dval.as<T>(), r,
                               nullptr, nullptr, z, xp);
        } else {
            sweep<T, T>(m.L, FWD_DIV_LAST, m.Lv.as<T>(), r, nullptr, nullptr, z);
        }
    }
    template <typename T> void bc(const T* r, T* z) {
        using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
        if (m.sfUs.kmax > 0) {
            T* xp = pos_buf<T>(m.sfUs, 0);
            if (!m.sfUs.by_level) launch_sf_prepare<T>(s, m.n, xp ? xp : z, nullptr, tickets());
            if (o.factor_precision == SSGPU_FACTOR_F32)
                sf_sweep<T, F>(m.sfUs, BWD_DIV_FIRST, true, tickets(), m.sfUs.eval32.as<F>(), m.sfUs.dval32.as<F>(),
                               r, nullptr, nullptr, z, xp);
            else
                sf_sweep<T, T>(m.sfUs, BWD_DIV_FIRST, true, tickets(), m.sfUs.eval.as<T>(), m.sfUs.dval.as<T>(), r,
                               nullptr, nullptr, z, xp);
        } else {
            sweep<T, T>(m.Us, BWD_DIV_FIRST, m.Uv.as<T>(), r, nullptr, nullptr, z);
        }
    }
    // load lower(av) / upper(av) values for the plain sweeps (and their FP32
    // copies for an FP32 preparation)
    template <typename T> void sgs_values(const T* av) {
        const bool f32 = o.factor_precision == SSGPU_FACTOR_F32;
        if (m.Lv.bytes < sizeof(T) * m.L.nnz) m.Lv.alloc(c, sizeof(T) * m.L.nnz);
        if (m.Uv.bytes < sizeof(T) * m.Us.nnz) m.Uv.alloc(c, sizeof(T) * m.Us.nnz);
        launch_gather<T, T>(s, m.L.nnz, m.L.map.as<int>(), av, m.Lv.as<T>());
        launch_gather<T, T>(s, m.Us.nnz, m.Us.map.as<int>(), av, m.Uv.as<T>());
        if (m.sfL.kmax > 0) sf_values<T>(m.sfL, av, false, f32);
        if (m.sfUs.kmax > 0) sf_values<T>(m.sfUs, av, false, f32);
        m.prep_precision = o.factor_precision;
        m.factor_valid = false;
        m.prep_kind = 2;
    }
    // throws if a sync-free sweep hit its spin limit (reads back, synchronizes)
    void check_sweep_error() { (void)read_scal(); }

    // --- shifted IC(0) ------------------------------------------------------
    // Update lists of the factorization (once per matrix); false if they would
    // not fit 32-bit offsets (the searching level kernel is used then).
    bool ic_symbolic() {
        if (m.icsym_built) return m.icsym_total >= 0;
        m.icsym_built = true;
        const int64_t nl = m.L.nnz;
        // every entry has fewer updates than its row has entries, so
        // nnzL * longest row bounds the list length; keep it in 32-bit offsets
        if ((nl + 1) * (int64_t)std::max(m.lmax_row, 1) > 2147483647LL) { m.icsym_total = -1; return false; }
        DevBuf cnt, tmp;
        cnt.alloc(c, sizeof(int) * (nl + 1));
        m.ic_po.alloc(c, sizeof(int) * (nl + 1));
        launch_fill<int>(s, nl + 1, cnt.as<int>(), 0);
        launch_ic_symbolic_count(s, m.n, m.L.prp.as<int>(), m.L.pci.as<int>(), m.L.prow.as<int>(),
                                 m.ppos.as<int>(), cnt.as<int>());
        const size_t tb = scan_tmp_bytes((int)(nl + 1));
        tmp.alloc(c, std::max<size_t>(tb, 8));
        launch_exclusive_scan(s, cnt.as<int>(), m.ic_po.as<int>(), (int)(nl + 1), tmp.p, tb);
        int total = 0;
        SSGPU_HIP(hipMemcpyAsync(&total, m.ic_po.as<int>() + nl, sizeof(int), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        m.icsym_total = total;
        m.ic_slot.alloc(c, std::max<size_t>((size_t)total, 1));
        m.ic_pos.alloc(c, sizeof(int) * std::max<size_t>((size_t)total, 1));
        m.ic_col.alloc(c, sizeof(int) * std::max<size_t>((size_t)total, 1));
        launch_ic_symbolic_fill(s, m.n, m.L.prp.as<int>(), m.L.pci.as<int>(), m.L.prow.as<int>(), m.ppos.as<int>(),
                                m.ic_po.as<int>(), m.ic_slot.p, m.ic_pos.as<int>(), m.ic_col.as<int>());
        if (c->d_count.bytes == 0) c->d_count.alloc(c, 64);
        m.ic_max_updates = ic_row_updates_max(s, m.n, m.L.prp.as<int>(), m.ic_po.as<int>(), c->d_count.as<int>());
        return true;
    }
    // Factor the matrix with values av (A or the scaled copy); returns pivot fixes.
    // Sweep launch shape, measured once per matrix on its first factor: the
    // forward sweep is timed sync-free with 128, 160 and 1024 waves, and one
    // launch per level when the levels are few; the fastest is kept for the
    // IC sweeps of this matrix.  Every shape gives the same bits.
    template <typename T> void tune_sweeps() {
        m.sweeps_tuned = true;
        const char* mode_env = std::getenv("SSGPU_SWEEP_MODE");
        const bool fix_mode = mode_env != nullptr;
        const bool fix_grid = std::getenv("SSGPU_SWEEP_GRID") != nullptr;
        const bool want_tile = fix_mode && !std::strcmp(mode_env, "tile");
        // a tiled shape exists when both directions have tiles whose shared
        // memory fits a workgroup (gfx950: 160 KB)
        const int pair = sizeof(T) == 16 ? 2 : 0;
        const bool have_tile = m.sfL.tile && m.sfUic.tile && m.sfL.tile->ntiles > 0 && m.sfUic.tile->ntiles > 0 &&
                               std::max(m.sfL.tile->lds[pair], m.sfUic.tile->lds[pair]) <= (160u << 10);
        struct Shape { int kind; int grid; };   // 0 sync-free chunks, 1 one launch per level, 2 tiles
        auto apply = [&](const Shape& sh) {
            m.sfL.by_level = m.sfUic.by_level = sh.kind == 1;
            m.sfL.use_tile = m.sfUic.use_tile = sh.kind == 2;
            if (sh.kind == 0) m.sweep_grid = sh.grid;
            if (sh.kind == 2) m.tile_grid = sh.grid;
        };
        std::vector<Shape> cand;
        if (want_tile) {
            if (!have_tile) return;
            if (const char* g = std::getenv("SSGPU_TILE_GRID")) {
                apply({2, std::max(1, std::atoi(g))});
                return;
            }
            for (int g : {256, 512}) cand.push_back({2, g});
        } else {
            if (fix_mode && fix_grid) return;
            if (!fix_mode || !m.sfL.by_level) {
                for (int g : {128, 160, 1024})
                    if (!fix_grid || g == m.sweep_grid) cand.push_back({0, fix_grid ? m.sweep_grid : g});
            }
            if ((!fix_mode || m.sfL.by_level) && m.L.nlev <= kTuneMaxLevels) cand.push_back({1, m.sweep_grid});
            if (!fix_mode && have_tile)
                for (int g : {256, 512}) cand.push_back({2, g});
        }
        if (cand.size() < 2) {
            if (cand.size() == 1 && want_tile) apply(cand[0]);
            return;
        }
        DevBuf r, out, xpl;   // freed on return: factorization leaves the context's pool as it was
        r.alloc(c, sizeof(T) * m.n);
        out.alloc(c, sizeof(T) * m.n);
        launch_fill<T>(s, m.n, r.as<T>(), one<T>());
        hipEvent_t e0, e1;
        SSGPU_HIP(hipEventCreate(&e0));
        SSGPU_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        Shape pick = cand[0];
        for (const Shape& sh : cand) {
            apply(sh);
            float t = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {   // the first run warms up
                T* xp = nullptr;   // the array a position-published sweep polls
                if (sh.kind == 0 && m.sfL.ecolp.bytes) {
                    if (xpl.bytes < out.bytes) xpl.alloc(c, out.bytes);
                    xp = xpl.as<T>();
                }
                if (sh.kind != 1) launch_sf_prepare<T>(s, m.n, xp ? xp : out.as<T>(), nullptr, tickets());
                SSGPU_HIP(hipEventRecord(e0, s));
                sf_sweep<T, T>(m.sfL, FWD_DIV_LAST, false, tickets(), m.sfL.eval.as<T>(), m.sfL.dval.as<T>(),
                               r.as<T>(), nullptr, nullptr, out.as<T>(), xp);
                SSGPU_HIP(hipEventRecord(e1, s));
                SSGPU_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                SSGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) t = std::min(t, ms);
            }
            if (t < best) { best = t; pick = sh; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        check_sweep_error();
        apply(pick);
    }
    template <typename T> int factor(const T* av, double accel) {
        const int64_t nl = m.L.nnz;
        if (m.Lv.bytes < sizeof(T) * nl) m.Lv.alloc(c, sizeof(T) * nl);
        if (m.Uv.bytes < sizeof(T) * m.Uic.nnz) m.Uv.alloc(c, sizeof(T) * m.Uic.nnz);
        if (m.diagD.bytes < sizeof(T) * m.n) m.diagD.alloc(c, sizeof(T) * m.n);
        if (c->d_count.bytes == 0) c->d_count.alloc(c, 64);
        launch_gather<T, T>(s, nl, m.L.map.as<int>(), av, m.Lv.as<T>());
        launch_fill<T>(s, m.n, m.diagD.as<T>(), Host<T>::zero());
        launch_fill<int>(s, 1, c->d_count.as<int>(), 0);
        if (m.lmax_row <= 48 && ic_symbolic()) {
            launch_fill<int>(s, 1, c->d_count.as<int>(), 0);
            for (int l = 0; l < m.L.nlev; ++l) {
                if (m.ic_max_updates <= 256)   // wave per row: one gather round trip per row
                    launch_ic_numeric_wave<T>(s, m.lmax_row, m.L.levptr[l], m.L.levptr[l + 1], m.L.prp.as<int>(),
                                              m.L.pci.as<int>(), m.L.prow.as<int>(), m.ic_po.as<int>(), m.ic_slot.p,
                                              m.ic_pos.as<int>(), m.ic_col.as<int>(), m.Lv.as<T>(), m.diagD.as<T>(),
                                              accel, c->d_count.as<int>());
                else
                    launch_ic_numeric<T>(s, m.lmax_row, m.L.levptr[l], m.L.levptr[l + 1], m.L.prp.as<int>(),
                                         m.L.pci.as<int>(), m.L.prow.as<int>(), m.ic_po.as<int>(), m.ic_slot.p,
                                         m.ic_pos.as<int>(), m.ic_col.as<int>(), m.Lv.as<T>(), m.diagD.as<T>(), accel,
                                         c->d_count.as<int>());
            }
        } else {
            for (int l = 0; l < m.L.nlev; ++l)
                launch_ic_level<T>(s, m.L.levptr[l], m.L.levptr[l + 1], m.L.prp.as<int>(), m.L.pci.as<int>(),
                                   m.Lv.as<T>(), m.L.prow.as<int>(), m.ppos.as<int>(), m.diagD.as<T>(), accel,
                                   c->d_count.as<int>());
        }
        launch_gather<T, T>(s, m.Uic.nnz, m.Uic.map.as<int>(), m.Lv.as<T>(), m.Uv.as<T>());
        if (sf_ok()) {
            const bool f32 = o.factor_precision == SSGPU_FACTOR_F32;
            sf_values<T>(m.sfL, m.Lv.as<T>(), true, f32);
            sf_values<T>(m.sfUic, m.Lv.as<T>(), true, f32);
            if (!m.sweeps_tuned) tune_sweeps<T>();
        }
        if (o.factor_precision == SSGPU_FACTOR_F32) {
            using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
            if (m.Lv32.bytes < sizeof(F) * nl) m.Lv32.alloc(c, sizeof(F) * nl);
            if (m.Uv32.bytes < sizeof(F) * m.Uic.nnz) m.Uv32.alloc(c, sizeof(F) * m.Uic.nnz);
            launch_narrow<T, F>(s, nl, m.Lv.as<T>(), m.Lv32.as<F>());
            launch_narrow<T, F>(s, m.Uic.nnz, m.Uv.as<T>(), m.Uv32.as<F>());
        }
        int fixes = 0;
        SSGPU_HIP(hipMemcpyAsync(&fixes, c->d_count.p, sizeof(int), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        m.factor_pivot_fixes = fixes;
        m.factor_accel = accel;
        m.factor_valid = true;
        m.prep_kind = 1;
        m.prep_precision = o.factor_precision;
        return fixes;
    }
    template <typename T> bool diag_positive() {
        launch_fill<int>(s, 1, c->d_count.as<int>(), 0);
        launch_count_nonpos<T>(s, m.n, m.diagD.as<T>(), c->d_count.as<int>());
        int bad = 0;
        SSGPU_HIP(hipMemcpyAsync(&bad, c->d_count.p, sizeof(int), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        return bad == 0;
    }
    // MatSolverBase.cpp:344-391
    template <typename T> double auto_accel(const T* av, double accel_ini) {
        double acc = std::fabs(accel_ini);
        if (acc < 0.9 || acc > 1.8) acc = 1;
        for (int k = 0; k < 80; ++k) {
            factor<T>(av, acc);
            if (diag_positive<T>()) return acc;
            acc += 0.01;
        }
        factor<T>(av, 1.8);
        return 1.8;
    }

    // --- convergence bookkeeping (identical to the reference) -----------------
    double abs_conv = 0.0, normalizer = 1.0, best = 1.0e+6;
    int bad = 0;
    static constexpr double kSmallAbs = 1.0e-20;
    // returns 1 (return true), -1 continue
    int prologue(double normB, double pure, double conv_cri) {
        if (normB <= 1.0e-30 && pure <= kSmallAbs) {
            if (o.is_save_log) { log.clear(); log.push_back(0.0); }
            return 1;
        }
        abs_conv = (normB * conv_cri * 0.9 < kSmallAbs ? kSmallAbs : normB * conv_cri * 0.9);
        const double first = pure / normB;
        if (first < conv_cri * 0.5 || pure < abs_conv * 0.5) return 1;
        if (o.conv_normalize_type == 1) normalizer = pure;
        else if (o.conv_normalize_type == 2) normalizer = o.conv_normalize_const;
        else normalizer = normB;
        if (o.is_save_log) log.push_back(pure / normalizer);
        return -1;
    }
    // 1 converged, 2 diverged, 0 continue; saves the best iterate when asked
    template <typename T> int monitor(double pure, double conv_cri, const T* x, T* best_x) {
        const double normR = pure / normalizer;
        final_res = normR;
        if (o.is_save_log) log.push_back(normR);
        if (normR < conv_cri || pure < abs_conv) return 1;
        if (normR < best) {
            best = normR;
            if (o.is_save_best) launch_copy<T>(s, m.n, x, best_x);
        }
        if (o.diverge_judge_type == 1) {
            if (normR < best * o.bad_div_val) bad = 0;
            if (normR >= best * o.bad_div_val) bad++;
            if (bad >= o.bad_div_count_thres) return 2;
        }
        return 0;
    }

    // --- ICCG / ICCOCG ---------------------------------------------------------
    template <typename T>
    bool iccg_main(double conv_cri, int max_ite, double normB, const T* av, const T* b, T* x) {
        const int n = m.n;
        T *P = vec<T>(0), *R = vec<T>(1), *Z = vec<T>(2), *AP = vec<T>(3), *tmp = vec<T>(4);
        T* best_x = o.is_save_best ? vec<T>(5) : nullptr;
        if (o.is_init) launch_fill<T>(s, n, x, Host<T>::zero());
        if (o.is_save_log) log.clear();
        if (best_x) launch_fill<T>(s, n, best_x, Host<T>::zero());
        if (!o.is_init) {
            spmv<T>(av, x, AP);
            launch_residual<T>(s, n, b, AP, R);
        } else {
            launch_residual<T>(s, n, b, nullptr, R);
        }
        const double pure = norm<T>(R);
        if (prologue(normB, pure, conv_cri) >= 0) return true;
        ic_apply<T>(R, P, tmp);
        launch_copy<T>(s, n, P, Z);
        T rur0 = dot<T>(R, Z);
        bool conv = false;
        for (int it = 0; it < max_ite; ++it) {
            const T t2 = spmv_dot<T>(av, P, AP);
            const T alpha = Host<T>::div(rur0, t2);
            launch_cg_update<T>(s, n, alpha, P, AP, x, R, partial());
            const double pn = std::sqrt(fetch(reduce_blocks(n), 1)[0]);
            iterations = it + 1;
            const int mo = monitor<T>(pn, conv_cri, x, best_x);
            if (mo == 1) { conv = true; break; }
            if (mo == 2) break;
            ic_apply<T>(R, Z, tmp);
            const T rur1 = dot<T>(R, Z);
            const T beta = Host<T>::div(rur1, rur0);
            rur0 = rur1;
            launch_cg_pupdate<T>(s, n, beta, Z, P);
        }
        if (!conv && o.is_save_best) launch_copy<T>(s, n, best_x, x);
        return conv;
    }

    // --- IC-MRTR ---------------------------------------------------------------
    template <typename T>
    bool mrtr_main(double conv_cri, int max_ite, double normB, const T* av, const T* b, T* xout) {
        const int n = m.n;
        T *P = vec<T>(0), *R = vec<T>(1), *U = vec<T>(2), *X = vec<T>(3), *Y = vec<T>(4), *Zv = vec<T>(5);
        T *V = vec<T>(6), *W = vec<T>(7), *tmp = vec<T>(8);
        T* best_x = o.is_save_best ? vec<T>(9) : nullptr;
        launch_fill<T>(s, n, P, Host<T>::zero());
        if (o.is_init) {
            launch_fill<T>(s, n, xout, Host<T>::zero());
            launch_fill<T>(s, n, X, Host<T>::zero());
        } else {
            launch_copy<T>(s, n, xout, X);
        }
        if (o.is_save_log) log.clear();
        if (best_x) launch_fill<T>(s, n, best_x, Host<T>::zero());
        if (o.is_init) {   // MatSolverMRTR_ICver.cpp:213-216 subtracts A*x only here
            spmv<T>(av, xout, V);
            launch_residual<T>(s, n, b, V, R);
        } else {
            launch_residual<T>(s, n, b, nullptr, R);
        }
        const double pure = norm<T>(R);
        if (prologue(normB, pure, conv_cri) >= 0) return true;
        ic_apply<T>(R, U, tmp);
        launch_negate<T>(s, n, R, Y);
        ic_apply<T>(Y, Zv, tmp);
        T zeta = one<T>(), zeta_old = one<T>(), eta = Host<T>::zero(), nu = one<T>();
        bool conv = false;
        for (int it = 0; it < max_ite; ++it) {
            spmv<T>(av, U, V);
            ic_apply<T>(V, W, tmp);
            launch_dot3<T>(s, n, W, R, V, it == 0 ? nullptr : Y, partial());
            const double* d = fetch(reduce_blocks(n), 3 * kparts<T>());
            const T w_r = Host<T>::from(d), v_w = Host<T>::from(d + kparts<T>());
            if (it == 0) {
                zeta = Host<T>::div(w_r, v_w);
                zeta_old = zeta;
                eta = Host<T>::zero();
            } else {
                const T w_y = Host<T>::from(d + 2 * kparts<T>());
                const T t0 = hdiv(1.0, nu * v_w - w_y * w_y);
                zeta = nu * w_r * t0;
                eta = -1.0 * w_y * w_r * t0;
            }
            nu = zeta * w_r;
            const T t = Host<T>::div(eta * zeta_old, zeta);
            zeta_old = zeta;
            launch_mrtr_update<T>(s, n, o.mrtr_direction_fix, t, zeta, eta, U, V, P, X, Y, R, partial());
            const double pn = std::sqrt(fetch(reduce_blocks(n), 1)[0]);
            iterations = it + 1;
            const int mo = monitor<T>(pn, conv_cri, X, best_x);
            if (mo == 1) { conv = true; break; }
            if (mo == 2) break;
            launch_mrtr_update2<T>(s, n, zeta, eta, W, Zv, U);
        }
        launch_copy<T>(s, n, X, xout);
        if (!conv && o.is_save_best) launch_copy<T>(s, n, best_x, xout);
        return conv;
    }

    // --- SGS-MRTR ----------------------------------------------------------------
    template <typename T>
    bool sgs_main(double conv_cri, int max_ite, double normB, const T* av, const T* b, T* xout) {
        const int n = m.n;
        T *P = vec<T>(0), *Rd = vec<T>(1), *U = vec<T>(2), *X = vec<T>(3), *Y = vec<T>(4), *ARd = vec<T>(5);
        T* tmpv = vec<T>(6);
        T* best_x = o.is_save_best ? vec<T>(9) : nullptr;
        launch_fill<T>(s, n, P, Host<T>::zero());
        if (o.is_init) {
            launch_fill<T>(s, n, xout, Host<T>::zero());
            launch_fill<T>(s, n, X, Host<T>::zero());
        } else {
            launch_copy<T>(s, n, xout, X);
        }
        if (o.is_save_log) log.clear();
        if (best_x) launch_fill<T>(s, n, best_x, Host<T>::zero());
        if (o.is_init) {
            spmv<T>(av, xout, ARd);
            launch_residual<T>(s, n, b, ARd, tmpv);
        } else {
            launch_residual<T>(s, n, b, nullptr, tmpv);
        }
        fr<T>(tmpv, Rd);
        const double pure = norm<T>(Rd);
        if (prologue(normB, pure, conv_cri) >= 0) return true;
        launch_negate<T>(s, n, Rd, Y);
        T zeta = one<T>(), zeta_old = one<T>(), eta = Host<T>::zero(), nu = one<T>();
        bool conv = false;
        // the preconditioned operator (I + L)^-1 S A S (I + U)^-1 on Rd.  FP64:
        // Eisenstat's splitting, U + fr(Rd - U) with U = bc(Rd), which is
        // exact only when the triangles are those of S A S.  FP32 triangles:
        // U = bc(Rd), then fr(S A S U) with the matrix in FP64, so only the
        // preconditioner is rounded and a singular consistent system stays
        // consistent (the splitting would apply the FP32-rounded matrix,
        // whose range differs, and stagnate).
        const bool f32 = o.factor_precision == SSGPU_FACTOR_F32;
        for (int it = 0; it < max_ite; ++it) {
            bc<T>(Rd, U);
            if (f32) {
                spmv<T>(av, U, tmpv);
                fr<T>(tmpv, ARd);
            } else {
                launch_sub<T>(s, n, Rd, U, tmpv);
                fr<T>(tmpv, ARd);
                launch_add_inplace<T>(s, n, U, ARd);
            }
            launch_dot3<T>(s, n, ARd, Rd, ARd, it == 0 ? nullptr : Y, partial());
            const double* d = fetch(reduce_blocks(n), 3 * kparts<T>());
            const T ar_r = Host<T>::from(d), ar_ar = Host<T>::from(d + kparts<T>());
            if (it == 0) {
                zeta = Host<T>::div(ar_r, ar_ar);
                zeta_old = zeta;
                eta = Host<T>::zero();
            } else {
                const T ar_y = Host<T>::from(d + 2 * kparts<T>());
                const T t = hdiv(1.0, nu * ar_ar - ar_y * ar_y);
                zeta = nu * ar_r * t;
                eta = -1.0 * ar_y * ar_r * t;
            }
            nu = zeta * ar_r;
            const T t2 = Host<T>::div(eta * zeta_old, zeta);
            zeta_old = zeta;
            launch_sgs_update<T>(s, n, t2, zeta, eta, U, ARd, P, X, Y, Rd, partial());
            const double pn = std::sqrt(fetch(reduce_blocks(n), 1)[0]);
            iterations = it + 1;
            const int mo = monitor<T>(pn, conv_cri, X, best_x);
            if (mo == 1) { conv = true; break; }
            if (mo == 2) break;
        }
        launch_copy<T>(s, n, X, xout);
        if (!conv && o.is_save_best) launch_copy<T>(s, n, best_x, xout);
        return conv;
    }

    template <typename T> static T one();
};
template <> double Solver::one<double>() { return 1.0; }
template <> zc Solver::one<zc>() { return zc{1.0, 0.0}; }

// Full solve with device-resident b, x (front ends MatSolverICCG.cpp:99-168,
// MatSolverMRTR_ICver.cpp:99-173, MatSolverMRTR_SGSver.cpp:42-69).
constexpr int kMaxRefine = 3;   // outer refinement passes of an FP32-factor solve

// True residual ||b - A x|| / normB of the returned x, recomputed in FP64.
// With an FP32 factor (and no diagonal scaling), outer refinement passes
// x += A^-1 r through the same factor until it meets conv_cri (at most
// kMaxRefine), for the solvers whose x is the solution of their recurrence.
template <typename T>
double finish_residual(Solver& S, int solver, const ssgpu_options& o, bool unscaled, bool conv, double conv_cri,
                       int max_ite, double normB, const T* b, T* x, int& refinements, int& refine_its) {
    Matrix& m = S.m;
    hipStream_t s = S.s;
    const int n = m.n;
    const double nb = normB > 0.0 ? normB : 1.0;
    T* rt = S.vec<T>(12);
    S.spmv<T>(m.val.as<T>(), x, rt);
    launch_residual<T>(s, n, b, rt, rt);
    double rn = S.norm<T>(rt);
    double true_res = rn / nb;
    refinements = refine_its = 0;
    const bool refinable = o.factor_precision == SSGPU_FACTOR_F32 && unscaled &&
                           (solver == SSGPU_ICCG || (solver == SSGPU_MRTR_IC && o.mrtr_direction_fix));
    while (refinable && conv && true_res > conv_cri && refinements < kMaxRefine) {
        ssgpu_options o2 = o;
        o2.is_init = 1;
        o2.is_save_log = 0;
        o2.is_save_best = 0;
        o2.diverge_judge_type = 0;
        o2.conv_normalize_type = 0;
        Solver R(m, S.c, o2);
        R.ws = S.ws;
        T* d = S.vec<T>(13);
        const double cc = std::min(0.5 * conv_cri * nb / rn, 0.5);
        const bool ok = solver == SSGPU_ICCG ? R.iccg_main<T>(cc, max_ite, rn, m.val.as<T>(), rt, d)
                                             : R.mrtr_main<T>(cc, max_ite, rn, m.val.as<T>(), rt, d);
        refine_its += R.iterations;
        launch_add_inplace<T>(s, n, d, x);
        ++refinements;
        S.spmv<T>(m.val.as<T>(), x, rt);
        launch_residual<T>(s, n, b, rt, rt);
        rn = S.norm<T>(rt);
        true_res = rn / nb;
        if (!ok) break;
    }
    return true_res;
}

// The end of a batched solve of one system, as solve_factored ends: the true
// residual and, for an FP32 factor, the refinement passes (context c's
// stream and work space).
template <typename T>
double batch_finish(Context* c, Matrix& m, int solver, const ssgpu_options& o, bool conv, double conv_cri,
                    int max_ite, double normB, const T* b, T* x, int& refinements, int& refine_its) {
    Solver S(m, c, o);
    return finish_residual<T>(S, solver, o, true, conv, conv_cri, max_ite, normB, b, x, refinements, refine_its);
}
template double batch_finish<double>(Context*, Matrix&, int, const ssgpu_options&, bool, double, int, double,
                                     const double*, double*, int&, int&);
template double batch_finish<zc>(Context*, Matrix&, int, const ssgpu_options&, bool, double, int, double, const zc*,
                                 zc*, int&, int&);

template <typename T>
int solve_device(Matrix& m, int solver, const ssgpu_options& o, double conv_cri, int max_ite, double normB,
                 const T* b, T* x, double* log, int log_cap, ssgpu_result* res) {
    Solver S(m, o);
    Context* c = m.ctx;
    hipStream_t s = c->stream;
    const int n = m.n;
    hipEvent_t e0, e1, e2;
    SSGPU_HIP(hipEventCreate(&e0));
    SSGPU_HIP(hipEventCreate(&e1));
    SSGPU_HIP(hipEventCreate(&e2));
    SSGPU_HIP(hipEventRecord(e0, s));
    const T* av = m.val.as<T>();
    const T* bb = b;
    T* xx = x;
    double nB = normB;
    bool scaled = false;
    bool conv = false;
    double accel_used = o.accel_factor;
    int fixes = 0;
    if (solver == SSGPU_MRTR_SGS) {
        ensure_upper(m);
        if (!m.lower_diag_ok || !m.upper_diag_ok)
            throw Error{NO_DIAGONAL, "SGS-MRTR needs every row's diagonal entry"};
        scaled = true;
    } else {
        if (!m.lower_diag_ok) throw Error{NO_DIAGONAL, "there is not diagonal element!"};
        scaled = o.is_diag_scale != 0;
    }
    if (scaled) {
        if (m.sc.bytes < sizeof(T) * n) m.sc.alloc(c, sizeof(T) * n);
        if (m.bs.bytes < sizeof(T) * n) m.bs.alloc(c, sizeof(T) * n);
        if (m.scaledA.bytes < sizeof(T) * m.nnz) m.scaledA.alloc(c, sizeof(T) * m.nnz);
        launch_diag_scale<T>(s, n, m.diag.as<int>(), av, b, m.sc.as<T>(), m.bs.as<T>());
        launch_scale_matrix<T>(s, n, m.rp.as<int>(), m.ci.as<int>(), av, m.sc.as<T>(), m.scaledA.as<T>());
        av = m.scaledA.as<T>();
        bb = m.bs.as<T>();
        nB = S.norm<T>(bb);
        T* xs = S.vec<T>(10);
        launch_copy<T>(s, n, x, xs);
        if (!o.is_init) launch_div_scale<T>(s, n, m.sc.as<T>(), xs);
        xx = xs;
    }
    if (solver == SSGPU_MRTR_SGS) {
        S.sgs_values<T>(av);
        SSGPU_HIP(hipEventRecord(e1, s));
        conv = S.sgs_main<T>(conv_cri, max_ite, nB, av, bb, xx);
    } else {
        if (o.accel_factor < 0.0) accel_used = S.auto_accel<T>(av, o.accel_factor);
        else S.factor<T>(av, o.accel_factor);
        fixes = m.factor_pivot_fixes;
        SSGPU_HIP(hipEventRecord(e1, s));
        if (solver == SSGPU_ICCG) conv = S.iccg_main<T>(conv_cri, max_ite, nB, av, bb, xx);
        else conv = S.mrtr_main<T>(conv_cri, max_ite, nB, av, bb, xx);
    }
    if (scaled) {
        launch_mul_scale<T>(s, n, m.sc.as<T>(), xx);
        launch_copy<T>(s, n, xx, x);
    }
    int refinements = 0, refine_its = 0;
    const double true_res = finish_residual<T>(S, solver, o, !scaled, conv, conv_cri, max_ite, normB, b, x,
                                               refinements, refine_its);
    SSGPU_HIP(hipEventRecord(e2, s));
    SSGPU_HIP(hipEventSynchronize(e2));
    S.check_sweep_error();
    float t_setup = 0.f, t_it = 0.f;
    (void)hipEventElapsedTime(&t_setup, e0, e1);
    (void)hipEventElapsedTime(&t_it, e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (res) {
        res->converged = conv ? 1 : 0;
        res->status = conv ? OK : NOT_CONVERGED;
        res->iterations = S.iterations;
        res->log_len = o.is_save_log ? (int)S.log.size() : 0;
        res->final_residual = S.final_res;
        res->accel_used = solver == SSGPU_MRTR_SGS ? 0.0 : accel_used;
        res->pivot_fixes = fixes;
        res->setup_ms = t_setup;
        res->iterate_ms = t_it;
        res->true_residual = true_res;
        res->refinements = refinements;
        res->refine_iterations = refine_its;
    }
    if (log && o.is_save_log)
        for (int i = 0; i < (int)S.log.size() && i < log_cap; ++i) log[i] = S.log[i];
    return conv ? OK : NOT_CONVERGED;
}

// Solve with the preconditioner built by ssgpu_factorize, on context c's
// stream and work space.  Nothing of the matrix is written, so solves on
// different contexts may run at once on one matrix.
// ws: take every device buffer of the solve from that work space (the
// allocation-free entry point; timed on the host clock then, no events).
template <typename T>
int solve_factored(Context* c, Matrix& m, int solver, const ssgpu_options& o, double conv_cri, int max_ite,
                   double normB, const T* b, T* x, double* log, int log_cap, ssgpu_result* res,
                   Workspace* ws = nullptr) {
    Solver S(m, c, o);
    S.ws = ws;
    hipStream_t s = c->stream;
    const int n = m.n;
    hipEvent_t e1 = nullptr, e2 = nullptr;
    const auto h0 = std::chrono::steady_clock::now();
    if (!ws) {
        SSGPU_HIP(hipEventCreate(&e1));
        SSGPU_HIP(hipEventCreate(&e2));
        SSGPU_HIP(hipEventRecord(e1, s));
    } else {
        launch_fill<int>(s, 2 * kScalSlots, reinterpret_cast<int*>(ws->scal), 0);   // the error flag starts clear
    }
    bool conv = false;
    if (solver == SSGPU_MRTR_SGS) {
        if (m.prep_kind != 2) throw Error{BAD_ARG, "no SGS preparation: call ssgpu_factorize with SSGPU_MRTR_SGS"};
        if (o.factor_precision == SSGPU_FACTOR_F32 && m.prep_precision != SSGPU_FACTOR_F32)
            throw Error{BAD_ARG, "the SGS preparation was built without its FP32 copy"};
        T* bs = S.vec<T>(14);
        T* xs = S.vec<T>(10);
        launch_copy<T>(s, n, b, bs);
        launch_mul_scale<T>(s, n, m.sc.as<T>(), bs);   // b~ = s o b (SprsPreconditioner.cpp:163-184)
        launch_copy<T>(s, n, x, xs);
        if (!o.is_init) launch_div_scale<T>(s, n, m.sc.as<T>(), xs);
        const double nB = S.norm<T>(bs);
        conv = S.sgs_main<T>(conv_cri, max_ite, nB, m.scaledA.as<T>(), bs, xs);
        launch_mul_scale<T>(s, n, m.sc.as<T>(), xs);
        launch_copy<T>(s, n, xs, x);
    } else {
        if (m.prep_kind != 1) throw Error{BAD_ARG, "no IC factor: call ssgpu_factorize first"};
        if (o.factor_precision == SSGPU_FACTOR_F32 && m.prep_precision != SSGPU_FACTOR_F32)
            throw Error{BAD_ARG, "the factor was built without its FP32 copy"};
        if (o.is_diag_scale) throw Error{UNSUPPORTED, "diagonal scaling is not available on a cached factor"};
        conv = solver == SSGPU_ICCG ? S.iccg_main<T>(conv_cri, max_ite, normB, m.val.as<T>(), b, x)
                                    : S.mrtr_main<T>(conv_cri, max_ite, normB, m.val.as<T>(), b, x);
    }
    int refinements = 0, refine_its = 0;
    const double true_res = finish_residual<T>(S, solver, o, solver != SSGPU_MRTR_SGS, conv, conv_cri, max_ite,
                                               normB, b, x, refinements, refine_its);
    float t_it = 0.f;
    if (!ws) {
        SSGPU_HIP(hipEventRecord(e2, s));
        SSGPU_HIP(hipEventSynchronize(e2));
        S.check_sweep_error();
        (void)hipEventElapsedTime(&t_it, e1, e2);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
    } else {
        S.check_sweep_error();
        t_it = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    }
    if (res) {
        res->converged = conv ? 1 : 0;
        res->status = conv ? OK : NOT_CONVERGED;
        res->iterations = S.iterations;
        res->log_len = o.is_save_log ? (int)S.log.size() : 0;
        res->final_residual = S.final_res;
        res->accel_used = solver == SSGPU_MRTR_SGS ? 0.0 : m.factor_accel;
        res->pivot_fixes = solver == SSGPU_MRTR_SGS ? 0 : m.factor_pivot_fixes;
        res->setup_ms = 0.0;
        res->iterate_ms = t_it;
        res->true_residual = true_res;
        res->refinements = refinements;
        res->refine_iterations = refine_its;
    }
    if (log && o.is_save_log)
        for (int i = 0; i < (int)S.log.size() && i < log_cap; ++i) log[i] = S.log[i];
    return conv ? OK : NOT_CONVERGED;
}

// Block solves: nrhs right-hand sides against the cached preparation, in
// kb >= nrhs padded columns, a row's columns side by side (entry (i, c) at
// i * kb + c; the caller's vectors are transposed in and out).
// Column j runs exactly the solve iccg_main (run), mrtr_main or sgs_main
// (run_mrtr) runs on it alone (same kernels per column, same reduction
// trees, same host recurrences and bookkeeping); a column that converges or
// diverges stops updating its x while the others go on.
template <typename T> struct BlockSolve {
    Solver& S;
    Context* c;
    Matrix& m;
    hipStream_t s;
    int n, kb, P;
    using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
    explicit BlockSolve(Solver& SS, int kbb) : S(SS), c(SS.c), m(SS.m), s(SS.s), n(SS.m.n), kb(kbb), P(kparts<T>()) {}
    // polling waves of a block sweep: 128 per column up to 2048 in all
    // (measured best at n=95 for 8, 16 and 32 columns)
    int sweep_waves() const { return m.block_grid > 0 ? m.block_grid : std::min(2048, 128 * kb); }

    T* bvec(size_t idx) {   // n x kb work array from the context's pool
        while (c->work.size() <= idx) c->work.emplace_back(new DevBuf());
        DevBuf& b = *c->work[idx];
        const size_t need = sizeof(T) * (size_t)std::max(n, 1) * kb;
        if (b.bytes < need) b.alloc(c, need);
        return b.as<T>();
    }
    double* bpartial() {
        // per SpMV block at most kb * P = 64 dot parts; per reduction block
        // at most three regions of kb * P (the MRTR phase's three dots)
        const size_t need = sizeof(double) * std::max(64 * (size_t)m.nblk, 192 * (size_t)kMaxReduceBlocks);
        if (c->d_bpartial.bytes < need) c->d_bpartial.alloc(c, need);
        return c->d_bpartial.as<double>();
    }
    int* btickets() {
        if (c->d_btickets.bytes == 0) c->d_btickets.alloc(c, 64 * sizeof(int));
        return c->d_btickets.as<int>();
    }
    void ensure_scal() {
        if (c->d_bscal.bytes == 0) c->d_bscal.alloc(c, 4096);
        if (!c->h_bscal) SSGPU_HIP(hipHostMalloc(&c->h_bscal, 4096, hipHostMallocDefault));
        (void)scal_buf(c);
    }
    // finalize kb*P*q partial sums over nb blocks (q per-column sums of
    // P parts each) and read them back (synchronizes)
    std::vector<double> fetch(int nb, int q = 1) { return fetch_k(nb, kb * P * q); }
    // the per-column squared norms (one part per column)
    std::vector<double> fetch_norms(int nb) { return fetch_k(nb, kb); }
    std::vector<double> fetch_k(int nb, int K) {
        launch_finalize(s, nb, K, c->d_bpartial.as<double>(), c->d_scal.as<double>());
        std::vector<double> out(K);
        const double* h = read_scalars(c);
        std::copy(h, h + K, out.begin());
        return out;
    }
    std::vector<T> dots(const std::vector<double>& v) {
        std::vector<T> out(kb);
        for (int j = 0; j < kb; ++j) out[j] = Host<T>::from(&v[(size_t)j * P]);
        return out;
    }
    // per-column scalars (at most kActOff bytes: three per column) and active
    // flags to the device (one copy)
    static constexpr size_t kActOff = 3072;
    void put_scalars(const std::vector<T>& a, const std::vector<int>& act) {
        char* h = static_cast<char*>(c->h_bscal);
        std::memcpy(h, a.data(), sizeof(T) * a.size());
        std::memcpy(h + kActOff, act.data(), sizeof(int) * kb);
        SSGPU_HIP(hipMemcpyAsync(c->d_bscal.p, h, kActOff + sizeof(int) * kb, hipMemcpyHostToDevice, s));
    }
    const T* d_scalars() { return c->d_bscal.as<T>(); }
    const int* d_act() { return reinterpret_cast<const int*>(static_cast<char*>(c->d_bscal.p) + kActOff); }

    // z = M^-1 r for every column (the interleaved block sweeps)
    void ic_apply(const T* r, T* z, T* tmp) {
        launch_sf_prepare<T>(s, n * kb, tmp, z, S.tickets());
        int* tk = btickets();
        launch_fill<int>(s, 64, tk, 0);
        const int grid = sweep_waves();
        if (S.o.factor_precision == SSGPU_FACTOR_F32) {
            launch_sweep_il<T, F>(s, FWD_DIV_LAST, m.sfL.kmax, n, grid, kb, tk, S.errflag(), S.spin_limit, m.sfL.srow,
                                  m.sfL.slice_off.as<int>(), m.sfL.rlen.as<int>(), m.sfL.ecol.as<int>(),
                                  m.sfL.eval32.as<F>(), m.sfL.dval32.as<F>(), r, nullptr, nullptr, tmp);
            launch_sweep_il<T, F>(s, BWD_IC, m.sfUic.kmax, n, grid, kb, tk + 1, S.errflag(), S.spin_limit,
                                  m.sfUic.srow, m.sfUic.slice_off.as<int>(), m.sfUic.rlen.as<int>(),
                                  m.sfUic.ecol.as<int>(), m.sfUic.eval32.as<F>(), nullptr, nullptr, tmp,
                                  m.diagD.as<T>(), z);
        } else {
            launch_sweep_il<T, T>(s, FWD_DIV_LAST, m.sfL.kmax, n, grid, kb, tk, S.errflag(), S.spin_limit, m.sfL.srow,
                                  m.sfL.slice_off.as<int>(), m.sfL.rlen.as<int>(), m.sfL.ecol.as<int>(),
                                  m.sfL.eval.as<T>(), m.sfL.dval.as<T>(), r, nullptr, nullptr, tmp);
            launch_sweep_il<T, T>(s, BWD_IC, m.sfUic.kmax, n, grid, kb, tk + 1, S.errflag(), S.spin_limit,
                                  m.sfUic.srow, m.sfUic.slice_off.as<int>(), m.sfUic.rlen.as<int>(),
                                  m.sfUic.ecol.as<int>(), m.sfUic.eval.as<T>(), nullptr, nullptr, tmp,
                                  m.diagD.as<T>(), z);
        }
    }

    // the plain sweeps of SGS-MRTR on S A S's triangles: z = fr(r) (forward,
    // divide by the row's last entry) or bc(r) (backward, divide by the first)
    void sgs_sweep(bool forward, const T* r, T* z) {
        launch_sf_prepare<T>(s, n * kb, z, nullptr, S.tickets());
        int* tk = btickets();
        launch_fill<int>(s, 64, tk, 0);
        const int grid = sweep_waves();
        const SfCsr& sf = forward ? m.sfL : m.sfUs;
        const int mode = forward ? FWD_DIV_LAST : BWD_DIV_FIRST;
        if (S.o.factor_precision == SSGPU_FACTOR_F32)
            launch_sweep_il<T, F>(s, mode, sf.kmax, n, grid, kb, tk, S.errflag(), S.spin_limit, sf.srow,
                                  sf.slice_off.as<int>(), sf.rlen.as<int>(), sf.ecol.as<int>(), sf.eval32.as<F>(),
                                  sf.dval32.as<F>(), r, nullptr, nullptr, z);
        else
            launch_sweep_il<T, T>(s, mode, sf.kmax, n, grid, kb, tk, S.errflag(), S.spin_limit, sf.srow,
                                  sf.slice_off.as<int>(), sf.rlen.as<int>(), sf.ecol.as<int>(), sf.eval.as<T>(),
                                  sf.dval.as<T>(), r, nullptr, nullptr, z);
    }

    struct Col {
        bool active = true, conv = false;
        int iterations = 0;
        double normalizer = 1.0, abs_conv = 0.0, best = 1.0e+6, final_res = 0.0;
        int bad = 0;
    };

    // B, X: kb columns of n entries (padding columns of B are zero)
    std::vector<Col> run(double conv_cri, int max_ite, const std::vector<double>& normB, const T* B, T* X) {
        const ssgpu_options& o = S.o;
        ensure_scal();
        T *P_ = bvec(20), *R = bvec(21), *Z = bvec(22), *AP = bvec(23), *TMP = bvec(24);
        T* BEST = o.is_save_best ? bvec(25) : nullptr;
        const int64_t nk = (int64_t)n * kb;
        std::vector<Col> col(kb);
        if (o.is_init) launch_fill<T>(s, nk, X, Host<T>::zero());
        if (BEST) launch_fill<T>(s, nk, BEST, Host<T>::zero());
        if (!o.is_init) {
            launch_spmv_il<T>(s, kb, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(), X,
                                AP, nullptr, nullptr);
            launch_residual<T>(s, (int)nk, B, AP, R);
        } else {
            launch_residual<T>(s, (int)nk, B, nullptr, R);
        }
        bpartial();
        launch_norm2_il<T>(s, kb, n, R, c->d_bpartial.as<double>());
        {
            const int K = kb;   // norms: one part per column
            launch_finalize(s, reduce_blocks(n), K, c->d_bpartial.as<double>(), c->d_scal.as<double>());
            (void)read_scalars(c);
        }
        std::vector<double> pure(c->h_scal, c->h_scal + kb);
        constexpr double kSmallAbs = 1.0e-20;
        for (int j = 0; j < kb; ++j) {   // the prologue of Solver::prologue, per column
            Col& q = col[j];
            pure[j] = std::sqrt(pure[j]);
            const double nB = normB[j];
            if (nB <= 1.0e-30 && pure[j] <= kSmallAbs) { q.active = false; q.conv = true; continue; }
            q.abs_conv = (nB * conv_cri * 0.9 < kSmallAbs ? kSmallAbs : nB * conv_cri * 0.9);
            if (pure[j] / nB < conv_cri * 0.5 || pure[j] < q.abs_conv * 0.5) { q.active = false; q.conv = true; continue; }
            if (o.conv_normalize_type == 1) q.normalizer = pure[j];
            else if (o.conv_normalize_type == 2) q.normalizer = o.conv_normalize_const;
            else q.normalizer = nB;
        }
        auto any_active = [&]() {
            for (auto& q : col) if (q.active) return true;
            return false;
        };
        if (!any_active()) return col;
        ic_apply(R, P_, TMP);
        launch_copy<T>(s, nk, P_, Z);
        launch_dot_il<T>(s, kb, n, R, Z, c->d_bpartial.as<double>());
        std::vector<T> rur0 = dots(fetch(reduce_blocks(n)));
        std::vector<T> sc(kb, Host<T>::zero());
        std::vector<int> act(kb, 0);
        for (int it = 0; it < max_ite && any_active(); ++it) {
            launch_spmv_il<T>(s, kb, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(),
                                P_, AP, P_, c->d_bpartial.as<double>());
            const std::vector<T> t2 = dots(fetch(m.nblk));
            for (int j = 0; j < kb; ++j) {
                act[j] = col[j].active ? 1 : 0;
                sc[j] = col[j].active ? Host<T>::div(rur0[j], t2[j]) : Host<T>::zero();
            }
            put_scalars(sc, act);
            launch_cg_update_il<T>(s, kb, n, d_scalars(), d_act(), P_, AP, X, R, c->d_bpartial.as<double>());
            std::vector<double> pn;
            {
                launch_finalize(s, reduce_blocks(n), kb, c->d_bpartial.as<double>(), c->d_scal.as<double>());
                const double* h = read_scalars(c);
                pn.assign(h, h + kb);
            }
            for (int j = 0; j < kb; ++j) {   // Solver::monitor, per column
                Col& q = col[j];
                if (!q.active) continue;
                q.iterations = it + 1;
                const double p = std::sqrt(pn[j]);
                const double normR = p / q.normalizer;
                q.final_res = normR;
                if (normR < conv_cri || p < q.abs_conv) { q.active = false; q.conv = true; continue; }
                if (normR < q.best) {
                    q.best = normR;
                    if (BEST) launch_copy_col_il<T>(s, kb, n, j, X, BEST);
                }
                if (o.diverge_judge_type == 1) {
                    if (normR < q.best * o.bad_div_val) q.bad = 0;
                    if (normR >= q.best * o.bad_div_val) q.bad++;
                    if (q.bad >= o.bad_div_count_thres) { q.active = false; continue; }
                }
            }
            if (!any_active()) break;
            ic_apply(R, Z, TMP);
            launch_dot_il<T>(s, kb, n, R, Z, c->d_bpartial.as<double>());
            const std::vector<T> rur1 = dots(fetch(reduce_blocks(n)));
            for (int j = 0; j < kb; ++j) {
                act[j] = col[j].active ? 1 : 0;
                sc[j] = col[j].active ? Host<T>::div(rur1[j], rur0[j]) : Host<T>::zero();
                if (col[j].active) rur0[j] = rur1[j];
            }
            put_scalars(sc, act);
            launch_cg_pupdate_il<T>(s, kb, n, d_scalars(), d_act(), Z, P_);
        }
        if (BEST)
            for (int j = 0; j < kb; ++j)
                if (!col[j].conv) launch_copy_col_il<T>(s, kb, n, j, BEST, X);
        return col;
    }

    // the prologue of Solver::prologue per column, from the norms of the
    // first residuals; false when no column goes on
    bool prologue(std::vector<Col>& col, double conv_cri, const std::vector<double>& normB,
                  const std::vector<double>& pure2) {
        constexpr double kSmallAbs = 1.0e-20;
        bool any = false;
        for (int j = 0; j < kb; ++j) {
            Col& q = col[j];
            const double pure = std::sqrt(pure2[j]);
            const double nB = normB[j];
            if (nB <= 1.0e-30 && pure <= kSmallAbs) { q.active = false; q.conv = true; continue; }
            q.abs_conv = (nB * conv_cri * 0.9 < kSmallAbs ? kSmallAbs : nB * conv_cri * 0.9);
            if (pure / nB < conv_cri * 0.5 || pure < q.abs_conv * 0.5) { q.active = false; q.conv = true; continue; }
            if (S.o.conv_normalize_type == 1) q.normalizer = pure;
            else if (S.o.conv_normalize_type == 2) q.normalizer = S.o.conv_normalize_const;
            else q.normalizer = nB;
            any = true;
        }
        return any;
    }

    // IC-MRTR (Solver::mrtr_main) or, with sgs, the SGS-MRTR iteration
    // (Solver::sgs_main) per column.  B, X: kb columns; for SGS they are the
    // scaled right-hand sides and iterates and normB the norms of the scaled
    // right-hand sides, as solve_factored hands them to sgs_main.
    std::vector<Col> run_mrtr(bool sgs, double conv_cri, int max_ite, const std::vector<double>& normB, const T* B,
                              T* X) {
        const ssgpu_options& o = S.o;
        ensure_scal();
        bpartial();
        // P, R (Rd), U, X (the iterate), Y, Zv (IC) and V, W (IC; ARd, tmp for SGS), tmp
        T *P_ = bvec(20), *R = bvec(21), *U = bvec(22), *Xi = bvec(23), *Y = bvec(24), *Zv = bvec(25);
        T *V = bvec(28), *W = bvec(29), *TMP = bvec(30);
        T* BEST = o.is_save_best ? bvec(31) : nullptr;
        const int64_t nk = (int64_t)n * kb;
        const T* av = sgs ? m.scaledA.as<T>() : m.val.as<T>();
        const bool f32 = o.factor_precision == SSGPU_FACTOR_F32;
        std::vector<Col> col(kb);
        launch_fill<T>(s, nk, P_, Host<T>::zero());
        if (o.is_init) {
            launch_fill<T>(s, nk, X, Host<T>::zero());
            launch_fill<T>(s, nk, Xi, Host<T>::zero());
        } else {
            launch_copy<T>(s, nk, X, Xi);
        }
        if (BEST) launch_fill<T>(s, nk, BEST, Host<T>::zero());
        // the first residual: MatSolverMRTR_ICver.cpp:213-216 and
        // MatSolverMRTR_SGSver.cpp:112-114 subtract A x only when is_init
        T* r0 = sgs ? TMP : R;
        if (o.is_init) {
            launch_spmv_il<T>(s, kb, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), av, X, V, nullptr,
                                nullptr);
            launch_residual<T>(s, (int)nk, B, V, r0);
        } else {
            launch_residual<T>(s, (int)nk, B, nullptr, r0);
        }
        if (sgs) sgs_sweep(true, TMP, R);   // Rd = fr(b - A x)
        launch_norm2_il<T>(s, kb, n, R, c->d_bpartial.as<double>());
        if (!prologue(col, conv_cri, normB, fetch_norms(reduce_blocks(n)))) return col;
        if (!sgs) ic_apply(R, U, TMP);
        launch_negate<T>(s, (int)nk, R, Y);
        if (!sgs) ic_apply(Y, Zv, TMP);
        std::vector<T> zeta(kb, Solver::one<T>()), zeta_old(kb, Solver::one<T>()), eta(kb, Host<T>::zero()),
            nu(kb, Solver::one<T>());
        std::vector<T> sc(3 * (size_t)kb, Host<T>::zero());
        std::vector<int> act(kb, 0);
        auto any_active = [&]() {
            for (auto& q : col) if (q.active) return true;
            return false;
        };
        // v = the preconditioned operator on u's direction, w its product
        // with the preconditioner (IC); for SGS both are ARd
        T *dv = V, *dw = W;
        for (int it = 0; it < max_ite && any_active(); ++it) {
            if (sgs) {
                sgs_sweep(false, R, U);   // U = bc(Rd)
                if (f32) {
                    launch_spmv_il<T>(s, kb, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), av, U,
                                        TMP, nullptr, nullptr);
                    sgs_sweep(true, TMP, V);
                } else {
                    launch_sub<T>(s, (int)nk, R, U, TMP);
                    sgs_sweep(true, TMP, V);
                    launch_add_inplace<T>(s, (int)nk, U, V);
                }
                dv = V;
                dw = V;
            } else {
                launch_spmv_il<T>(s, kb, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), av, U, V,
                                    nullptr, nullptr);
                ic_apply(V, W, TMP);
            }
            // the three dots of k_dot3 (w.r, w.v and, from the second
            // iteration on, w.y), each over its own partials and finalized
            // into its own third of the scalars (the same trees per value)
            {
                const int nb = reduce_blocks(n);
                double* part = c->d_bpartial.as<double>();
                double* out = c->d_scal.as<double>();
                const size_t reg = (size_t)nb * kb * P;
                const int nd = it == 0 ? 2 : 3;
                const T* rhs3[3] = {R, dv, Y};
                for (int k = 0; k < nd; ++k) launch_dot_il<T>(s, kb, n, dw, rhs3[k], part + k * reg);
                for (int k = 0; k < nd; ++k) launch_finalize(s, nb, kb * P, part + k * reg, out + (size_t)k * kb * P);
            }
            const double* d3 = read_scalars(c);
            for (int j = 0; j < kb; ++j) {
                act[j] = col[j].active ? 1 : 0;
                if (!col[j].active) continue;
                const T w_r = Host<T>::from(d3 + (size_t)j * P), v_w = Host<T>::from(d3 + (size_t)(kb + j) * P);
                if (it == 0) {
                    zeta[j] = Host<T>::div(w_r, v_w);
                    zeta_old[j] = zeta[j];
                    eta[j] = Host<T>::zero();
                } else {
                    const T w_y = Host<T>::from(d3 + (size_t)(2 * kb + j) * P);
                    const T t0 = hdiv(1.0, nu[j] * v_w - w_y * w_y);
                    zeta[j] = nu[j] * w_r * t0;
                    eta[j] = -1.0 * w_y * w_r * t0;
                }
                nu[j] = zeta[j] * w_r;
                sc[j] = Host<T>::div(eta[j] * zeta_old[j], zeta[j]);
                sc[kb + j] = zeta[j];
                sc[2 * kb + j] = eta[j];
                zeta_old[j] = zeta[j];
            }
            put_scalars(sc, act);
            launch_mrtr_update_il<T>(s, kb, n, sgs, o.mrtr_direction_fix, d_scalars(), d_act(), U, dv, P_, Xi, Y,
                                       R, c->d_bpartial.as<double>());
            const std::vector<double> pn = fetch_norms(reduce_blocks(n));
            for (int j = 0; j < kb; ++j) {   // Solver::monitor, per column
                Col& q = col[j];
                if (!q.active) continue;
                q.iterations = it + 1;
                const double p = std::sqrt(pn[j]);
                const double normR = p / q.normalizer;
                q.final_res = normR;
                if (normR < conv_cri || p < q.abs_conv) { q.active = false; q.conv = true; continue; }
                if (normR < q.best) {
                    q.best = normR;
                    if (BEST) launch_copy_col_il<T>(s, kb, n, j, Xi, BEST);
                }
                if (o.diverge_judge_type == 1) {
                    if (normR < q.best * o.bad_div_val) q.bad = 0;
                    if (normR >= q.best * o.bad_div_val) q.bad++;
                    if (q.bad >= o.bad_div_count_thres) { q.active = false; continue; }
                }
            }
            if (!any_active()) break;
            if (!sgs) {
                for (int j = 0; j < kb; ++j) act[j] = col[j].active ? 1 : 0;
                put_scalars(sc, act);
                launch_mrtr_update2_il<T>(s, kb, n, d_scalars(), d_act(), W, Zv, U);
            }
        }
        launch_copy<T>(s, nk, Xi, X);
        if (BEST)
            for (int j = 0; j < kb; ++j)
                if (!col[j].conv) launch_copy_col_il<T>(s, kb, n, j, BEST, X);
        return col;
    }
};

// Block solve entry: copy to kb padded columns, run, copy back, true
// residuals (and, for an FP32 factor, the refinement passes of
// solve_factored on each column that needs them).  Column j returns the bits
// of solve_factored on right-hand side j.
template <typename T>
int solve_block(Context* c, Matrix& m, int solver, const ssgpu_options& o, double conv_cri, int max_ite, int nrhs,
                const double* normB, const T* b, T* x, ssgpu_result* res) {
    const bool sgs = solver == SSGPU_MRTR_SGS;
    if (solver != SSGPU_ICCG && solver != SSGPU_MRTR_IC && !sgs) throw Error{BAD_ARG, "unknown solver"};
    if (sgs) {
        if (m.prep_kind != 2) throw Error{BAD_ARG, "no SGS preparation: call ssgpu_factorize with SSGPU_MRTR_SGS"};
        if (o.factor_precision == SSGPU_FACTOR_F32 && m.prep_precision != SSGPU_FACTOR_F32)
            throw Error{BAD_ARG, "the SGS preparation was built without its FP32 copy"};
        if (!m.sfL.kmax || !m.sfUs.kmax || m.sfL.kmax > 32 || m.sfUs.kmax > 32)
            throw Error{UNSUPPORTED, "block solves need rows of at most 32 off-diagonal entries"};
    } else {
        if (m.prep_kind != 1) throw Error{BAD_ARG, "no IC factor: call ssgpu_factorize first"};
        if (o.factor_precision == SSGPU_FACTOR_F32 && m.prep_precision != SSGPU_FACTOR_F32)
            throw Error{BAD_ARG, "the factor was built without its FP32 copy"};
        if (o.is_diag_scale) throw Error{UNSUPPORTED, "diagonal scaling is not available on a cached factor"};
        if (!m.sfL.kmax || !m.sfUic.kmax || m.sfL.kmax > 32 || m.sfUic.kmax > 32)
            throw Error{UNSUPPORTED, "block solves need rows of at most 32 off-diagonal entries"};
    }
    int kb = 1;
    while (kb < nrhs) kb *= 2;
    if (kb > 32) throw Error{BAD_ARG, "at most 32 right-hand sides per block solve"};
    if ((int64_t)m.n * kb >= INT_MAX) throw Error{BAD_ARG, "n x the padded column count must stay below 2^31"};
    Solver S(m, c, o);
    BlockSolve<T> B(S, kb);
    hipStream_t s = c->stream;
    const int n = m.n;
    hipEvent_t e1, e2;
    SSGPU_HIP(hipEventCreate(&e1));
    SSGPU_HIP(hipEventCreate(&e2));
    SSGPU_HIP(hipEventRecord(e1, s));
    // the caller's vectors (one after another) to the block layout, the kb
    // columns of a row side by side (padding columns zero)
    T* Bp = B.bvec(26);
    T* Xp = B.bvec(27);
    launch_to_il<T>(s, kb, n, nrhs, b, Bp);
    launch_to_il<T>(s, kb, n, nrhs, x, Xp);
    std::vector<double> nB(kb, 0.0);
    std::vector<typename BlockSolve<T>::Col> col;
    if (sgs) {
        // as solve_factored around sgs_main: b~ = s o b, x~ = x / s for a
        // warm start, the norms of b~ for the checks, x = s o x~ at the end
        T* Bs = B.bvec(32);
        T* Xs = B.bvec(33);
        launch_copy<T>(s, (int64_t)n * kb, Bp, Bs);
        launch_copy<T>(s, (int64_t)n * kb, Xp, Xs);
        launch_scale_il<T>(s, kb, n, m.sc.as<T>(), Bs, false);
        if (!o.is_init) launch_scale_il<T>(s, kb, n, m.sc.as<T>(), Xs, true);
        B.ensure_scal();
        B.bpartial();
        launch_norm2_il<T>(s, kb, n, Bs, c->d_bpartial.as<double>());
        const std::vector<double> n2 = B.fetch_norms(reduce_blocks(n));
        for (int j = 0; j < kb; ++j) nB[j] = std::sqrt(n2[j]);
        col = B.run_mrtr(true, conv_cri, max_ite, nB, Bs, Xs);
        launch_scale_il<T>(s, kb, n, m.sc.as<T>(), Xs, false);
        launch_copy<T>(s, (int64_t)n * kb, Xs, Xp);
    } else {
        for (int j = 0; j < nrhs; ++j) nB[j] = normB[j];
        col = solver == SSGPU_ICCG ? B.run(conv_cri, max_ite, nB, Bp, Xp)
                                   : B.run_mrtr(false, conv_cri, max_ite, nB, Bp, Xp);
    }
    launch_from_il<T>(s, kb, n, nrhs, Xp, x);
    // true residuals ||b_j - A x_j|| / normB_j
    T* Rt = B.bvec(21);
    launch_spmv_il<T>(s, kb, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(), Xp, Rt,
                        nullptr, nullptr);
    launch_residual<T>(s, n * kb, Bp, Rt, Rt);
    B.bpartial();
    launch_norm2_il<T>(s, kb, n, Rt, c->d_bpartial.as<double>());
    launch_finalize(s, reduce_blocks(n), kb, c->d_bpartial.as<double>(), c->d_scal.as<double>());
    SSGPU_HIP(hipMemcpyAsync(c->h_scal, c->d_scal.p, sizeof(double) * kb, hipMemcpyDeviceToHost, s));
    SSGPU_HIP(hipEventRecord(e2, s));
    SSGPU_HIP(hipEventSynchronize(e2));
    S.check_sweep_error();
    std::vector<double> true_res(nrhs);
    for (int j = 0; j < nrhs; ++j) true_res[j] = std::sqrt(c->h_scal[j]) / (normB[j] > 0.0 ? normB[j] : 1.0);
    // an FP32 factor's refinement passes, column by column, exactly as
    // solve_factored runs them (finish_residual recomputes the same true
    // residual first)
    const bool refinable = o.factor_precision == SSGPU_FACTOR_F32 &&
                           (solver == SSGPU_ICCG || (solver == SSGPU_MRTR_IC && o.mrtr_direction_fix));
    std::vector<int> refinements(nrhs, 0), refine_its(nrhs, 0);
    if (refinable)
        for (int j = 0; j < nrhs; ++j)
            if (col[j].conv && true_res[j] > conv_cri)
                true_res[j] = finish_residual<T>(S, solver, o, true, true, conv_cri, max_ite, normB[j],
                                                 b + (size_t)j * n, x + (size_t)j * n, refinements[j], refine_its[j]);
    SSGPU_HIP(hipEventRecord(e2, s));
    SSGPU_HIP(hipEventSynchronize(e2));
    S.check_sweep_error();
    float t_it = 0.f;
    (void)hipEventElapsedTime(&t_it, e1, e2);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    int all_conv = 1;
    for (int j = 0; j < nrhs; ++j) {
        ssgpu_result& r = res[j];
        r = ssgpu_result{};
        r.converged = col[j].conv ? 1 : 0;
        r.status = col[j].conv ? OK : NOT_CONVERGED;
        r.iterations = col[j].iterations;
        r.final_residual = col[j].final_res;
        r.accel_used = sgs ? 0.0 : m.factor_accel;
        r.pivot_fixes = sgs ? 0 : m.factor_pivot_fixes;
        r.iterate_ms = t_it;
        r.true_residual = true_res[j];
        r.refinements = refinements[j];
        r.refine_iterations = refine_its[j];
        all_conv &= r.converged;
    }
    return all_conv ? OK : NOT_CONVERGED;
}

// Build the preconditioner a cached solve uses: the shifted IC(0) factor
// (automatic shift when accel < 0) or, for SGS-MRTR, S A S and its triangles.
template <typename T> void factorize(Matrix& m, int solver, double accel, int precision, double* accel_used,
                                     int* fixes) {
    ssgpu_options o;
    ssgpu_options_default(&o);
    o.factor_precision = precision;
    Solver S(m, o);
    hipStream_t s = m.ctx->stream;
    const int n = m.n;
    if (solver == SSGPU_MRTR_SGS) {
        ensure_upper(m);
        if (!m.lower_diag_ok || !m.upper_diag_ok)
            throw Error{NO_DIAGONAL, "SGS-MRTR needs every row's diagonal entry"};
        if (m.sc.bytes < sizeof(T) * n) m.sc.alloc(m.ctx, sizeof(T) * n);
        if (m.scaledA.bytes < sizeof(T) * m.nnz) m.scaledA.alloc(m.ctx, sizeof(T) * m.nnz);
        launch_diag_scale<T>(s, n, m.diag.as<int>(), m.val.as<T>(), nullptr, m.sc.as<T>(), nullptr);
        launch_scale_matrix<T>(s, n, m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(), m.sc.as<T>(),
                               m.scaledA.as<T>());
        S.sgs_values<T>(m.scaledA.as<T>());
        if (accel_used) *accel_used = 0.0;
        if (fixes) *fixes = 0;
    } else {
        if (!m.lower_diag_ok) throw Error{NO_DIAGONAL, "there is not diagonal element!"};
        const double a = accel < 0.0 ? S.auto_accel<T>(m.val.as<T>(), accel) : (S.factor<T>(m.val.as<T>(), accel), accel);
        if (accel_used) *accel_used = a;
        if (fixes) *fixes = m.factor_pivot_fixes;
    }
    SSGPU_HIP(hipStreamSynchronize(s));
}

// SpMV staging, measured once per matrix: both LDS footprints give the same
// bits; the faster one on this matrix's gather pattern is kept.
template <typename T> void tune_spmv(Matrix& m, T one) {
    if (const char* e = std::getenv("SSGPU_SPMV_STAGE_KB")) {
        m.spmv_stage_kb = std::atoi(e) >= 32 ? 32 : std::atoi(e) >= 16 ? 16 : 8;
        return;
    }
    Context* c = m.ctx;
    DevBuf x, y;
    x.alloc(c, sizeof(T) * m.n);
    y.alloc(c, sizeof(T) * m.n);
    launch_fill<T>(c->stream, m.n, x.as<T>(), one);
    hipEvent_t e0, e1;
    SSGPU_HIP(hipEventCreate(&e0));
    SSGPU_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    for (int kb : {8, 16, 32}) {
        float t = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {   // the first run warms up
            SSGPU_HIP(hipEventRecord(e0, c->stream));
            launch_spmv<T>(c->stream, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(),
                           x.as<T>(), y.as<T>(), nullptr, nullptr, kb);
            SSGPU_HIP(hipEventRecord(e1, c->stream));
            SSGPU_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            SSGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) t = std::min(t, ms);
        }
        if (t < best) {
            best = t;
            m.spmv_stage_kb = kb;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// Solves on an ABMC-reordered matrix: b and x (ncol vectors of n) are
// gathered into the matrix's order, solved there, and x gathered back.
// The permuted copies live in the context's work pool (indices 40 and 41,
// beside the solvers' own work vectors), so repeated solves allocate nothing.
struct PermutedIO {
    Context* c;
    Matrix& m;
    DevBuf& bp;
    DevBuf& xp;
    const void* b;
    void* x;
    int ncol;
    static DevBuf& pool(Context* c, size_t idx, size_t bytes) {
        while (c->work.size() <= idx) c->work.emplace_back(new DevBuf());
        DevBuf& d = *c->work[idx];
        if (d.bytes < bytes) d.alloc(c, bytes);
        return d;
    }
    PermutedIO(Context* cc, Matrix& mm, const void* bb, void* xx, int nc)
        : c(cc), m(mm), bp(pool(cc, 40, mm.vsize * (size_t)mm.n * nc)), xp(pool(cc, 41, mm.vsize * (size_t)mm.n * nc)),
          b(bb), x(xx), ncol(nc) {
        const size_t stride = m.vsize * (size_t)m.n;
        for (int j = 0; j < nc; ++j) {
            gather(static_cast<const char*>(b) + stride * j, static_cast<char*>(bp.p) + stride * j, m.perm);
            gather(static_cast<const char*>(x) + stride * j, static_cast<char*>(xp.p) + stride * j, m.perm);
        }
    }
    void gather(const void* src, void* dst, const DevBuf& map) {
        if (m.dtype == SSGPU_R64)
            launch_gather<double, double>(c->stream, m.n, map.as<int>(), (const double*)src, (double*)dst);
        else
            launch_gather<zc, zc>(c->stream, m.n, map.as<int>(), (const zc*)src, (zc*)dst);
    }
    void finish() {   // x <- xp in the caller's order; waits for the copies
        const size_t stride = m.vsize * (size_t)m.n;
        for (int j = 0; j < ncol; ++j)
            gather(static_cast<const char*>(xp.p) + stride * j, static_cast<char*>(x) + stride * j, m.iperm);
        SSGPU_HIP(hipStreamSynchronize(c->stream));
    }
};

// ------------------------------------------------------------------ scratch
// Device memory of the allocation-free entry point: one array per GPU, held
// by the library's code object until the process ends.  Host code never uses
// its name as an address: the address is looked up once per device
// (hipGetSymbolAddress, with that device current) and kept in scratch_dev.
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

// this device's scratch address (the device must be current)
void* scratch_dev(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error{BAD_ARG, "device index beyond the scratch table"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error{HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch_dev[device] = p;
    }
    return g_scratch_dev[device];
}

// Lay out a solve's work space in the scratch array: the work vectors the
// solver, its options and the true-residual pass use, the two arrays the
// position-published sweeps poll, the reduction partials, the scalar block
// and the tickets.  Throws BAD_ARG when it does
// not fit.
void plan_workspace(Workspace& w, char* base, const Matrix& m, int solver, const ssgpu_options& o) {
    std::vector<int> idx;
    if (solver == SSGPU_ICCG) idx = {0, 1, 2, 3, 4};
    else if (solver == SSGPU_MRTR_IC) idx = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    else idx = {0, 1, 2, 3, 4, 5, 6, 10, 14};
    if (o.is_save_best) idx.push_back(9);
    if (solver == SSGPU_ICCG && o.is_save_best) idx.push_back(5);
    idx.push_back(12);   // true residual
    const bool refinable = o.factor_precision == SSGPU_FACTOR_F32 &&
                           (solver == SSGPU_ICCG || (solver == SSGPU_MRTR_IC && o.mrtr_direction_fix));
    if (refinable) idx.push_back(13);
    for (int k = 0; k < 16; ++k) w.slot_of[k] = -1;
    int ns = 0;
    for (int k : idx)
        if (w.slot_of[k] < 0) w.slot_of[k] = ns++;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t scal_b = up(kScalSlots * sizeof(double)), tick_b = 256;
    const size_t part_b = up(sizeof(double) * 8 * (size_t)std::max(m.nblk, kMaxReduceBlocks));
    w.vec_bytes = up(m.vsize * (size_t)std::max(m.n, 1));
    // plus the two arrays the position-published sweeps poll
    const size_t need = scal_b + tick_b + part_b + w.vec_bytes * (size_t)(ns + 2);
    if (need > kScratchBytes)
        throw Error{BAD_ARG, "the allocation-free solve needs " + std::to_string(need) + " bytes of work space for "
                             "this system (n = " + std::to_string(m.n) + ", " + std::to_string(ns) +
                             " vectors); the library's scratch holds " + std::to_string(kScratchBytes)};
    w.scal = reinterpret_cast<double*>(base);
    w.tickets = reinterpret_cast<int*>(base + scal_b);
    w.partial = reinterpret_cast<double*>(base + scal_b + tick_b);
    w.vecs = base + scal_b + tick_b + part_b;
    w.xpos = w.vecs + w.vec_bytes * (size_t)ns;
}

}  // namespace ssgpu

// ==================================================================== C-ABI
using namespace ssgpu;

extern "C" {

void ssgpu_options_default(ssgpu_options* o) {
    if (!o) return;
    o->is_init = 1;
    o->is_diag_scale = 0;
    o->is_save_best = 0;
    o->is_save_log = 0;
    o->diverge_judge_type = 0;
    o->bad_div_val = 1000.0;
    o->bad_div_count_thres = 1000;
    o->conv_normalize_type = 0;
    o->conv_normalize_const = 1.0;
    o->accel_factor = 1.05;
    o->factor_precision = SSGPU_FACTOR_F64;
    o->mrtr_direction_fix = 0;
}

const char* ssgpu_status_string(int st) {
    switch (st) {
        case SSGPU_OK: return "ok";
        case SSGPU_NOT_CONVERGED: return "not converged";
        case SSGPU_NO_DIAGONAL: return "missing diagonal entry";
        case SSGPU_BAD_ARG: return "bad argument";
        case SSGPU_OOM: return "out of memory";
        case SSGPU_HIP_ERROR: return "HIP runtime error";
        case SSGPU_NOT_SQUARE: return "matrix is not square";
        case SSGPU_UNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

const char* ssgpu_last_error(void) { return g_last_error.c_str(); }

int ssgpu_context_create(int device, ssgpu_context* out) {
    if (!out) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    int ndev = 0;
    SSGPU_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error{BAD_ARG, "device index out of range"};
    std::unique_ptr<ssgpu_context_s> h(new ssgpu_context_s());
    h->c.device = device;
    SSGPU_HIP(hipSetDevice(device));
    SSGPU_HIP(hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking));
    SSGPU_HIP(hipHostMalloc((void**)&h->c.h_scal, kScalSlots * sizeof(double), hipHostMallocDefault));
    h->c.owner = h.get();
    *out = h.release();
    return SSGPU_OK;
    SSGPU_GUARD_END
}

static std::mutex g_lifetime_mu;   // guards live_matrices / destroy_pending of every context

static void context_free(ssgpu_context ctx) {
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    ctx->c.d_scal.release();
    ctx->c.d_count.release();
    ctx->c.d_tickets.release();
    if (ctx->c.h_scal) (void)hipHostFree(ctx->c.h_scal);
    if (ctx->c.h_bscal) (void)hipHostFree(ctx->c.h_bscal);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

int ssgpu_context_destroy(ssgpu_context ctx) {
    if (!ctx) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    {
        std::lock_guard<std::mutex> lk(g_lifetime_mu);
        if (ctx->c.live_matrices > 0) {   // freed when its last matrix is destroyed
            ctx->c.destroy_pending = true;
            return SSGPU_OK;
        }
    }
    context_free(ctx);
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_context_synchronize(ssgpu_context ctx) {
    if (!ctx) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    SSGPU_HIP(hipStreamSynchronize(ctx->c.stream));
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_context_resident_bytes(ssgpu_context ctx, int64_t* bytes) {
    if (!ctx || !bytes) return SSGPU_BAD_ARG;
    *bytes = ctx->c.resident;
    return SSGPU_OK;
}

int ssgpu_context_trim(ssgpu_context ctx) {
    if (!ctx) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Context& c = ctx->c;
    std::lock_guard<std::mutex> lk(c.mu);
    SSGPU_HIP(hipSetDevice(c.device));
    SSGPU_HIP(hipStreamSynchronize(c.stream));   // nothing queued may still use the pool
    c.work.clear();
    c.partial.release();
    c.d_bpartial.release();
    c.d_btickets.release();
    c.d_bscal.release();
    c.d_xpos[0].release();
    c.d_xpos[1].release();
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_memory(int64_t* resident, int64_t* peak, int reset_peak) {
    if (resident) *resident = g_resident.load();
    if (peak) *peak = g_peak.load();
    if (reset_peak) g_peak.store(g_resident.load());
    return SSGPU_OK;
}

// the common tail of matrix creation (ctx's mutex held): the structural
// analysis, the SpMV tuner, the handle
static void finish_create(ssgpu_context ctx, std::unique_ptr<ssgpu_matrix_s>& h, ssgpu_matrix* out) {
    Matrix& m = h->m;
    analyse(m);
    if (m.dtype == SSGPU_R64) tune_spmv<double>(m, 1.0);
    else tune_spmv<zc>(m, zc{1.0, 0.0});
    SSGPU_HIP(hipStreamSynchronize(ctx->c.stream));
    {
        std::lock_guard<std::mutex> lk(g_lifetime_mu);
        ++ctx->c.live_matrices;
    }
    *out = h.release();
}

int ssgpu_matrix_create(ssgpu_context ctx, int dtype, int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                        const void* values, ssgpu_matrix* out) {
    if (!ctx || !out || !row_ptr || n <= 0 || (dtype != SSGPU_R64 && dtype != SSGPU_C128)) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    std::unique_ptr<ssgpu_matrix_s> h(new ssgpu_matrix_s());
    Matrix& m = h->m;
    m.ctx = &ctx->c;
    m.dtype = dtype;
    m.n = n;
    m.vsize = dtype == SSGPU_R64 ? 8 : 16;
    if (row_ptr[0] != 0) throw Error{BAD_ARG, "row_ptr[0] must be 0"};
    for (int i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) throw Error{BAD_ARG, "row_ptr not monotone"};
    m.nnz = row_ptr[n];
    if (m.nnz > 0 && (!col_idx || !values)) throw Error{BAD_ARG, "null col_idx/values"};
    m.h_rp.assign(row_ptr, row_ptr + n + 1);
    m.h_ci.assign(col_idx, col_idx + m.nnz);
    for (int i = 0; i < n; ++i)
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            if (col_idx[k] < 0 || col_idx[k] >= n) throw Error{NOT_SQUARE, "column index outside the square matrix"};
            if (k > row_ptr[i] && col_idx[k] <= col_idx[k - 1])
                throw Error{BAD_ARG, "column indices must be strictly ascending within a row"};
        }
    // the context's stream carries the upload and the SpMV tuner: one call at a time
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    SSGPU_HIP(hipSetDevice(ctx->c.device));
    upload(m.ctx, m.rp, m.h_rp);
    upload(m.ctx, m.ci, m.h_ci);
    m.val.alloc(m.ctx, m.vsize * (size_t)std::max<int64_t>(m.nnz, 1));
    if (m.nnz) SSGPU_HIP(hipMemcpyAsync(m.val.p, values, m.vsize * m.nnz, hipMemcpyHostToDevice, ctx->c.stream));
    finish_create(ctx, h, out);
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_matrix_create_triplets(ssgpu_context ctx, int dtype, int32_t n, int64_t ntrip, const int32_t* rows,
                                 const int32_t* cols, const void* values, int mode, int on_device,
                                 ssgpu_matrix* out) {
    if (!ctx || !out || n <= 0 || ntrip < 0 || (dtype != SSGPU_R64 && dtype != SSGPU_C128)) return SSGPU_BAD_ARG;
    if (mode < SSGPU_ASSEMBLE_COO || mode > SSGPU_ASSEMBLE_BUILDER_SQUARE) return SSGPU_BAD_ARG;
    if (ntrip > 0 && (!rows || !cols || !values)) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    std::unique_ptr<ssgpu_matrix_s> h(new ssgpu_matrix_s());
    Matrix& m = h->m;
    m.ctx = &ctx->c;
    m.dtype = dtype;
    m.n = n;
    m.vsize = dtype == SSGPU_R64 ? 8 : 16;
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    SSGPU_HIP(hipSetDevice(ctx->c.device));
    DevBuf dr, dc, dv;   // host triplets are staged in device memory first
    if (!on_device && ntrip > 0) {
        dr.alloc(m.ctx, sizeof(int32_t) * (size_t)ntrip);
        dc.alloc(m.ctx, sizeof(int32_t) * (size_t)ntrip);
        dv.alloc(m.ctx, m.vsize * (size_t)ntrip);
        SSGPU_HIP(hipMemcpy(dr.p, rows, sizeof(int32_t) * (size_t)ntrip, hipMemcpyHostToDevice));
        SSGPU_HIP(hipMemcpy(dc.p, cols, sizeof(int32_t) * (size_t)ntrip, hipMemcpyHostToDevice));
        SSGPU_HIP(hipMemcpy(dv.p, values, m.vsize * (size_t)ntrip, hipMemcpyHostToDevice));
        rows = dr.as<int32_t>();
        cols = dc.as<int32_t>();
        values = dv.p;
    }
    assemble_triplets(m, ntrip, rows, cols, values, mode);
    dr.release();
    dc.release();
    dv.release();
    finish_create(ctx, h, out);
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_matrix_get_csr(ssgpu_matrix mh, int32_t* row_ptr, int32_t* col_idx, void* values) {
    if (!mh) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    std::lock_guard<std::mutex> lk(m.ctx->mu);
    SSGPU_HIP(hipSetDevice(m.ctx->device));
    if (row_ptr) std::memcpy(row_ptr, m.h_rp.data(), sizeof(int32_t) * ((size_t)m.n + 1));
    if (col_idx && m.nnz) std::memcpy(col_idx, m.h_ci.data(), sizeof(int32_t) * (size_t)m.nnz);
    if (values && m.nnz) {
        SSGPU_HIP(hipMemcpyAsync(values, m.val.p, m.vsize * (size_t)m.nnz, hipMemcpyDeviceToHost, m.ctx->stream));
        SSGPU_HIP(hipStreamSynchronize(m.ctx->stream));
    }
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_abmc_order(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int block_size, int32_t* perm,
                     int32_t* ncolors) {
    if (n <= 0 || !row_ptr || !col_idx || !perm || block_size < 1) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    std::vector<int> p;
    const int nc = abmc_order(n, row_ptr, col_idx, block_size, p);
    std::copy(p.begin(), p.end(), perm);
    if (ncolors) *ncolors = nc;
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_matrix_create_ordered(ssgpu_context ctx, int dtype, int32_t n, const int32_t* row_ptr,
                                const int32_t* col_idx, const void* values, int block_size, ssgpu_matrix* out,
                                int32_t* ncolors) {
    if (!ctx || !out || !row_ptr || !col_idx || !values || n <= 0 || block_size < 1 ||
        (dtype != SSGPU_R64 && dtype != SSGPU_C128))
        return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    for (int i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) throw Error{BAD_ARG, "row_ptr not monotone"};
    const int64_t nnz = row_ptr[n];
    for (int64_t k = 0; k < nnz; ++k)
        if (col_idx[k] < 0 || col_idx[k] >= n) throw Error{NOT_SQUARE, "column index outside the square matrix"};
    std::vector<int> perm;
    const int nc = abmc_order(n, row_ptr, col_idx, block_size, perm);
    std::vector<int> iperm(n);
    for (int r = 0; r < n; ++r) iperm[perm[r]] = r;
    // P A P^T with ascending columns per row
    const size_t vs = dtype == SSGPU_R64 ? 8 : 16;
    std::vector<int> prp(n + 1, 0), pci((size_t)nnz);
    std::vector<char> pv(vs * (size_t)nnz);
    std::vector<int64_t> vmap((size_t)nnz);
    for (int r = 0; r < n; ++r) prp[r + 1] = prp[r] + (row_ptr[perm[r] + 1] - row_ptr[perm[r]]);
    parallel_for(n, [&](int r) {
        thread_local std::vector<std::pair<int, int64_t>> row;
        const int i = perm[r];
        row.clear();
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) row.emplace_back(iperm[col_idx[k]], k);
        std::sort(row.begin(), row.end());
        int q = prp[r];
        for (const auto& e : row) {
            pci[q] = e.first;
            vmap[q] = e.second;
            std::memcpy(&pv[vs * q], static_cast<const char*>(values) + vs * e.second, vs);
            ++q;
        }
    });
    ssgpu_matrix h = nullptr;
    const int st = ssgpu_matrix_create(ctx, dtype, n, prp.data(), pci.data(), pv.data(), &h);
    if (st != SSGPU_OK) return st;
    Matrix& m = h->m;
    m.ordered = true;
    m.ncolors = nc;
    m.vmap = std::move(vmap);
    upload(m.ctx, m.perm, perm);
    upload(m.ctx, m.iperm, iperm);
    if (ncolors) *ncolors = nc;
    *out = h;
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_matrix_update_values(ssgpu_matrix mh, const void* values) {
    if (!mh || (!values && mh->m.nnz)) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    std::lock_guard<std::mutex> lk(m.ctx->mu);
    SSGPU_HIP(hipSetDevice(m.ctx->device));
    std::vector<char> pv;
    if (m.ordered) {   // caller's entry order -> the stored P A P^T order
        pv.resize(m.vsize * (size_t)m.nnz);
        for (int64_t q = 0; q < m.nnz; ++q)
            std::memcpy(&pv[m.vsize * q], static_cast<const char*>(values) + m.vsize * m.vmap[q], m.vsize);
        values = pv.data();
    }
    if (m.nnz) SSGPU_HIP(hipMemcpyAsync(m.val.p, values, m.vsize * m.nnz, hipMemcpyHostToDevice, m.ctx->stream));
    SSGPU_HIP(hipStreamSynchronize(m.ctx->stream));
    // the cached preconditioner (IC factor, or S A S for SGS) belongs to the
    // old values: cached solves refuse until ssgpu_factorize runs again
    m.factor_valid = false;
    m.prep_kind = 0;
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_matrix_destroy(ssgpu_matrix mh) {
    if (!mh) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Context* c = mh->m.ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete mh;
    bool free_ctx = false;
    {
        std::lock_guard<std::mutex> lk(g_lifetime_mu);
        free_ctx = --c->live_matrices == 0 && c->destroy_pending;
    }
    if (free_ctx) context_free(static_cast<ssgpu_context>(c->owner));
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_matrix_info(ssgpu_matrix mh, int32_t* n, int64_t* nnz, int* dtype, int32_t* fwd_levels,
                      int32_t* bwd_levels) {
    if (!mh) return SSGPU_BAD_ARG;
    if (n) *n = mh->m.n;
    if (nnz) *nnz = mh->m.nnz;
    if (dtype) *dtype = mh->m.dtype;
    if (fwd_levels) *fwd_levels = mh->m.L.nlev;
    if (bwd_levels) *bwd_levels = mh->m.Uic.nlev;
    return SSGPU_OK;
}

int ssgpu_matrix_sweep_info(ssgpu_matrix mh, int* mode, int* grid, int32_t* tiles) {
    if (!mh) return SSGPU_BAD_ARG;
    const Matrix& m = mh->m;
    const int k = m.sfL.use_tile ? 2 : m.sfL.by_level ? 1 : 0;
    if (mode) *mode = k;
    if (grid) *grid = k == 2 ? m.tile_grid : m.sweep_grid;
    if (tiles) *tiles = m.sfL.tile ? m.sfL.tile->ntiles : 0;
    return SSGPU_OK;
}

static int check_opts(const ssgpu_options* o, int solver) {
    if (!o) return SSGPU_BAD_ARG;
    if (solver < SSGPU_ICCG || solver > SSGPU_MRTR_SGS) return SSGPU_BAD_ARG;
    if (o->factor_precision != SSGPU_FACTOR_F64 && o->factor_precision != SSGPU_FACTOR_F32) return SSGPU_BAD_ARG;
    return SSGPU_OK;
}

int ssgpu_solve_device(ssgpu_matrix mh, int solver, const ssgpu_options* opt, double conv_cri, int max_ite,
                       double normB, const void* d_b, void* d_x, double* log, int log_cap, ssgpu_result* res) {
    if (!mh || !d_b || !d_x) return SSGPU_BAD_ARG;
    const int ck = check_opts(opt, solver);
    if (ck != SSGPU_OK) return ck;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    std::lock_guard<std::mutex> lk(m.ctx->mu);
    SSGPU_HIP(hipSetDevice(m.ctx->device));
    std::unique_ptr<PermutedIO> pio;
    if (m.ordered) {
        pio.reset(new PermutedIO(m.ctx, m, d_b, d_x, 1));
        d_b = pio->bp.p;
        d_x = pio->xp.p;
    }
    const int st = m.dtype == SSGPU_R64
                       ? solve_device<double>(m, solver, *opt, conv_cri, max_ite, normB, (const double*)d_b,
                                              (double*)d_x, log, log_cap, res)
                       : solve_device<zc>(m, solver, *opt, conv_cri, max_ite, normB, (const zc*)d_b, (zc*)d_x, log,
                                          log_cap, res);
    if (pio) pio->finish();
    return st;
    SSGPU_GUARD_END
}

int ssgpu_factorize(ssgpu_matrix mh, int solver, double accel, int factor_precision, double* accel_used,
                    int* pivot_fixes) {
    if (!mh || solver < SSGPU_ICCG || solver > SSGPU_MRTR_SGS) return SSGPU_BAD_ARG;
    if (factor_precision != SSGPU_FACTOR_F64 && factor_precision != SSGPU_FACTOR_F32) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    std::lock_guard<std::mutex> lk(m.ctx->mu);
    SSGPU_HIP(hipSetDevice(m.ctx->device));
    if (m.dtype == SSGPU_R64) factorize<double>(m, solver, accel, factor_precision, accel_used, pivot_fixes);
    else factorize<zc>(m, solver, accel, factor_precision, accel_used, pivot_fixes);
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_solve_factored(ssgpu_context ctx, ssgpu_matrix mh, int solver, const ssgpu_options* opt, double conv_cri,
                         int max_ite, double normB, const void* d_b, void* d_x, double* log, int log_cap,
                         ssgpu_result* res) {
    if (!ctx || !mh || !d_b || !d_x) return SSGPU_BAD_ARG;
    const int ck = check_opts(opt, solver);
    if (ck != SSGPU_OK) return ck;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = &ctx->c;
    if (c->device != m.ctx->device) throw Error{BAD_ARG, "context and matrix are on different devices"};
    std::lock_guard<std::mutex> lk(c->mu);
    SSGPU_HIP(hipSetDevice(c->device));
    std::unique_ptr<PermutedIO> pio;
    if (m.ordered) {
        pio.reset(new PermutedIO(c, m, d_b, d_x, 1));
        d_b = pio->bp.p;
        d_x = pio->xp.p;
    }
    const int st = m.dtype == SSGPU_R64
                       ? solve_factored<double>(c, m, solver, *opt, conv_cri, max_ite, normB, (const double*)d_b,
                                                (double*)d_x, log, log_cap, res)
                       : solve_factored<zc>(c, m, solver, *opt, conv_cri, max_ite, normB, (const zc*)d_b, (zc*)d_x,
                                            log, log_cap, res);
    if (pio) pio->finish();
    return st;
    SSGPU_GUARD_END
}

int ssgpu_solve_factored_noalloc(ssgpu_context ctx, ssgpu_matrix mh, int solver, const ssgpu_options* opt,
                                 double conv_cri, int max_ite, double normB, const void* d_b, void* d_x, double* log,
                                 int log_cap, ssgpu_result* res) {
    if (!ctx || !mh || !d_b || !d_x) return SSGPU_BAD_ARG;
    const int ck = check_opts(opt, solver);
    if (ck != SSGPU_OK) return ck;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = &ctx->c;
    if (c->device != m.ctx->device) throw Error{BAD_ARG, "context and matrix are on different devices"};
    if (m.ordered) throw Error{UNSUPPORTED, "the allocation-free solve takes matrices in their own order only"};
    std::lock_guard<std::mutex> lk(c->mu);
    SSGPU_HIP(hipSetDevice(c->device));
    Workspace w;
    plan_workspace(w, static_cast<char*>(scratch_dev(c->device)), m, solver, *opt);
    return m.dtype == SSGPU_R64
               ? solve_factored<double>(c, m, solver, *opt, conv_cri, max_ite, normB, (const double*)d_b,
                                        (double*)d_x, log, log_cap, res, &w)
               : solve_factored<zc>(c, m, solver, *opt, conv_cri, max_ite, normB, (const zc*)d_b, (zc*)d_x, log,
                                    log_cap, res, &w);
    SSGPU_GUARD_END
}

int ssgpu_solve_block(ssgpu_context ctx, ssgpu_matrix mh, int solver, const ssgpu_options* opt, double conv_cri,
                      int max_ite, int nrhs, const double* normB, const void* d_B, void* d_X, ssgpu_result* res) {
    if (!ctx || !mh || !opt || !normB || !d_B || !d_X || !res || nrhs < 1 || nrhs > 32) return SSGPU_BAD_ARG;
    const int ck = check_opts(opt, solver);
    if (ck != SSGPU_OK) return ck;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = &ctx->c;
    if (c->device != m.ctx->device) throw Error{BAD_ARG, "context and matrix are on different devices"};
    std::lock_guard<std::mutex> lk(c->mu);
    SSGPU_HIP(hipSetDevice(c->device));
    std::unique_ptr<PermutedIO> pio;
    if (m.ordered) {
        pio.reset(new PermutedIO(c, m, d_B, d_X, nrhs));
        d_B = pio->bp.p;
        d_X = pio->xp.p;
    }
    const int st = m.dtype == SSGPU_R64
                       ? solve_block<double>(c, m, solver, *opt, conv_cri, max_ite, nrhs, normB,
                                             (const double*)d_B, (double*)d_X, res)
                       : solve_block<zc>(c, m, solver, *opt, conv_cri, max_ite, nrhs, normB, (const zc*)d_B,
                                         (zc*)d_X, res);
    if (pio) pio->finish();
    return st;
    SSGPU_GUARD_END
}

int ssgpu_solve(ssgpu_matrix mh, int solver, const ssgpu_options* opt, double conv_cri, int max_ite, double normB,
                const void* b, void* x, double* log, int log_cap, ssgpu_result* res) {
    if (!mh || !b || !x) return SSGPU_BAD_ARG;
    const int ck = check_opts(opt, solver);
    if (ck != SSGPU_OK) return ck;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = m.ctx;
    SSGPU_HIP(hipSetDevice(c->device));
    DevBuf db, dx;
    db.alloc(c, m.vsize * m.n);
    dx.alloc(c, m.vsize * m.n);
    SSGPU_HIP(hipMemcpyAsync(db.p, b, m.vsize * m.n, hipMemcpyHostToDevice, c->stream));
    SSGPU_HIP(hipMemcpyAsync(dx.p, x, m.vsize * m.n, hipMemcpyHostToDevice, c->stream));
    const int st = ssgpu_solve_device(mh, solver, opt, conv_cri, max_ite, normB, db.p, dx.p, log, log_cap, res);
    if (st != SSGPU_OK && st != SSGPU_NOT_CONVERGED) return st;
    SSGPU_HIP(hipMemcpyAsync(x, dx.p, m.vsize * m.n, hipMemcpyDeviceToHost, c->stream));
    SSGPU_HIP(hipStreamSynchronize(c->stream));
    return st;
    SSGPU_GUARD_END
}

int ssgpu_spmv_device(ssgpu_matrix mh, const void* d_x, void* d_y) {
    if (!mh || !d_x || !d_y) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    SSGPU_HIP(hipSetDevice(m.ctx->device));
    if (m.dtype == SSGPU_R64)
        launch_spmv<double>(m.ctx->stream, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<double>(),
                            (const double*)d_x, (double*)d_y, nullptr, nullptr, m.spmv_stage_kb);
    else
        launch_spmv<zc>(m.ctx->stream, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<zc>(),
                        (const zc*)d_x, (zc*)d_y, nullptr, nullptr, m.spmv_stage_kb);
    SSGPU_HIP(hipGetLastError());
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_ic_factor_host(ssgpu_matrix mh, double accel, int32_t* L_row_ptr, int32_t* L_col_idx, void* L_values,
                         void* diagD, int* pivot_fixes) {
    if (!mh) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = m.ctx;
    SSGPU_HIP(hipSetDevice(c->device));
    if (!m.lower_diag_ok) throw Error{NO_DIAGONAL, "there is not diagonal element!"};
    ssgpu_options o;
    ssgpu_options_default(&o);
    Solver S(m, o);
    int fixes;
    if (m.dtype == SSGPU_R64) fixes = S.factor<double>(m.val.as<double>(), accel);
    else fixes = S.factor<zc>(m.val.as<zc>(), accel);
    if (pivot_fixes) *pivot_fixes = fixes;
    // copy back in lower(A) (original row) order
    const int n = m.n;
    std::vector<int> prp(n + 1), prow(n);
    std::vector<char> pv(m.vsize * m.L.nnz);
    SSGPU_HIP(hipMemcpy(prp.data(), m.L.prp.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    SSGPU_HIP(hipMemcpy(prow.data(), m.L.prow.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int> pci(m.L.nnz);
    SSGPU_HIP(hipMemcpy(pci.data(), m.L.pci.p, sizeof(int) * m.L.nnz, hipMemcpyDeviceToHost));
    SSGPU_HIP(hipMemcpy(pv.data(), m.Lv.p, m.vsize * m.L.nnz, hipMemcpyDeviceToHost));
    std::vector<int> pos(n);
    for (int p = 0; p < n; ++p) pos[prow[p]] = p;
    if (L_row_ptr) {
        L_row_ptr[0] = 0;
        for (int i = 0; i < n; ++i) L_row_ptr[i + 1] = L_row_ptr[i] + (prp[pos[i] + 1] - prp[pos[i]]);
        for (int i = 0; i < n; ++i) {
            const int a = prp[pos[i]], e = prp[pos[i] + 1];
            for (int k = a; k < e; ++k) {
                const int d = L_row_ptr[i] + (k - a);
                if (L_col_idx) L_col_idx[d] = pci[k];
                if (L_values) std::memcpy((char*)L_values + m.vsize * d, pv.data() + m.vsize * k, m.vsize);
            }
        }
    }
    if (diagD) SSGPU_HIP(hipMemcpy(diagD, m.diagD.p, m.vsize * n, hipMemcpyDeviceToHost));
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_ic_apply_device(ssgpu_matrix mh, const void* d_r, void* d_z) {
    if (!mh || !d_r || !d_z) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    if (!m.factor_valid) throw Error{BAD_ARG, "no factor: call ssgpu_ic_factor_host first"};
    SSGPU_HIP(hipSetDevice(m.ctx->device));
    ssgpu_options o;
    ssgpu_options_default(&o);
    Solver S(m, o);
    if (m.dtype == SSGPU_R64) S.ic_apply<double>((const double*)d_r, (double*)d_z, S.vec<double>(11));
    else S.ic_apply<zc>((const zc*)d_r, (zc*)d_z, S.vec<zc>(11));
    S.check_sweep_error();
    SSGPU_HIP(hipGetLastError());
    return SSGPU_OK;
    SSGPU_GUARD_END
}

static int plain_sweep(ssgpu_matrix mh, const void* d_r, void* d_z, bool lower) {
    if (!mh || !d_r || !d_z) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = m.ctx;
    SSGPU_HIP(hipSetDevice(c->device));
    ensure_upper(m);
    if (lower ? !m.lower_diag_ok : !m.upper_diag_ok) throw Error{NO_DIAGONAL, "missing diagonal entry"};
    ssgpu_options o;
    ssgpu_options_default(&o);
    Solver S(m, o);
    if (m.dtype == SSGPU_R64) {
        S.sgs_values<double>(m.val.as<double>());
        if (lower) S.fr<double>((const double*)d_r, (double*)d_z);
        else S.bc<double>((const double*)d_r, (double*)d_z);
    } else {
        S.sgs_values<zc>(m.val.as<zc>());
        if (lower) S.fr<zc>((const zc*)d_r, (zc*)d_z);
        else S.bc<zc>((const zc*)d_r, (zc*)d_z);
    }
    S.check_sweep_error();
    SSGPU_HIP(hipGetLastError());
    return SSGPU_OK;
    SSGPU_GUARD_END
}
int ssgpu_lower_sweep_device(ssgpu_matrix mh, const void* d_r, void* d_z) { return plain_sweep(mh, d_r, d_z, true); }
int ssgpu_upper_sweep_device(ssgpu_matrix mh, const void* d_r, void* d_z) { return plain_sweep(mh, d_r, d_z, false); }

int ssgpu_dot_device(ssgpu_context ctx, int dtype, int32_t n, const void* d_a, const void* d_b, double* out) {
    if (!ctx || !d_a || !d_b || !out || n < 0) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Context* c = &ctx->c;
    SSGPU_HIP(hipSetDevice(c->device));
    DevBuf part;
    part.alloc(c, sizeof(double) * 2 * kMaxReduceBlocks);
    (void)scal_buf(c);
    const int K = dtype == SSGPU_R64 ? 1 : 2;
    if (dtype == SSGPU_R64) launch_dot<double>(c->stream, n, (const double*)d_a, (const double*)d_b, part.as<double>());
    else launch_dot<zc>(c->stream, n, (const zc*)d_a, (const zc*)d_b, part.as<double>());
    launch_finalize(c->stream, reduce_blocks(n), K, part.as<double>(), c->d_scal.as<double>());
    SSGPU_HIP(hipMemcpyAsync(out, c->d_scal.p, sizeof(double) * K, hipMemcpyDeviceToHost, c->stream));
    SSGPU_HIP(hipStreamSynchronize(c->stream));
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_norm_device(ssgpu_context ctx, int dtype, int32_t n, const void* d_a, double* out) {
    if (!ctx || !d_a || !out || n < 0) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Context* c = &ctx->c;
    SSGPU_HIP(hipSetDevice(c->device));
    DevBuf part;
    part.alloc(c, sizeof(double) * kMaxReduceBlocks);
    (void)scal_buf(c);
    if (dtype == SSGPU_R64) launch_norm2<double>(c->stream, n, (const double*)d_a, part.as<double>());
    else launch_norm2<zc>(c->stream, n, (const zc*)d_a, part.as<double>());
    launch_finalize(c->stream, reduce_blocks(n), 1, part.as<double>(), c->d_scal.as<double>());
    double v = 0.0;
    SSGPU_HIP(hipMemcpyAsync(&v, c->d_scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SSGPU_HIP(hipStreamSynchronize(c->stream));
    *out = std::sqrt(v);
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_time_kernel(ssgpu_matrix mh, int which, int reps, double accel, int factor_precision, const void* d_in,
                      void* d_out, double* avg_ms) {
    if (!mh || !avg_ms || reps <= 0 || which < 0 || which > 4) return SSGPU_BAD_ARG;
    if (which <= 3 && (!d_in || !d_out)) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Matrix& m = mh->m;
    Context* c = m.ctx;
    SSGPU_HIP(hipSetDevice(c->device));
    ssgpu_options o;
    ssgpu_options_default(&o);
    o.factor_precision = factor_precision;
    Solver S(m, o);
    if (which >= 1 && which <= 3 && !m.factor_valid) throw Error{BAD_ARG, "no factor"};
    hipEvent_t e0, e1;
    SSGPU_HIP(hipEventCreate(&e0));
    SSGPU_HIP(hipEventCreate(&e1));
    auto run = [&]() {
        if (m.dtype == SSGPU_R64) {
            double* tmp = S.vec<double>(11);
            if (which == 0) S.spmv<double>(m.val.as<double>(), (const double*)d_in, (double*)d_out);
            else if (which == 3) S.ic_apply<double>((const double*)d_in, (double*)d_out, tmp);
            else if (which == 4) S.factor<double>(m.val.as<double>(), accel);
            else S.ic_sweep_one<double>(which == 1, (const double*)d_in, (double*)d_out, tmp);
        } else {
            zc* tmp = S.vec<zc>(11);
            if (which == 0) S.spmv<zc>(m.val.as<zc>(), (const zc*)d_in, (zc*)d_out);
            else if (which == 3) S.ic_apply<zc>((const zc*)d_in, (zc*)d_out, tmp);
            else if (which == 4) S.factor<zc>(m.val.as<zc>(), accel);
            else S.ic_sweep_one<zc>(which == 1, (const zc*)d_in, (zc*)d_out, tmp);
        }
    };
    run();   // warm
    SSGPU_HIP(hipStreamSynchronize(c->stream));
    SSGPU_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) run();
    SSGPU_HIP(hipEventRecord(e1, c->stream));
    SSGPU_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SSGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    S.check_sweep_error();
    return SSGPU_OK;
    SSGPU_GUARD_END
}

}  // extern "C"
