// Host-side state of the gfx950 solver library shared by its translation
// units: device buffers, contexts (device + stream + work space), device
// matrices with their structural analysis, and the C-ABI handle types.
#pragma once

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "sparsesolv_amd.h"
#include "ssgpu_kernels.h"

namespace ssgpu {

// ------------------------------------------------------------------ memory
struct Context;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    Context* ctx = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release();
    void alloc(Context* c, size_t b);
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::atomic<int64_t> resident{0};   // bytes of the DevBufs allocated on it (any thread may free one)
    double* h_scal = nullptr;   // pinned scalar read-back (kScalSlots doubles)
    DevBuf d_scal;              // device scalar results; the last slot holds the sweep error flag
    DevBuf d_count;
    DevBuf d_tickets;           // sync-free sweep tickets (2 ints)
    DevBuf d_xpos[2];           // the arrays the position-published sweeps poll (n values each)
    std::mutex mu;
    // matrices created on this context and not yet destroyed; a context whose
    // destroy arrives while some are alive is freed with the last of them
    int live_matrices = 0;
    bool destroy_pending = false;
    void* owner = nullptr;   // the C-ABI handle holding this context
    // Krylov work vectors and reduction partials of the solves run on this
    // stream (grown on demand): solves on different contexts never share them
    std::vector<std::unique_ptr<DevBuf>> work;
    DevBuf partial;
    // block solves: per-column scalars (device + pinned staging), ticket counters
    DevBuf d_bscal, d_btickets, d_bpartial;
    void* h_bscal = nullptr;
};



// Device scalars of a context: reduction results in the first slots, and in
// the last slot the flag a sync-free sweep sets when it exceeds its spin
// limit.  Every scalar read-back copies the whole block, so a solve sees the
// flag at its next convergence check and stops there.  The largest read-back
// is a block MRTR's three complex dots for 32 columns (192 values).
constexpr int kScalSlots = 256;
constexpr int kErrInt = 2 * (kScalSlots - 1);   // the flag, as an int index into the block
DevBuf& scal_buf(Context* c);
// copy the scalar block to the pinned host copy and synchronize; a set sweep
// error flag is cleared and reported
const double* read_scalars(Context* c);

// Host arrays of the analysis that are written in full after allocation:
// elements are default-initialized (not zeroed), so a large array costs no
// single-threaded fill before the (parallel) loop that writes it.
template <typename T> struct NoInit {
    using value_type = T;
    NoInit() = default;
    template <typename U> NoInit(const NoInit<U>&) {}
    T* allocate(size_t n) { return std::allocator<T>().allocate(n); }
    void deallocate(T* p, size_t n) { std::allocator<T>().deallocate(p, n); }
    template <typename U> void construct(U* p) noexcept { ::new ((void*)p) U; }
    template <typename U, typename... A> void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
    template <typename U> bool operator==(const NoInit<U>&) const { return true; }
    template <typename U> bool operator!=(const NoInit<U>&) const { return false; }
};
template <typename T> using hvec = std::vector<T, NoInit<T>>;

template <typename T, typename A> void upload(Context* c, DevBuf& d, const std::vector<T, A>& h) {
    d.alloc(c, sizeof(T) * std::max<size_t>(h.size(), 1));
    // synchronous: h may be a temporary of the analysis
    if (!h.empty()) SSGPU_HIP(hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

// Host loops of the analysis whose iterations write disjoint outputs: split
// [0, n) into contiguous ranges over a few threads (results do not depend on
// the split).
template <typename F> void parallel_for(int n, F&& f) {
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min({hw > 0 ? hw : 1, 16, n / 65536}));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            const int a = (int)((int64_t)n * t / nt), b = (int)((int64_t)n * (t + 1) / nt);
            for (int i = a; i < b; ++i) f(i);
        });
    for (auto& x : th) x.join();
}

// ------------------------------------------------------------------ structure
// Rows of a triangular structure stored in level order: level l is the row
// range levptr[l] .. levptr[l+1] of the permuted CSR (prp, pci), whose row p
// is original row prow[p]; entries keep the original ascending column order.
struct LevelCsr {
    int nlev = 0;
    int64_t nnz = 0;
    std::vector<int> levptr;
    DevBuf prp, pci, prow, map;   // map: source index of each entry
};

// Sliced-ELL copy of a triangular structure for the sync-free sweeps: rows in
// level order (position p holds row srow[p]), 64-position slices padded to the
// slice's longest row, entry k of position (64*s + l) at slice_off[s] + 64*k + l.
// Off-diagonal entries only, in summation order; the diagonal is kept per
// position.  Two value sources: the level-ordered factor (f) and the matrix
// values (a).
// Tiles of a sweep (k_sweep_tile): device arrays of TileArgs and the largest
// tile's shared memory for each value-type pair (FP64/FP64, FP64/FP32,
// complex/complex, complex/complex64).
struct TileCsr {
    int ntiles = 0;
    size_t lds[4] = {0, 0, 0, 0};
    DevBuf tile_rp, tile_sp, step_rp, tile_xp, ext_rows, tpos, tent, tslot, teidx;
    TileArgs args() const {
        return TileArgs{ntiles, tile_rp.as<int>(), tile_sp.as<int>(), step_rp.as<int>(), tile_xp.as<int>(),
                        ext_rows.as<int>(), tpos.as<int>(), tent.as<int>(), tslot.as<unsigned short>(),
                        teidx.as<int>()};
    }
};

struct SfCsr {
    std::unique_ptr<TileCsr> tile;   // tiles of this sweep (built for long level chains), or null
    bool use_tile = false;           // sweeps through the tiles (chosen by the tuner)
    int kmax = 0;          // longest row; 0 when a row exceeds 64 entries (level path used)
    int64_t nell = 0;
    const int* srow = nullptr;   // level position -> row (the LevelCsr's prow)
    const LevelCsr* lv = nullptr;   // its levels (host level starts)
    bool by_level = false;       // one launch per level instead of the sync-free sweep
    DevBuf slice_off, rlen, ecol, emap_f, dmap_f, emap_a, dmap_a;
    DevBuf ecolp;   // ecol as level-order positions (sync-free structures: the position-published sweep)
    DevBuf eval, dval, eval32, dval32;
    std::vector<int> h_soff;     // slice offsets (host copy, for the tile layout's entry indices)
};

struct Matrix {
    Context* ctx = nullptr;
    int dtype = SSGPU_R64;
    int n = 0;
    int64_t nnz = 0;
    size_t vsize = 8;   // bytes per value
    std::vector<int> h_rp, h_ci;
    DevBuf rp, ci, val, blk, diag;
    int nblk = 0;
    bool lower_diag_ok = true, upper_diag_ok = true;
    // IC structure: lower(A) in forward-level order; transpose(lower) in backward-level order
    LevelCsr L, Uic, Us;
    DevBuf ppos;   // original row -> position in L
    SfCsr sfL, sfUic, sfUs;   // sync-free sweep copies of lower(A), transpose(lower), upper(A)
    // IC(0) update lists (built on the first factorization)
    int lmax_row = 0;          // longest row of lower(A)
    bool icsym_built = false;
    int64_t icsym_total = 0;
    int ic_max_updates = 0;    // largest update list of a row
    DevBuf ic_po, ic_slot, ic_pos, ic_col;
    int sweep_grid = 128;     // persistent one-wave workgroups per sweep
    int tile_grid = 256;      // persistent one-wave workgroups of a tiled sweep
    bool sweeps_tuned = false;   // sweep launch shape measured (first factorization)
    int spmv_stage_kb = 16;      // SpMV staging (16 or 32 KB), measured at creation
    int sweep_idle = 2;       // longest back-off step of a waiting wave (2: 128 clocks ... 5: 1024)
    int block_grid = 0;       // block sweeps: polling waves in all (0: 128 per column, at most 2048)
    bool uic_built = false, us_built = false;
    // factor / preconditioner values and work space (grown on demand)
    DevBuf Lv, Uv, Lv32, Uv32, diagD, scaledA, sc, bs;
    int factor_pivot_fixes = 0;
    double factor_accel = 0.0;
    bool factor_valid = false;
    // what the sweep values currently hold: 0 nothing, 1 an IC factor
    // (prep_precision: FP64 only, or FP64 + FP32 copies), 2 SGS values of S A S
    int prep_kind = 0;
    int prep_precision = 0;
    // ABMC-reordered matrix: the library holds P A P^T; solves gather b and x
    // through perm (new -> old) and return x through iperm (old -> new)
    bool ordered = false;
    int ncolors = 0;
    DevBuf perm, iperm;
    std::vector<int64_t> vmap;   // stored entry q of P A P^T = caller's entry vmap[q]
};

// CSR of m.n rows assembled on the device from ntrip device triplets in the
// order given (ssgpu_assemble.hip): fills m.rp, m.ci, m.val, m.h_rp, m.h_ci and
// m.nnz.  mode 0: duplicates summed from the first value; 1: from zero; 2: as
// 1 with the to_square padding entry.
void assemble_triplets(Matrix& m, int64_t ntrip, const int* rows, const int* cols, const void* vals, int mode);

// true residual (and FP32-factor refinement) of a finished batched solve
template <typename T>
double batch_finish(Context* c, Matrix& m, int solver, const ssgpu_options& o, bool conv, double conv_cri,
                    int max_ite, double normB, const T* b, T* x, int& refinements, int& refine_its);

}  // namespace ssgpu

struct ssgpu_context_s {
    ssgpu::Context c;
};
struct ssgpu_matrix_s {
    ssgpu::Matrix m;
};

#define SSGPU_GUARD_BEGIN try {
#define SSGPU_GUARD_END                                  \
    }                                                    \
    catch (const ::ssgpu::Error& e) {                    \
        ::ssgpu::set_last_error(e.msg);                  \
        return e.code;                                   \
    }                                                    \
    catch (const std::bad_alloc&) {                      \
        ::ssgpu::set_last_error("host allocation failed"); \
        return SSGPU_OOM;                                \
    }                                                    \
    catch (const std::exception& e) {                    \
        ::ssgpu::set_last_error(e.what());               \
        return SSGPU_HIP_ERROR;                          \
    }
