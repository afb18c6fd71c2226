// Device kernels of the gfx950 solver library (declarations of the launchers).
//
// All kernels are deterministic: every reduction uses a fixed grid and a
// fixed tree, so the same inputs give the same bits on every run.  Per-row
// arithmetic (SpMV, sweeps, IC factorization) is sequential within the row in
// the reference's order, so those results are bit-identical to the
// reference's CPU build.
#pragma once

#include "ssgpu_common.h"
#include "ssgpu_tiles.h"

#include <tuple>
#include <utility>

namespace ssgpu {

constexpr int kBS = 256;          // threads per block of the streaming kernels
constexpr int kSpmvCap = 4096;    // products staged in LDS per SpMV block
constexpr int kSpmvRows = 128;    // rows per SpMV block (two threads per row)
constexpr int kMaxReduceBlocks = 1024;

// --- SpMV --------------------------------------------------------------------
// y = A x over the row blocks blk[0..nblk] (<= 256 rows, <= kSpmvCap entries
// unless the block is one long row).  With dotv != nullptr, also writes the
// per-block partial of sum(dotv[i] * y[i]) into partial (1 or 2 doubles per
// block, non-conjugated).
// stage_kb: LDS staging of the products, 16 or 32 KB (same results)
template <typename T>
void launch_spmv(hipStream_t s, int nblk, const int* blk, const int* rp, const int* ci, const T* v,
                 const T* x, T* y, const T* dotv, double* partial, int stage_kb = 16);

// --- level-scheduled sweeps ----------------------------------------------------
enum SweepMode { FWD_DIV_LAST = 0, BWD_IC = 1, BWD_DIV_FIRST = 2 };
// Rows [p0, p1) of a permuted CSR (prp/pci/pv/prow).
//  FWD_DIV_LAST : x[i] = (rhs[i] - sum_{k<last} v_k x[c_k]) / v_last
//  BWD_IC       : x[i] = aux[i] + dd[i] * (0 - sum_{k>first} v_k x[c_k])
//  BWD_DIV_FIRST: x[i] = (rhs[i] - sum_{k>first} v_k x[c_k]) / v_first
template <typename T, typename F>
void launch_sweep(hipStream_t s, int mode, int p0, int p1, const int* prp, const int* pci, const F* pv,
                  const int* prow, const T* rhs, const T* aux, const T* dd, T* x);

// --- shifted IC(0) on one level -------------------------------------------------
template <typename T>
void launch_ic_level(hipStream_t s, int p0, int p1, const int* prp, const int* pci, T* pv, const int* prow,
                     const int* ppos, T* diagD, double accel, int* nfix);

// --- shifted IC(0) with a precomputed update list (ssgpu_icfact.hip) -----------------
// per factor entry, the number of updates (cnt) / the updates themselves
// (slot of L_ik in its row, position of L_jk, column k) in the reference's order
void launch_ic_symbolic_count(hipStream_t s, int n, const int* prp, const int* pci, const int* prow, const int* ppos,
                              int* cnt);
void launch_ic_symbolic_fill(hipStream_t s, int n, const int* prp, const int* pci, const int* prow, const int* ppos,
                             const int* po, void* u_slot, int* u_pos, int* u_col);
size_t scan_tmp_bytes(int nitems);
void launch_exclusive_scan(hipStream_t s, const int* in, int* out, int nitems, void* tmp, size_t tmp_bytes);
// one level; kmax_row >= longest row of lower(A) (<= 48)
template <typename T>
void launch_ic_numeric(hipStream_t s, int kmax_row, int p0, int p1, const int* prp, const int* pci, const int* prow,
                       const int* po, const void* u_slot, const int* u_pos, const int* u_col, T* fv, T* diagD,
                       double accel, int* nfix);

// wave-per-row variant (rows of at most 64 * 4 updates): same factor, one
// gather round trip per row
template <typename T>
void launch_ic_numeric_wave(hipStream_t s, int kmax_row, int p0, int p1, const int* prp, const int* pci,
                            const int* prow, const int* po, const void* u_slot, const int* u_pos, const int* u_col,
                            T* fv, T* diagD, double accel, int* nfix);
// largest update count of a row (synchronizes; scratch: one device int)
int ic_row_updates_max(hipStream_t s, int n, const int* prp, const int* po, int* scratch);

// --- reductions -----------------------------------------------------------------
// Grid size used for n elements (fixed function of n).
int reduce_blocks(int64_t n);
// out[k] = sum over blocks of partial[b*K + k], fixed order.
void launch_finalize(hipStream_t s, int nb, int K, const double* partial, double* out);

template <typename T> void launch_dot(hipStream_t s, int n, const T* a, const T* b, double* partial);
template <typename T> void launch_norm2(hipStream_t s, int n, const T* a, double* partial);
// three non-conjugated dots a.b, a.c, a.d (d may be null: then 2 dots)
template <typename T>
void launch_dot3(hipStream_t s, int n, const T* a, const T* b, const T* c, const T* d, double* partial);

// --- vector kernels ---------------------------------------------------------------
template <typename T> void launch_fill(hipStream_t s, int64_t n, T* x, T val);
template <typename T> void launch_copy(hipStream_t s, int64_t n, const T* a, T* b);
// ICCG: x += alpha p; r -= alpha q; partial |r|^2
template <typename T>
void launch_cg_update(hipStream_t s, int n, T alpha, const T* p, const T* q, T* x, T* r, double* partial);
// ICCG: p = p*beta + z
template <typename T> void launch_cg_pupdate(hipStream_t s, int n, T beta, const T* z, T* p);
// r = b - q (or r = b if q null)
template <typename T> void launch_residual(hipStream_t s, int n, const T* b, const T* q, T* r);
// y = -1*x
template <typename T> void launch_negate(hipStream_t s, int n, const T* x, T* y);
// MRTR update group 1: p = u + t*u (fix=0) or p = u + t*p (fix=1); x += zeta p; y = y*eta + zeta v;
// r -= y; partial |r|^2
template <typename T>
void launch_mrtr_update(hipStream_t s, int n, int fix, T t, T zeta, T eta, const T* u, const T* v, T* p, T* x,
                        T* y, T* r, double* partial);
// MRTR update group 2: z = z*eta + zeta w; u -= z
template <typename T>
void launch_mrtr_update2(hipStream_t s, int n, T zeta, T eta, const T* w, T* z, T* u);
// SGS: tmp = rd - u
template <typename T> void launch_sub(hipStream_t s, int n, const T* a, const T* b, T* c);
// SGS: ard += u
template <typename T> void launch_add_inplace(hipStream_t s, int n, const T* u, T* a);
// SGS update: p = p*t2 + u; x += zeta p; y = y*eta + zeta ard; rd -= y; partial |rd|^2
template <typename T>
void launch_sgs_update(hipStream_t s, int n, T t2, T zeta, T eta, const T* u, const T* ard, T* p, T* x, T* y,
                       T* rd, double* partial);

// --- scaling and gathers ---------------------------------------------------------
// s_i = 1/sqrt|a_ii| (1/DBL_EPSILON below 1e-12), bs_i = b_i * s_i
template <typename T>
void launch_diag_scale(hipStream_t s, int n, const int* diag_idx, const T* v, const T* b, T* sc, T* bs);
// dst[k] = v[k] * (s_row * s_col) over all entries
template <typename T>
void launch_scale_matrix(hipStream_t s, int n, const int* rp, const int* ci, const T* v, const T* sc, T* dst);
// dst[k] = src[map[k]]
template <typename T, typename F>
void launch_gather(hipStream_t s, int64_t n, const int* map, const T* src, F* dst);
// x = x / s  and x = x * s (s real-valued even for complex)
// dst[k] = narrow(src[k]) (FP64 -> FP32 factor copies)
template <typename T, typename F> void launch_narrow(hipStream_t s, int64_t n, const T* src, F* dst);
template <typename T> void launch_div_scale(hipStream_t s, int n, const T* sc, T* x);
template <typename T> void launch_mul_scale(hipStream_t s, int n, const T* sc, T* x);
// count of diagonal entries failing the auto-shift test (real > 0 / Re > 0)
template <typename T> void launch_count_nonpos(hipStream_t s, int n, const T* d, int* count);

// --- sync-free sweeps (ssgpu_sweep.hip) -------------------------------------------
constexpr int kSfRows = 64;    // rows per chunk = threads per workgroup (one wave)
// not-ready value of a sync-free sweep's outputs: a signalling-NaN pattern that
// arithmetic never produces (it only yields quiet NaNs)
constexpr unsigned long long kSweepSentinel = 0xFFF40000DEADBEEFull;
// polling passes a sync-free sweep's wave may make without progress before it
// sets the error flag and leaves (every wave exits, so the grid always drains)
constexpr long kSweepSpinLimit = 1L << 23;
// fill a and b (either may be null) with the not-ready sentinel; zero two tickets
template <typename T> void launch_sf_prepare(hipStream_t s, int n, T* a, T* b, int* tickets);
// one sweep over all rows in level order (srow: level position -> row); kmax >=
// longest row (off-diagonals), <= 64; grid = persistent one-wave workgroups
template <typename T, typename F>
void launch_sweep_sf(hipStream_t s, int mode, int kmax, int n, int grid, int* ticket, int* err, long spin_limit,
                     const int* srow,
                     const int* slice_off, const int* rlen, const int* ecol, const F* eval, const F* dval,
                     const T* rhs, const T* aux, const T* dd, int max_idle, T* x);
// Tiled sweep: the rows of a band of consecutive levels, cut into intervals of
// row index, form a tile (layout: ssgpu_tiles.h); a one-wave workgroup takes
// tiles in ticket order, waits for the tile's external inputs, stages its
// entries in LDS and walks the tile's levels ("steps") with every in-tile
// input handed over through LDS.  Packed rows are tile by tile, step by step.
struct TileArgs {
    int ntiles;
    const int* tile_rp;    // ntiles + 1: packed rows of each tile
    const int* tile_sp;    // ntiles + 1: steps of each tile in step_rp
    const int* step_rp;    // first packed row of each step
    const int* tile_xp;    // ntiles + 1: external inputs of each tile in ext_rows
    const int* ext_rows;   // rows a tile reads from earlier tiles
    const int* tpos;       // packed row -> its level-order position
    const int* tent;       // packed row -> first packed entry (packed rows + 1 entries)
    const unsigned short* tslot;   // packed entry -> LDS slot of its input (in-tile rows, then externals)
    const int* teidx;      // packed entry -> its index in the sweep's sliced-ELL value arrays
};
// one tiled sweep; lds_bytes = the largest tile's shared-memory footprint
template <typename T, typename F>
void launch_sweep_tile(hipStream_t s, int mode, int kmax, int grid, size_t lds_bytes, const TileArgs& ta,
                       int* ticket, int* err, long spin_limit, const int* srow, const int* slice_off,
                       const int* rlen, const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd,
                       int max_idle, T* x);
// the same sweep one level per launch (levptr: host array of nlev + 1 level
// starts in the level order); for orderings with few, wide levels
template <typename T, typename F>
void launch_sweep_levels(hipStream_t s, int mode, int kmax, int nlev, const int* levptr, const int* srow,
                         const int* slice_off, const int* rlen, const int* ecol, const F* eval, const F* dval,
                         const T* rhs, const T* aux, const T* dd, T* x);
// the same sweep publishing results at level positions: xp (n values,
// sentinel-filled) is polled, ecolp gives the inputs' positions, and xo
// receives the results by row
template <typename T, typename F>
void launch_sweep_sf_pos(hipStream_t s, int mode, int kmax, int n, int grid, int* ticket, int* err, long spin_limit,
                         const int* srow, const int* slice_off, const int* rlen, const int* ecolp, const F* eval,
                         const F* dval, const T* rhs, const T* aux, const T* dd, int max_idle, T* xp, T* xo);
// block sweep over kb right-hand sides with entry (i, c) at i * kb + c (kb a
// power of two <= 32; rhs, aux and x in that layout); grid: waves in all;
// ticket: one zeroed counter
template <typename T, typename F>
void launch_sweep_il(hipStream_t s, int mode, int kmax, int n, int grid, int kb, int* ticket, int* err,
                     long spin_limit, const int* srow, const int* slice_off, const int* rlen, const int* ecol,
                     const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd, T* x);
// --- block solves (ssgpu_block.hip): kb columns side by side, entry (i, c) at i * kb + c ----
// (kb a power of two <= 32); every column computed as the one-column kernel computes it
template <typename T>
void launch_spmv_il(hipStream_t s, int kb, int nblk, const int* blk, const int* rp, const int* ci, const T* v,
                    const T* x, T* y, const T* dotv, double* partial);
// partials: kb * parts per block ((block * kb + c) * parts), finalized by launch_finalize
template <typename T> void launch_dot_il(hipStream_t s, int kb, int n, const T* a, const T* b, double* partial);
template <typename T> void launch_norm2_il(hipStream_t s, int kb, int n, const T* a, double* partial);
template <typename T>
void launch_cg_update_il(hipStream_t s, int kb, int n, const T* alpha, const int* act, const T* p, const T* q,
                         T* x, T* r, double* partial);
template <typename T>
void launch_cg_pupdate_il(hipStream_t s, int kb, int n, const T* beta, const int* act, const T* z, T* p);
// sc: t[kb], zeta[kb], eta[kb]; sgs: the SGS-MRTR update, else IC-MRTR's
template <typename T>
void launch_mrtr_update_il(hipStream_t s, int kb, int n, bool sgs, int fix, const T* sc, const int* act, const T* u,
                           const T* v, T* p, T* x, T* y, T* r, double* partial);
template <typename T>
void launch_mrtr_update2_il(hipStream_t s, int kb, int n, const T* sc, const int* act, const T* w, T* z, T* u);
// x(i, c) *= s_i (div: /= s_i)
template <typename T> void launch_scale_il(hipStream_t s, int kb, int n, const T* sc, T* x, bool div);
template <typename T> void launch_copy_col_il(hipStream_t s, int kb, int n, int c, const T* src, T* dst);
// nr vectors one after another <-> kb interleaved columns (columns >= nr zero)
template <typename T> void launch_to_il(hipStream_t s, int kb, int n, int nr, const T* src, T* dst);
template <typename T> void launch_from_il(hipStream_t s, int kb, int n, int nr, const T* src, T* dst);
// --- batched solves (many independent systems, one launch per phase) ----------------
// Each system keeps its own vectors and scalar state; a launch covers every
// system of the batch whose status equals the launch's `want` (the others are
// no-ops).  Every per-system sum runs over the same blocks and the same tree
// as the one-system kernels above, so a system's iterates are bit-identical to
// its solo solve.
// per-system vectors: ICCG uses the first eight; IC-MRTR also U, Y, W and its
// internal iterate XI (its x is copied to X at the end); Z holds MRTR's z and Q
// its v = A u.  SGS-MRTR: B the scaled b, Z the caller's b (for the true
// residual), XI the scaled iterate, R its rd, Q its A rd, T the sweeps' input
enum BatchVec { BV_B = 0, BV_X, BV_P, BV_R, BV_Z, BV_Q, BV_T, BV_BEST, BV_U, BV_Y, BV_W, BV_XI, kBatchVecs };
constexpr int kBatchVecsIccg = 8;
enum BatchStatus {
    BS_ACTIVE = 0,      // iterating
    BS_CONVERGED = 1,   // met the tolerance in the loop
    BS_STOPPED = 2,     // diverged or reached max_ite
    BS_EARLY = 3,       // returned from the prologue (b = 0 or the initial residual already small)
    BS_PENDING = 4,     // added, prologue not run
    BS_PRIMED = 5,      // prologue passed, first preconditioner apply pending
    BS_IDLE = 6         // slot holds no system
};
struct BatchSys {               // device-resident descriptor of one system
    int n, nblk, rblk;          // rows; SpMV row blocks; reduction blocks (reduce_blocks(n))
    const int *blk, *rp, *ci;
    const void* val;
    const int *f_srow, *f_soff, *f_rlen, *f_ecol;   // forward sweep: sliced-ELL lower(A) with factor values
    const void *f_eval, *f_dval;
    const int *b_srow, *b_soff, *b_rlen, *b_ecol;   // backward sweep: the factor's transpose (SGS: upper(S A S))
    const void* b_eval;
    const void* b_dval;         // SGS: the upper triangle's diagonal (divide-first backward sweep)
    const void* diagD;
    void* vec[kBatchVecs];
    double* part;               // this system's partial sums (max(2 nblk, 6 rblk) doubles)
};
struct BatchState {             // Krylov scalars and bookkeeping of one system (device)
    double rz[2], alpha[2], beta[2];   // ICCG: (r, z), alpha, beta; complex as (re, im)
    double zeta[2], zeta_old[2], eta[2], nu[2], t[2];   // MRTR recurrences
    double normB, normalizer, abs_conv, best, final_res;
    double conv_cri, bad_div_val, normalize_const;
    int max_ite, iterations, bad, status, copy_best;
    int normalize_type, save_best, diverge_type, bad_div_thres;
    int mrtr_fix;               // MRTR direction: 0 the reference's p = u + t u, 1 p = u + t p
};
// vector phases of a batched ICCG iteration and prologue
enum BatchOp {
    BO_FILL0 = 0,     // vec[a] = 0
    BO_RESID,         // vec[c] = vec[a] - vec[b]   (b < 0: vec[c] = vec[a])
    BO_NORM2,         // partial |vec[a]|^2
    BO_DOT,           // partial vec[a] . vec[b] (non-conjugated)
    BO_COPY,          // vec[b] = vec[a]
    BO_SENT,          // vec[a], vec[b] = the sweeps' not-ready sentinel
    BO_CG_UPDATE,     // x += alpha p; r -= alpha q; partial |r|^2; t, z = sentinel
    BO_PUPDATE,       // best = x when asked; p = p*beta + z
    BO_NEG,           // vec[b] = -1 * vec[a]
    BO_DOT3,          // partials w.r, w.v, w.y (w.y only after the first iteration), w = vec[W]
    BO_MRTR_UPDATE,   // p = u + t u (or u + t p); xi += zeta p; y = y eta + zeta v; r -= y; partial |r|^2
    BO_MRTR_UPDATE2,  // best = xi when asked; z = z eta + zeta w; u -= z; t, w = sentinel
    BO_SGS_SUB,       // best = xi when asked; q = sentinel; a != 0: t = r - u (Eisenstat's splitting)
    BO_SGS_DOT3,      // a != 0: q += u; partials q.r, q.q, q.y (q.y only after the first iteration)
    BO_SGS_UPDATE     // p = p t + u; xi += zeta p; y = y eta + zeta q; r -= y; partial |r|^2; u = sentinel
};
// scalar phases (one workgroup per system: the fixed-tree finalize of its partials, then)
enum BatchPhase {
    BP_PROLOGUE = 0,  // ||r0||: the reference's prologue; PENDING -> PRIMED or EARLY
    BP_RZ0,           // (r, z); PRIMED -> ACTIVE
    BP_ALPHA,         // alpha = rz / (p, Ap)  (partials of the SpMV blocks)
    BP_MONITOR,       // ||r||: convergence, best iterate, divergence, max_ite
    BP_BETA,          // beta = rz' / rz; rz = rz'
    BP_ACTIVATE,      // (no sums) PRIMED -> ACTIVE, sweep tickets zeroed
    BP_MRTR_COEF,     // zeta, eta, nu, t from w.r, w.v, w.y (MatSolverMRTR_ICver.cpp:279-300; the same
                      // recurrence as MatSolverMRTR_SGSver.cpp:196-216)
    BP_NORMB          // normB = ||vec||, sweep tickets zeroed (SGS-MRTR: the scaled b's norm)
};
// A graph built node by node, without stream capture (so no stream is ever
// in capture mode, whatever other threads do meanwhile): a batched launcher
// given a recorder appends a kernel node depending on the previous node
// instead of launching.
struct GraphRec {
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
};
// launch kernel(args...) on s, or append it to rec when rec is given
template <typename... KArgs, size_t... I>
inline void launch_or_record_impl(GraphRec* rec, void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t s,
                                  std::tuple<KArgs...>& vals, std::index_sequence<I...>) {
    void* params[] = {static_cast<void*>(&std::get<I>(vals))..., nullptr};
    if (!rec) {
        (void)hipLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, params, 0, s);
        return;
    }
    if (rec->err != hipSuccess) return;
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(kernel);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = params;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    rec->err = hipGraphAddKernelNode(&node, rec->graph, rec->last ? &rec->last : nullptr, rec->last ? 1 : 0, &p);
    if (rec->err == hipSuccess) rec->last = node;
}
template <typename... KArgs, typename... Args>
inline void launch_or_record(GraphRec* rec, void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t s,
                             Args... args) {
    std::tuple<KArgs...> vals(args...);
    launch_or_record_impl(rec, kernel, grid, block, s, vals, std::index_sequence_for<KArgs...>{});
}

// items: (system, block) pairs; fin: systems, one workgroup each; rec: append
// to a graph instead of launching
template <typename T>
void launch_batch_spmv(hipStream_t s, int nitems, const int2* items, const BatchSys* sys, const BatchState* st,
                       int want, int in, int out, int with_dot, GraphRec* rec = nullptr);
// tickets (may be null): zeroed first by the launch's first workgroup
template <typename T>
void launch_batch_vec(hipStream_t s, int op, int nitems, const int2* items, const BatchSys* sys, BatchState* st,
                      int want, int a, int b, int c, int* tickets = nullptr, GraphRec* rec = nullptr);
// active: count of systems still iterating (decremented as they finish);
// tickets: the sweep tickets, zeroed by the PROLOGUE, ACTIVATE and MONITOR phases
template <typename T>
void launch_batch_finalize(hipStream_t s, int phase, int nfin, const int* fin, const BatchSys* sys, BatchState* st,
                           int want, int* active, int* tickets, GraphRec* rec = nullptr);
// persistent sync-free sweep over the chunks of every system in `order`
// (system, chunk) ticket order; mode FWD_DIV_LAST (vec[in] -> vec[out]) or
// BWD_IC (aux vec[in] -> vec[out]); F the factor's storage type
template <typename T, typename F>
void launch_batch_sweep(hipStream_t s, int mode, int kmax, int grid, int norder, const int2* order, int* ticket,
                        int* err, long spin_limit, const BatchSys* sys, const BatchState* st, int want, int in,
                        int out, int max_idle, GraphRec* rec = nullptr);
// dst[e] = map[e] >= 0 ? src[map[e]] : 0 (narrowing to F)
template <typename T, typename F>
void launch_gather_pad(hipStream_t s, int64_t n, const int* map, const T* src, F* dst);

}  // namespace ssgpu
