// Sync-free triangular sweeps for gfx950.
//
// One launch per sweep.  Rows are taken in level order (a row's level is one
// more than the highest level among the rows it reads) and cut into chunks of
// 64 rows, one lane per row.  A persistent grid of one-wave workgroups takes
// chunk tickets from an atomic counter in that order.  Every input of a row
// lies at a lower level, hence in a chunk with a smaller ticket (or in the same
// chunk, handled by the same wave), and a ticket is only taken by a running
// workgroup: the sweep cannot deadlock however workgroups are dispatched, and
// the chunks in flight are the ones at the level front.
//
// A lane keeps its row's off-diagonal entries in registers (loaded from a
// sliced-ELL copy in level order, entry k of slot l of chunk c at
// slice_off[c] + 64*k + l: each load instruction is one coalesced 512-byte
// access).  Readiness needs no flags: the output vector is pre-filled with a
// NaN sentinel (arithmetic only produces quiet NaNs, never this bit pattern),
// results are published with agent-scope stores and read back with
// agent-scope loads.  The row's result is then formed in the reference's
// sequential summation order.  A spin limit bounds every wait: a wave that
// exceeds it sets the error flag and leaves, so the grid always drains.
//
// The cost of a level is the latency of publishing a value and seeing it
// (most of it the agent-scope load round trip) plus the length of one polling
// pass, not bytes: a pass reloads only the inputs a lane still misses, and the
// register bucket (row length rounded up) keeps the pass short.
#include <algorithm>
#include <type_traits>

#include "ssgpu_kernels.h"

namespace ssgpu {

namespace {

constexpr unsigned long long kSentinel = kSweepSentinel;
constexpr int kLvlWaves = 4;   // waves (64-row chunks) per workgroup of the per-level sweep

__device__ inline bool sent(double v) { return (unsigned long long)__double_as_longlong(v) == kSentinel; }
__device__ inline bool sent(zc v) { return sent(v.re) || sent(v.im); }
__device__ inline double sent_val(double*) { return __longlong_as_double((long long)kSentinel); }
__device__ inline zc sent_val(zc*) {
    const double s = __longlong_as_double((long long)kSentinel);
    return zc{s, s};
}

__device__ inline double gload(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load((unsigned long long*)p, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline zc gload(const zc* p) {
    const double* d = reinterpret_cast<const double*>(p);
    return zc{gload(d), gload(d + 1)};
}
__device__ inline void gstore(double* p, double v) {
    __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void gstore(zc* p, zc v) {
    double* d = reinterpret_cast<double*>(p);
    gstore(d, v.re);
    gstore(d + 1, v.im);
}
template <typename T, typename F> __device__ inline T wid(F a);
template <> __device__ inline double wid<double, double>(double a) { return a; }
template <> __device__ inline double wid<double, float>(float a) { return (double)a; }
template <> __device__ inline zc wid<zc, zc>(zc a) { return a; }
template <> __device__ inline zc wid<zc, fc>(fc a) { return zc{(double)a.re, (double)a.im}; }

// Row state: the lane's row of the chunk, its ELL entries in registers.
template <typename T, typename F, int MODE, int KMAX> struct RowState {
    int i = 0, len = 0;
    unsigned cols[KMAX];   // unsigned: widening to an address offset needs no instruction on the loaded value
    F vals[KMAX];
    T r_i, a_i, d_i;
    F dg;
    __device__ void load(int p, bool active, int c, int lane, const int* __restrict__ srow,
                         const int* __restrict__ slice_off,
                         const int* __restrict__ rlen, const int* __restrict__ ecol, const F* __restrict__ eval,
                         const F* __restrict__ dval, const T* __restrict__ rhs, const T* __restrict__ aux,
                         const T* __restrict__ dd) {
        r_i = a_i = d_i = Num<T>::zero();
        dg = F();
        i = active ? srow[p] : 0;
        len = active ? rlen[p] : 0;
        const int base = active ? slice_off[c] + lane : 0;
        // Every lane issues all KMAX loads back to back, with no branch and
        // no wait between them: slots past the row's length (and inactive
        // lanes) read entry 0, which exists in every sliced-ELL copy (its
        // arrays hold at least one element).  A per-slot branch here let the
        // compiler finish each load before issuing the next: KMAX serial
        // memory round trips per chunk.
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int at = k < len ? base + kSfRows * k : 0;
            cols[k] = (unsigned)ecol[at];
            vals[k] = eval[at];
        }
        if (active) {
            if (MODE == BWD_IC) {
                a_i = aux[i];
                d_i = dd[i];
            } else {
                r_i = rhs[i];
                dg = dval[p];
            }
        }
    }
    // the row's result from its inputs g, in the reference's summation order
    __device__ T result(const T (&g)[KMAX]) const {
        T s = (MODE == BWD_IC) ? Num<T>::zero() : r_i;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (k < len) s = Num<T>::sub(s, Num<T>::mul(wid<T, F>(vals[k]), g[k]));
        if (MODE == BWD_IC) return Num<T>::add(Num<T>::mul(s, d_i), a_i);
        return Num<T>::div(s, wid<T, F>(dg));
    }
};

__device__ inline int take_ticket(int* ticket) {
    __shared__ int s_chunk;
    if (threadIdx.x == 0) s_chunk = atomicAdd(ticket, 1);
    __syncthreads();
    const int c = s_chunk;
    __syncthreads();
    return c;
}

// Chunk c of a sync-free sweep, one lane per row; false when the spin limit
// stopped the wave (the error flag is set).  spins counts the wave's polling
// passes since its last progress.
template <typename T, typename F, int MODE, int KMAX>
__device__ inline bool sf_chunk(int c, int n, int* __restrict__ err, long spin_limit, const int* __restrict__ srow,
                                const int* __restrict__ slice_off, const int* __restrict__ rlen,
                                const int* __restrict__ ecol, const F* __restrict__ eval, const F* __restrict__ dval,
                                const T* __restrict__ rhs, const T* __restrict__ aux, const T* __restrict__ dd,
                                int max_idle, T* __restrict__ x, long& spins, T* __restrict__ xo = nullptr) {
    const int lane = threadIdx.x;
    const int p = c * kSfRows + lane;   // position in level order
    const bool active = p < n;
    RowState<T, F, MODE, KMAX> row;
    row.load(p, active, c, lane, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd);
    T g[KMAX];
    using Mask = typename std::conditional<(KMAX <= 32), unsigned, unsigned long long>::type;
    constexpr Mask kOne = 1;
    Mask miss = active ? (row.len >= (int)(8 * sizeof(Mask)) ? ~Mask(0) : ((kOne << row.len) - kOne)) : Mask(0);
    bool done = !active;
    int idle = 0;
    while (true) {
        __asm__ __volatile__("" ::: "memory");
        bool progress = false;
        if (!done && miss) {
            const Mask want = miss;
#pragma unroll
            for (int k = 0; k < KMAX; ++k)
                if (want & (kOne << k)) g[k] = gload(x + row.cols[k]);
#pragma unroll
            for (int k = 0; k < KMAX; ++k)
                if ((want & (kOne << k)) && !sent(g[k])) miss &= ~(kOne << k);
            progress = miss != want;
        }
        if (!done && miss == 0ull) {
            const T res = row.result(g);
            if (xo) {   // published at the level position, and stored by row
                gstore(x + p, res);
                xo[row.i] = res;
            } else {
                gstore(x + row.i, res);
            }
            done = true;
            progress = true;
        }
        if (__all(done)) return true;
        // the limit bounds a stall, not the sweep: a pass with progress
        // restarts the count (a limit of 0 gives up at the first pass
        // that does not finish the chunk)
        const bool any_progress = __any(progress);
        if (any_progress && spin_limit > 0) spins = 0;
        if (++spins > spin_limit) {
            atomicOr(err, 1);
            return false;
        }
        if (any_progress) {
            idle = 0;
        } else {
            // back off while nothing arrives (waves far ahead of the level
            // front go quiet): 64, 128, then up to 64 * 2^(max_idle-2) clocks
            if (idle < max_idle) ++idle;
            if (idle <= 1) __builtin_amdgcn_s_sleep(1);
            else if (idle == 2) __builtin_amdgcn_s_sleep(2);
            else if (idle == 3) __builtin_amdgcn_s_sleep(4);
            else if (idle == 4) __builtin_amdgcn_s_sleep(8);
            else __builtin_amdgcn_s_sleep(16);
        }
    }
}

template <typename T, typename F, int MODE, int KMAX>
__global__ __launch_bounds__(kSfRows) void k_sweep_sf(int n, int nchunks, int* __restrict__ ticket,
                                                      int* __restrict__ err, long spin_limit,
                                                      const int* __restrict__ srow,
                                                      const int* __restrict__ slice_off, const int* __restrict__ rlen,
                                                      const int* __restrict__ ecol, const F* __restrict__ eval,
                                                      const F* __restrict__ dval, const T* __restrict__ rhs,
                                                      const T* __restrict__ aux, const T* __restrict__ dd,
                                                      int max_idle, T* __restrict__ x) {
    long spins = 0;
    while (true) {
        const int c = take_ticket(ticket);
        if (c >= nchunks) return;
        if (!sf_chunk<T, F, MODE, KMAX>(c, n, err, spin_limit, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd,
                                        max_idle, x, spins))
            return;
    }
}

// The same sweep publishing each result at its level-order position in xp
// (the array the waves poll, inputs given by their positions in ecol) and
// storing it by row in xo.  A chunk's rows then read, per input direction,
// adjacent positions of the previous levels, so a wave's polls touch a few
// cache lines instead of one per lane; the arithmetic is unchanged.
template <typename T, typename F, int MODE, int KMAX>
__global__ __launch_bounds__(kSfRows) void k_sweep_sf_pos(int n, int nchunks, int* __restrict__ ticket,
                                                          int* __restrict__ err, long spin_limit,
                                                          const int* __restrict__ srow,
                                                          const int* __restrict__ slice_off, const int* __restrict__ rlen,
                                                          const int* __restrict__ ecol, const F* __restrict__ eval,
                                                          const F* __restrict__ dval, const T* __restrict__ rhs,
                                                          const T* __restrict__ aux, const T* __restrict__ dd,
                                                          int max_idle, T* __restrict__ xp, T* __restrict__ xo) {
    long spins = 0;
    while (true) {
        const int c = take_ticket(ticket);
        if (c >= nchunks) return;
        if (!sf_chunk<T, F, MODE, KMAX>(c, n, err, spin_limit, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd,
                                        max_idle, xp, spins, xo))
            return;
    }
}

// The sweeps of many systems in one persistent launch (batched solves).
// order[t] = (system, chunk) is the ticket order: each system's chunks in its
// own level order, systems interleaved by the level of a chunk's first row.
// A chunk's inputs lie in chunks of the same system earlier in its level
// order, hence at smaller tickets, so the launch is deadlock-free like the
// one-system sweep; the interleave keeps the waves in flight on the level
// fronts of all systems at once.  Chunks of systems whose status is not
// `want` are skipped.
template <typename T, typename F, int MODE, int KMAX>
__global__ __launch_bounds__(kSfRows) void k_batch_sweep(int norder, const int2* __restrict__ order,
                                                         int* __restrict__ ticket, int* __restrict__ err,
                                                         long spin_limit, const BatchSys* __restrict__ sys,
                                                         const BatchState* __restrict__ st, int want, int in, int out,
                                                         int max_idle) {
    long spins = 0;
    while (true) {
        const int t = take_ticket(ticket);
        if (t >= norder) return;
        const int2 sc = order[t];
        if (st[sc.x].status != want) continue;
        const BatchSys& S = sys[sc.x];
        bool ok;
        if (MODE == FWD_DIV_LAST)
            ok = sf_chunk<T, F, MODE, KMAX>(sc.y, S.n, err, spin_limit, S.f_srow, S.f_soff, S.f_rlen, S.f_ecol,
                                            static_cast<const F*>(S.f_eval), static_cast<const F*>(S.f_dval),
                                            static_cast<const T*>(S.vec[in]), nullptr, nullptr, max_idle,
                                            static_cast<T*>(S.vec[out]), spins);
        else if (MODE == BWD_DIV_FIRST)
            ok = sf_chunk<T, F, MODE, KMAX>(sc.y, S.n, err, spin_limit, S.b_srow, S.b_soff, S.b_rlen, S.b_ecol,
                                            static_cast<const F*>(S.b_eval), static_cast<const F*>(S.b_dval),
                                            static_cast<const T*>(S.vec[in]), nullptr, nullptr, max_idle,
                                            static_cast<T*>(S.vec[out]), spins);
        else
            ok = sf_chunk<T, F, MODE, KMAX>(sc.y, S.n, err, spin_limit, S.b_srow, S.b_soff, S.b_rlen, S.b_ecol,
                                            static_cast<const F*>(S.b_eval), nullptr, nullptr,
                                            static_cast<const T*>(S.vec[in]), static_cast<const T*>(S.diagD),
                                            max_idle, static_cast<T*>(S.vec[out]), spins);
        if (!ok) return;
    }
}

// Tiled sweep.  A tile is the rows of a band of consecutive levels within an
// interval of row index; a row's inputs lie at lower levels and, inside its
// band, on the near side of its interval, so they come from the same tile or
// from tiles with smaller tickets (deadlock-free like the chunked sweep).
// One wave per tile:
//   1. it loads the tile's external inputs (agent-scope polls, repeated until
//      none is the sentinel) and meanwhile stages the tile's rows (row,
//      right-hand side, diagonal) and entries (value, LDS slot of the input)
//      in LDS;
//   2. it walks the tile's levels: a lane per row, inputs read from LDS, the
//      result written to LDS for the rows after it and published with an
//      agent-scope store for later tiles.
// Inside a tile a level costs an LDS round trip instead of a global one.
// Each row is formed by the same code, in the same order, as in the other
// sweeps (RowState::result).
template <typename T, typename F, int MODE, int KMAX>
__global__ __launch_bounds__(kTileThreads) void k_sweep_tile(TileArgs ta, int* __restrict__ ticket,
                                                             int* __restrict__ err, long spin_limit,
                                                             const int* __restrict__ srow,
                                                             const int* __restrict__ slice_off,
                                                             const int* __restrict__ rlen,
                                                             const F* __restrict__ eval, const F* __restrict__ dval,
                                                             const T* __restrict__ rhs, const T* __restrict__ aux,
                                                             const T* __restrict__ dd, int max_idle,
                                                             T* __restrict__ x) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    constexpr int NT = kTileThreads;
    constexpr int KX = kTileExtPerLane;
    long spins = 0;
    while (true) {
        const int t = take_ticket(ticket);
        if (t >= ta.ntiles) return;
        const int r0 = ta.tile_rp[t], R = ta.tile_rp[t + 1] - r0;
        const int x0 = ta.tile_xp[t], X = ta.tile_xp[t + 1] - x0;
        const int e0 = ta.tent[r0], E = ta.tent[r0 + R] - e0;
        const int s0 = ta.tile_sp[t], ns = ta.tile_sp[t + 1] - s0;
        T* sx = reinterpret_cast<T*>(smem);   // R in-tile rows, then X externals
        T* sb = sx + (R + X);                 // right-hand side (aux for BWD_IC)
        T* sd = sb + R;                       // diagonal (dd for BWD_IC)
        F* sv = reinterpret_cast<F*>(sd + R);
        int* srw = reinterpret_cast<int*>(sv + E);
        int* sof = srw + R;
        unsigned short* ss = reinterpret_cast<unsigned short*>(sof + R + 1);
        // the external inputs' rows (a slot past X reads row 0, never stored)
        int xr[KX];
#pragma unroll
        for (int b = 0; b < KX; ++b) {
            const int k = lane + NT * b;
            xr[b] = k < X ? ta.ext_rows[x0 + k] : -1;
        }
        // first packed row of each step, one per lane
        const int step_q = lane < ns ? ta.step_rp[s0 + lane] - r0 : R;
        // Stage rows and entries in batches: every load of a batch goes out
        // before the first is waited for (a load per loop trip would cost a
        // memory round trip each).  Slots past the end read entry / row 0.
        constexpr int JR = 8;
        for (int q0 = 0; q0 < R; q0 += NT * JR) {
            int pp[JR], off[JR], ii[JR];
#pragma unroll
            for (int j = 0; j < JR; ++j) {
                const int q = q0 + lane + NT * j;
                const int qq = q < R ? q : 0;
                pp[j] = ta.tpos[r0 + qq];
                off[j] = ta.tent[r0 + qq] - e0;
            }
            F dg[JR];
#pragma unroll
            for (int j = 0; j < JR; ++j) {
                ii[j] = srow[pp[j]];
                dg[j] = MODE == BWD_IC ? F() : dval[pp[j]];
            }
            T bv[JR], dv[JR];
#pragma unroll
            for (int j = 0; j < JR; ++j) {
                bv[j] = MODE == BWD_IC ? aux[ii[j]] : rhs[ii[j]];
                dv[j] = MODE == BWD_IC ? dd[ii[j]] : wid<T, F>(dg[j]);
            }
#pragma unroll
            for (int j = 0; j < JR; ++j) {
                const int q = q0 + lane + NT * j;
                if (q < R) {
                    srw[q] = ii[j];
                    sof[q] = off[j];
                    sb[q] = bv[j];
                    sd[q] = dv[j];
                }
            }
        }
        constexpr int JE = 32;
        for (int b0 = 0; b0 < E; b0 += NT * JE) {
            int ei[JE];
            unsigned short sl[JE];
#pragma unroll
            for (int j = 0; j < JE; ++j) {
                const int e = b0 + lane + NT * j;
                const int ee = e < E ? e : 0;
                ei[j] = ta.teidx[e0 + ee];
                sl[j] = ta.tslot[e0 + ee];
            }
            F ev[JE];
#pragma unroll
            for (int j = 0; j < JE; ++j) ev[j] = eval[ei[j]];
#pragma unroll
            for (int j = 0; j < JE; ++j) {
                const int e = b0 + lane + NT * j;
                if (e < E) {
                    sv[e] = ev[j];
                    ss[e] = sl[j];
                }
            }
        }
        if (lane == 0) sof[R] = E;
        // poll the external inputs until every one is present
        unsigned miss = 0;
#pragma unroll
        for (int b = 0; b < KX; ++b)
            if (xr[b] >= 0) miss |= 1u << b;
        int idle = 0;
        bool ok = true;
        while (true) {
            __asm__ __volatile__("" ::: "memory");
            const unsigned want = miss;
            T v[KX];
#pragma unroll
            for (int b = 0; b < KX; ++b) v[b] = gload(x + (xr[b] >= 0 ? xr[b] : 0));
#pragma unroll
            for (int b = 0; b < KX; ++b)
                if (((want >> b) & 1u) && !sent(v[b])) {
                    sx[R + lane + NT * b] = v[b];
                    miss &= ~(1u << b);
                }
            if (__all(miss == 0u)) break;
            const bool any_progress = __any(miss != want);
            if (any_progress && spin_limit > 0) spins = 0;
            if (++spins > spin_limit) {
                atomicOr(err, 1);
                ok = false;
                break;
            }
            if (any_progress) {
                idle = 0;
            } else {
                if (idle < max_idle) ++idle;
                if (idle <= 1) __builtin_amdgcn_s_sleep(1);
                else if (idle == 2) __builtin_amdgcn_s_sleep(2);
                else if (idle == 3) __builtin_amdgcn_s_sleep(4);
                else if (idle == 4) __builtin_amdgcn_s_sleep(8);
                else __builtin_amdgcn_s_sleep(16);
            }
        }
        if (!ok) return;
        __syncthreads();   // staging and externals in LDS
        for (int st = 0; st < ns; ++st) {
            const int qa = __shfl(step_q, st);
            const int qb = st + 1 < ns ? __shfl(step_q, st + 1) : R;
            for (int q = qa + lane; q < qb; q += NT) {
                const int ka = sof[q], len = sof[q + 1] - ka;
                F vals[KMAX];
                unsigned short sl[KMAX];
#pragma unroll
                for (int k = 0; k < KMAX; ++k) {
                    sl[k] = k < len ? ss[ka + k] : (unsigned short)0;
                    vals[k] = k < len ? sv[ka + k] : F();
                }
                T g[KMAX];
#pragma unroll
                for (int k = 0; k < KMAX; ++k) g[k] = sx[sl[k]];
                T acc = (MODE == BWD_IC) ? Num<T>::zero() : sb[q];
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if (k < len) acc = Num<T>::sub(acc, Num<T>::mul(wid<T, F>(vals[k]), g[k]));
                const T res = (MODE == BWD_IC) ? Num<T>::add(Num<T>::mul(acc, sd[q]), sb[q]) : Num<T>::div(acc, sd[q]);
                sx[q] = res;
                gstore(x + srw[q], res);
            }
            // the next level reads this one's results from LDS: wait for the
            // LDS writes only (lgkmcnt 0; vmcnt and expcnt left at their
            // maxima), not for the global stores, which only later tiles read
            __builtin_amdgcn_s_waitcnt(0xC07F);
            __builtin_amdgcn_s_barrier();
        }
    }
}

// One level of a sweep, for orderings whose levels are wide (few levels of
// many rows each, as a multi-color order has): every input of a row of level
// l was written by an earlier launch on the same stream, so plain loads see
// it and no readiness test is needed.  Positions [p0, p1) of the level order;
// one workgroup per 64-position chunk, lanes outside the level idle.  The
// row's result is formed by the same code, in the same order, as in the
// sync-free sweep.
template <typename T, typename F, int MODE, int KMAX, int WPB>
__global__ __launch_bounds__(kSfRows * WPB) void k_sweep_lvl(int p0, int p1, const int* __restrict__ srow,
                                                       const int* __restrict__ slice_off,
                                                       const int* __restrict__ rlen, const int* __restrict__ ecol,
                                                       const F* __restrict__ eval, const F* __restrict__ dval,
                                                       const T* __restrict__ rhs, const T* __restrict__ aux,
                                                       const T* __restrict__ dd, T* __restrict__ x) {
    const int lane = (int)threadIdx.x % kSfRows;
    const int c = p0 / kSfRows + (int)blockIdx.x * WPB + (int)threadIdx.x / kSfRows;
    const int p = c * kSfRows + lane;
    const bool active = p >= p0 && p < p1;
    RowState<T, F, MODE, KMAX> row;
    row.load(p, active, c, lane, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd);
    if (!active) return;
    // all KMAX gathers at once: a padding slot holds a valid column (that of
    // entry 0), and result() ignores slots past the row's length
    T g[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) g[k] = x[row.cols[k]];
    x[row.i] = row.result(g);
}


template <typename T>
__global__ void k_sf_prepare(int n, T* __restrict__ a, T* __restrict__ b, int* __restrict__ tickets) {
    const T s = sent_val((T*)nullptr);
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        if (a) a[i] = s;
        if (b) b[i] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2) tickets[threadIdx.x] = 0;
}

template <typename T, typename F>
__global__ void k_gather_pad(int64_t n, const int* __restrict__ map, const T* __restrict__ src, F* __restrict__ dst);

template <>
__global__ void k_gather_pad<double, double>(int64_t n, const int* __restrict__ map, const double* __restrict__ src,
                                             double* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBS) {
        const int m = map[e];
        dst[e] = m >= 0 ? src[m] : 0.0;
    }
}
template <>
__global__ void k_gather_pad<double, float>(int64_t n, const int* __restrict__ map, const double* __restrict__ src,
                                            float* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBS) {
        const int m = map[e];
        dst[e] = m >= 0 ? (float)src[m] : 0.0f;
    }
}
template <>
__global__ void k_gather_pad<zc, zc>(int64_t n, const int* __restrict__ map, const zc* __restrict__ src,
                                     zc* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBS) {
        const int m = map[e];
        dst[e] = m >= 0 ? src[m] : zc{0.0, 0.0};
    }
}
template <>
__global__ void k_gather_pad<zc, fc>(int64_t n, const int* __restrict__ map, const zc* __restrict__ src,
                                     fc* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBS) {
        const int m = map[e];
        dst[e] = m >= 0 ? fc{(float)src[m].re, (float)src[m].im} : fc{0.0f, 0.0f};
    }
}

// Several right-hand sides with the columns of a row side by side (entry
// (i, c) at i * kb + c, kb a power of two <= 64).  A wave covers 64 / kb
// consecutive level-order positions, lane = (row slot, column): the lanes
// of one row load the same entries (one broadcast access), poll their
// columns of an input row together (kb adjacent values) and store a row's
// results side by side.  A workgroup of blockDim / 64 waves takes one
// ticket for that many consecutive groups of positions (tickets in level
// order, so the sweep is deadlock-free as k_sweep_sf is; fewer tickets than
// one per wave, whose atomics on one counter would otherwise bound the
// sweep).  Each column's result is formed exactly as the one-column sweep
// forms it.  A wave that reaches the spin limit sets the error flag; its
// workgroup leaves together at the next ticket.
template <typename T, typename F, int MODE, int KMAX>
__global__ __launch_bounds__(sizeof(T) == 8 ? 512 : 256) void k_sweep_il(int n, int nticket, int kb_log2, int* __restrict__ ticket,
                                                   int* __restrict__ err, long spin_limit,
                                                   const int* __restrict__ srow, const int* __restrict__ slice_off,
                                                   const int* __restrict__ rlen, const int* __restrict__ ecol,
                                                   const F* __restrict__ eval, const F* __restrict__ dval,
                                                   const T* __restrict__ rhs, const T* __restrict__ aux,
                                                   const T* __restrict__ dd, T* __restrict__ x) {
    __shared__ int s_ticket, s_abort;
    const int lane = threadIdx.x & (kSfRows - 1), wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    const int kb = 1 << kb_log2;
    const int col = lane & (kb - 1);
    const int rows = kSfRows >> kb_log2;   // positions per wave
    const int slot = lane >> kb_log2;
    if (threadIdx.x == 0) s_abort = 0;
    long spins = 0;
    while (true) {
        __syncthreads();   // every wave is done with the previous ticket
        if (threadIdx.x == 0) s_ticket = s_abort ? nticket : atomicAdd(ticket, 1);
        __syncthreads();
        const int t = s_ticket;
        if (t >= nticket) return;
        const int p = (t * waves + wave) * rows + slot;
        const bool active = p < n;
        const int i = active ? srow[p] : 0;
        const int len = active ? rlen[p] : 0;
        const int base = active ? slice_off[p / kSfRows] + (p % kSfRows) : 0;
        unsigned cols[KMAX];
        F vals[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {   // branch-free, as in RowState::load
            const int at = k < len ? base + kSfRows * k : 0;
            cols[k] = (unsigned)ecol[at];
            vals[k] = eval[at];
        }
        T rv = Num<T>::zero(), d_i = Num<T>::zero();
        F dg = F();
        if (active) {
            if (MODE == BWD_IC) {
                d_i = dd[i];
                rv = aux[((size_t)i << kb_log2) + col];
            } else {
                dg = dval[p];
                rv = rhs[((size_t)i << kb_log2) + col];
            }
        }
        T g[KMAX];
        unsigned miss = active ? (len >= 32 ? ~0u : ((1u << len) - 1u)) : 0u;
        bool done = !active;
        int idle = 0;
        while (true) {
            __asm__ __volatile__("" ::: "memory");
            bool progress = false;
            if (!done && miss) {
                const unsigned want = miss;
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if (want & (1u << k)) g[k] = gload(x + ((size_t)cols[k] << kb_log2) + col);
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if ((want & (1u << k)) && !sent(g[k])) miss &= ~(1u << k);
                progress = miss != want;
            }
            if (!done && miss == 0u) {
                T sum = (MODE == BWD_IC) ? Num<T>::zero() : rv;
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if (k < len) sum = Num<T>::sub(sum, Num<T>::mul(wid<T, F>(vals[k]), g[k]));
                const T out = (MODE == BWD_IC) ? Num<T>::add(Num<T>::mul(sum, d_i), rv)
                                               : Num<T>::div(sum, wid<T, F>(dg));
                gstore(x + ((size_t)i << kb_log2) + col, out);
                done = true;
                progress = true;
            }
            if (__all(done)) break;
            const bool any_progress = __any(progress);   // as in k_sweep_sf: the limit bounds a stall
            if (any_progress && spin_limit > 0) spins = 0;
            if (++spins > spin_limit) {
                if (lane == 0) {
                    atomicOr(err, 1);
                    s_abort = 1;
                }
                break;
            }
            if (any_progress) {
                idle = 0;
            } else {
                if (idle < 2) ++idle;
                if (idle == 1) __builtin_amdgcn_s_sleep(1);
                else __builtin_amdgcn_s_sleep(2);
            }
        }
    }
}

inline int grid_for(int64_t n) {
    int64_t g = (n + kBS - 1) / kBS;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (int)g;
}

// register buckets: the row length rounded up (fewer predicated loads per poll
// pass shortens the pass, and the pass length is part of every level's cost)
template <typename T, typename F, int MODE>
void sweep_kmax(hipStream_t s, int kmax, int g, int max_idle, int n, int nchunks, int* ticket, int* err,
                long spin_limit, const int* srow, const int* slice_off, const int* rlen, const int* ecol,
                const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd, T* x) {
#define SSGPU_SF(K)                                                                                               \
    hipLaunchKernelGGL((k_sweep_sf<T, F, MODE, K>), dim3(g), dim3(kSfRows), 0, s, n, nchunks, ticket, err,        \
                       spin_limit, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd, max_idle, x)
    if (kmax <= 8) SSGPU_SF(8);
    else if (kmax <= 12) SSGPU_SF(12);
    else if (kmax <= 16) SSGPU_SF(16);
    else if (kmax <= 20) SSGPU_SF(20);
    else if (kmax <= 24) SSGPU_SF(24);
    else if (kmax <= 32) SSGPU_SF(32);
    else SSGPU_SF(64);
#undef SSGPU_SF
}

template <typename T, typename F, int MODE>
void batch_sweep_kmax(hipStream_t s, int kmax, int g, int norder, const int2* order, int* ticket, int* err,
                      long spin_limit, const BatchSys* sys, const BatchState* st, int want, int in, int out,
                      int max_idle, GraphRec* rec) {
#define SSGPU_BS(K)                                                                                                \
    launch_or_record(rec, k_batch_sweep<T, F, MODE, K>, dim3(g), dim3(kSfRows), s, norder, order, ticket, err, \
                     spin_limit, sys, st, want, in, out, max_idle)
    if (kmax <= 8) SSGPU_BS(8);
    else if (kmax <= 12) SSGPU_BS(12);
    else if (kmax <= 16) SSGPU_BS(16);
    else if (kmax <= 20) SSGPU_BS(20);
    else if (kmax <= 24) SSGPU_BS(24);
    else if (kmax <= 32) SSGPU_BS(32);
    else SSGPU_BS(64);
#undef SSGPU_BS
}


}  // namespace

template <typename T, typename F>
void launch_sweep_sf(hipStream_t s, int mode, int kmax, int n, int grid, int* ticket, int* err, long spin_limit,
                     const int* srow, const int* slice_off, const int* rlen, const int* ecol, const F* eval,
                     const F* dval, const T* rhs, const T* aux, const T* dd, int max_idle, T* x) {
    const int nchunks = (n + kSfRows - 1) / kSfRows;
    const int g = std::max(1, std::min(grid, nchunks));
    if (mode == FWD_DIV_LAST)
        sweep_kmax<T, F, FWD_DIV_LAST>(s, kmax, g, max_idle, n, nchunks, ticket, err, spin_limit, srow,
                                       slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x);
    else if (mode == BWD_IC)
        sweep_kmax<T, F, BWD_IC>(s, kmax, g, max_idle, n, nchunks, ticket, err, spin_limit, srow,
                                 slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x);
    else
        sweep_kmax<T, F, BWD_DIV_FIRST>(s, kmax, g, max_idle, n, nchunks, ticket, err, spin_limit, srow,
                                        slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x);
}

template <typename T, typename F, int MODE>
void sweep_lvl_kmax(hipStream_t s, int kmax, int p0, int p1, const int* srow, const int* slice_off, const int* rlen,
                    const int* ecol, const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd, T* x) {
    // chunks of the level, kLvlWaves per workgroup (a chunk past the level's
    // last one only exists inside the last workgroup: its lanes are inactive)
    const int nch = (p1 - 1) / kSfRows - p0 / kSfRows + 1;
    const int g = (nch + kLvlWaves - 1) / kLvlWaves;
#define SSGPU_LV(K)                                                                                              \
    hipLaunchKernelGGL((k_sweep_lvl<T, F, MODE, K, kLvlWaves>), dim3(g), dim3(kSfRows * kLvlWaves), 0, s, p0, p1, \
                       srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x)
    if (kmax <= 8) SSGPU_LV(8);
    else if (kmax <= 12) SSGPU_LV(12);
    else if (kmax <= 16) SSGPU_LV(16);
    else if (kmax <= 20) SSGPU_LV(20);
    else if (kmax <= 24) SSGPU_LV(24);
    else if (kmax <= 32) SSGPU_LV(32);
    else SSGPU_LV(64);
#undef SSGPU_LV
}

template <typename T, typename F>
void launch_sweep_levels(hipStream_t s, int mode, int kmax, int nlev, const int* levptr, const int* srow,
                         const int* slice_off, const int* rlen, const int* ecol, const F* eval, const F* dval,
                         const T* rhs, const T* aux, const T* dd, T* x) {
    for (int l = 0; l < nlev; ++l) {
        const int p0 = levptr[l], p1 = levptr[l + 1];
        if (p1 <= p0) continue;
        if (mode == FWD_DIV_LAST)
            sweep_lvl_kmax<T, F, FWD_DIV_LAST>(s, kmax, p0, p1, srow, slice_off, rlen, ecol, eval, dval, rhs, aux,
                                               dd, x);
        else if (mode == BWD_IC)
            sweep_lvl_kmax<T, F, BWD_IC>(s, kmax, p0, p1, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x);
        else
            sweep_lvl_kmax<T, F, BWD_DIV_FIRST>(s, kmax, p0, p1, srow, slice_off, rlen, ecol, eval, dval, rhs, aux,
                                                dd, x);
    }
}



template <typename T, typename F, int MODE>
void sweep_il_kmax(hipStream_t s, int kmax, int g, int wpb, int n, int nticket, int kb_log2, int* ticket, int* err,
                   long spin_limit, const int* srow, const int* slice_off, const int* rlen, const int* ecol,
                   const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd, T* x) {
#define SSGPU_IL(K)                                                                                               \
    hipLaunchKernelGGL((k_sweep_il<T, F, MODE, K>), dim3(g), dim3(kSfRows * wpb), 0, s, n, nticket, kb_log2, ticket, \
                       err, spin_limit, srow, slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x)
    if (kmax <= 8) SSGPU_IL(8);
    else if (kmax <= 12) SSGPU_IL(12);
    else if (kmax <= 16) SSGPU_IL(16);
    else if (kmax <= 20) SSGPU_IL(20);
    else if (kmax <= 24) SSGPU_IL(24);
    else SSGPU_IL(32);
#undef SSGPU_IL
}

template <typename T, typename F>
void launch_sweep_il(hipStream_t s, int mode, int kmax, int n, int grid, int kb, int* ticket, int* err,
                     long spin_limit, const int* srow, const int* slice_off, const int* rlen, const int* ecol,
                     const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd, T* x) {
    int kb_log2 = 0;
    while ((1 << kb_log2) < kb) ++kb_log2;
    const int rows = kSfRows >> kb_log2;          // positions per wave
    // waves per workgroup: up to 32 positions per ticket, at most 8 waves
    // (4 complex: their registers)
    const int wpb = std::max(1, std::min(sizeof(T) == 8 ? 8 : 4, 32 / rows));
    const int ngroups = (n + rows - 1) / rows;
    const int nticket = (ngroups + wpb - 1) / wpb;
    const int g = std::max(1, std::min(grid / wpb, nticket));   // grid: waves in all
    if (mode == FWD_DIV_LAST)
        sweep_il_kmax<T, F, FWD_DIV_LAST>(s, kmax, g, wpb, n, nticket, kb_log2, ticket, err, spin_limit, srow,
                                          slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x);
    else if (mode == BWD_IC)
        sweep_il_kmax<T, F, BWD_IC>(s, kmax, g, wpb, n, nticket, kb_log2, ticket, err, spin_limit, srow, slice_off,
                                    rlen, ecol, eval, dval, rhs, aux, dd, x);
    else
        sweep_il_kmax<T, F, BWD_DIV_FIRST>(s, kmax, g, wpb, n, nticket, kb_log2, ticket, err, spin_limit, srow,
                                           slice_off, rlen, ecol, eval, dval, rhs, aux, dd, x);
}

template <typename T, typename F, int MODE>
void sweep_pos_kmax(hipStream_t s, int kmax, int g, int max_idle, int n, int nchunks, int* ticket, int* err,
                    long spin_limit, const int* srow, const int* slice_off, const int* rlen, const int* ecolp,
                    const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd, T* xp, T* xo) {
#define SSGPU_SP(K)                                                                                               \
    hipLaunchKernelGGL((k_sweep_sf_pos<T, F, MODE, K>), dim3(g), dim3(kSfRows), 0, s, n, nchunks, ticket, err,      \
                       spin_limit, srow, slice_off, rlen, ecolp, eval, dval, rhs, aux, dd, max_idle, xp, xo)
    if (kmax <= 8) SSGPU_SP(8);
    else if (kmax <= 12) SSGPU_SP(12);
    else if (kmax <= 16) SSGPU_SP(16);
    else if (kmax <= 20) SSGPU_SP(20);
    else if (kmax <= 24) SSGPU_SP(24);
    else if (kmax <= 32) SSGPU_SP(32);
    else SSGPU_SP(64);
#undef SSGPU_SP
}

template <typename T, typename F>
void launch_sweep_sf_pos(hipStream_t s, int mode, int kmax, int n, int grid, int* ticket, int* err, long spin_limit,
                         const int* srow, const int* slice_off, const int* rlen, const int* ecolp, const F* eval,
                         const F* dval, const T* rhs, const T* aux, const T* dd, int max_idle, T* xp, T* xo) {
    const int nchunks = (n + kSfRows - 1) / kSfRows;
    const int g = std::max(1, std::min(grid, nchunks));
    if (mode == FWD_DIV_LAST)
        sweep_pos_kmax<T, F, FWD_DIV_LAST>(s, kmax, g, max_idle, n, nchunks, ticket, err, spin_limit, srow,
                                           slice_off, rlen, ecolp, eval, dval, rhs, aux, dd, xp, xo);
    else if (mode == BWD_IC)
        sweep_pos_kmax<T, F, BWD_IC>(s, kmax, g, max_idle, n, nchunks, ticket, err, spin_limit, srow, slice_off,
                                     rlen, ecolp, eval, dval, rhs, aux, dd, xp, xo);
    else
        sweep_pos_kmax<T, F, BWD_DIV_FIRST>(s, kmax, g, max_idle, n, nchunks, ticket, err, spin_limit, srow,
                                            slice_off, rlen, ecolp, eval, dval, rhs, aux, dd, xp, xo);
}

template <typename T, typename F>
void launch_batch_sweep(hipStream_t s, int mode, int kmax, int grid, int norder, const int2* order, int* ticket,
                        int* err, long spin_limit, const BatchSys* sys, const BatchState* st, int want, int in,
                        int out, int max_idle, GraphRec* rec) {
    if (norder <= 0) return;
    const int g = std::max(1, std::min(grid, norder));
    if (mode == FWD_DIV_LAST)
        batch_sweep_kmax<T, F, FWD_DIV_LAST>(s, kmax, g, norder, order, ticket, err, spin_limit, sys, st, want, in,
                                             out, max_idle, rec);
    else if (mode == BWD_IC)
        batch_sweep_kmax<T, F, BWD_IC>(s, kmax, g, norder, order, ticket, err, spin_limit, sys, st, want, in, out,
                                       max_idle, rec);
    else
        batch_sweep_kmax<T, F, BWD_DIV_FIRST>(s, kmax, g, norder, order, ticket, err, spin_limit, sys, st, want, in,
                                              out, max_idle, rec);
}

template <typename T, typename F, int MODE>
void sweep_tile_kmax(hipStream_t s, int kmax, int g, size_t lds, const TileArgs& ta, int* ticket, int* err,
                     long spin_limit, const int* srow, const int* slice_off, const int* rlen, const F* eval,
                     const F* dval, const T* rhs, const T* aux, const T* dd, int max_idle, T* x) {
#define SSGPU_TL(K)                                                                                             \
    do {                                                                                                        \
        auto fn = k_sweep_tile<T, F, MODE, K>;                                                                  \
        if (lds > 65536)                                                                                        \
            SSGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(fn),                                    \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));              \
        hipLaunchKernelGGL(fn, dim3(g), dim3(kTileThreads), lds, s, ta, ticket, err, spin_limit, srow,          \
                           slice_off, rlen, eval, dval, rhs, aux, dd, max_idle, x);                             \
    } while (0)
    if (kmax <= 8) SSGPU_TL(8);
    else if (kmax <= 12) SSGPU_TL(12);
    else if (kmax <= 16) SSGPU_TL(16);
    else if (kmax <= 20) SSGPU_TL(20);
    else if (kmax <= 24) SSGPU_TL(24);
    else if (kmax <= 32) SSGPU_TL(32);
    else SSGPU_TL(64);
#undef SSGPU_TL
}

template <typename T, typename F>
void launch_sweep_tile(hipStream_t s, int mode, int kmax, int grid, size_t lds_bytes, const TileArgs& ta,
                       int* ticket, int* err, long spin_limit, const int* srow, const int* slice_off,
                       const int* rlen, const F* eval, const F* dval, const T* rhs, const T* aux, const T* dd,
                       int max_idle, T* x) {
    if (ta.ntiles <= 0) return;
    const int g = std::max(1, std::min(grid, ta.ntiles));
    if (mode == FWD_DIV_LAST)
        sweep_tile_kmax<T, F, FWD_DIV_LAST>(s, kmax, g, lds_bytes, ta, ticket, err, spin_limit, srow, slice_off, rlen,
                                           eval, dval, rhs, aux, dd, max_idle, x);
    else if (mode == BWD_IC)
        sweep_tile_kmax<T, F, BWD_IC>(s, kmax, g, lds_bytes, ta, ticket, err, spin_limit, srow, slice_off, rlen,
                                     eval, dval, rhs, aux, dd, max_idle, x);
    else
        sweep_tile_kmax<T, F, BWD_DIV_FIRST>(s, kmax, g, lds_bytes, ta, ticket, err, spin_limit, srow, slice_off,
                                            rlen, eval, dval, rhs, aux, dd, max_idle, x);
}

template <typename T>
void launch_sf_prepare(hipStream_t s, int n, T* a, T* b, int* tickets) {
    hipLaunchKernelGGL(k_sf_prepare<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, a, b, tickets);
}

template <typename T, typename F>
void launch_gather_pad(hipStream_t s, int64_t n, const int* map, const T* src, F* dst) {
    hipLaunchKernelGGL((k_gather_pad<T, F>), dim3(grid_for(n)), dim3(kBS), 0, s, n, map, src, dst);
}

template void launch_sf_prepare<double>(hipStream_t, int, double*, double*, int*);
template void launch_sf_prepare<zc>(hipStream_t, int, zc*, zc*, int*);
template void launch_sweep_sf<double, double>(hipStream_t, int, int, int, int, int*, int*, long, const int*, const int*,
                                              const int*, const int*, const double*, const double*, const double*,
                                              const double*, const double*, int, double*);
template void launch_sweep_sf<double, float>(hipStream_t, int, int, int, int, int*, int*, long, const int*, const int*,
                                             const int*, const int*, const float*, const float*, const double*,
                                             const double*, const double*, int, double*);
template void launch_sweep_sf<zc, zc>(hipStream_t, int, int, int, int, int*, int*, long, const int*, const int*,
                                      const int*, const int*, const zc*, const zc*, const zc*, const zc*, const zc*,
                                      int, zc*);
template void launch_sweep_sf<zc, fc>(hipStream_t, int, int, int, int, int*, int*, long, const int*, const int*,
                                      const int*, const int*, const fc*, const fc*, const zc*, const zc*, const zc*,
                                      int, zc*);
template void launch_sweep_levels<double, double>(hipStream_t, int, int, int, const int*, const int*, const int*,
                                             const int*, const int*, const double*, const double*, const double*, const double*,
                                             const double*, double*);
template void launch_sweep_levels<double, float>(hipStream_t, int, int, int, const int*, const int*, const int*,
                                             const int*, const int*, const float*, const float*, const double*, const double*,
                                             const double*, double*);
template void launch_sweep_levels<zc, zc>(hipStream_t, int, int, int, const int*, const int*, const int*,
                                             const int*, const int*, const zc*, const zc*, const zc*, const zc*,
                                             const zc*, zc*);
template void launch_sweep_levels<zc, fc>(hipStream_t, int, int, int, const int*, const int*, const int*,
                                             const int*, const int*, const fc*, const fc*, const zc*, const zc*,
                                             const zc*, zc*);
#define SSGPU_TLI(T, F)                                                                                      \
    template void launch_sweep_tile<T, F>(hipStream_t, int, int, int, size_t, const TileArgs&, int*, int*, long, \
                                          const int*, const int*, const int*, const F*, const F*, const T*,     \
                                          const T*, const T*, int, T*);
SSGPU_TLI(double, double)
SSGPU_TLI(double, float)
SSGPU_TLI(zc, zc)
SSGPU_TLI(zc, fc)
#define SSGPU_SPI(T, F)                                                                                       \
    template void launch_sweep_sf_pos<T, F>(hipStream_t, int, int, int, int, int*, int*, long, const int*,         \
                                            const int*, const int*, const int*, const F*, const F*, const T*,     \
                                            const T*, const T*, int, T*, T*);
SSGPU_SPI(double, double)
SSGPU_SPI(double, float)
SSGPU_SPI(zc, zc)
SSGPU_SPI(zc, fc)
#define SSGPU_ILI(T, F)                                                                                     \
    template void launch_sweep_il<T, F>(hipStream_t, int, int, int, int, int, int*, int*, long, const int*,      \
                                        const int*, const int*, const int*, const F*, const F*, const T*, const T*, \
                                        const T*, T*);
SSGPU_ILI(double, double)
SSGPU_ILI(double, float)
SSGPU_ILI(zc, zc)
SSGPU_ILI(zc, fc)
#define SSGPU_BSI(T, F)                                                                                   \
    template void launch_batch_sweep<T, F>(hipStream_t, int, int, int, int, const int2*, int*, int*, long,       \
                                           const BatchSys*, const BatchState*, int, int, int, int, GraphRec*);
SSGPU_BSI(double, double)
SSGPU_BSI(double, float)
SSGPU_BSI(zc, zc)
SSGPU_BSI(zc, fc)
template void launch_gather_pad<double, double>(hipStream_t, int64_t, const int*, const double*, double*);
template void launch_gather_pad<double, float>(hipStream_t, int64_t, const int*, const double*, float*);
template void launch_gather_pad<zc, zc>(hipStream_t, int64_t, const int*, const zc*, zc*);
template void launch_gather_pad<zc, fc>(hipStream_t, int64_t, const int*, const zc*, fc*);

}  // namespace ssgpu
