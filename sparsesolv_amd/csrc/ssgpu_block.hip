// Kernels of block solves: kb right-hand sides at once, the kb columns of a
// row side by side (entry (i, c) at i * kb + c, kb a power of two <= 32), so
// that a gather, a poll or a store touches a row's columns in one access.
// Column c of every kernel computes exactly what the one-column kernel
// computes: the same products, the same summation order per row, the same
// reduction trees (the same rows per thread, the same grid, the same wave
// shuffles and wave order, the block partials finalized the same way), so a
// block solve returns, column for column, the bits of the separate solves.
// A thread of the reductions takes CQ adjacent columns of its rows
// (workgroup column group blockIdx.y); the updates take one entry per thread.
#include <cfloat>

#include "ssgpu_kernels.h"

namespace ssgpu {

namespace {

template <typename T> struct BParts;
template <> struct BParts<double> {
    static constexpr int K = 1;
    __device__ static void put(double* acc, double v) { acc[0] += v; }
};
template <> struct BParts<zc> {
    static constexpr int K = 2;
    __device__ static void put(double* acc, zc v) {
        acc[0] += v.re;
        acc[1] += v.im;
    }
};

// NA per-thread values reduced over the block like block_sum (wave shuffles
// down 32..1, then the 4 wave sums in order); thread a < NA writes value a.
template <int NA> __device__ inline void block_sum_multi(double* acc, double* out) {
    __shared__ double red[NA][kBS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        double v = acc[a];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[a][wave] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < NA) {
        double v = red[threadIdx.x][0];
        for (int w = 1; w < kBS / 64; ++w) v += red[threadIdx.x][w];
        out[threadIdx.x] = v;
    }
}

__device__ inline double ntl(const double* p) { return __builtin_nontemporal_load(p); }
__device__ inline zc ntl(const zc* p) {
    const double* d = reinterpret_cast<const double*>(p);
    return zc{__builtin_nontemporal_load(d), __builtin_nontemporal_load(d + 1)};
}

// y = A x for the kb columns, one lane per (row, column): the kb lanes of a
// row read an input row's columns in one access (kb adjacent values).  The
// block's entries are staged in LDS once; a lane sums its row's even- and
// odd-position chains (the one-column kernel's two threads per row) in
// order and stores y = ta + tb, so every column is summed exactly as the
// one-column kernel sums it.
template <typename T>
__global__ __launch_bounds__(kBS) void k_spmv_rc(int kb_log2, const int* __restrict__ blk, const int* __restrict__ rp,
                                                 const int* __restrict__ ci, const T* __restrict__ v,
                                                 const T* __restrict__ x, T* __restrict__ y) {
    __shared__ int s_ci[kSpmvCap];
    __shared__ T s_v[kSpmvCap];
    const int b = blockIdx.x;
    const int r0 = blk[b], r1 = blk[b + 1];
    const int e0 = rp[r0], e1 = rp[r1];
    const int ne = e1 - e0;
    const int kb = 1 << kb_log2;
    const int c = (int)threadIdx.x & (kb - 1);
    const int rows_per_pass = kBS >> kb_log2;
    if (ne <= kSpmvCap) {
        for (int k = threadIdx.x; k < ne; k += kBS) {
            s_ci[k] = __builtin_nontemporal_load(ci + e0 + k);
            s_v[k] = ntl(v + e0 + k);
        }
        __syncthreads();
        for (int r = r0 + ((int)threadIdx.x >> kb_log2); r < r1; r += rows_per_pass) {
            const int e = rp[r + 1] - e0;
            int k = rp[r] - e0;
            T ta = Num<T>::zero(), tb = Num<T>::zero();
            // both chains in order, a few entries' gathers issued together
            for (; k + 7 < e; k += 8) {
                T g[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) g[u] = x[((size_t)s_ci[k + u] << kb_log2) + c];
#pragma unroll
                for (int u = 0; u < 8; u += 2) {
                    ta = Num<T>::add(ta, Num<T>::mul(s_v[k + u], g[u]));
                    tb = Num<T>::add(tb, Num<T>::mul(s_v[k + u + 1], g[u + 1]));
                }
            }
            for (; k < e; ++k) {
                const T prod = Num<T>::mul(s_v[k], x[((size_t)s_ci[k] << kb_log2) + c]);
                if (((k - (rp[r] - e0)) & 1) == 0) ta = Num<T>::add(ta, prod);
                else tb = Num<T>::add(tb, prod);
            }
            y[((size_t)r << kb_log2) + c] = Num<T>::add(ta, tb);
        }
    } else if ((int)threadIdx.x < kb && r0 < r1) {   // one long row per block
        const int e = rp[r0 + 1];
        int k = rp[r0];
        T ta = Num<T>::zero(), tb = Num<T>::zero();
        while (k < e) {
            ta = Num<T>::add(ta, Num<T>::mul(v[k], x[((size_t)ci[k] << kb_log2) + c]));
            ++k;
            if (k < e) {
                tb = Num<T>::add(tb, Num<T>::mul(v[k], x[((size_t)ci[k] << kb_log2) + c]));
                ++k;
            }
        }
        y[((size_t)r0 << kb_log2) + c] = Num<T>::add(ta, tb);
    }
}

// the fused p.Ap of the one-column SpMV: per SpMV block and column, the
// block tree (row j's term in thread slot j) of dotv_c(r) * y_c(r)
template <typename T, int KB>
__global__ __launch_bounds__(kBS) void k_spmv_dot_rc(const int* __restrict__ blk, const T* __restrict__ dotv,
                                                     const T* __restrict__ y, double* __restrict__ partial) {
    constexpr int P = BParts<T>::K;
    const int b = blockIdx.x;
    const int r0 = blk[b], rows = blk[b + 1] - r0;
    double acc[KB * P] = {};
    if ((int)threadIdx.x < rows) {
        const size_t o = (size_t)(r0 + (int)threadIdx.x) * KB;
#pragma unroll
        for (int c = 0; c < KB; ++c) BParts<T>::put(acc + c * P, Num<T>::mul(dotv[o + c], y[o + c]));
    }
    block_sum_multi<KB * P>(acc, partial + (size_t)b * KB * P);
}

// one partial per block and column of a_c . b_c (or |a_c|^2), reduced
// exactly as the one-column k_dot / k_norm2 reduce it
template <typename T, int CQ, bool NORM>
__global__ __launch_bounds__(kBS) void k_reduce_il(int n, int kb, const T* __restrict__ a, const T* __restrict__ b,
                                                   double* __restrict__ partial) {
    constexpr int P = NORM ? 1 : BParts<T>::K;
    const int c0 = blockIdx.y * CQ;
    double acc[CQ * P] = {};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const size_t o = (size_t)i * kb + c0;
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
            if (NORM) acc[q] += Num<T>::abs2(a[o + q]);
            else BParts<T>::put(acc + q * P, Num<T>::mul(a[o + q], b[o + q]));
        }
    }
    block_sum_multi<CQ * P>(acc, partial + ((size_t)blockIdx.x * kb + c0) * P);
}

// Updates of the active columns, one thread per entry (coalesced); column c
// takes its scalars from sc.  The norms of the updated residuals are formed
// afterwards by k_reduce_il, over the one-column kernels' tree.
//   CG:   x += alpha_c p; r -= alpha_c q          (sc: alpha)
//   CGP:  p = p beta_c + z                         (sc: beta)
//   MRTR: p = u + t_c p (fix) or u + t_c u; x += zeta_c p; y = y eta_c + zeta_c v;
//         r -= y                                   (sc: t, zeta, eta)
//   SGS:  p = p t_c + u, then as MRTR              (k_sgs_update)
//   MRTR2: z = z eta_c + zeta_c w; u -= z
enum { UPD_CG, UPD_CGP, UPD_MRTR, UPD_SGS, UPD_MRTR2 };
template <typename T, int OP>
__global__ void k_update_il(int64_t total, int kb_log2, int fix, const T* __restrict__ sc,
                            const int* __restrict__ act, const T* __restrict__ a, const T* __restrict__ b,
                            T* __restrict__ p, T* __restrict__ x, T* __restrict__ y, T* __restrict__ r) {
    const int kb = 1 << kb_log2;
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBS) {
        const int c = (int)(e & (kb - 1));
        if (!act[c]) continue;
        if (OP == UPD_CG) {   // a = p, b = q
            const T al = sc[c];
            x[e] = Num<T>::add(x[e], Num<T>::mul(al, a[e]));
            r[e] = Num<T>::sub(r[e], Num<T>::mul(al, b[e]));
        } else if (OP == UPD_CGP) {   // a = z
            p[e] = Num<T>::add(Num<T>::mul(p[e], sc[c]), a[e]);
        } else if (OP == UPD_MRTR || OP == UPD_SGS) {   // a = u, b = v
            const T t = sc[c], zeta = sc[kb + c], eta = sc[2 * kb + c];
            const T ui = a[e];
            T pi;
            if (OP == UPD_SGS) pi = Num<T>::add(Num<T>::mul(p[e], t), ui);
            else pi = fix ? Num<T>::add(ui, Num<T>::mul(t, p[e])) : Num<T>::add(ui, Num<T>::mul(t, ui));
            p[e] = pi;
            x[e] = Num<T>::add(x[e], Num<T>::mul(zeta, pi));
            const T yi = Num<T>::add(Num<T>::mul(y[e], eta), Num<T>::mul(zeta, b[e]));
            y[e] = yi;
            r[e] = Num<T>::sub(r[e], yi);
        } else {   // UPD_MRTR2: a = w, y = z, r = u
            const T zi = Num<T>::add(Num<T>::mul(y[e], sc[2 * kb + c]), Num<T>::mul(sc[kb + c], a[e]));
            y[e] = zi;
            r[e] = Num<T>::sub(r[e], zi);
        }
    }
}

// x(i, c) = x(i, c) * s_i (or / s_i), as k_mul_scale / k_div_scale per column
template <typename T, bool DIV>
__global__ void k_scale_il(int64_t total, int kb_log2, const T* __restrict__ sc, T* __restrict__ x) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBS) {
        const T s = sc[e >> kb_log2];
        x[e] = DIV ? Num<T>::div(x[e], s) : Num<T>::mul(x[e], s);
    }
}

// column c of an interleaved array
template <typename T>
__global__ void k_copy_col_il(int n, int kb, int c, const T* __restrict__ src, T* __restrict__ dst) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS)
        dst[(size_t)i * kb + c] = src[(size_t)i * kb + c];
}

// nr vectors one after another (vector c at c * n) -> kb interleaved columns
// (columns >= nr zero), and back
template <typename T>
__global__ void k_to_il(int64_t total, int n, int kb_log2, int nr, const T* __restrict__ src, T* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBS) {
        const int64_t i = e >> kb_log2;
        const int c = (int)(e & ((1 << kb_log2) - 1));
        dst[e] = c < nr ? src[(size_t)c * n + i] : Num<T>::zero();
    }
}
template <typename T>
__global__ void k_from_il(int64_t total, int n, int kb_log2, const T* __restrict__ src, T* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)kBS + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBS) {
        const int64_t c = e / n, i = e - c * n;
        dst[e] = src[((size_t)i << kb_log2) + c];
    }
}

inline int grid_of(int64_t n, int cap) {
    int64_t g = (n + kBS - 1) / kBS;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}
inline int log2_of(int kb) {
    int l = 0;
    while ((1 << l) < kb) ++l;
    return l;
}
// adjacent columns per thread of the reductions: 128 bytes of a row (16 real
// or 8 complex values)
template <typename T> inline int cq_of(int kb) { return std::min(kb, sizeof(T) == 8 ? 16 : 8); }

}  // namespace

template <typename T>
void launch_spmv_il(hipStream_t s, int kb, int nblk, const int* blk, const int* rp, const int* ci, const T* v,
                    const T* x, T* y, const T* dotv, double* partial) {
    hipLaunchKernelGGL(k_spmv_rc<T>, dim3(nblk), dim3(kBS), 0, s, log2_of(kb), blk, rp, ci, v, x, y);
    if (!dotv) return;
#define SSGPU_SD(K) hipLaunchKernelGGL((k_spmv_dot_rc<T, K>), dim3(nblk), dim3(kBS), 0, s, blk, dotv, y, partial)
    if (kb == 32) SSGPU_SD(32);
    else if (kb == 16) SSGPU_SD(16);
    else if (kb == 8) SSGPU_SD(8);
    else if (kb == 4) SSGPU_SD(4);
    else if (kb == 2) SSGPU_SD(2);
    else SSGPU_SD(1);
#undef SSGPU_SD
}

// launch KERNEL(CQ) over (GX, kb / CQ) workgroups
#define SSGPU_CQ(KERNEL, GX, ...)                                                                      \
    do {                                                                                               \
        const int cq_ = cq_of<T>(kb);                                                                  \
        if (cq_ == 16) hipLaunchKernelGGL(KERNEL(16), dim3(GX, kb / 16), dim3(kBS), 0, s, __VA_ARGS__); \
        else if (cq_ == 8) hipLaunchKernelGGL(KERNEL(8), dim3(GX, kb / 8), dim3(kBS), 0, s, __VA_ARGS__); \
        else if (cq_ == 4) hipLaunchKernelGGL(KERNEL(4), dim3(GX, kb / 4), dim3(kBS), 0, s, __VA_ARGS__); \
        else if (cq_ == 2) hipLaunchKernelGGL(KERNEL(2), dim3(GX, kb / 2), dim3(kBS), 0, s, __VA_ARGS__); \
        else hipLaunchKernelGGL(KERNEL(1), dim3(GX, kb), dim3(kBS), 0, s, __VA_ARGS__);                \
    } while (0)

template <typename T> void launch_dot_il(hipStream_t s, int kb, int n, const T* a, const T* b, double* partial) {
#define K_(Q) (k_reduce_il<T, Q, false>)
    SSGPU_CQ(K_, reduce_blocks(n), n, kb, a, b, partial);
#undef K_
}
template <typename T> void launch_norm2_il(hipStream_t s, int kb, int n, const T* a, double* partial) {
#define K_(Q) (k_reduce_il<T, Q, true>)
    SSGPU_CQ(K_, reduce_blocks(n), n, kb, a, (const T*)nullptr, partial);
#undef K_
}
template <typename T>
void launch_cg_update_il(hipStream_t s, int kb, int n, const T* alpha, const int* act, const T* p, const T* q,
                         T* x, T* r, double* partial) {
    const int64_t total = (int64_t)n * kb;
    hipLaunchKernelGGL((k_update_il<T, UPD_CG>), dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total, log2_of(kb), 0,
                       alpha, act, p, q, (T*)nullptr, x, (T*)nullptr, r);
    launch_norm2_il<T>(s, kb, n, r, partial);
}
template <typename T>
void launch_cg_pupdate_il(hipStream_t s, int kb, int n, const T* beta, const int* act, const T* z, T* p) {
    const int64_t total = (int64_t)n * kb;
    hipLaunchKernelGGL((k_update_il<T, UPD_CGP>), dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total, log2_of(kb), 0,
                       beta, act, z, (const T*)nullptr, p, (T*)nullptr, (T*)nullptr, (T*)nullptr);
}
template <typename T>
void launch_mrtr_update_il(hipStream_t s, int kb, int n, bool sgs, int fix, const T* sc, const int* act, const T* u,
                           const T* v, T* p, T* x, T* y, T* r, double* partial) {
    const int64_t total = (int64_t)n * kb;
    if (sgs)
        hipLaunchKernelGGL((k_update_il<T, UPD_SGS>), dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total,
                           log2_of(kb), fix, sc, act, u, v, p, x, y, r);
    else
        hipLaunchKernelGGL((k_update_il<T, UPD_MRTR>), dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total,
                           log2_of(kb), fix, sc, act, u, v, p, x, y, r);
    launch_norm2_il<T>(s, kb, n, r, partial);
}
template <typename T>
void launch_mrtr_update2_il(hipStream_t s, int kb, int n, const T* sc, const int* act, const T* w, T* z, T* u) {
    const int64_t total = (int64_t)n * kb;
    hipLaunchKernelGGL((k_update_il<T, UPD_MRTR2>), dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total, log2_of(kb),
                       0, sc, act, w, (const T*)nullptr, (T*)nullptr, (T*)nullptr, z, u);
}
#undef SSGPU_CQ

template <typename T> void launch_scale_il(hipStream_t s, int kb, int n, const T* sc, T* x, bool div) {
    const int64_t total = (int64_t)n * kb;
    const int g = grid_of(total, 8192);
    if (div) hipLaunchKernelGGL((k_scale_il<T, true>), dim3(g), dim3(kBS), 0, s, total, log2_of(kb), sc, x);
    else hipLaunchKernelGGL((k_scale_il<T, false>), dim3(g), dim3(kBS), 0, s, total, log2_of(kb), sc, x);
}
template <typename T> void launch_copy_col_il(hipStream_t s, int kb, int n, int c, const T* src, T* dst) {
    hipLaunchKernelGGL(k_copy_col_il<T>, dim3(grid_of(n, 4096)), dim3(kBS), 0, s, n, kb, c, src, dst);
}
template <typename T> void launch_to_il(hipStream_t s, int kb, int n, int nr, const T* src, T* dst) {
    const int64_t total = (int64_t)n * kb;
    hipLaunchKernelGGL(k_to_il<T>, dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total, n, log2_of(kb), nr, src, dst);
}
template <typename T> void launch_from_il(hipStream_t s, int kb, int n, int nr, const T* src, T* dst) {
    const int64_t total = (int64_t)n * nr;
    hipLaunchKernelGGL(k_from_il<T>, dim3(grid_of(total, 8192)), dim3(kBS), 0, s, total, n, log2_of(kb), src, dst);
}

#define SSGPU_INST(T)                                                                                              \
    template void launch_spmv_il<T>(hipStream_t, int, int, const int*, const int*, const int*, const T*, const T*, \
                                    T*, const T*, double*);                                                       \
    template void launch_dot_il<T>(hipStream_t, int, int, const T*, const T*, double*);                           \
    template void launch_norm2_il<T>(hipStream_t, int, int, const T*, double*);                                   \
    template void launch_cg_update_il<T>(hipStream_t, int, int, const T*, const int*, const T*, const T*, T*, T*, \
                                         double*);                                                                \
    template void launch_cg_pupdate_il<T>(hipStream_t, int, int, const T*, const int*, const T*, T*);             \
    template void launch_mrtr_update_il<T>(hipStream_t, int, int, bool, int, const T*, const int*, const T*,     \
                                           const T*, T*, T*, T*, T*, double*);                                    \
    template void launch_mrtr_update2_il<T>(hipStream_t, int, int, const T*, const int*, const T*, T*, T*);       \
    template void launch_scale_il<T>(hipStream_t, int, int, const T*, T*, bool);                                  \
    template void launch_copy_col_il<T>(hipStream_t, int, int, int, const T*, T*);                                \
    template void launch_to_il<T>(hipStream_t, int, int, int, const T*, T*);                                      \
    template void launch_from_il<T>(hipStream_t, int, int, int, const T*, T*);
SSGPU_INST(double)
SSGPU_INST(zc)
#undef SSGPU_INST

}  // namespace ssgpu
