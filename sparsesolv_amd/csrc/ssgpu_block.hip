// Kernels of block solves: kb right-hand sides at once, each vector stored
// whole (column c at offset c * ld).  Column c of every kernel computes
// exactly what the one-column kernel computes: the same products, the same
// summation order per row, the same reduction trees (the same rows per
// thread, the same wave shuffles and wave order, the block partials
// finalized the same way), so a block solve returns, column for column, the
// bits of the separate solves.
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

__device__ inline double shfl_xor1(double v) { return __shfl_xor(v, 1); }
__device__ inline zc shfl_xor1(zc v) { return zc{__shfl_xor(v.re, 1), __shfl_xor(v.im, 1)}; }
__device__ inline double ntl(const double* p) { return __builtin_nontemporal_load(p); }
__device__ inline zc ntl(const zc* p) {
    const double* d = reinterpret_cast<const double*>(p);
    return zc{__builtin_nontemporal_load(d), __builtin_nontemporal_load(d + 1)};
}

// y_c = A x_c for the kb columns (column c at offset c * ld); the block's
// entries are staged in LDS once and every column is summed over them as the
// one-column kernel sums it (two threads per row: the even- and odd-position
// chains, y = ta + tb).  Q columns at a time: their gathers are issued
// together, each column's sum still in its own order.  With dotv, one
// partial per SpMV block and column of dotv_c . y_c, row j's term in thread
// slot j as in the one-column kernel (the Q columns' partials reduced in one
// pass, each over the same tree).
template <typename T, int Q>
__global__ __launch_bounds__(kBS) void k_spmv_cols(int kb, int ld, const int* __restrict__ blk,
                                                   const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const T* __restrict__ v, const T* __restrict__ x,
                                                   T* __restrict__ y, const T* __restrict__ dotv,
                                                   double* __restrict__ partial) {
    constexpr int P = BParts<T>::K;
    __shared__ int s_ci[kSpmvCap];
    __shared__ T s_v[kSpmvCap];
    __shared__ T s_term[Q][kSpmvRows];
    const int b = blockIdx.x;
    const int r0 = blk[b], r1 = blk[b + 1];
    const int e0 = rp[r0], e1 = rp[r1];
    const int ne = e1 - e0;
    const int rows = r1 - r0;
    const int j = (int)threadIdx.x >> 1, par = (int)threadIdx.x & 1;
    const int r = r0 + j;
    const bool staged = ne <= kSpmvCap;
    if (staged) {
        for (int k = threadIdx.x; k < ne; k += kBS) {
            s_ci[k] = __builtin_nontemporal_load(ci + e0 + k);
            s_v[k] = ntl(v + e0 + k);
        }
        __syncthreads();
    }
    for (int c0 = 0; c0 < kb; c0 += Q) {   // kb is a multiple of Q
        T yv[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) yv[q] = Num<T>::zero();
        if (staged) {
            T acc[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q] = Num<T>::zero();
            if (r < r1) {
                const int e = rp[r + 1] - e0;
                int k = rp[r] - e0 + par;
                // gathers of a batch issued together, then summed in order
                for (; k + 14 < e; k += 16) {
                    T g[8][Q];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
#pragma unroll
                        for (int q = 0; q < Q; ++q) g[u][q] = x[(size_t)(c0 + q) * ld + s_ci[k + 2 * u]];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
#pragma unroll
                        for (int u = 0; u < 8; ++u) acc[q] = Num<T>::add(acc[q], Num<T>::mul(s_v[k + 2 * u], g[u][q]));
                }
                for (; k < e; k += 2) {
                    T g[Q];
#pragma unroll
                    for (int q = 0; q < Q; ++q) g[q] = x[(size_t)(c0 + q) * ld + s_ci[k]];
#pragma unroll
                    for (int q = 0; q < Q; ++q) acc[q] = Num<T>::add(acc[q], Num<T>::mul(s_v[k], g[q]));
                }
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const T other = shfl_xor1(acc[q]);
                if (r < r1 && par == 0) {
                    yv[q] = Num<T>::add(acc[q], other);
                    y[(size_t)(c0 + q) * ld + r] = yv[q];
                }
            }
        } else if (threadIdx.x == 0 && r0 < r1) {   // one long row per block
            for (int q = 0; q < Q; ++q) {
                const T* xc = x + (size_t)(c0 + q) * ld;
                const int e = rp[r0 + 1];
                int k = rp[r0];
                T ta = Num<T>::zero(), tb = Num<T>::zero();
                while (k < e) {
                    ta = Num<T>::add(ta, Num<T>::mul(v[k], xc[ci[k]]));
                    ++k;
                    if (k < e) { tb = Num<T>::add(tb, Num<T>::mul(v[k], xc[ci[k]])); ++k; }
                }
                yv[q] = Num<T>::add(ta, tb);
                y[(size_t)(c0 + q) * ld + r0] = yv[q];
            }
        }
        if (dotv) {
            if (r < r1 && par == 0)
#pragma unroll
                for (int q = 0; q < Q; ++q) s_term[q][j] = Num<T>::mul(dotv[(size_t)(c0 + q) * ld + r], yv[q]);
            __syncthreads();
            double acc[Q * P] = {};
            if ((int)threadIdx.x < rows)
#pragma unroll
                for (int q = 0; q < Q; ++q) BParts<T>::put(acc + q * P, s_term[q][threadIdx.x]);
            // partials of columns c0 .. c0+Q-1 of this block are adjacent
            block_sum_multi<Q * P>(acc, partial + ((size_t)b * kb + c0) * P);
            __syncthreads();
        }
    }
}

// one partial per block and column of a_c . b_c (or |a_c|^2): column c =
// blockIdx.y, reduced exactly as the one-column k_dot / k_norm2 reduce it
template <typename T, bool NORM>
__global__ __launch_bounds__(kBS) void k_reduce_cols(int n, int ld, const T* __restrict__ a,
                                                     const T* __restrict__ b, double* __restrict__ partial) {
    constexpr int P = NORM ? 1 : BParts<T>::K;
    const int c = blockIdx.y, kb = gridDim.y;
    const T* ac = a + (size_t)c * ld;
    const T* bc = NORM ? nullptr : b + (size_t)c * ld;
    double acc[P] = {};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        if (NORM) acc[0] += Num<T>::abs2(ac[i]);
        else BParts<T>::put(acc, Num<T>::mul(ac[i], bc[i]));
    }
    block_sum_multi<P>(acc, partial + ((size_t)blockIdx.x * kb + c) * P);
}

// x_c += alpha_c p_c; r_c -= alpha_c q_c for the active columns; partials of |r_c|^2
template <typename T>
__global__ __launch_bounds__(kBS) void k_cg_update_cols(int n, int ld, const T* __restrict__ alpha,
                                                        const int* __restrict__ act, const T* __restrict__ p,
                                                        const T* __restrict__ q, T* __restrict__ x,
                                                        T* __restrict__ r, double* __restrict__ partial) {
    const int c = blockIdx.y, kb = gridDim.y;
    if (!act[c]) return;   // its partial is not read
    const T al = alpha[c];
    const size_t o = (size_t)c * ld;
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        x[o + i] = Num<T>::add(x[o + i], Num<T>::mul(al, p[o + i]));
        const T ri = Num<T>::sub(r[o + i], Num<T>::mul(al, q[o + i]));
        r[o + i] = ri;
        acc[0] += Num<T>::abs2(ri);
    }
    block_sum_multi<1>(acc, partial + (size_t)blockIdx.x * kb + c);
}

// p_c = p_c beta_c + z_c for the active columns
template <typename T>
__global__ void k_cg_pupdate_cols(int n, int ld, const T* __restrict__ beta, const int* __restrict__ act,
                                  const T* __restrict__ z, T* __restrict__ p) {
    const int c = blockIdx.y;
    if (!act[c]) return;
    const T be = beta[c];
    const size_t o = (size_t)c * ld;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS)
        p[o + i] = Num<T>::add(Num<T>::mul(p[o + i], be), z[o + i]);
}

// one partial per block and column of (a_c . b_c, a_c . c_c, a_c . d_c), as
// k_dot3 forms them (the third zero without d)
template <typename T>
__global__ __launch_bounds__(kBS) void k_dot3_cols(int n, int ld, const T* __restrict__ a, const T* __restrict__ b,
                                                   const T* __restrict__ cc, const T* __restrict__ d,
                                                   double* __restrict__ partial) {
    constexpr int P = BParts<T>::K;
    const int c = blockIdx.y, kb = gridDim.y;
    const size_t o = (size_t)c * ld;
    double acc[3 * P] = {};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T ai = a[o + i];
        BParts<T>::put(acc, Num<T>::mul(ai, b[o + i]));
        BParts<T>::put(acc + P, Num<T>::mul(ai, cc[o + i]));
        if (d) BParts<T>::put(acc + 2 * P, Num<T>::mul(ai, d[o + i]));
    }
    block_sum_multi<3 * P>(acc, partial + ((size_t)blockIdx.x * kb + c) * 3 * P);
}

// MRTR updates of the active columns, column c with the scalars sc[c]
// (t), sc[kb + c] (zeta), sc[2 kb + c] (eta).  SGS: the Eisenstat form
// (k_sgs_update); else IC-MRTR's (k_mrtr_update, fix: keep p's recurrence).
// Partials of |r_c|^2.
template <typename T, bool SGS>
__global__ __launch_bounds__(kBS) void k_mrtr_update_cols(int n, int ld, int fix, const T* __restrict__ sc,
                                                          const int* __restrict__ act, const T* __restrict__ u,
                                                          const T* __restrict__ v, T* __restrict__ p,
                                                          T* __restrict__ x, T* __restrict__ y, T* __restrict__ r,
                                                          double* __restrict__ partial) {
    const int c = blockIdx.y, kb = gridDim.y;
    if (!act[c]) return;   // its partial is not read
    const T t = sc[c], zeta = sc[kb + c], eta = sc[2 * kb + c];
    const size_t o = (size_t)c * ld;
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T ui = u[o + i];
        T pi;
        if (SGS) pi = Num<T>::add(Num<T>::mul(p[o + i], t), ui);
        else pi = fix ? Num<T>::add(ui, Num<T>::mul(t, p[o + i])) : Num<T>::add(ui, Num<T>::mul(t, ui));
        p[o + i] = pi;
        x[o + i] = Num<T>::add(x[o + i], Num<T>::mul(zeta, pi));
        const T yi = Num<T>::add(Num<T>::mul(y[o + i], eta), Num<T>::mul(zeta, v[o + i]));
        y[o + i] = yi;
        const T ri = Num<T>::sub(r[o + i], yi);
        r[o + i] = ri;
        acc[0] += Num<T>::abs2(ri);
    }
    block_sum_multi<1>(acc, partial + (size_t)blockIdx.x * kb + c);
}

// z_c = z_c eta_c + zeta_c w_c; u_c -= z_c for the active columns (k_mrtr_update2)
template <typename T>
__global__ void k_mrtr_update2_cols(int n, int ld, const T* __restrict__ sc, const int* __restrict__ act,
                                    const T* __restrict__ w, T* __restrict__ z, T* __restrict__ u) {
    const int c = blockIdx.y, kb = gridDim.y;
    if (!act[c]) return;
    const T zeta = sc[kb + c], eta = sc[2 * kb + c];
    const size_t o = (size_t)c * ld;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T zi = Num<T>::add(Num<T>::mul(z[o + i], eta), Num<T>::mul(zeta, w[o + i]));
        z[o + i] = zi;
        u[o + i] = Num<T>::sub(u[o + i], zi);
    }
}

inline int grid_of(int64_t n, int cap) {
    int64_t g = (n + kBS - 1) / kBS;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

}  // namespace

template <typename T>
void launch_spmv_cols(hipStream_t s, int kb, int ld, int nblk, const int* blk, const int* rp, const int* ci,
                      const T* v, const T* x, T* y, const T* dotv, double* partial) {
    // columns summed together: 4 real or 2 complex (kb is a power of two)
    constexpr int QM = sizeof(T) == 8 ? 4 : 2;
    if (kb % QM == 0)
        hipLaunchKernelGGL((k_spmv_cols<T, QM>), dim3(nblk), dim3(kBS), 0, s, kb, ld, blk, rp, ci, v, x, y, dotv,
                           partial);
    else if (kb % 2 == 0)
        hipLaunchKernelGGL((k_spmv_cols<T, 2>), dim3(nblk), dim3(kBS), 0, s, kb, ld, blk, rp, ci, v, x, y, dotv,
                           partial);
    else
        hipLaunchKernelGGL((k_spmv_cols<T, 1>), dim3(nblk), dim3(kBS), 0, s, kb, ld, blk, rp, ci, v, x, y, dotv,
                           partial);
}
template <typename T>
void launch_dot_cols(hipStream_t s, int kb, int n, int ld, const T* a, const T* b, double* partial) {
    hipLaunchKernelGGL((k_reduce_cols<T, false>), dim3(reduce_blocks(n), kb), dim3(kBS), 0, s, n, ld, a, b, partial);
}
template <typename T> void launch_norm2_cols(hipStream_t s, int kb, int n, int ld, const T* a, double* partial) {
    hipLaunchKernelGGL((k_reduce_cols<T, true>), dim3(reduce_blocks(n), kb), dim3(kBS), 0, s, n, ld, a,
                       (const T*)nullptr, partial);
}
template <typename T>
void launch_cg_update_cols(hipStream_t s, int kb, int n, int ld, const T* alpha, const int* act, const T* p,
                           const T* q, T* x, T* r, double* partial) {
    hipLaunchKernelGGL(k_cg_update_cols<T>, dim3(reduce_blocks(n), kb), dim3(kBS), 0, s, n, ld, alpha, act, p, q, x,
                       r, partial);
}
template <typename T>
void launch_cg_pupdate_cols(hipStream_t s, int kb, int n, int ld, const T* beta, const int* act, const T* z, T* p) {
    hipLaunchKernelGGL(k_cg_pupdate_cols<T>, dim3(grid_of(n, 4096), kb), dim3(kBS), 0, s, n, ld, beta, act, z, p);
}

template <typename T>
void launch_dot3_cols(hipStream_t s, int kb, int n, int ld, const T* a, const T* b, const T* c, const T* d,
                      double* partial) {
    hipLaunchKernelGGL(k_dot3_cols<T>, dim3(reduce_blocks(n), kb), dim3(kBS), 0, s, n, ld, a, b, c, d, partial);
}
template <typename T>
void launch_mrtr_update_cols(hipStream_t s, int kb, int n, int ld, bool sgs, int fix, const T* sc, const int* act,
                             const T* u, const T* v, T* p, T* x, T* y, T* r, double* partial) {
    if (sgs)
        hipLaunchKernelGGL((k_mrtr_update_cols<T, true>), dim3(reduce_blocks(n), kb), dim3(kBS), 0, s, n, ld, fix, sc,
                           act, u, v, p, x, y, r, partial);
    else
        hipLaunchKernelGGL((k_mrtr_update_cols<T, false>), dim3(reduce_blocks(n), kb), dim3(kBS), 0, s, n, ld, fix,
                           sc, act, u, v, p, x, y, r, partial);
}
template <typename T>
void launch_mrtr_update2_cols(hipStream_t s, int kb, int n, int ld, const T* sc, const int* act, const T* w, T* z,
                              T* u) {
    hipLaunchKernelGGL(k_mrtr_update2_cols<T>, dim3(grid_of(n, 4096), kb), dim3(kBS), 0, s, n, ld, sc, act, w, z, u);
}

#define SSGPU_INST(T)                                                                                              \
    template void launch_spmv_cols<T>(hipStream_t, int, int, int, const int*, const int*, const int*, const T*,     \
                                      const T*, T*, const T*, double*);                                         \
    template void launch_dot_cols<T>(hipStream_t, int, int, int, const T*, const T*, double*);                    \
    template void launch_norm2_cols<T>(hipStream_t, int, int, int, const T*, double*);                            \
    template void launch_cg_update_cols<T>(hipStream_t, int, int, int, const T*, const int*, const T*, const T*, T*,  \
                                           T*, double*);                                                        \
    template void launch_cg_pupdate_cols<T>(hipStream_t, int, int, int, const T*, const int*, const T*, T*);      \
    template void launch_dot3_cols<T>(hipStream_t, int, int, int, const T*, const T*, const T*, const T*, double*); \
    template void launch_mrtr_update_cols<T>(hipStream_t, int, int, int, bool, int, const T*, const int*, const T*, \
                                             const T*, T*, T*, T*, T*, double*);                                \
    template void launch_mrtr_update2_cols<T>(hipStream_t, int, int, int, const T*, const int*, const T*, T*, T*);
SSGPU_INST(double)
SSGPU_INST(zc)
#undef SSGPU_INST

}  // namespace ssgpu
