// Device kernels of the gfx950 solver library.
//
// Compiled with -ffp-contract=off: every product and sum is rounded on its
// own, as in the reference's x86-64 build, so per-row results match it bit
// for bit.
#include <cfloat>

#include "ssgpu_kernels.h"

namespace ssgpu {

namespace {

__device__ inline double re_part(double a) { return a; }
__device__ inline double re_part(zc a) { return a.re; }

template <typename T> struct Parts;
template <> struct Parts<double> {
    static constexpr int K = 1;
    __device__ static void put(double* acc, double v) { acc[0] += v; }
};
template <> struct Parts<zc> {
    static constexpr int K = 2;
    __device__ static void put(double* acc, zc v) { acc[0] += v.re; acc[1] += v.im; }
};

__device__ inline double to_d(double a) { return a; }
__device__ inline double to_d(float a) { return (double)a; }
__device__ inline zc to_z(zc a) { return a; }
__device__ inline zc to_z(fc a) { return zc{(double)a.re, (double)a.im}; }
template <typename T, typename F> __device__ inline T widen(F a);
template <> __device__ inline double widen<double, double>(double a) { return a; }
template <> __device__ inline double widen<double, float>(float a) { return (double)a; }
template <> __device__ inline zc widen<zc, zc>(zc a) { return a; }
template <> __device__ inline zc widen<zc, fc>(fc a) { return zc{(double)a.re, (double)a.im}; }

template <typename F, typename T> __device__ inline F narrow(T a);
template <> __device__ inline double narrow<double, double>(double a) { return a; }
template <> __device__ inline float narrow<float, double>(double a) { return (float)a; }
template <> __device__ inline zc narrow<zc, zc>(zc a) { return a; }
template <> __device__ inline fc narrow<fc, zc>(zc a) { return fc{(float)a.re, (float)a.im}; }

// Block-wide sum of K doubles per thread; result valid in thread 0.
// Fixed order: wave shuffles (down 32,16,...,1), then the 4 wave sums in order.
template <int K> __device__ inline void block_sum(double* acc) {
    __shared__ double red[K][kBS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double v = red[k][0];
            for (int w = 1; w < kBS / 64; ++w) v += red[k][w];
            acc[k] = v;
        }
    }
}

// ---------------------------------------------------------------- SpMV
__device__ inline double ntload(const double* p) { return __builtin_nontemporal_load(p); }
__device__ inline zc ntload(const zc* p) {
    const double* d = reinterpret_cast<const double*>(p);
    return zc{__builtin_nontemporal_load(d), __builtin_nontemporal_load(d + 1)};
}
__device__ inline double shfl_xor1(double v) { return __shfl_xor(v, 1); }
__device__ inline zc shfl_xor1(zc v) { return zc{__shfl_xor(v.re, 1), __shfl_xor(v.im, 1)}; }

// Workgroups are handed to the 8 XCDs in turn (workgroup w starts on XCD
// w % 8).  Give each XCD one contiguous range of row blocks: the band of x
// that a range gathers then stays in that XCD's L2, instead of every XCD
// streaming all of x.  A bijection of [0, nb).
__device__ inline int xcd_block(int w, int nb) {
    constexpr int kXcd = 8;
    const int q = nb / kXcd, rem = nb % kXcd;
    const int xcd = w % kXcd, local = w / kXcd;
    return xcd * q + min(xcd, rem) + local;
}

// Row block b of y = A x (and its dot partial): the body of the one-system
// and the batched SpMV kernels.
template <typename T, int STAGE_BYTES>
__device__ inline void spmv_block(int b, const int* __restrict__ blk, const int* __restrict__ rp,
                                  const int* __restrict__ ci, const T* __restrict__ v, const T* __restrict__ x,
                                  T* __restrict__ y, const T* __restrict__ dotv, double* __restrict__ partial) {
    // products are staged through LDS in pieces of kStage entries: 16 KB
    // keeps 8 workgroups per CU in flight, 32 KB (4 per CU) suits orders whose
    // gathers are scattered (fewer of them compete for L2 at once)
    constexpr int kStage = STAGE_BYTES / (int)sizeof(T) < kSpmvCap ? STAGE_BYTES / (int)sizeof(T) : kSpmvCap;
    static_assert(kStage % kBS == 0 && kSpmvCap % kStage == 0 && kStage >= kSpmvRows, "staging");
    __shared__ T prod[kStage];
    const int r0 = blk[b], r1 = blk[b + 1];
    const int e0 = rp[r0], e1 = rp[r1];
    const int ne = e1 - e0;
    // two threads per row: the even- and odd-position chains of the
    // reference's two alternating accumulators, added at the end (ta + tb)
    const int j = (int)threadIdx.x >> 1, par = (int)threadIdx.x & 1;
    const int r = r0 + j;
    T yv = Num<T>::zero();
    if (ne <= kSpmvCap) {
        // each piece: a thread's kStage / kBS entries are loaded together
        // (indices and values, then the gathers) so their latencies overlap;
        // then every chain adds the piece's products at its positions, in
        // order, so a chain's sum runs exactly as over the whole row
        constexpr int U = kStage / kBS;
        T acc = Num<T>::zero();
        int k = r < r1 ? rp[r] - e0 + par : 0;
        const int e = r < r1 ? rp[r + 1] - e0 : 0;
        for (int base = 0; base < ne; base += kStage) {
            int cc[U];
            T vv[U], xx[U];
            // branch-free, so that all U loads (then all U gathers) are in
            // flight at once: a slot past the block's entries re-reads its
            // last entry (ne >= 1 here) and its product is not stored
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = base + u * kBS + (int)threadIdx.x;
                const int qc = q < ne ? q : ne - 1;
                // the matrix streams through once: keep it out of the caches' way
                cc[u] = __builtin_nontemporal_load(ci + e0 + qc);
                vv[u] = ntload(v + e0 + qc);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xx[u] = x[cc[u]];
            if (base > 0) __syncthreads();   // the previous piece has been summed
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = base + u * kBS + (int)threadIdx.x;
                if (q < ne) prod[q - base] = Num<T>::mul(vv[u], xx[u]);
            }
            __syncthreads();
            const int lim = e < base + kStage ? e : base + kStage;
#pragma unroll 4
            for (; k < lim; k += 2) acc = Num<T>::add(acc, prod[k - base]);
        }
        const T other = shfl_xor1(acc);
        if (r < r1 && par == 0) {
            yv = Num<T>::add(acc, other);
            y[r] = yv;
        }
    } else if (threadIdx.x == 0 && r0 < r1) {   // one long row per block
        const int e = rp[r0 + 1];
        int k = rp[r0];
        T ta = Num<T>::zero(), tb = Num<T>::zero();
        while (k < e) {
            ta = Num<T>::add(ta, Num<T>::mul(v[k], x[ci[k]]));
            ++k;
            if (k < e) { tb = Num<T>::add(tb, Num<T>::mul(v[k], x[ci[k]])); ++k; }
        }
        yv = Num<T>::add(ta, tb);
        y[r0] = yv;
    }
    if (dotv) {
        // row j's term at thread slot j, then the block tree
        __syncthreads();
        T* term = prod;
        if (r < r1 && par == 0) term[j] = Num<T>::mul(dotv[r], yv);
        __syncthreads();
        constexpr int K = Parts<T>::K;
        double acc[K] = {};
        if ((int)threadIdx.x < r1 - r0) Parts<T>::put(acc, term[threadIdx.x]);
        block_sum<K>(acc);
        if (threadIdx.x == 0)
            for (int k = 0; k < K; ++k) partial[b * K + k] = acc[k];
    }
}

template <typename T, int STAGE_BYTES>
__global__ __launch_bounds__(kBS) void k_spmv(const int* __restrict__ blk, const int* __restrict__ rp,
                                              const int* __restrict__ ci, const T* __restrict__ v,
                                              const T* __restrict__ x, T* __restrict__ y,
                                              const T* __restrict__ dotv, double* __restrict__ partial) {
    spmv_block<T, STAGE_BYTES>(xcd_block((int)blockIdx.x, (int)gridDim.x), blk, rp, ci, v, x, y, dotv, partial);
}

// ---------------------------------------------------------------- sweeps
template <typename T, typename F, int MODE>
__global__ __launch_bounds__(kBS) void k_sweep(int p0, int p1, const int* __restrict__ prp,
                                               const int* __restrict__ pci, const F* __restrict__ pv,
                                               const int* __restrict__ prow, const T* __restrict__ rhs,
                                               const T* __restrict__ aux, const T* __restrict__ dd,
                                               T* __restrict__ x) {
    const int p = p0 + blockIdx.x * kBS + threadIdx.x;
    if (p >= p1) return;
    const int row = prow[p];
    const int a = prp[p], e = prp[p + 1];
    if (MODE == FWD_DIV_LAST) {
        T s = rhs[row];
        for (int k = a; k < e - 1; ++k) s = Num<T>::sub(s, Num<T>::mul(widen<T, F>(pv[k]), x[pci[k]]));
        x[row] = Num<T>::div(s, widen<T, F>(pv[e - 1]));
    } else if (MODE == BWD_IC) {
        T s = Num<T>::zero();
        for (int k = a + 1; k < e; ++k) s = Num<T>::sub(s, Num<T>::mul(widen<T, F>(pv[k]), x[pci[k]]));
        s = Num<T>::mul(s, dd[row]);
        x[row] = Num<T>::add(s, aux[row]);
    } else {
        T s = rhs[row];
        for (int k = a + 1; k < e; ++k) s = Num<T>::sub(s, Num<T>::mul(widen<T, F>(pv[k]), x[pci[k]]));
        x[row] = Num<T>::div(s, widen<T, F>(pv[a]));
    }
}

// ---------------------------------------------------------------- IC(0)
__device__ inline double absval(double a) { return fabs(a); }
__device__ inline double absval(zc a) { return hypot(a.re, a.im); }
__device__ inline double mulacc(double v, double a) { return v * a; }
__device__ inline zc mulacc(zc v, double a) { return v * a; }
__device__ inline double recip(double s) { return 1.0 / s; }
__device__ inline zc recip(zc s) { return zdiv(zc{1.0, 0.0}, s); }
template <typename T> __device__ inline T from_real(double a);
template <> __device__ inline double from_real<double>(double a) { return a; }
template <> __device__ inline zc from_real<zc>(double a) { return zc{a, 0.0}; }

template <typename T>
__global__ __launch_bounds__(kBS) void k_ic_level(int p0, int p1, const int* __restrict__ prp,
                                                  const int* __restrict__ pci, T* __restrict__ pv,
                                                  const int* __restrict__ prow, const int* __restrict__ ppos,
                                                  T* __restrict__ diagD, double accel, int* __restrict__ nfix) {
    const int p = p0 + blockIdx.x * kBS + threadIdx.x;
    if (p >= p1) return;
    const int i = prow[p];
    const int rs = prp[p], re = prp[p + 1];
    for (int jj = rs; jj < re; ++jj) {
        const int j = pci[jj];
        if (j >= i) break;
        T s = pv[jj];
        const int q = ppos[j];
        int ll = prp[q];
        const int le = prp[q + 1];
        for (int kk = rs; kk < re; ++kk) {
            const int ck = pci[kk];
            if (!(ck <= j - 1)) break;
            while (ll < le && pci[ll] < ck) ++ll;
            if (ll < le && pci[ll] == ck)
                s = Num<T>::sub(s, Num<T>::mul(Num<T>::mul(pv[kk], pv[ll]), diagD[ck]));
        }
        pv[jj] = s;
    }
    const int last = re - 1;
    T s = mulacc(pv[last], accel);
    for (int kk = rs; kk < last; ++kk) {
        if (pci[kk] <= i - 1) s = Num<T>::sub(s, Num<T>::mul(Num<T>::mul(pv[kk], pv[kk]), diagD[pci[kk]]));
        else break;
    }
    pv[last] = s;
    if (absval(s) < 1.0e-12) {
        diagD[i] = from_real<T>(1.0 / DBL_EPSILON);
        atomicAdd(nfix, 1);
    } else {
        diagD[i] = recip(s);
    }
}

// ---------------------------------------------------------------- reductions
template <typename T>
__global__ __launch_bounds__(kBS) void k_dot(int n, const T* __restrict__ a, const T* __restrict__ b,
                                             double* __restrict__ partial) {
    constexpr int K = Parts<T>::K;
    double acc[K] = {};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) Parts<T>::put(acc, Num<T>::mul(a[i], b[i]));
    block_sum<K>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) partial[blockIdx.x * K + k] = acc[k];
}

template <typename T>
__global__ __launch_bounds__(kBS) void k_norm2(int n, const T* __restrict__ a, double* __restrict__ partial) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) acc[0] += Num<T>::abs2(a[i]);
    block_sum<1>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

template <typename T>
__global__ __launch_bounds__(kBS) void k_dot3(int n, const T* __restrict__ a, const T* __restrict__ b,
                                              const T* __restrict__ c, const T* __restrict__ d,
                                              double* __restrict__ partial) {
    constexpr int K = Parts<T>::K;
    double acc[3 * K] = {};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T ai = a[i];
        Parts<T>::put(acc, Num<T>::mul(ai, b[i]));
        Parts<T>::put(acc + K, Num<T>::mul(ai, c[i]));
        if (d) Parts<T>::put(acc + 2 * K, Num<T>::mul(ai, d[i]));
    }
    block_sum<3 * K>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3 * K; ++k) partial[blockIdx.x * 3 * K + k] = acc[k];
}

// out[k] (thread 0) = the fixed-tree sum of partial[b * K + k] over b < nb
__device__ inline void finalize_parts(int nb, int K, const double* __restrict__ partial, double* __restrict__ out) {
    for (int k = 0; k < K; ++k) {
        double acc[1] = {0.0};
        for (int b = threadIdx.x; b < nb; b += kBS) acc[0] += partial[b * K + k];
        block_sum<1>(acc);
        if (threadIdx.x == 0) out[k] = acc[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBS) void k_finalize(int nb, int K, const double* __restrict__ partial,
                                                  double* __restrict__ out) {
    finalize_parts(nb, K, partial, out);
}

// ---------------------------------------------------------------- vectors
template <typename T> __global__ void k_fill(int64_t n, T* x, T val) {
    for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) x[i] = val;
}
template <typename T> __global__ void k_copy(int64_t n, const T* __restrict__ a, T* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) b[i] = a[i];
}

template <typename T>
__global__ __launch_bounds__(kBS) void k_cg_update(int n, T alpha, const T* __restrict__ p, const T* __restrict__ q,
                                                   T* __restrict__ x, T* __restrict__ r, double* __restrict__ partial) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        x[i] = Num<T>::add(x[i], Num<T>::mul(alpha, p[i]));
        const T ri = Num<T>::sub(r[i], Num<T>::mul(alpha, q[i]));
        r[i] = ri;
        acc[0] += Num<T>::abs2(ri);
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

template <typename T> __global__ void k_cg_pupdate(int n, T beta, const T* __restrict__ z, T* __restrict__ p) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS)
        p[i] = Num<T>::add(Num<T>::mul(p[i], beta), z[i]);
}

template <typename T> __global__ void k_residual(int n, const T* __restrict__ b, const T* __restrict__ q, T* __restrict__ r) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) r[i] = q ? Num<T>::sub(b[i], q[i]) : b[i];
}

__device__ inline double neg(double a) { return -1.0 * a; }
__device__ inline zc neg(zc a) { return a * -1.0; }
template <typename T> __global__ void k_negate(int n, const T* __restrict__ x, T* __restrict__ y) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) y[i] = neg(x[i]);
}

template <typename T>
__global__ __launch_bounds__(kBS) void k_mrtr_update(int n, int fix, T t, T zeta, T eta, const T* __restrict__ u,
                                                     const T* __restrict__ v, T* __restrict__ p, T* __restrict__ x,
                                                     T* __restrict__ y, T* __restrict__ r, double* __restrict__ partial) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T ui = u[i];
        const T pi = fix ? Num<T>::add(ui, Num<T>::mul(t, p[i])) : Num<T>::add(ui, Num<T>::mul(t, ui));
        p[i] = pi;
        x[i] = Num<T>::add(x[i], Num<T>::mul(zeta, pi));
        const T yi = Num<T>::add(Num<T>::mul(y[i], eta), Num<T>::mul(zeta, v[i]));
        y[i] = yi;
        const T ri = Num<T>::sub(r[i], yi);
        r[i] = ri;
        acc[0] += Num<T>::abs2(ri);
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

template <typename T>
__global__ void k_mrtr_update2(int n, T zeta, T eta, const T* __restrict__ w, T* __restrict__ z, T* __restrict__ u) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T zi = Num<T>::add(Num<T>::mul(z[i], eta), Num<T>::mul(zeta, w[i]));
        z[i] = zi;
        u[i] = Num<T>::sub(u[i], zi);
    }
}

template <typename T> __global__ void k_sub(int n, const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ c) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) c[i] = Num<T>::sub(a[i], b[i]);
}
template <typename T> __global__ void k_add_inplace(int n, const T* __restrict__ u, T* __restrict__ a) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) a[i] = Num<T>::add(a[i], u[i]);
}

template <typename T>
__global__ __launch_bounds__(kBS) void k_sgs_update(int n, T t2, T zeta, T eta, const T* __restrict__ u,
                                                    const T* __restrict__ ard, T* __restrict__ p, T* __restrict__ x,
                                                    T* __restrict__ y, T* __restrict__ rd, double* __restrict__ partial) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const T pi = Num<T>::add(Num<T>::mul(p[i], t2), u[i]);
        p[i] = pi;
        x[i] = Num<T>::add(x[i], Num<T>::mul(zeta, pi));
        const T yi = Num<T>::add(Num<T>::mul(y[i], eta), Num<T>::mul(zeta, ard[i]));
        y[i] = yi;
        const T ri = Num<T>::sub(rd[i], yi);
        rd[i] = ri;
        acc[0] += Num<T>::abs2(ri);
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

// ---------------------------------------------------------------- scaling
__device__ inline double diag_mag(double a) { return fabs(a); }
__device__ inline double diag_mag(zc a) { return a.re * a.re + a.im * a.im; }   // std::norm, then sqrt

template <typename T>
__global__ void k_diag_scale(int n, const int* __restrict__ diag_idx, const T* __restrict__ v, const T* __restrict__ b,
                             T* __restrict__ sc, T* __restrict__ bs) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const int k = diag_idx[i];
        double t2 = 0.0;
        if (k >= 0) {
            const double t1b = sqrt(diag_mag(v[k]));
            t2 = (t1b < 1.0e-12) ? 1.0 / DBL_EPSILON : 1.0 / t1b;
        }
        const T s = from_real<T>(t2);
        sc[i] = s;
        if (b) bs[i] = Num<T>::mul(b[i], s);
    }
}

__device__ inline double sval(double a) { return a; }
__device__ inline double sval(zc a) { return a.re; }
__device__ inline double scl(double v, double f) { return v * f; }
__device__ inline zc scl(zc v, double f) { return zc{v.re * f, v.im * f}; }

template <typename T>
__global__ void k_scale_matrix(int n, const int* __restrict__ rp, const int* __restrict__ ci, const T* __restrict__ v,
                               const T* __restrict__ sc, T* __restrict__ dst) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const double di = sval(sc[i]);
        for (int k = rp[i]; k < rp[i + 1]; ++k) dst[k] = scl(v[k], di * sval(sc[ci[k]]));
    }
}

template <typename T, typename F>
__global__ void k_gather(int64_t n, const int* __restrict__ map, const T* __restrict__ src, F* __restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
        dst[i] = narrow<F, T>(src[map[i]]);
}

template <typename T, typename F>
__global__ void k_narrow(int64_t n, const T* __restrict__ src, F* __restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
        dst[i] = narrow<F, T>(src[i]);
}

template <typename T> __global__ void k_div_scale(int n, const T* __restrict__ sc, T* __restrict__ x) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) x[i] = Num<T>::div(x[i], sc[i]);
}
template <typename T> __global__ void k_mul_scale(int n, const T* __restrict__ sc, T* __restrict__ x) {
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) x[i] = Num<T>::mul(x[i], sc[i]);
}

__device__ inline bool nonpos(double d) { return !(d > 0); }
__device__ inline bool nonpos(zc d) { return !(d.re > 0); }
template <typename T> __global__ void k_count_nonpos(int n, const T* __restrict__ d, int* count) {
    int c = 0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) c += nonpos(d[i]) ? 1 : 0;
    if (c) atomicAdd(count, c);
}

// ---------------------------------------------------------------- batched solves
// Every kernel below covers many systems: a workgroup reads its (system,
// block) item, leaves at once when the system's status is not `want`, and
// otherwise runs exactly what the one-system kernel runs for that block (the
// same rows, the same grid stride reduce_blocks(n) * kBS, the same tree).
template <typename T> __device__ inline T* bvec(const BatchSys& S, int k) { return static_cast<T*>(S.vec[k]); }
template <typename T> __device__ inline T from_parts(const double* d);
template <> __device__ inline double from_parts<double>(const double* d) { return d[0]; }
template <> __device__ inline zc from_parts<zc>(const double* d) { return zc{d[0], d[1]}; }
__device__ inline void to_parts(double v, double* d) { d[0] = v; d[1] = 0.0; }
__device__ inline void to_parts(zc v, double* d) { d[0] = v.re; d[1] = v.im; }
template <typename T> __device__ inline T sentinel_value();
template <> __device__ inline double sentinel_value<double>() { return __longlong_as_double((long long)kSweepSentinel); }
template <> __device__ inline zc sentinel_value<zc>() {
    const double s = __longlong_as_double((long long)kSweepSentinel);
    return zc{s, s};
}

template <typename T, int STAGE_BYTES>
__global__ __launch_bounds__(kBS) void k_batch_spmv(const int2* __restrict__ items, const BatchSys* __restrict__ sys,
                                                    const BatchState* __restrict__ st, int want, int in, int out,
                                                    int with_dot) {
    const int2 it = items[xcd_block((int)blockIdx.x, (int)gridDim.x)];
    if (st[it.x].status != want) return;
    const BatchSys& S = sys[it.x];
    const T* x = bvec<T>(S, in);
    spmv_block<T, STAGE_BYTES>(it.y, S.blk, S.rp, S.ci, static_cast<const T*>(S.val), x, bvec<T>(S, out),
                               with_dot ? x : nullptr, S.part);
}

template <typename T, int OP>
__global__ __launch_bounds__(kBS) void k_batch_vec(const int2* __restrict__ items, const BatchSys* __restrict__ sys,
                                                   const BatchState* __restrict__ st, int want, int a, int b, int c,
                                                   int* __restrict__ tickets) {
    if (tickets && blockIdx.x == 0 && threadIdx.x < 2) tickets[threadIdx.x] = 0;
    const int2 it = items[blockIdx.x];
    const BatchState& q = st[it.x];
    if (q.status != want) return;
    const BatchSys& S = sys[it.x];
    const int n = S.n, stride = S.rblk * kBS;
    const int i0 = it.y * kBS + (int)threadIdx.x;
    constexpr int K = Parts<T>::K;
    if (OP == BO_FILL0) {
        T* v = bvec<T>(S, a);
        for (int i = i0; i < n; i += stride) v[i] = Num<T>::zero();
    } else if (OP == BO_RESID) {
        const T* bb = bvec<T>(S, a);
        const T* qq = b >= 0 ? bvec<T>(S, b) : nullptr;
        T* r = bvec<T>(S, c);
        for (int i = i0; i < n; i += stride) r[i] = qq ? Num<T>::sub(bb[i], qq[i]) : bb[i];
    } else if (OP == BO_COPY) {
        const T* u = bvec<T>(S, a);
        T* v = bvec<T>(S, b);
        for (int i = i0; i < n; i += stride) v[i] = u[i];
    } else if (OP == BO_SENT) {
        const T sv = sentinel_value<T>();
        T* u = bvec<T>(S, a);
        T* v = bvec<T>(S, b);
        for (int i = i0; i < n; i += stride) {
            u[i] = sv;
            v[i] = sv;
        }
    } else if (OP == BO_NORM2) {
        const T* u = bvec<T>(S, a);
        double acc[1] = {0.0};
        for (int i = i0; i < n; i += stride) acc[0] += Num<T>::abs2(u[i]);
        block_sum<1>(acc);
        if (threadIdx.x == 0) S.part[it.y] = acc[0];
    } else if (OP == BO_DOT) {
        const T* u = bvec<T>(S, a);
        const T* v = bvec<T>(S, b);
        double acc[K] = {};
        for (int i = i0; i < n; i += stride) Parts<T>::put(acc, Num<T>::mul(u[i], v[i]));
        block_sum<K>(acc);
        if (threadIdx.x == 0)
            for (int k = 0; k < K; ++k) S.part[it.y * K + k] = acc[k];
    } else if (OP == BO_CG_UPDATE) {
        // as k_cg_update; the sweeps' outputs t and z are reset to the
        // not-ready sentinel on the way
        const T alpha = from_parts<T>(q.alpha);
        const T* p = bvec<T>(S, BV_P);
        const T* qv = bvec<T>(S, BV_Q);
        T* x = bvec<T>(S, BV_X);
        T* r = bvec<T>(S, BV_R);
        T* t = bvec<T>(S, BV_T);
        T* z = bvec<T>(S, BV_Z);
        const T sv = sentinel_value<T>();
        double acc[1] = {0.0};
        for (int i = i0; i < n; i += stride) {
            x[i] = Num<T>::add(x[i], Num<T>::mul(alpha, p[i]));
            const T ri = Num<T>::sub(r[i], Num<T>::mul(alpha, qv[i]));
            r[i] = ri;
            acc[0] += Num<T>::abs2(ri);
            t[i] = sv;
            z[i] = sv;
        }
        block_sum<1>(acc);
        if (threadIdx.x == 0) S.part[it.y] = acc[0];
    } else if (OP == BO_PUPDATE) {
        // as k_cg_pupdate; first the best-iterate copy the monitor asked for
        const T beta = from_parts<T>(q.beta);
        const T* z = bvec<T>(S, BV_Z);
        T* p = bvec<T>(S, BV_P);
        if (q.copy_best) {
            const T* x = bvec<T>(S, BV_X);
            T* best = bvec<T>(S, BV_BEST);
            for (int i = i0; i < n; i += stride) best[i] = x[i];
        }
        for (int i = i0; i < n; i += stride) p[i] = Num<T>::add(Num<T>::mul(p[i], beta), z[i]);
    } else if (OP == BO_NEG) {
        const T* u = bvec<T>(S, a);
        T* v = bvec<T>(S, b);
        for (int i = i0; i < n; i += stride) v[i] = neg(u[i]);
    } else if (OP == BO_DOT3) {
        // as k_dot3 (w; r, v, y), y from the second iteration on (MatSolverMRTR_ICver.cpp:276-281)
        const T* w = bvec<T>(S, BV_W);
        const T* r = bvec<T>(S, BV_R);
        const T* v = bvec<T>(S, BV_Q);
        const T* y = q.iterations > 0 ? bvec<T>(S, BV_Y) : nullptr;
        double acc[3 * K] = {};
        for (int i = i0; i < n; i += stride) {
            const T wi = w[i];
            Parts<T>::put(acc, Num<T>::mul(wi, r[i]));
            Parts<T>::put(acc + K, Num<T>::mul(wi, v[i]));
            if (y) Parts<T>::put(acc + 2 * K, Num<T>::mul(wi, y[i]));
        }
        block_sum<3 * K>(acc);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3 * K; ++k) S.part[it.y * 3 * K + k] = acc[k];
    } else if (OP == BO_MRTR_UPDATE) {
        // as k_mrtr_update on the internal iterate xi
        const T tt = from_parts<T>(q.t), zeta = from_parts<T>(q.zeta), eta = from_parts<T>(q.eta);
        const int fix = q.mrtr_fix;
        const T* u = bvec<T>(S, BV_U);
        const T* v = bvec<T>(S, BV_Q);
        T* p = bvec<T>(S, BV_P);
        T* x = bvec<T>(S, BV_XI);
        T* y = bvec<T>(S, BV_Y);
        T* r = bvec<T>(S, BV_R);
        double acc[1] = {0.0};
        for (int i = i0; i < n; i += stride) {
            const T ui = u[i];
            const T pi = fix ? Num<T>::add(ui, Num<T>::mul(tt, p[i])) : Num<T>::add(ui, Num<T>::mul(tt, ui));
            p[i] = pi;
            x[i] = Num<T>::add(x[i], Num<T>::mul(zeta, pi));
            const T yi = Num<T>::add(Num<T>::mul(y[i], eta), Num<T>::mul(zeta, v[i]));
            y[i] = yi;
            const T ri = Num<T>::sub(r[i], yi);
            r[i] = ri;
            acc[0] += Num<T>::abs2(ri);
        }
        block_sum<1>(acc);
        if (threadIdx.x == 0) S.part[it.y] = acc[0];
    } else if (OP == BO_MRTR_UPDATE2) {
        // the best-iterate copy the monitor asked for, then k_mrtr_update2; the
        // sweeps' outputs t and w go back to the sentinel once w is read
        const T zeta = from_parts<T>(q.zeta), eta = from_parts<T>(q.eta);
        if (q.copy_best) {
            const T* x = bvec<T>(S, BV_XI);
            T* best = bvec<T>(S, BV_BEST);
            for (int i = i0; i < n; i += stride) best[i] = x[i];
        }
        const T sv = sentinel_value<T>();
        T* w = bvec<T>(S, BV_W);
        T* z = bvec<T>(S, BV_Z);
        T* u = bvec<T>(S, BV_U);
        T* t = bvec<T>(S, BV_T);
        for (int i = i0; i < n; i += stride) {
            const T zi = Num<T>::add(Num<T>::mul(z[i], eta), Num<T>::mul(zeta, w[i]));
            z[i] = zi;
            u[i] = Num<T>::sub(u[i], zi);
            w[i] = sv;
            t[i] = sv;
        }
    } else if (OP == BO_SGS_SUB) {
        // SGS-MRTR, between the backward and the forward sweep: the best-iterate
        // copy the monitor asked for, the forward sweep's output q back to the
        // sentinel, and with a != 0 t = r - u as k_sub (Eisenstat's splitting)
        if (q.copy_best) {
            const T* x = bvec<T>(S, BV_XI);
            T* best = bvec<T>(S, BV_BEST);
            for (int i = i0; i < n; i += stride) best[i] = x[i];
        }
        const T sv = sentinel_value<T>();
        const T* r = bvec<T>(S, BV_R);
        const T* u = bvec<T>(S, BV_U);
        T* t = bvec<T>(S, BV_T);
        T* v = bvec<T>(S, BV_Q);
        for (int i = i0; i < n; i += stride) {
            if (a) t[i] = Num<T>::sub(r[i], u[i]);
            v[i] = sv;
        }
    } else if (OP == BO_SGS_DOT3) {
        // with a != 0 q += u as k_add_inplace, then k_dot3 (q; r, q, y), y from
        // the second iteration on (MatSolverMRTR_SGSver.cpp:186-204)
        const T* u = bvec<T>(S, BV_U);
        const T* r = bvec<T>(S, BV_R);
        const T* y = q.iterations > 0 ? bvec<T>(S, BV_Y) : nullptr;
        T* v = bvec<T>(S, BV_Q);
        double acc[3 * K] = {};
        for (int i = i0; i < n; i += stride) {
            T vi = v[i];
            if (a) {
                vi = Num<T>::add(vi, u[i]);
                v[i] = vi;
            }
            Parts<T>::put(acc, Num<T>::mul(vi, r[i]));
            Parts<T>::put(acc + K, Num<T>::mul(vi, vi));
            if (y) Parts<T>::put(acc + 2 * K, Num<T>::mul(vi, y[i]));
        }
        block_sum<3 * K>(acc);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3 * K; ++k) S.part[it.y * 3 * K + k] = acc[k];
    } else if (OP == BO_SGS_UPDATE) {
        // as k_sgs_update (MatSolverMRTR_SGSver.cpp:216-229) on the scaled
        // iterate xi; the backward sweep's output u goes back to the sentinel
        // once it is read
        const T tt = from_parts<T>(q.t), zeta = from_parts<T>(q.zeta), eta = from_parts<T>(q.eta);
        const T sv = sentinel_value<T>();
        const T* v = bvec<T>(S, BV_Q);
        T* u = bvec<T>(S, BV_U);
        T* p = bvec<T>(S, BV_P);
        T* x = bvec<T>(S, BV_XI);
        T* y = bvec<T>(S, BV_Y);
        T* r = bvec<T>(S, BV_R);
        double acc[1] = {0.0};
        for (int i = i0; i < n; i += stride) {
            const T pi = Num<T>::add(Num<T>::mul(p[i], tt), u[i]);
            p[i] = pi;
            u[i] = sv;
            x[i] = Num<T>::add(x[i], Num<T>::mul(zeta, pi));
            const T yi = Num<T>::add(Num<T>::mul(y[i], eta), Num<T>::mul(zeta, v[i]));
            y[i] = yi;
            const T ri = Num<T>::sub(r[i], yi);
            r[i] = ri;
            acc[0] += Num<T>::abs2(ri);
        }
        block_sum<1>(acc);
        if (threadIdx.x == 0) S.part[it.y] = acc[0];
    }
}

// One workgroup per system: the fixed-tree finalize of the system's partials
// (as k_finalize), then the scalar step of the reference's loop for that
// system, in FP64 / complex FP64 on the device (the same operations the
// one-system drivers evaluate on the host: MatSolverICCG.cpp:212-237 for the
// prologue, :259-306 for the loop).
template <typename T, int PH>
__global__ __launch_bounds__(kBS) void k_batch_finalize(const int* __restrict__ fin, const BatchSys* __restrict__ sys,
                                                        BatchState* __restrict__ st, int want, int* __restrict__ active,
                                                        int* __restrict__ tickets) {
    if ((PH == BP_PROLOGUE || PH == BP_MONITOR || PH == BP_ACTIVATE || PH == BP_NORMB) && blockIdx.x == 0 &&
        threadIdx.x < 2)
        tickets[threadIdx.x] = 0;
    const int s = fin[blockIdx.x];
    if (st[s].status != want) return;
    const BatchSys& S = sys[s];
    constexpr int K = Parts<T>::K;
    constexpr int KP = (PH == BP_PROLOGUE || PH == BP_MONITOR || PH == BP_NORMB) ? 1 : PH == BP_MRTR_COEF ? 3 * K : K;
    __shared__ double sum[6];
    if (PH != BP_ACTIVATE) finalize_parts(PH == BP_ALPHA ? S.nblk : S.rblk, KP, S.part, sum);
    if (threadIdx.x != 0) return;
    BatchState& q = st[s];
    constexpr double kSmallAbs = 1.0e-20;
    if (PH == BP_PROLOGUE) {
        const double pure = sqrt(sum[0]);
        const double normB = q.normB, conv = q.conv_cri;
        if (normB <= 1.0e-30 && pure <= kSmallAbs) {
            q.status = BS_EARLY;
            atomicSub(active, 1);
            return;
        }
        q.abs_conv = (normB * conv * 0.9 < kSmallAbs ? kSmallAbs : normB * conv * 0.9);
        const double first = pure / normB;
        if (first < conv * 0.5 || pure < q.abs_conv * 0.5) {
            q.status = BS_EARLY;
            atomicSub(active, 1);
            return;
        }
        if (q.normalize_type == 1) q.normalizer = pure;
        else if (q.normalize_type == 2) q.normalizer = q.normalize_const;
        else q.normalizer = normB;
        q.status = BS_PRIMED;
    } else if (PH == BP_NORMB) {
        q.normB = sqrt(sum[0]);   // as Solver::norm of the scaled b (the status stays pending)
    } else if (PH == BP_RZ0) {
        q.rz[0] = sum[0];
        q.rz[1] = KP > 1 ? sum[1] : 0.0;
        if (q.max_ite <= 0) {
            q.status = BS_STOPPED;
            atomicSub(active, 1);
        } else {
            q.status = BS_ACTIVE;
        }
    } else if (PH == BP_ALPHA) {
        to_parts(Num<T>::div(from_parts<T>(q.rz), from_parts<T>(sum)), q.alpha);
    } else if (PH == BP_BETA) {
        const T rz1 = from_parts<T>(sum);
        to_parts(Num<T>::div(rz1, from_parts<T>(q.rz)), q.beta);
        to_parts(rz1, q.rz);
    } else if (PH == BP_ACTIVATE) {
        if (q.max_ite <= 0) {
            q.status = BS_STOPPED;
            atomicSub(active, 1);
        } else {
            q.status = BS_ACTIVE;
        }
    } else if (PH == BP_MRTR_COEF) {
        // the scalar recurrence of MatSolverMRTR_ICver.cpp:279-300, in the
        // one-system driver's operation order
        const T w_r = from_parts<T>(sum), v_w = from_parts<T>(sum + K);
        T zeta_old = from_parts<T>(q.zeta_old), nu = from_parts<T>(q.nu), zeta, eta;
        if (q.iterations == 0) {
            zeta = Num<T>::div(w_r, v_w);
            zeta_old = zeta;
            eta = Num<T>::zero();
        } else {
            const T w_y = from_parts<T>(sum + 2 * K);
            const T t0 = Num<T>::div(from_real<T>(1.0), Num<T>::sub(Num<T>::mul(nu, v_w), Num<T>::mul(w_y, w_y)));
            zeta = Num<T>::mul(Num<T>::mul(nu, w_r), t0);
            eta = Num<T>::mul(Num<T>::mul(neg(w_y), w_r), t0);
        }
        nu = Num<T>::mul(zeta, w_r);
        const T tt = Num<T>::div(Num<T>::mul(eta, zeta_old), zeta);
        zeta_old = zeta;
        to_parts(zeta, q.zeta);
        to_parts(zeta_old, q.zeta_old);
        to_parts(eta, q.eta);
        to_parts(nu, q.nu);
        to_parts(tt, q.t);
    } else {   // BP_MONITOR
        const double pn = sqrt(sum[0]);
        q.iterations += 1;
        const double normR = pn / q.normalizer;
        q.final_res = normR;
        q.copy_best = 0;
        if (normR < q.conv_cri || pn < q.abs_conv) {
            q.status = BS_CONVERGED;
            atomicSub(active, 1);
            return;
        }
        if (normR < q.best) {
            q.best = normR;
            if (q.save_best) q.copy_best = 1;
        }
        if (q.diverge_type == 1) {
            if (normR < q.best * q.bad_div_val) q.bad = 0;
            if (normR >= q.best * q.bad_div_val) q.bad++;
            if (q.bad >= q.bad_div_thres) {
                q.status = BS_STOPPED;
                atomicSub(active, 1);
                return;
            }
        }
        if (q.iterations >= q.max_ite) {
            q.status = BS_STOPPED;
            atomicSub(active, 1);
        }
    }
}

inline int grid_for(int64_t n, int cap = 4096) {
    int64_t g = (n + kBS - 1) / kBS;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

int reduce_blocks(int64_t n) { return grid_for(n, kMaxReduceBlocks); }

void launch_finalize(hipStream_t s, int nb, int K, const double* partial, double* out) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, s, nb, K, partial, out);
}

template <typename T>
void launch_spmv(hipStream_t s, int nblk, const int* blk, const int* rp, const int* ci, const T* v, const T* x, T* y,
                 const T* dotv, double* partial, int stage_kb) {
    if (stage_kb >= 32)
        hipLaunchKernelGGL((k_spmv<T, 32768>), dim3(nblk), dim3(kBS), 0, s, blk, rp, ci, v, x, y, dotv, partial);
    else if (stage_kb >= 16)
        hipLaunchKernelGGL((k_spmv<T, 16384>), dim3(nblk), dim3(kBS), 0, s, blk, rp, ci, v, x, y, dotv, partial);
    else
        hipLaunchKernelGGL((k_spmv<T, 8192>), dim3(nblk), dim3(kBS), 0, s, blk, rp, ci, v, x, y, dotv, partial);
}

template <typename T>
void launch_batch_spmv(hipStream_t s, int nitems, const int2* items, const BatchSys* sys, const BatchState* st,
                       int want, int in, int out, int with_dot, GraphRec* rec) {
    if (nitems <= 0) return;
    launch_or_record(rec, k_batch_spmv<T, 16384>, dim3(nitems), dim3(kBS), s, items, sys, st, want, in, out,
                     with_dot);
}

template <typename T>
void launch_batch_vec(hipStream_t s, int op, int nitems, const int2* items, const BatchSys* sys, BatchState* st,
                      int want, int a, int b, int c, int* tickets, GraphRec* rec) {
    if (nitems <= 0) return;
    const BatchState* cst = st;
#define SSGPU_BV(OP) \
    launch_or_record(rec, k_batch_vec<T, OP>, dim3(nitems), dim3(kBS), s, items, sys, cst, want, a, b, c, tickets)
    switch (op) {
        case BO_NEG: SSGPU_BV(BO_NEG); break;
        case BO_DOT3: SSGPU_BV(BO_DOT3); break;
        case BO_MRTR_UPDATE: SSGPU_BV(BO_MRTR_UPDATE); break;
        case BO_MRTR_UPDATE2: SSGPU_BV(BO_MRTR_UPDATE2); break;
        case BO_SGS_SUB: SSGPU_BV(BO_SGS_SUB); break;
        case BO_SGS_DOT3: SSGPU_BV(BO_SGS_DOT3); break;
        case BO_SGS_UPDATE: SSGPU_BV(BO_SGS_UPDATE); break;
        case BO_FILL0: SSGPU_BV(BO_FILL0); break;
        case BO_RESID: SSGPU_BV(BO_RESID); break;
        case BO_NORM2: SSGPU_BV(BO_NORM2); break;
        case BO_DOT: SSGPU_BV(BO_DOT); break;
        case BO_COPY: SSGPU_BV(BO_COPY); break;
        case BO_SENT: SSGPU_BV(BO_SENT); break;
        case BO_CG_UPDATE: SSGPU_BV(BO_CG_UPDATE); break;
        default: SSGPU_BV(BO_PUPDATE); break;
    }
#undef SSGPU_BV
}

template <typename T>
void launch_batch_finalize(hipStream_t s, int phase, int nfin, const int* fin, const BatchSys* sys, BatchState* st,
                           int want, int* active, int* tickets, GraphRec* rec) {
    if (nfin <= 0) return;
#define SSGPU_BF(PH) \
    launch_or_record(rec, k_batch_finalize<T, PH>, dim3(nfin), dim3(kBS), s, fin, sys, st, want, active, tickets)
    switch (phase) {
        case BP_PROLOGUE: SSGPU_BF(BP_PROLOGUE); break;
        case BP_RZ0: SSGPU_BF(BP_RZ0); break;
        case BP_ALPHA: SSGPU_BF(BP_ALPHA); break;
        case BP_BETA: SSGPU_BF(BP_BETA); break;
        case BP_ACTIVATE: SSGPU_BF(BP_ACTIVATE); break;
        case BP_MRTR_COEF: SSGPU_BF(BP_MRTR_COEF); break;
        case BP_NORMB: SSGPU_BF(BP_NORMB); break;
        default: SSGPU_BF(BP_MONITOR); break;
    }
#undef SSGPU_BF
}

template <typename T, typename F>
void launch_sweep(hipStream_t s, int mode, int p0, int p1, const int* prp, const int* pci, const F* pv,
                  const int* prow, const T* rhs, const T* aux, const T* dd, T* x) {
    const int nr = p1 - p0;
    if (nr <= 0) return;
    const dim3 g((nr + kBS - 1) / kBS);
    if (mode == FWD_DIV_LAST)
        hipLaunchKernelGGL((k_sweep<T, F, FWD_DIV_LAST>), g, dim3(kBS), 0, s, p0, p1, prp, pci, pv, prow, rhs, aux, dd, x);
    else if (mode == BWD_IC)
        hipLaunchKernelGGL((k_sweep<T, F, BWD_IC>), g, dim3(kBS), 0, s, p0, p1, prp, pci, pv, prow, rhs, aux, dd, x);
    else
        hipLaunchKernelGGL((k_sweep<T, F, BWD_DIV_FIRST>), g, dim3(kBS), 0, s, p0, p1, prp, pci, pv, prow, rhs, aux, dd, x);
}

template <typename T>
void launch_ic_level(hipStream_t s, int p0, int p1, const int* prp, const int* pci, T* pv, const int* prow,
                     const int* ppos, T* diagD, double accel, int* nfix) {
    const int nr = p1 - p0;
    if (nr <= 0) return;
    hipLaunchKernelGGL(k_ic_level<T>, dim3((nr + kBS - 1) / kBS), dim3(kBS), 0, s, p0, p1, prp, pci, pv, prow, ppos,
                       diagD, accel, nfix);
}

template <typename T> void launch_dot(hipStream_t s, int n, const T* a, const T* b, double* partial) {
    hipLaunchKernelGGL(k_dot<T>, dim3(reduce_blocks(n)), dim3(kBS), 0, s, n, a, b, partial);
}
template <typename T> void launch_norm2(hipStream_t s, int n, const T* a, double* partial) {
    hipLaunchKernelGGL(k_norm2<T>, dim3(reduce_blocks(n)), dim3(kBS), 0, s, n, a, partial);
}
template <typename T>
void launch_dot3(hipStream_t s, int n, const T* a, const T* b, const T* c, const T* d, double* partial) {
    hipLaunchKernelGGL(k_dot3<T>, dim3(reduce_blocks(n)), dim3(kBS), 0, s, n, a, b, c, d, partial);
}

template <typename T> void launch_fill(hipStream_t s, int64_t n, T* x, T val) {
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, x, val);
}
template <typename T> void launch_copy(hipStream_t s, int64_t n, const T* a, T* b) {
    hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, a, b);
}
template <typename T>
void launch_cg_update(hipStream_t s, int n, T alpha, const T* p, const T* q, T* x, T* r, double* partial) {
    hipLaunchKernelGGL(k_cg_update<T>, dim3(reduce_blocks(n)), dim3(kBS), 0, s, n, alpha, p, q, x, r, partial);
}
template <typename T> void launch_cg_pupdate(hipStream_t s, int n, T beta, const T* z, T* p) {
    hipLaunchKernelGGL(k_cg_pupdate<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, beta, z, p);
}
template <typename T> void launch_residual(hipStream_t s, int n, const T* b, const T* q, T* r) {
    hipLaunchKernelGGL(k_residual<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, b, q, r);
}
template <typename T> void launch_negate(hipStream_t s, int n, const T* x, T* y) {
    hipLaunchKernelGGL(k_negate<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, x, y);
}
template <typename T>
void launch_mrtr_update(hipStream_t s, int n, int fix, T t, T zeta, T eta, const T* u, const T* v, T* p, T* x, T* y,
                        T* r, double* partial) {
    hipLaunchKernelGGL(k_mrtr_update<T>, dim3(reduce_blocks(n)), dim3(kBS), 0, s, n, fix, t, zeta, eta, u, v, p, x, y,
                       r, partial);
}
template <typename T> void launch_mrtr_update2(hipStream_t s, int n, T zeta, T eta, const T* w, T* z, T* u) {
    hipLaunchKernelGGL(k_mrtr_update2<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, zeta, eta, w, z, u);
}
template <typename T> void launch_sub(hipStream_t s, int n, const T* a, const T* b, T* c) {
    hipLaunchKernelGGL(k_sub<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, a, b, c);
}
template <typename T> void launch_add_inplace(hipStream_t s, int n, const T* u, T* a) {
    hipLaunchKernelGGL(k_add_inplace<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, u, a);
}
template <typename T>
void launch_sgs_update(hipStream_t s, int n, T t2, T zeta, T eta, const T* u, const T* ard, T* p, T* x, T* y, T* rd,
                       double* partial) {
    hipLaunchKernelGGL(k_sgs_update<T>, dim3(reduce_blocks(n)), dim3(kBS), 0, s, n, t2, zeta, eta, u, ard, p, x, y, rd,
                       partial);
}
template <typename T>
void launch_diag_scale(hipStream_t s, int n, const int* diag_idx, const T* v, const T* b, T* sc, T* bs) {
    hipLaunchKernelGGL(k_diag_scale<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, diag_idx, v, b, sc, bs);
}
template <typename T>
void launch_scale_matrix(hipStream_t s, int n, const int* rp, const int* ci, const T* v, const T* sc, T* dst) {
    hipLaunchKernelGGL(k_scale_matrix<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, rp, ci, v, sc, dst);
}
template <typename T, typename F> void launch_gather(hipStream_t s, int64_t n, const int* map, const T* src, F* dst) {
    hipLaunchKernelGGL((k_gather<T, F>), dim3(grid_for(n)), dim3(kBS), 0, s, n, map, src, dst);
}
template <typename T, typename F> void launch_narrow(hipStream_t s, int64_t n, const T* src, F* dst) {
    hipLaunchKernelGGL((k_narrow<T, F>), dim3(grid_for(n)), dim3(kBS), 0, s, n, src, dst);
}
template <typename T> void launch_div_scale(hipStream_t s, int n, const T* sc, T* x) {
    hipLaunchKernelGGL(k_div_scale<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, sc, x);
}
template <typename T> void launch_mul_scale(hipStream_t s, int n, const T* sc, T* x) {
    hipLaunchKernelGGL(k_mul_scale<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, sc, x);
}
template <typename T> void launch_count_nonpos(hipStream_t s, int n, const T* d, int* count) {
    hipLaunchKernelGGL(k_count_nonpos<T>, dim3(grid_for(n)), dim3(kBS), 0, s, n, d, count);
}

// explicit instantiations
#define SSGPU_INST(T)                                                                                          \
    template void launch_spmv<T>(hipStream_t, int, const int*, const int*, const int*, const T*, const T*, T*,  \
                                 const T*, double*, int);                                                      \
    template void launch_ic_level<T>(hipStream_t, int, int, const int*, const int*, T*, const int*, const int*, \
                                     T*, double, int*);                                                        \
    template void launch_dot<T>(hipStream_t, int, const T*, const T*, double*);                                \
    template void launch_norm2<T>(hipStream_t, int, const T*, double*);                                        \
    template void launch_dot3<T>(hipStream_t, int, const T*, const T*, const T*, const T*, double*);           \
    template void launch_fill<T>(hipStream_t, int64_t, T*, T);                                                 \
    template void launch_copy<T>(hipStream_t, int64_t, const T*, T*);                                          \
    template void launch_cg_update<T>(hipStream_t, int, T, const T*, const T*, T*, T*, double*);               \
    template void launch_cg_pupdate<T>(hipStream_t, int, T, const T*, T*);                                     \
    template void launch_residual<T>(hipStream_t, int, const T*, const T*, T*);                                \
    template void launch_negate<T>(hipStream_t, int, const T*, T*);                                            \
    template void launch_mrtr_update<T>(hipStream_t, int, int, T, T, T, const T*, const T*, T*, T*, T*, T*,     \
                                        double*);                                                              \
    template void launch_mrtr_update2<T>(hipStream_t, int, T, T, const T*, T*, T*);                            \
    template void launch_sub<T>(hipStream_t, int, const T*, const T*, T*);                                     \
    template void launch_add_inplace<T>(hipStream_t, int, const T*, T*);                                       \
    template void launch_sgs_update<T>(hipStream_t, int, T, T, T, const T*, const T*, T*, T*, T*, T*, double*); \
    template void launch_diag_scale<T>(hipStream_t, int, const int*, const T*, const T*, T*, T*);              \
    template void launch_scale_matrix<T>(hipStream_t, int, const int*, const int*, const T*, const T*, T*);    \
    template void launch_div_scale<T>(hipStream_t, int, const T*, T*);                                         \
    template void launch_mul_scale<T>(hipStream_t, int, const T*, T*);                                         \
    template void launch_count_nonpos<T>(hipStream_t, int, const T*, int*);

SSGPU_INST(double)
SSGPU_INST(zc)
#define SSGPU_BINST(T)                                                                                      \
    template void launch_batch_spmv<T>(hipStream_t, int, const int2*, const BatchSys*, const BatchState*, int, int, \
                                       int, int, GraphRec*);                                                    \
    template void launch_batch_vec<T>(hipStream_t, int, int, const int2*, const BatchSys*, BatchState*, int, int, \
                                      int, int, int*, GraphRec*);                                               \
    template void launch_batch_finalize<T>(hipStream_t, int, int, const int*, const BatchSys*, BatchState*, int, \
                                           int*, int*, GraphRec*);
SSGPU_BINST(double)
SSGPU_BINST(zc)
template void launch_sweep<double, double>(hipStream_t, int, int, int, const int*, const int*, const double*,
                                           const int*, const double*, const double*, const double*, double*);
template void launch_sweep<double, float>(hipStream_t, int, int, int, const int*, const int*, const float*,
                                          const int*, const double*, const double*, const double*, double*);
template void launch_sweep<zc, zc>(hipStream_t, int, int, int, const int*, const int*, const zc*, const int*,
                                   const zc*, const zc*, const zc*, zc*);
template void launch_sweep<zc, fc>(hipStream_t, int, int, int, const int*, const int*, const fc*, const int*,
                                   const zc*, const zc*, const zc*, zc*);
template void launch_narrow<double, float>(hipStream_t, int64_t, const double*, float*);
template void launch_narrow<zc, fc>(hipStream_t, int64_t, const zc*, fc*);
template void launch_fill<int>(hipStream_t, int64_t, int*, int);
template void launch_gather<double, double>(hipStream_t, int64_t, const int*, const double*, double*);
template void launch_gather<double, float>(hipStream_t, int64_t, const int*, const double*, float*);
template void launch_gather<zc, zc>(hipStream_t, int64_t, const int*, const zc*, zc*);
template void launch_gather<zc, fc>(hipStream_t, int64_t, const int*, const zc*, fc*);

}  // namespace ssgpu
