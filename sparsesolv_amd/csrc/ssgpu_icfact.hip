// Shifted IC(0) factorization on gfx950 with a precomputed update list.
//
// The reference computes each entry of row i of lower(A) as
//   L_ij = a_ij - sum_{k<j} L_ik L_jk d_k      (j < i)
//   L_ii = alpha a_ii - sum_{k<i} L_ik^2 d_k,  d_i = 1/L_ii
// by searching row j for every k of row i (SprsPreconditioner.cpp:59-110).
// Which (k) match depends only on the sparsity pattern, so a symbolic pass
// records, for every entry, its list of updates in the reference's order:
// (slot of L_ik within row i, position of L_jk, column k).  The symbolic pass
// needs no ordering between rows and runs as one launch over all rows.  The
// numeric pass then runs level by level (rows of a level are independent), a
// thread per row; each update list is consumed in batches whose loads are all
// in flight together, and the subtractions are applied in list order, so the
// factor is bit-identical to the reference's.
#include <cfloat>

#include <hipcub/hipcub.hpp>

#include "ssgpu_kernels.h"

namespace ssgpu {

namespace {

constexpr int kFactBS = 128;     // threads per block of the numeric kernel
constexpr int kFactBatch = 8;    // updates fetched per batch

// count (po == nullptr) or fill the update list of every entry of row p
__global__ __launch_bounds__(kBS) void k_ic_symbolic(int n, const int* __restrict__ prp, const int* __restrict__ pci,
                                                     const int* __restrict__ prow, const int* __restrict__ ppos,
                                                     int* __restrict__ cnt, const int* __restrict__ po,
                                                     unsigned char* __restrict__ u_slot, int* __restrict__ u_pos,
                                                     int* __restrict__ u_col) {
    const int p = blockIdx.x * kBS + threadIdx.x;
    if (p >= n) return;
    const int i = prow[p];
    const int rs = prp[p], re = prp[p + 1];
    for (int jj = rs; jj < re; ++jj) {
        const int j = pci[jj];
        int c = 0;
        const int base = po ? po[jj] : 0;
        if (j < i) {
            const int q = ppos[j];
            int ll = prp[q];
            const int le = prp[q + 1];
            for (int kk = rs; kk < re; ++kk) {
                const int ck = pci[kk];
                if (!(ck <= j - 1)) break;
                while (ll < le && pci[ll] < ck) ++ll;
                if (ll < le && pci[ll] == ck) {
                    if (po) {
                        u_slot[base + c] = (unsigned char)(kk - rs);
                        u_pos[base + c] = ll;
                        u_col[base + c] = ck;
                    }
                    ++c;
                }
            }
        } else if (j == i) {
            for (int kk = rs; kk < jj; ++kk) {
                if (!(pci[kk] <= i - 1)) break;
                if (po) {
                    u_slot[base + c] = (unsigned char)(kk - rs);
                    u_pos[base + c] = kk;
                    u_col[base + c] = pci[kk];
                }
                ++c;
            }
        }
        if (!po) cnt[jj] = c;
    }
}

__device__ inline double absval(double a) { return fabs(a); }
__device__ inline double shfl_idx(double v, int src) { return __shfl(v, src, 64); }
__device__ inline zc shfl_idx(zc v, int src) { return zc{__shfl(v.re, src, 64), __shfl(v.im, src, 64)}; }
__device__ inline double absval(zc a) { return hypot(a.re, a.im); }
__device__ inline double mulr(double v, double a) { return v * a; }
__device__ inline zc mulr(zc v, double a) { return v * a; }
__device__ inline double recip(double s) { return 1.0 / s; }
__device__ inline zc recip(zc s) { return zdiv(zc{1.0, 0.0}, s); }
template <typename T> __device__ inline T from_real(double a);
template <> __device__ inline double from_real<double>(double a) { return a; }
template <> __device__ inline zc from_real<zc>(double a) { return zc{a, 0.0}; }

// rows [p0, p1) of one level; fv holds lower(A) on entry and the factor on exit
template <typename T, int KROW>
__global__ __launch_bounds__(kFactBS) void k_ic_numeric(int p0, int p1, const int* __restrict__ prp,
                                                        const int* __restrict__ pci, const int* __restrict__ prow,
                                                        const int* __restrict__ po,
                                                        const unsigned char* __restrict__ u_slot,
                                                        const int* __restrict__ u_pos, const int* __restrict__ u_col,
                                                        T* __restrict__ fv, T* __restrict__ diagD, double accel,
                                                        int* __restrict__ nfix) {
    __shared__ T own_s[KROW][kFactBS];   // this row's finished entries, slot-major (no bank conflicts)
    const int p = p0 + blockIdx.x * kFactBS + threadIdx.x;
    if (p >= p1) return;
    const int i = prow[p];
    const int rs = prp[p], re = prp[p + 1];
    for (int jj = rs; jj < re; ++jj) {
        const int j = pci[jj];
        T s = (j == i) ? mulr(fv[jj], accel) : fv[jj];
        const int u0 = po[jj], u1 = po[jj + 1];
        for (int u = u0; u < u1; u += kFactBatch) {
            T lj[kFactBatch], dk[kFactBatch];
            int sl[kFactBatch];
#pragma unroll
            for (int b = 0; b < kFactBatch; ++b)
                if (u + b < u1) {
                    sl[b] = u_slot[u + b];
                    lj[b] = fv[u_pos[u + b]];
                    dk[b] = diagD[u_col[u + b]];
                }
#pragma unroll
            for (int b = 0; b < kFactBatch; ++b)
                if (u + b < u1) {
                    const T lik = own_s[sl[b]][threadIdx.x];
                    // (L_ik * L_jk) * d_k, SprsPreconditioner.cpp:75 / :98
                    s = Num<T>::sub(s, Num<T>::mul(Num<T>::mul(lik, (j == i) ? lik : lj[b]), dk[b]));
                }
        }
        fv[jj] = s;
        own_s[jj - rs][threadIdx.x] = s;
        if (j == i) {
            if (absval(s) < 1.0e-12) {
                diagD[i] = from_real<T>(1.0 / DBL_EPSILON);
                atomicAdd(nfix, 1);
            } else {
                diagD[i] = recip(s);
            }
        }
    }
}

// Wave per row: the row's update list (at most 64 * kWaveUpd entries) and
// its gathers of L_jk and d_k are loaded by the lanes at once (one memory
// round trip per row instead of one per entry); then, entry by entry, the
// lanes holding that entry's updates form their products in parallel and
// lane 0 subtracts them in list order, so the factor is bit-identical to the
// thread-per-row kernel's.  One wave per workgroup (its barriers are wave
// barriers).
constexpr int kWaveUpd = 4;   // updates per lane

template <typename T, int KROW>
__global__ __launch_bounds__(64) void k_ic_numeric_wave(int p0, int p1, const int* __restrict__ prp,
                                                        const int* __restrict__ pci, const int* __restrict__ prow,
                                                        const int* __restrict__ po,
                                                        const unsigned char* __restrict__ u_slot,
                                                        const int* __restrict__ u_pos, const int* __restrict__ u_col,
                                                        T* __restrict__ fv, T* __restrict__ diagD, double accel,
                                                        int* __restrict__ nfix) {
    __shared__ T own_s[KROW];            // the row's finished entries
    __shared__ T prod_s[64 * kWaveUpd];  // products of the current entry's updates
    const int p = p0 + blockIdx.x;
    if (p >= p1) return;
    const int lane = threadIdx.x;
    const int i = prow[p];
    const int rs = prp[p], re = prp[p + 1];
    const int ub = po[rs];
    const int nu = po[re] - ub;
    // gathers of the whole row, issued together
    int sl[kWaveUpd];
    T lj[kWaveUpd], dk[kWaveUpd];
#pragma unroll
    for (int q = 0; q < kWaveUpd; ++q) {
        const int t = q * 64 + lane;
        if (t < nu) {
            sl[q] = u_slot[ub + t];
            lj[q] = fv[u_pos[ub + t]];
            dk[q] = diagD[u_col[ub + t]];
        }
    }
    T a = Num<T>::zero();
    if (lane < re - rs) a = fv[rs + lane];
    for (int jj = rs; jj < re; ++jj) {
        const int j = pci[jj];
        const int t0 = po[jj] - ub, t1 = po[jj + 1] - ub;
        // products of this entry's updates (own-row factors are final: k < j)
#pragma unroll
        for (int q = 0; q < kWaveUpd; ++q) {
            const int t = q * 64 + lane;
            if (t >= t0 && t < t1) {
                const T lik = own_s[sl[q]];
                prod_s[t - t0] = Num<T>::mul(Num<T>::mul(lik, (j == i) ? lik : lj[q]), dk[q]);
            }
        }
        const T aj = shfl_idx(a, jj - rs);
        __syncthreads();
        if (lane == 0) {
            T sum = (j == i) ? mulr(aj, accel) : aj;
            for (int t = 0; t < t1 - t0; ++t) sum = Num<T>::sub(sum, prod_s[t]);   // SprsPreconditioner.cpp:75 / :98
            fv[jj] = sum;
            own_s[jj - rs] = sum;
            if (j == i) {
                if (absval(sum) < 1.0e-12) {
                    diagD[i] = from_real<T>(1.0 / DBL_EPSILON);
                    atomicAdd(nfix, 1);
                } else {
                    diagD[i] = recip(sum);
                }
            }
        }
        __syncthreads();
    }
}

// largest update count of a row (the wave kernel's capacity check)
__global__ void k_row_updates_max(int n, const int* __restrict__ prp, const int* __restrict__ po,
                                  int* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) atomicMax(out, po[prp[p + 1]] - po[prp[p]]);
}

inline int grid_of(int64_t n, int bs) { return (int)std::max<int64_t>(1, (n + bs - 1) / bs); }

__global__ void k_fill_one_int(int* p) { *p = 0; }

}  // namespace

void launch_ic_symbolic_count(hipStream_t s, int n, const int* prp, const int* pci, const int* prow, const int* ppos,
                              int* cnt) {
    hipLaunchKernelGGL(k_ic_symbolic, dim3(grid_of(n, kBS)), dim3(kBS), 0, s, n, prp, pci, prow, ppos, cnt, nullptr,
                       nullptr, nullptr, nullptr);
}

size_t scan_tmp_bytes(int nitems) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int*)nullptr, (int*)nullptr, nitems);
    return bytes;
}

void launch_exclusive_scan(hipStream_t s, const int* in, int* out, int nitems, void* tmp, size_t tmp_bytes) {
    (void)hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, nitems, s);
}

void launch_ic_symbolic_fill(hipStream_t s, int n, const int* prp, const int* pci, const int* prow, const int* ppos,
                             const int* po, void* u_slot, int* u_pos, int* u_col) {
    hipLaunchKernelGGL(k_ic_symbolic, dim3(grid_of(n, kBS)), dim3(kBS), 0, s, n, prp, pci, prow, ppos, nullptr, po,
                       (unsigned char*)u_slot, u_pos, u_col);
}

template <typename T>
void launch_ic_numeric(hipStream_t s, int kmax_row, int p0, int p1, const int* prp, const int* pci, const int* prow,
                       const int* po, const void* u_slot, const int* u_pos, const int* u_col, T* fv, T* diagD,
                       double accel, int* nfix) {
    const int nr = p1 - p0;
    if (nr <= 0) return;
    const dim3 g((nr + kFactBS - 1) / kFactBS);
    const unsigned char* sl = (const unsigned char*)u_slot;
    if (kmax_row <= 24)
        hipLaunchKernelGGL((k_ic_numeric<T, 24>), g, dim3(kFactBS), 0, s, p0, p1, prp, pci, prow, po, sl, u_pos, u_col,
                           fv, diagD, accel, nfix);
    else
        hipLaunchKernelGGL((k_ic_numeric<T, 48>), g, dim3(kFactBS), 0, s, p0, p1, prp, pci, prow, po, sl, u_pos, u_col,
                           fv, diagD, accel, nfix);
}

int ic_row_updates_max(hipStream_t s, int n, const int* prp, const int* po, int* scratch) {
    hipLaunchKernelGGL(k_fill_one_int, dim3(1), dim3(1), 0, s, scratch);
    hipLaunchKernelGGL(k_row_updates_max, dim3(grid_of(n, kBS)), dim3(kBS), 0, s, n, prp, po, scratch);
    int h = 0;
    (void)hipMemcpyAsync(&h, scratch, sizeof(int), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    return h;
}

template <typename T>
void launch_ic_numeric_wave(hipStream_t s, int kmax_row, int p0, int p1, const int* prp, const int* pci,
                            const int* prow, const int* po, const void* u_slot, const int* u_pos, const int* u_col,
                            T* fv, T* diagD, double accel, int* nfix) {
    const int nr = p1 - p0;
    if (nr <= 0) return;
    const unsigned char* sl = (const unsigned char*)u_slot;
    if (kmax_row <= 24)
        hipLaunchKernelGGL((k_ic_numeric_wave<T, 24>), dim3(nr), dim3(64), 0, s, p0, p1, prp, pci, prow, po, sl,
                           u_pos, u_col, fv, diagD, accel, nfix);
    else
        hipLaunchKernelGGL((k_ic_numeric_wave<T, 48>), dim3(nr), dim3(64), 0, s, p0, p1, prp, pci, prow, po, sl,
                           u_pos, u_col, fv, diagD, accel, nfix);
}
template void launch_ic_numeric_wave<double>(hipStream_t, int, int, int, const int*, const int*, const int*,
                                             const int*, const void*, const int*, const int*, double*, double*,
                                             double, int*);
template void launch_ic_numeric_wave<zc>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                         const void*, const int*, const int*, zc*, zc*, double, int*);

template void launch_ic_numeric<double>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                        const void*, const int*, const int*, double*, double*, double, int*);
template void launch_ic_numeric<zc>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                    const void*, const int*, const int*, zc*, zc*, double, int*);

}  // namespace ssgpu
