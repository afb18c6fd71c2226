// Matrix assembly from (row, col, value) triplets on the device.
//
// The reference assembles on the host: SparseBuilder::add accumulates into a
// std::map per row (SparseBuilder.hpp:75), build() turns it into triplets
// (SparseBuilder.cpp:107-141) and SparseMat hands them to Eigen's
// setFromTriplets (SparseMat.cpp:113-122; the COO constructor,
// SparseMat.cpp:79-104, does the same directly).  Here a stable radix sort of
// (row * n + col) keys with the triplet index as payload puts every entry's
// triplets next to each other in the order they were given; one thread per
// entry then sums them sequentially in that order, which is the reference's
// arithmetic exactly:
//   builder modes: 0 + v1 + v2 + ...  (each map entry starts value-initialized)
//   COO mode:      v1 + v2 + ...      (setFromTriplets keeps the first value)
// Row pointers come from the entry heads; columns are ascending by
// construction.  Traffic per triplet: the keys and payloads through the sort,
// one gather of the value; no atomics on values.
#include <hipcub/hipcub.hpp>

#include <climits>

#include "ssgpu_host.h"

namespace ssgpu {

namespace {

constexpr int kAsmBS = 256;

inline int asm_grid(int64_t n) {
    const int64_t g = (n + kAsmBS - 1) / kAsmBS;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

// keys, payloads, and per block: the count of indices outside [0, n) and
// the largest column (stats[0] += bad, stats[1] = max(col))
__global__ __launch_bounds__(kAsmBS) void k_asm_keys(int64_t ntrip, int n, const int* __restrict__ rows,
                                                     const int* __restrict__ cols,
                                                     unsigned long long* __restrict__ keys, int* __restrict__ idx,
                                                     int* __restrict__ stats) {
    int bad = 0, maxc = -1;
    for (int64_t i = blockIdx.x * (int64_t)kAsmBS + threadIdx.x; i < ntrip; i += (int64_t)gridDim.x * kAsmBS) {
        const int r = rows[i], c = cols[i];
        const bool ok = r >= 0 && r < n && c >= 0 && c < n;
        bad += ok ? 0 : 1;
        maxc = ok && c > maxc ? c : maxc;
        keys[i] = ok ? (unsigned long long)r * (unsigned long long)n + (unsigned long long)c : 0ull;
        idx[i] = (int)i;
    }
    __shared__ int s_bad[kAsmBS], s_max[kAsmBS];
    s_bad[threadIdx.x] = bad;
    s_max[threadIdx.x] = maxc;
    __syncthreads();
    for (int w = kAsmBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_bad[threadIdx.x] += s_bad[threadIdx.x + w];
            s_max[threadIdx.x] = max(s_max[threadIdx.x], s_max[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (s_bad[0]) atomicAdd(stats, s_bad[0]);
        atomicMax(stats + 1, s_max[0]);
    }
}

// head[i] = 1 where sorted position i starts a new (row, col) entry
__global__ __launch_bounds__(kAsmBS) void k_asm_heads(int64_t ntrip, const unsigned long long* __restrict__ keys,
                                                      int* __restrict__ head) {
    for (int64_t i = blockIdx.x * (int64_t)kAsmBS + threadIdx.x; i < ntrip; i += (int64_t)gridDim.x * kAsmBS)
        head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

template <typename T> __device__ inline T asm_add(T a, T b);
template <> __device__ inline double asm_add<double>(double a, double b) { return a + b; }
template <> __device__ inline zc asm_add<zc>(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }

// One thread per entry head: the entry's triplets summed in the order given
// (sorted positions i, i+1, ... carry ascending triplet indices: the sort is
// stable), its column, and the row pointers of the rows it opens.
template <typename T>
__global__ __launch_bounds__(kAsmBS) void k_asm_sum(int64_t ntrip, int n, const unsigned long long* __restrict__ keys,
                                                    const int* __restrict__ idx, const T* __restrict__ vals,
                                                    const int* __restrict__ head, const int* __restrict__ pos,
                                                    int zero_start, int nnz, int* __restrict__ rp,
                                                    int* __restrict__ ci, T* __restrict__ v) {
    for (int64_t i = blockIdx.x * (int64_t)kAsmBS + threadIdx.x; i < ntrip; i += (int64_t)gridDim.x * kAsmBS) {
        if (!head[i]) continue;
        const unsigned long long key = keys[i];
        T s = vals[idx[i]];
        if (zero_start) s = asm_add<T>(T{}, s);
        int64_t j = i + 1;
        while (j < ntrip && keys[j] == key) s = asm_add<T>(s, vals[idx[j++]]);
        const int p = pos[i];
        const int r = (int)(key / (unsigned long long)n);
        ci[p] = (int)(key % (unsigned long long)n);
        v[p] = s;
        // rows (previous entry's row, r] start at p; after the last entry,
        // rows (r, n] end at nnz
        const int r0 = i == 0 ? -1 : (int)(keys[i - 1] / (unsigned long long)n);
        for (int q = r0 + 1; q <= r; ++q) rp[q] = p;
        if (j == ntrip)
            for (int q = r + 1; q <= n; ++q) rp[q] = nnz;
    }
}

}  // namespace

void assemble_triplets(Matrix& m, int64_t ntrip, const int* rows, const int* cols, const void* vals, int mode) {
    Context* c = m.ctx;
    hipStream_t s = c->stream;
    const int n = m.n;
    if (ntrip < 0 || ntrip >= INT_MAX) throw Error{BAD_ARG, "triplet count outside [0, 2^31 - 1)"};
    const size_t vs = m.vsize;
    m.rp.alloc(c, sizeof(int) * ((size_t)n + 1));
    int nnz = 0, maxcol = -1;
    // every temporary stays allocated until the final synchronization (a
    // free could otherwise meet a kernel still reading the buffer)
    DevBuf keys, keys2, idx, idx2, head, pos, sort_tmp, scan_tmp, stats;
    int st[2] = {0, -1}, last[2] = {0, 0};
    if (ntrip > 0) {
        keys.alloc(c, 8 * (size_t)ntrip);
        keys2.alloc(c, 8 * (size_t)ntrip);
        idx.alloc(c, 4 * (size_t)ntrip);
        idx2.alloc(c, 4 * (size_t)ntrip);
        stats.alloc(c, 2 * sizeof(int));
        SSGPU_HIP(hipMemcpyAsync(stats.p, st, sizeof(st), hipMemcpyHostToDevice, s));
        const int g = asm_grid(ntrip);
        hipLaunchKernelGGL(k_asm_keys, dim3(g), dim3(kAsmBS), 0, s, ntrip, n, rows, cols,
                           keys.as<unsigned long long>(), idx.as<int>(), stats.as<int>());
        SSGPU_HIP(hipGetLastError());
        SSGPU_HIP(hipMemcpyAsync(st, stats.p, sizeof(st), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        if (st[0]) throw Error{BAD_ARG, "a triplet index lies outside [0, n)"};
        maxcol = st[1];
        // sort on the bits the keys use (row * n + col < n^2)
        const unsigned long long kmax = (unsigned long long)n * (unsigned long long)n - 1ull;
        int end_bit = 1;
        while (end_bit < 64 && (kmax >> end_bit) != 0ull) ++end_bit;
        size_t tb = 0;
        SSGPU_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<unsigned long long>(),
                                                     keys2.as<unsigned long long>(), idx.as<int>(), idx2.as<int>(),
                                                     (int)ntrip, 0, end_bit, s));
        sort_tmp.alloc(c, std::max<size_t>(tb, 8));
        SSGPU_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, tb, keys.as<unsigned long long>(),
                                                     keys2.as<unsigned long long>(), idx.as<int>(), idx2.as<int>(),
                                                     (int)ntrip, 0, end_bit, s));
        head.alloc(c, 4 * (size_t)ntrip);
        pos.alloc(c, 4 * (size_t)ntrip);
        hipLaunchKernelGGL(k_asm_heads, dim3(g), dim3(kAsmBS), 0, s, ntrip, keys2.as<unsigned long long>(),
                           head.as<int>());
        SSGPU_HIP(hipGetLastError());
        size_t sb = 0;
        SSGPU_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, head.as<int>(), pos.as<int>(), (int)ntrip, s));
        scan_tmp.alloc(c, std::max<size_t>(sb, 8));
        SSGPU_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, sb, head.as<int>(), pos.as<int>(), (int)ntrip, s));
        SSGPU_HIP(hipMemcpyAsync(&last[0], pos.as<int>() + ntrip - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipMemcpyAsync(&last[1], head.as<int>() + ntrip - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        nnz = last[0] + last[1];
    }
    // SparseBuilder.cpp:134-139: to_square appends an explicit zero at
    // (n-1, n-1) when no column reaches n-1 (the largest column counts as 0
    // when there is none); it sorts last, after every other entry
    const bool pad = mode == 2 && std::max(maxcol, 0) + 1 < n;
    const int total = nnz + (pad ? 1 : 0);
    const int pad_col = n - 1;
    const char pad_zero[16] = {};
    m.ci.alloc(c, sizeof(int) * (size_t)std::max(total, 1));
    m.val.alloc(c, vs * (size_t)std::max(total, 1));
    if (ntrip > 0) {
        const int g = asm_grid(ntrip);
        if (m.dtype == SSGPU_R64)
            hipLaunchKernelGGL(k_asm_sum<double>, dim3(g), dim3(kAsmBS), 0, s, ntrip, n,
                               keys2.as<unsigned long long>(), idx2.as<int>(), static_cast<const double*>(vals),
                               head.as<int>(), pos.as<int>(), mode != 0 ? 1 : 0, nnz, m.rp.as<int>(), m.ci.as<int>(),
                               m.val.as<double>());
        else
            hipLaunchKernelGGL(k_asm_sum<zc>, dim3(g), dim3(kAsmBS), 0, s, ntrip, n, keys2.as<unsigned long long>(),
                               idx2.as<int>(), static_cast<const zc*>(vals), head.as<int>(), pos.as<int>(),
                               mode != 0 ? 1 : 0, nnz, m.rp.as<int>(), m.ci.as<int>(), m.val.as<zc>());
        SSGPU_HIP(hipGetLastError());
    } else {
        launch_fill<int>(s, (int64_t)n + 1, m.rp.as<int>(), 0);
    }
    if (pad) {
        SSGPU_HIP(hipMemcpyAsync(m.ci.as<int>() + nnz, &pad_col, sizeof(int), hipMemcpyHostToDevice, s));
        SSGPU_HIP(hipMemcpyAsync(static_cast<char*>(m.val.p) + vs * nnz, pad_zero, vs, hipMemcpyHostToDevice, s));
        SSGPU_HIP(hipMemcpyAsync(m.rp.as<int>() + n, &total, sizeof(int), hipMemcpyHostToDevice, s));
    }
    m.nnz = total;
    m.h_rp.resize((size_t)n + 1);
    m.h_ci.resize((size_t)total);
    SSGPU_HIP(hipMemcpyAsync(m.h_rp.data(), m.rp.p, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    if (total) SSGPU_HIP(hipMemcpyAsync(m.h_ci.data(), m.ci.p, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, s));
    SSGPU_HIP(hipStreamSynchronize(s));   // the temporaries are freed on return
}

}  // namespace ssgpu
