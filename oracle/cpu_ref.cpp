// CPU oracle: restatement of the reference's shifted-IC Krylov path.
//
// TEST INFRASTRUCTURE ONLY.  Tests, the smoke check and the benchmark's
// CPU-baseline leg load this library as the checker / the timed CPU baseline;
// the product path (sparsesolv_amd) never calls it.
//
// Each routine cites the reference code it restates.  Arithmetic follows the
// reference build (g++ -O3, SSE2, no FMA): row-major SpMV with two alternating
// accumulators (Eigen SparseCore/SparseDenseProduct.h:66-80), real dot with
// four packet accumulators (Eigen Core/InnerProduct.h inner_product_impl),
// sums / squared norms with two packet accumulators (Core/Redux.h:275-325,
// Core/Dot.h:21-27), and the reference's own sequential loops elsewhere.
// The IC inner search of SprsPreconditioner.cpp:69-84 is done as a merge of
// the two sorted rows: it visits the same (k) in the same ascending order and
// performs the same subtractions, so the factor is bit-identical.

#include "cpu_ref.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;

template <typename T> struct Csr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<T> v;
};

template <typename T> inline const T* as(const double* p) { return reinterpret_cast<const T*>(p); }
template <typename T> inline T* as(double* p) { return reinterpret_cast<T*>(p); }

template <typename T> Csr<T> from_raw(int n, const int* rp, const int* ci, const double* v) {
    Csr<T> a;
    a.n = n;
    a.rp.assign(rp, rp + n + 1);
    a.ci.assign(ci + rp[0], ci + rp[n]);
    const T* tv = as<T>(v);
    a.v.assign(tv + rp[0], tv + rp[n]);
    if (rp[0] != 0)
        for (auto& x : a.rp) x -= rp[0];
    return a;
}

template <typename T> void to_raw(const Csr<T>& a, int* rp, int* ci, double* v) {
    std::memcpy(rp, a.rp.data(), sizeof(int) * (a.n + 1));
    std::memcpy(ci, a.ci.data(), sizeof(int) * a.ci.size());
    std::memcpy(v, a.v.data(), sizeof(T) * a.v.size());
}

// ---- structure (SprsTrans.cpp:312-394) ------------------------------------
template <typename T> Csr<T> lower(const Csr<T>& a) {   // MatLower: col <= row, order kept
    Csr<T> l;
    l.n = a.n;
    l.rp.assign(a.n + 1, 0);
    for (int i = 0; i < a.n; ++i) {
        for (int k = a.rp[i]; k < a.rp[i + 1]; ++k)
            if (a.ci[k] <= i) { l.ci.push_back(a.ci[k]); l.v.push_back(a.v[k]); }
        l.rp[i + 1] = (int)l.ci.size();
    }
    return l;
}
template <typename T> Csr<T> upper(const Csr<T>& a) {   // MatUpper: col >= row
    Csr<T> u;
    u.n = a.n;
    u.rp.assign(a.n + 1, 0);
    for (int i = 0; i < a.n; ++i) {
        for (int k = a.rp[i]; k < a.rp[i + 1]; ++k)
            if (a.ci[k] >= i) { u.ci.push_back(a.ci[k]); u.v.push_back(a.v[k]); }
        u.rp[i + 1] = (int)u.ci.size();
    }
    return u;
}
template <typename T> Csr<T> transpose(const Csr<T>& a) {  // trans(): rows of A^T ascending
    Csr<T> t;
    t.n = a.n;
    t.rp.assign(a.n + 1, 0);
    for (int k = 0; k < a.rp[a.n]; ++k) t.rp[a.ci[k] + 1]++;
    for (int i = 0; i < a.n; ++i) t.rp[i + 1] += t.rp[i];
    t.ci.resize(a.rp[a.n]);
    t.v.resize(a.rp[a.n]);
    std::vector<int> pos(t.rp.begin(), t.rp.end() - 1);
    for (int i = 0; i < a.n; ++i)
        for (int k = a.rp[i]; k < a.rp[i + 1]; ++k) {
            const int p = pos[a.ci[k]]++;
            t.ci[p] = i;
            t.v[p] = a.v[k];
        }
    return t;
}

// ---- Eigen-order kernels ----------------------------------------------------
// y = A x (SparseDenseProduct.h:42-80): res = 0 + 1*(tmp_a + tmp_b).
template <typename T> void spmv(const Csr<T>& a, const T* x, T* y) {
    const long long nnz = a.rp[a.n];
    auto row = [&](int i) {
        T ta(0), tb(0);
        int k = a.rp[i];
        const int e = a.rp[i + 1];
        while (k < e) {
            ta += a.v[k] * x[a.ci[k]];
            ++k;
            if (k < e) { tb += a.v[k] * x[a.ci[k]]; ++k; }
        }
        y[i] = ta + tb;
    };
#ifdef _OPENMP
    const int threads = omp_get_max_threads();   // Eigen::nbThreads() default
    if (threads > 1 && nnz > 20000) {
        const int chunk = (a.n + threads * 4 - 1) / (threads * 4);
#pragma omp parallel for schedule(dynamic, chunk) num_threads(threads)
        for (int i = 0; i < a.n; ++i) row(i);
        return;
    }
#endif
    (void)nnz;
    for (int i = 0; i < a.n; ++i) row(i);
}

// Real dot, Eigen inner_product_impl with Packet2d (4 packet accumulators,
// pmadd = a*b + c without FMA).
double dot_real(int n, const double* a, const double* b) {
    if (n < 2) {
        if (n == 0) return 0.0;
        return a[0] * b[0];
    }
    const long long packetEnd = (n / 2) * 2, quadEnd = (n / 8) * 8;
    const long long numPackets = n / 2, numRem = (packetEnd - quadEnd) / 2;
    double p[4][2];
    auto load = [&](int j, long long k) { p[j][0] = a[k] * b[k]; p[j][1] = a[k + 1] * b[k + 1]; };
    auto madd = [&](int j, long long k) { p[j][0] = a[k] * b[k] + p[j][0]; p[j][1] = a[k + 1] * b[k + 1] + p[j][1]; };
    auto add = [&](int d, int s) { p[d][0] = p[d][0] + p[s][0]; p[d][1] = p[d][1] + p[s][1]; };
    load(0, 0);
    if (numPackets >= 2) load(1, 2);
    if (numPackets >= 3) load(2, 4);
    if (numPackets >= 4) {
        load(3, 6);
        for (long long k = 8; k < quadEnd; k += 8) { madd(0, k); madd(1, k + 2); madd(2, k + 4); madd(3, k + 6); }
        if (numRem >= 1) madd(0, quadEnd);
        if (numRem >= 2) madd(1, quadEnd + 2);
        if (numRem == 3) madd(2, quadEnd + 4);
        add(2, 3);
    }
    if (numPackets >= 3) add(1, 2);
    if (numPackets >= 2) add(0, 1);
    double r = p[0][0] + p[0][1];
    for (long long k = packetEnd; k < n; ++k) r = a[k] * b[k] + r;
    return r;
}

// Real sum of e(i) with Redux.h LinearVectorizedTraversal, Packet2d.
template <typename E> double redux_real(int n, E e) {
    const long long aS2 = (n / 4) * 4, aS = (n / 2) * 2;
    double r;
    if (aS) {
        double p0[2] = {e(0), e(1)};
        if (aS > 2) {
            double p1[2] = {e(2), e(3)};
            for (long long k = 4; k < aS2; k += 4) {
                p0[0] = p0[0] + e(k); p0[1] = p0[1] + e(k + 1);
                p1[0] = p1[0] + e(k + 2); p1[1] = p1[1] + e(k + 3);
            }
            p0[0] = p0[0] + p1[0]; p0[1] = p0[1] + p1[1];
            if (aS > aS2) { p0[0] = p0[0] + e(aS2); p0[1] = p0[1] + e(aS2 + 1); }
        }
        r = p0[0] + p0[1];
        for (long long k = aS; k < n; ++k) r = r + e(k);
    } else {
        r = e(0);
    }
    return r;
}

// Complex sum of e(i) with Redux.h LinearVectorizedTraversal, Packet1cd.
template <typename E> cplx redux_cplx(int n, E e) {
    if (n == 1) return e(0);
    const long long aE2 = (n / 2) * 2;
    cplx p0 = e(0), p1 = e(1);
    for (long long k = 2; k < aE2; k += 2) {
        const cplx q0 = e(k), q1 = e(k + 1);
        p0 = cplx(p0.real() + q0.real(), p0.imag() + q0.imag());
        p1 = cplx(p1.real() + q1.real(), p1.imag() + q1.imag());
    }
    p0 = cplx(p0.real() + p1.real(), p0.imag() + p1.imag());
    if (n > aE2) { const cplx q = e(aE2); p0 = cplx(p0.real() + q.real(), p0.imag() + q.imag()); }
    return p0;
}

// Packet1cd pmul (SSE/Complex.h:286-303, SSE2 branch).
inline cplx pmul_c(const cplx& a, const cplx& b) {
    return cplx(a.real() * b.real() + (-(a.imag() * b.imag())), a.real() * b.imag() + a.imag() * b.real());
}

// ---- optional: the GPU library's fixed-tree reduction order ---------------------
// Mode 1 reproduces the device reductions exactly: a grid of
// G = clamp(ceil(n/256), 1, 1024) blocks of 256 threads, thread g summing
// elements g, g + 256G, ... from 0.0; per 64-lane wave a shuffle-down tree
// (offsets 32..1); the 4 wave sums added in order; the G block partials then
// reduced the same way by one 256-thread block.  Used to check that the GPU
// path differs from the reference only through reduction order.
int g_red_mode = 0;   // 0 = reference (Eigen) order, 1 = GPU tree order
constexpr int kTB = 256;

inline int tree_blocks(long long n) {
    long long g = (n + kTB - 1) / kTB;
    if (g < 1) g = 1;
    if (g > 1024) g = 1024;
    return (int)g;
}
// block_sum of 256 per-thread values
inline double block_tree(const double* v) {
    double w[4];
    for (int wave = 0; wave < 4; ++wave) {
        double lane[64];
        for (int l = 0; l < 64; ++l) lane[l] = v[wave * 64 + l];
        for (int off = 32; off > 0; off >>= 1)
            for (int l = 0; l + off < 64; ++l) lane[l] = lane[l] + lane[l + off];
        w[wave] = lane[0];
    }
    double r = w[0];
    for (int k = 1; k < 4; ++k) r = r + w[k];
    return r;
}
template <int K> void finalize_tree(int nb, const double* part, double* out) {
    for (int k = 0; k < K; ++k) {
        double th[kTB];
        for (int t = 0; t < kTB; ++t) {
            double acc = 0.0;
            for (int b = t; b < nb; b += kTB) acc += part[(size_t)b * K + k];
            th[t] = acc;
        }
        out[k] = block_tree(th);
    }
}
// K interleaved accumulators, element contribution f(i, acc)
template <int K, typename F> void tree_reduce(long long n, F f, double* out) {
    const int G = tree_blocks(n);
    std::vector<double> part((size_t)G * K);
    std::vector<double> th((size_t)kTB * K);
    for (int b = 0; b < G; ++b) {
        for (int t = 0; t < kTB; ++t) {
            double acc[K] = {};
            for (long long i = (long long)b * kTB + t; i < n; i += (long long)G * kTB) f(i, acc);
            for (int k = 0; k < K; ++k) th[(size_t)k * kTB + t] = acc[k];
        }
        for (int k = 0; k < K; ++k) part[(size_t)b * K + k] = block_tree(&th[(size_t)k * kTB]);
    }
    finalize_tree<K>(G, part.data(), out);
}

inline double dotv(int n, const double* a, const double* b) {
    if (g_red_mode == 1) {
        double r;
        tree_reduce<1>(n, [&](long long i, double* acc) { acc[0] += a[i] * b[i]; }, &r);
        return r;
    }
    return dot_real(n, a, b);
}
inline cplx dotv(int n, const cplx* a, const cplx* b) {   // (a.array()*b.array()).sum()
    if (g_red_mode == 1) {
        double r[2];
        tree_reduce<2>(n, [&](long long i, double* acc) {
            const cplx p = a[i] * b[i];
            acc[0] += p.real();
            acc[1] += p.imag();
        }, r);
        return cplx(r[0], r[1]);
    }
    if (n == 0) return cplx(0, 0);
    return redux_cplx(n, [&](long long k) { return pmul_c(a[k], b[k]); });
}
inline double normv(int n, const double* a) {
    if (g_red_mode == 1) {
        double r;
        tree_reduce<1>(n, [&](long long i, double* acc) { acc[0] += a[i] * a[i]; }, &r);
        return std::sqrt(r);
    }
    if (n == 0) return 0.0;
    return std::sqrt(redux_real(n, [&](long long k) { return a[k] * a[k]; }));
}
inline double normv(int n, const cplx* a) {
    if (g_red_mode == 1) {
        double r;
        tree_reduce<1>(n, [&](long long i, double* acc) {
            acc[0] += a[i].real() * a[i].real() + a[i].imag() * a[i].imag();
        }, &r);
        return std::sqrt(r);
    }
    if (n == 0) return 0.0;
    const cplx s = redux_cplx(n, [&](long long k) {
        return cplx(a[k].real() * a[k].real(), a[k].imag() * a[k].imag());
    });
    return std::sqrt(s.real() + s.imag());
}

// p.(A p) as the GPU's fused SpMV computes it: one partial per SpMV row block
// (<= 128 rows, <= 4096 entries unless one long row; row j's term in thread
// slot j of a 256-thread block tree), then the finalize tree.
template <typename T> T dot_spmv_order(const Csr<T>& A, const T* p, const T* ap) {
    if (g_red_mode != 1) return dotv(A.n, p, ap);
    constexpr int K = sizeof(T) / sizeof(double);
    std::vector<double> part;
    int r = 0;
    while (r < A.n) {
        int r1 = r + 1;
        while (r1 < A.n && r1 - r < 128 && A.rp[r1 + 1] - A.rp[r] <= 4096) ++r1;
        double th[K][kTB] = {};
        for (int t = 0; t < kTB && r + t < r1; ++t) {
            const T prod = p[r + t] * ap[r + t];
            const double* q = reinterpret_cast<const double*>(&prod);
            for (int k = 0; k < K; ++k) th[k][t] = 0.0 + q[k];
        }
        for (int k = 0; k < K; ++k) part.push_back(block_tree(th[k]));
        r = r1;
    }
    double out[K];
    finalize_tree<K>((int)(part.size() / K), part.data(), out);
    T res;
    std::memcpy(&res, out, sizeof(T));
    return res;
}

inline double re_of(double x) { return x; }
inline double re_of(const cplx& x) { return x.real(); }

// ---- shifted IC(0), SprsPreconditioner.cpp:42-111 ---------------------------
// returns false on a missing diagonal
template <typename T> bool ic_decomp_sub(Csr<T>& m, T* diagD, double accela) {
    const int n = m.n;
    for (int i = 0; i < n; ++i) diagD[i] = 0;
    const int* cp = m.ci.data();
    T* vp = m.v.data();
    for (int i = 0; i < n; ++i) {
        const int rs = m.rp[i], re = m.rp[i + 1];
        for (int jj = rs; jj < re; ++jj) {
            const int j = cp[jj];
            if (j >= i) break;
            T s = vp[jj];
            // merge of row i (cols < j, ascending) with row j
            int ll = m.rp[j];
            const int le = m.rp[j + 1];
            for (int kk = rs; kk < re; ++kk) {
                const int ck = cp[kk];
                if (!(ck <= j - 1)) break;
                while (ll < le && cp[ll] < ck) ++ll;
                if (ll < le && cp[ll] == ck) s -= vp[kk] * vp[ll] * diagD[ck];
            }
            vp[jj] = s;
        }
        const int last = re - 1;
        if (re <= rs || cp[last] != i) return false;
        T s = vp[last] * accela;
        for (int kk = rs; kk < last; ++kk) {
            if (cp[kk] <= i - 1) s -= vp[kk] * vp[kk] * diagD[cp[kk]];
            else break;
        }
        vp[last] = s;
        if (std::abs(s) < 1.0e-12) diagD[i] = 1.0 / (DBL_EPSILON);
        else diagD[i] = 1.0 / s;
    }
    return true;
}

template <typename T> bool ic_decomp(const Csr<T>& a, T* diagD, double accel, Csr<T>& L) {
    L = lower(a);
    return ic_decomp_sub(L, diagD, accel);
}

// MatSolverBase.cpp:344-391
template <typename T> bool diag_ok(int n, const T* d);
template <> bool diag_ok<double>(int n, const double* d) {
    bool ok = true;
    for (int i = 0; i < n; ++i) ok &= (d[i] > 0);
    return ok;
}
template <> bool diag_ok<cplx>(int n, const cplx* d) {
    bool ok = true;
    for (int i = 0; i < n; ++i) ok &= (d[i].real() > 0);
    return ok;
}
template <typename T> double auto_accel(const Csr<T>& a, double accel_ini, T* diagD, Csr<T>& L, bool& okdiag) {
    double acc = std::fabs(accel_ini);
    if (acc < 0.9 || acc > 1.8) acc = 1;
    okdiag = true;
    for (int kkk = 0; kkk < 80; ++kkk) {
        Csr<T> Lt;
        if (!ic_decomp(a, diagD, acc, Lt)) { okdiag = false; return acc; }
        if (diag_ok<T>(a.n, diagD)) { L = std::move(Lt); return acc; }
        acc += 0.01;
    }
    Csr<T> L0;
    if (!ic_decomp(a, diagD, 1.8, L0)) okdiag = false;
    L = std::move(L0);
    return 1.8;
}

// ---- sweeps, MatSolverBase.cpp:177-327 ---------------------------------------
template <typename T>
void ic_apply(const Csr<T>& L, const Csr<T>& U, const T* diagD, const T* r, T* z, std::vector<T>& tmp) {
    const int n = L.n;
    tmp.resize(n);
    for (int i = 0; i < n; ++i) {
        T s = r[i];
        const int c = L.rp[i + 1] - 1;
        for (int jj = L.rp[i]; jj < c; ++jj) s -= L.v[jj] * tmp[L.ci[jj]];
        tmp[i] = s / L.v[c];
    }
    for (int i = n; i-- > 0;) {
        T s = 0;
        const int c = U.rp[i + 1];
        for (int j = U.rp[i] + 1; j < c; ++j) s -= U.v[j] * z[U.ci[j]];
        s *= diagD[i];
        z[i] = s + tmp[i];
    }
}
template <typename T> void fr(const Csr<T>& L, const T* r, T* z) {
    for (int i = 0; i < L.n; ++i) {
        T s = r[i];
        const int c = L.rp[i + 1] - 1;
        for (int jj = L.rp[i]; jj < c; ++jj) s -= L.v[jj] * z[L.ci[jj]];
        z[i] = s / L.v[c];
    }
}
template <typename T> void bc(const Csr<T>& U, const T* r, T* z) {
    for (int i = U.n; i-- > 0;) {
        T s = r[i];
        const int c = U.rp[i + 1];
        for (int j = U.rp[i] + 1; j < c; ++j) s -= U.v[j] * z[U.ci[j]];
        z[i] = s / U.v[U.rp[i]];
    }
}

// ---- diagonal scaling, SprsPreconditioner.cpp:150-231 / MatSolverBase.cpp:99-141
inline double diag_mag(double v) { return std::fabs(v); }
inline double diag_mag(const cplx& v) { return std::norm(v); }   // complex: |a|^2, then sqrt
template <typename T> void diag_scaling(const Csr<T>& a, const T* b, std::vector<T>& s, std::vector<T>& bs) {
    s.assign(a.n, T(0));
    bs.assign(a.n, T(0));
    for (int i = 0; i < a.n; ++i)
        for (int jj = a.rp[i]; jj < a.rp[i + 1]; ++jj)
            if (a.ci[jj] == i) {
                const double t1 = diag_mag(a.v[jj]);
                const double t1b = std::sqrt(t1);
                const double t2 = (t1b < 1.0e-12) ? 1.0 / (DBL_EPSILON) : 1.0 / t1b;
                s[i] = T(t2);
                bs[i] = b[i] * T(t2);
                break;
            }
}
template <typename T> Csr<T> diag_mat_trans(const Csr<T>& a, const std::vector<T>& s) {
    Csr<T> o = a;
    for (int i = 0; i < a.n; ++i) {
        const T di = s[i];
        for (int k = o.rp[i]; k < o.rp[i + 1]; ++k) o.v[k] *= di * s[o.ci[k]];
    }
    return o;
}

// ---- solvers -----------------------------------------------------------------
struct State {
    ref_options o;
    std::vector<double> log;
};

inline double set_normalizer(const ref_options& o, double normB, double pure) {
    if (o.conv_normalize_type == 0) return normB;
    if (o.conv_normalize_type == 1) return pure;
    if (o.conv_normalize_type == 2) return o.conv_normalize_const;
    return normB;
}
constexpr double kSmallAbs = 1.0e-20;

// Convergence prologue shared by every solve_main (MatSolverICCG.cpp:212-237).
// Returns 1 / 0 to return that value immediately, or -1 to continue.
inline int prologue(State& st, double normB, double pure, double conv_cri, double& abs_conv, double& normalizer) {
    const double rhs_eps = 1.0e-30;
    if (normB <= rhs_eps) {
        if (pure <= kSmallAbs) {
            if (st.o.is_save_log) { st.log.clear(); st.log.push_back(0.0); }
            return 1;
        }
    }
    abs_conv = (normB * conv_cri * 0.9 < kSmallAbs ? kSmallAbs : normB * conv_cri * 0.9);
    const double first = pure / normB;
    if (first < conv_cri * 0.5 || pure < abs_conv * 0.5) return 1;
    normalizer = set_normalizer(st.o, normB, pure);
    if (st.o.is_save_log) st.log.push_back(pure / normalizer);
    return -1;
}

// Per-iteration bookkeeping after the residual norm (MatSolverICCG.cpp:270-305).
// Returns 1 = converged (break), 2 = diverged (break), 0 = continue.
template <typename T>
inline int monitor(State& st, double pure, double normalizer, double conv_cri, double abs_conv, double& best,
                   int& bad, std::vector<T>& best_x, const T* x, int n) {
    const double normR = pure / normalizer;
    if (st.o.is_save_log) st.log.push_back(normR);
    if (normR < conv_cri || pure < abs_conv) return 1;
    if (normR < best) {
        best = normR;
        if (st.o.is_save_best) best_x.assign(x, x + n);
    }
    if (st.o.diverge_judge_type == 1) {
        if (normR < best * st.o.bad_div_val) bad = 0;
        if (normR >= best * st.o.bad_div_val) bad++;
        if (bad >= st.o.bad_div_count_thres) return 2;
    }
    return 0;
}

// ICCG / ICCOCG main loop, MatSolverICCG.cpp:178-328 / 477-626
template <typename T>
bool iccg_main(State& st, double conv_cri, int max_ite, double normB, const T* diagD, const Csr<T>& A,
               const Csr<T>& L, const Csr<T>& U, const T* b, T* x) {
    const int n = A.n;
    std::vector<T> P(n), R(n), Z(n), AP(n), tmp, best_x;
    if (st.o.is_init) std::fill(x, x + n, T(0));
    if (st.o.is_save_log) st.log.clear();
    double best = 1.0e+6;
    if (st.o.is_save_best) best_x.assign(n, T(0));
    for (int i = 0; i < n; ++i) R[i] = b[i];
    if (!st.o.is_init) {
        spmv(A, x, AP.data());
        for (int i = 0; i < n; ++i) R[i] -= AP[i];
    }
    const double pure = normv(n, R.data());
    double abs_conv = 0, normalizer = 1;
    const int pr = prologue(st, normB, pure, conv_cri, abs_conv, normalizer);
    if (pr >= 0) return pr == 1;
    ic_apply(L, U, diagD, R.data(), P.data(), tmp);
    Z = P;
    T rur0 = dotv(n, R.data(), Z.data());
    bool conv = false;
    int bad = 0;
    for (int it = 0; it < max_ite; ++it) {
        spmv(A, P.data(), AP.data());
        const T t2 = dot_spmv_order(A, P.data(), AP.data());
        const T alpha = rur0 / t2;
        for (int i = 0; i < n; ++i) x[i] += alpha * P[i];
        for (int i = 0; i < n; ++i) R[i] -= alpha * AP[i];
        const double pn = normv(n, R.data());
        const int m = monitor(st, pn, normalizer, conv_cri, abs_conv, best, bad, best_x, x, n);
        if (m == 1) { conv = true; break; }
        if (m == 2) { conv = false; break; }
        ic_apply(L, U, diagD, R.data(), Z.data(), tmp);
        const T rur1 = dotv(n, R.data(), Z.data());
        const T beta = rur1 / rur0;
        rur0 = rur1;
        for (int i = 0; i < n; ++i) P[i] *= beta;
        for (int i = 0; i < n; ++i) P[i] += Z[i];
    }
    if (!conv && st.o.is_save_best) std::copy(best_x.begin(), best_x.end(), x);
    return conv;
}

// IC-MRTR main loop, MatSolverMRTR_ICver.cpp:183-373 / 529-722
// (search direction p = u + t*u exactly as the reference writes it, :302-304).
template <typename T>
bool mrtr_main(State& st, double conv_cri, int max_ite, double normB, const T* diagD, const Csr<T>& A,
               const Csr<T>& L, const Csr<T>& U, const T* b, T* x) {
    const int n = A.n;
    std::vector<T> P(n, T(0)), R(n), Uv(n), X(n, T(0)), Y(n), Zv(n), V(n), W(n), tmp, best_x;
    if (st.o.is_init) std::fill(x, x + n, T(0));
    else X.assign(x, x + n);
    if (st.o.is_save_log) st.log.clear();
    double best = 1.0e+6;
    if (st.o.is_save_best) best_x.assign(n, T(0));
    for (int i = 0; i < n; ++i) R[i] = b[i];
    if (st.o.is_init) {   // subtracts A*x only when x was just zeroed (:213-216)
        spmv(A, x, V.data());
        for (int i = 0; i < n; ++i) R[i] -= V[i];
    }
    const double pure = normv(n, R.data());
    double abs_conv = 0, normalizer = 1;
    const int pr = prologue(st, normB, pure, conv_cri, abs_conv, normalizer);
    if (pr >= 0) return pr == 1;
    ic_apply(L, U, diagD, R.data(), Uv.data(), tmp);
    for (int i = 0; i < n; ++i) Y[i] = -1.0 * R[i];
    ic_apply(L, U, diagD, Y.data(), Zv.data(), tmp);
    T zeta = 1.0, zeta_old = 1.0, eta = 0.0, nu = 1.0;
    bool conv = false;
    int bad = 0;
    for (int it = 0; it < max_ite; ++it) {
        spmv(A, Uv.data(), V.data());
        ic_apply(L, U, diagD, V.data(), W.data(), tmp);
        const T w_r = dotv(n, W.data(), R.data());
        const T v_w = dotv(n, V.data(), W.data());
        if (it == 0) {
            zeta = w_r / v_w;
            zeta_old = zeta;
            eta = 0.0;
        } else {
            const T w_y = dotv(n, W.data(), Y.data());
            const T t0 = 1.0 / (nu * v_w - w_y * w_y);
            zeta = nu * w_r * t0;
            eta = -1.0 * w_y * w_r * t0;
        }
        nu = zeta * w_r;
        const T t = eta * zeta_old / zeta;
        if (st.o.mrtr_direction_fix) {   // standard MRTR direction (labelled correction)
            for (int i = 0; i < n; ++i) P[i] = Uv[i] + t * P[i];
        } else {                          // reference form, MatSolverMRTR_ICver.cpp:302-304
            P = Uv;
            for (int i = 0; i < n; ++i) P[i] += t * P[i];
        }
        zeta_old = zeta;
        for (int i = 0; i < n; ++i) X[i] += zeta * P[i];
        for (int i = 0; i < n; ++i) Y[i] *= eta;
        for (int i = 0; i < n; ++i) Y[i] += zeta * V[i];
        for (int i = 0; i < n; ++i) R[i] -= Y[i];
        const double pn = normv(n, R.data());
        const int m = monitor(st, pn, normalizer, conv_cri, abs_conv, best, bad, best_x, X.data(), n);
        if (m == 1) { conv = true; break; }
        if (m == 2) { conv = false; break; }
        for (int i = 0; i < n; ++i) Zv[i] *= eta;
        for (int i = 0; i < n; ++i) Zv[i] += zeta * W[i];
        for (int i = 0; i < n; ++i) Uv[i] -= Zv[i];
    }
    std::copy(X.begin(), X.end(), x);
    if (!conv && st.o.is_save_best) std::copy(best_x.begin(), best_x.end(), x);
    return conv;
}

// SGS-MRTR main loop, MatSolverMRTR_SGSver.cpp:81-282 / 340-539
template <typename T>
bool sgs_main(State& st, double conv_cri, int max_ite, double normB, const Csr<T>& A, const Csr<T>& L,
              const Csr<T>& U, const T* b, T* x) {
    const int n = A.n;
    std::vector<T> P(n, T(0)), Rd(n), Uv(n), X(n, T(0)), Y(n), ARd(n), tmpv(n), best_x;
    if (st.o.is_init) std::fill(x, x + n, T(0));
    else X.assign(x, x + n);
    if (st.o.is_save_log) st.log.clear();
    double best = 1.0e+6;
    if (st.o.is_save_best) best_x.assign(n, T(0));
    for (int i = 0; i < n; ++i) Rd[i] = b[i];
    if (st.o.is_init) {
        spmv(A, x, ARd.data());
        for (int i = 0; i < n; ++i) Rd[i] -= ARd[i];
    }
    tmpv = Rd;
    fr(L, tmpv.data(), Rd.data());
    const double pure = normv(n, Rd.data());
    double abs_conv = 0, normalizer = 1;
    const int pr = prologue(st, normB, pure, conv_cri, abs_conv, normalizer);
    if (pr >= 0) return pr == 1;
    for (int i = 0; i < n; ++i) Y[i] = -1.0 * Rd[i];
    T zeta = 1.0, zeta_old = 1.0, eta = 0.0, nu = 1.0;
    bool conv = false;
    int bad = 0;
    for (int it = 0; it < max_ite; ++it) {
        bc(U, Rd.data(), Uv.data());
        for (int i = 0; i < n; ++i) tmpv[i] = Rd[i] - Uv[i];
        fr(L, tmpv.data(), ARd.data());
        for (int i = 0; i < n; ++i) ARd[i] += Uv[i];
        const T ar_r = dotv(n, ARd.data(), Rd.data());
        const T ar_ar = dotv(n, ARd.data(), ARd.data());
        if (it == 0) {
            zeta = ar_r / ar_ar;
            zeta_old = zeta;
            eta = 0.0;
        } else {
            const T ar_y = dotv(n, ARd.data(), Y.data());
            const T t = 1.0 / (nu * ar_ar - ar_y * ar_y);
            zeta = nu * ar_r * t;
            eta = -1.0 * ar_y * ar_r * t;
        }
        nu = zeta * ar_r;
        const T t2 = eta * zeta_old / zeta;
        for (int i = 0; i < n; ++i) P[i] *= t2;
        for (int i = 0; i < n; ++i) P[i] += Uv[i];
        zeta_old = zeta;
        for (int i = 0; i < n; ++i) X[i] += zeta * P[i];
        for (int i = 0; i < n; ++i) Y[i] *= eta;
        for (int i = 0; i < n; ++i) Y[i] += zeta * ARd[i];
        for (int i = 0; i < n; ++i) Rd[i] -= Y[i];
        const double pn = normv(n, Rd.data());
        const int m = monitor(st, pn, normalizer, conv_cri, abs_conv, best, bad, best_x, X.data(), n);
        if (m == 1) { conv = true; break; }
        if (m == 2) { conv = false; break; }
    }
    std::copy(X.begin(), X.end(), x);
    if (!conv && st.o.is_save_best) std::copy(best_x.begin(), best_x.end(), x);
    return conv;
}

// Factorization front end shared by ICCG and IC-MRTR (MatSolverICCG.cpp:99-168).
template <typename T>
int ic_solve(State& st, int solver, double conv_cri, int max_ite, double normB, const Csr<T>& A0, const T* b,
             T* x) {
    const int n = A0.n;
    const Csr<T>* A = &A0;
    Csr<T> DAD;
    std::vector<T> s, bs, xs;
    const T* bb = b;
    T* xx = x;
    double nB = normB;
    if (st.o.is_diag_scale) {
        diag_scaling(A0, b, s, bs);
        DAD = diag_mat_trans(A0, s);
        A = &DAD;
        nB = normv(n, bs.data());
        bb = bs.data();
        xs.assign(x, x + n);
        if (!st.o.is_init)
            for (int i = 0; i < n; ++i) xs[i] /= s[i];
        xx = xs.data();
    }
    std::vector<T> diagD(n);
    Csr<T> L;
    if (st.o.accel_factor < 0.0) {
        bool ok;
        auto_accel(*A, st.o.accel_factor, diagD.data(), L, ok);
        if (!ok) return -2;
    } else if (!ic_decomp(*A, diagD.data(), st.o.accel_factor, L)) {
        return -2;
    }
    const Csr<T> U = transpose(L);
    bool c;
    if (solver == 0) c = iccg_main(st, conv_cri, max_ite, nB, diagD.data(), *A, L, U, bb, xx);
    else c = mrtr_main(st, conv_cri, max_ite, nB, diagD.data(), *A, L, U, bb, xx);
    if (st.o.is_diag_scale)
        for (int i = 0; i < n; ++i) x[i] = xs[i] * s[i];
    return c ? 1 : 0;
}

// SGS front end, MatSolverMRTR_SGSver.cpp:42-69
template <typename T>
int sgs_solve(State& st, double conv_cri, int max_ite, const Csr<T>& A0, const T* b, T* x) {
    const int n = A0.n;
    std::vector<T> s, bs;
    diag_scaling(A0, b, s, bs);
    const Csr<T> DAD = diag_mat_trans(A0, s);
    const double nB = normv(n, bs.data());
    const Csr<T> L = lower(DAD), U = upper(DAD);
    std::vector<T> xs(x, x + n);
    if (!st.o.is_init)
        for (int i = 0; i < n; ++i) xs[i] /= s[i];
    const bool c = sgs_main(st, conv_cri, max_ite, nB, DAD, L, U, bs.data(), xs.data());
    for (int i = 0; i < n; ++i) x[i] = xs[i] * s[i];
    return c ? 1 : 0;
}

template <typename T>
int solve_t(int solver, const ref_options* opt, double conv_cri, int max_ite, double normB, int n, const int* rp,
            const int* ci, const double* v, const double* b, double* x, double* log, int log_cap, int* log_len) {
    State st;
    st.o = *opt;
    const Csr<T> A = from_raw<T>(n, rp, ci, v);
    int r;
    if (solver == 1 && opt->precondition_type == 1) r = sgs_solve(st, conv_cri, max_ite, A, as<T>(b), as<T>(x));
    else r = ic_solve(st, solver, conv_cri, max_ite, normB, A, as<T>(b), as<T>(x));
    if (log_len) *log_len = st.o.is_save_log ? (int)st.log.size() : 0;
    if (log && st.o.is_save_log)
        for (int i = 0; i < (int)st.log.size() && i < log_cap; ++i) log[i] = st.log[i];
    return r;
}

// Triplet assembly (SparseMat.cpp:79-104 for mode 0; SparseBuilder.hpp:75,
// SparseBuilder.cpp:107-141 and SparseMat.cpp:113-122 for modes 1 and 2):
// entries ordered by (row, col), each entry's triplets summed in the order
// they were given -- from the first value (Eigen's setFromTriplets), or from
// a value-initialized zero (the builder's std::map += per add).
template <typename T>
long long assemble(int mode, int n, long long ntrip, const int* rows, const int* cols, const T* vals, int* rp,
                   int* ci, T* v) {
    for (long long i = 0; i < ntrip; ++i)
        if (rows[i] < 0 || rows[i] >= n || cols[i] < 0 || cols[i] >= n) return -1;
    std::vector<long long> ord((size_t)ntrip);
    for (long long i = 0; i < ntrip; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](long long a, long long b) {
        return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
    });
    for (int i = 0; i <= n; ++i) rp[i] = 0;
    long long nnz = 0;
    int maxcol = 0;
    for (long long k = 0; k < ntrip;) {
        const int r = rows[ord[k]], c = cols[ord[k]];
        T s = mode == 0 ? vals[ord[k]] : T{} + vals[ord[k]];
        long long e = k + 1;
        while (e < ntrip && rows[ord[e]] == r && cols[ord[e]] == c) s += vals[ord[e++]];
        ci[nnz] = c;
        v[nnz] = s;
        ++nnz;
        ++rp[r + 1];
        maxcol = std::max(maxcol, c);
        k = e;
    }
    if (mode == 2 && maxcol + 1 < n) {   // SparseBuilder.cpp:134-139
        ci[nnz] = n - 1;
        v[nnz] = T{};
        ++nnz;
        ++rp[n];
    }
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    return nnz;
}

}  // namespace

extern "C" {

long long ref_assemble(int cplxf, int mode, int n, long long ntrip, const int* rows, const int* cols,
                       const double* vals, int* rp, int* ci, double* v) {
    if (cplxf) return assemble<cplx>(mode, n, ntrip, rows, cols, as<cplx>(vals), rp, ci, as<cplx>(v));
    return assemble<double>(mode, n, ntrip, rows, cols, vals, rp, ci, v);
}

void ref_set_reduction_order(int mode) { g_red_mode = mode; }

void ref_options_default(ref_options* o) {
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
    o->precondition_type = 1;
    o->mrtr_direction_fix = 0;
}

long long ref_lower_nnz(int n, const int* rp, const int* ci) {
    long long c = 0;
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] <= i) ++c;
    return c;
}

void ref_spmv(int cplxf, int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    if (cplxf) { auto A = from_raw<cplx>(n, rp, ci, v); spmv(A, as<cplx>(x), as<cplx>(y)); }
    else { auto A = from_raw<double>(n, rp, ci, v); spmv(A, x, y); }
}

void ref_dot(int cplxf, int n, const double* a, const double* b, double* out) {
    if (cplxf) {
        const cplx r = dotv(n, as<cplx>(a), as<cplx>(b));
        out[0] = r.real();
        out[1] = r.imag();
    } else {
        out[0] = dotv(n, a, b);
    }
}

double ref_norm(int cplxf, int n, const double* a) {
    return cplxf ? normv(n, as<cplx>(a)) : normv(n, a);
}

double ref_norm_naive(int n, const double* b) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += b[i] * b[i];
    return std::sqrt(s);
}

int ref_ic_decomp(int cplxf, int n, const int* rp, const int* ci, const double* v, double accel, int* Lrp, int* Lci,
                  double* Lv, double* diagD) {
    if (cplxf) {
        auto A = from_raw<cplx>(n, rp, ci, v);
        Csr<cplx> L;
        const bool ok = ic_decomp(A, as<cplx>(diagD), accel, L);
        to_raw(L, Lrp, Lci, Lv);
        return ok ? 0 : 2;
    }
    auto A = from_raw<double>(n, rp, ci, v);
    Csr<double> L;
    const bool ok = ic_decomp(A, diagD, accel, L);
    to_raw(L, Lrp, Lci, Lv);
    return ok ? 0 : 2;
}

double ref_auto_accel(int cplxf, int n, const int* rp, const int* ci, const double* v, double accel_ini, int* Lrp,
                      int* Lci, double* Lv, double* diagD) {
    bool ok;
    double a;
    if (cplxf) {
        auto A = from_raw<cplx>(n, rp, ci, v);
        Csr<cplx> L;
        a = auto_accel(A, accel_ini, as<cplx>(diagD), L, ok);
        if (ok) to_raw(L, Lrp, Lci, Lv);
    } else {
        auto A = from_raw<double>(n, rp, ci, v);
        Csr<double> L;
        a = auto_accel(A, accel_ini, diagD, L, ok);
        if (ok) to_raw(L, Lrp, Lci, Lv);
    }
    return ok ? a : -1.0;
}

void ref_transpose(int cplxf, int n, const int* rp, const int* ci, const double* v, int* Trp, int* Tci, double* Tv) {
    if (cplxf) to_raw(transpose(from_raw<cplx>(n, rp, ci, v)), Trp, Tci, Tv);
    else to_raw(transpose(from_raw<double>(n, rp, ci, v)), Trp, Tci, Tv);
}

void ref_ic_apply(int cplxf, int n, const int* Lrp, const int* Lci, const double* Lv, const int* Urp, const int* Uci,
                  const double* Uv, const double* diagD, const double* r, double* z) {
    if (cplxf) {
        std::vector<cplx> tmp;
        ic_apply(from_raw<cplx>(n, Lrp, Lci, Lv), from_raw<cplx>(n, Urp, Uci, Uv), as<cplx>(diagD), as<cplx>(r),
                 as<cplx>(z), tmp);
    } else {
        std::vector<double> tmp;
        ic_apply(from_raw<double>(n, Lrp, Lci, Lv), from_raw<double>(n, Urp, Uci, Uv), diagD, r, z, tmp);
    }
}

void ref_fr(int cplxf, int n, const int* Lrp, const int* Lci, const double* Lv, const double* r, double* z) {
    if (cplxf) fr(from_raw<cplx>(n, Lrp, Lci, Lv), as<cplx>(r), as<cplx>(z));
    else fr(from_raw<double>(n, Lrp, Lci, Lv), r, z);
}

void ref_bc(int cplxf, int n, const int* Urp, const int* Uci, const double* Uv, const double* r, double* z) {
    if (cplxf) bc(from_raw<cplx>(n, Urp, Uci, Uv), as<cplx>(r), as<cplx>(z));
    else bc(from_raw<double>(n, Urp, Uci, Uv), r, z);
}

int ref_diag_scaling(int cplxf, int n, const int* rp, const int* ci, const double* v, const double* b, double* s,
                     double* bs) {
    if (cplxf) {
        std::vector<cplx> ss, bb;
        diag_scaling(from_raw<cplx>(n, rp, ci, v), as<cplx>(b), ss, bb);
        std::memcpy(s, ss.data(), sizeof(cplx) * n);
        std::memcpy(bs, bb.data(), sizeof(cplx) * n);
    } else {
        std::vector<double> ss, bb;
        diag_scaling(from_raw<double>(n, rp, ci, v), b, ss, bb);
        std::memcpy(s, ss.data(), sizeof(double) * n);
        std::memcpy(bs, bb.data(), sizeof(double) * n);
    }
    return 0;
}

int ref_solve(int solver, int cplxf, const ref_options* opt, double conv_cri, int max_ite, double normB, int n,
              const int* rp, const int* ci, const double* v, const double* b, double* x, double* log, int log_cap,
              int* log_len) {
    if (cplxf) return solve_t<cplx>(solver, opt, conv_cri, max_ite, normB, n, rp, ci, v, b, x, log, log_cap, log_len);
    return solve_t<double>(solver, opt, conv_cri, max_ite, normB, n, rp, ci, v, b, x, log, log_cap, log_len);
}

}  // extern "C"
