// Host C++ layer of the solver API (libSparseSolv.so): the reference's
// SparseBuilder / SparseMat / MatSolverICCG / MatSolverMRTR classes in
// namespace JPMRspace::SparseSolv, running every solve on the GPU through the
// C-ABI of libsparsesolv_amd.so.  Matrices are uploaded once per version and
// cached on the SparseMat; each solve rebuilds the factor, as the reference
// does (MatSolverICCG.cpp:109-118).
#include <SparseSolv/MatSolverICCG.hpp>
#include <SparseSolv/MatSolverMRTR.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "sparsesolv_amd.h"

namespace JPMRspace {
namespace SparseSolv {

// ============================================================ SparseBuilder
template <typename DType> void SparseBuilderTMPL<DType>::resetMat() {
    for (auto& row : tempMat)
        for (auto& kv : row) kv.second = DType{};
}

// SparseBuilder.cpp:107-141
template <typename DType>
slv_int SparseBuilderTMPL<DType>::build(std::vector<Triplet<DType>>& tripletList, bool to_square) {
    slv_int max_retu = 0;
    tripletList.clear();
    tripletList.reserve((size_t)size * 5);
    for (slv_int i = 0; i < size; i++) {
        for (auto& itr : tempMat[i]) {
            tripletList.push_back(Triplet<DType>{i, itr.first, itr.second});
            if (itr.first > max_retu) max_retu = itr.first;
        }
        std::map<slv_int, DType>().swap(tempMat[i]);
    }
    max_retu++;
    if (to_square && max_retu < size) {
        tripletList.push_back(Triplet<DType>{size - 1, size - 1, DType(0.0)});
        max_retu = size;
    }
    return max_retu;
}

template <typename DType> void SparseBuilderTMPL<DType>::printTempMat() const {
    for (slv_int i = 0; i < size; i++) {
        std::cout << i << "-th size = " << tempMat[i].size() << " ----- ";
        for (const auto& kv : tempMat[i]) std::cout << kv.first << ", ";
        std::cout << std::endl;
    }
    std::cout << "----" << std::endl;
    for (slv_int i = 0; i < size; i++) {
        for (const auto& kv : tempMat[i]) std::cout << kv.second << ", ";
        std::cout << std::endl;
    }
}

template class SparseBuilderTMPL<double>;
template class SparseBuilderTMPL<dcomplex>;

// ============================================================ device cache
struct DeviceMatrixCache {
    ssgpu_matrix h = nullptr;
    uint64_t version = 0;
    int device = -1;
    int abmc = 0;
    ~DeviceMatrixCache() {
        if (h) ssgpu_matrix_destroy(h);
    }
};

namespace {

// One context per device for the process's lifetime.  They are not destroyed
// from a static destructor: at process exit the HIP runtime may already be
// gone, and the driver reclaims the memory anyway.
struct ContextRegistry {
    std::mutex mu;
    std::map<int, ssgpu_context> ctx;
};
ContextRegistry& registry() {
    static ContextRegistry r;
    return r;
}

ssgpu_context context_for(int device) {
    auto& r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    auto it = r.ctx.find(device);
    if (it != r.ctx.end()) return it->second;
    ssgpu_context c = nullptr;
    const int st = ssgpu_context_create(device, &c);
    if (st != SSGPU_OK)
        throw std::runtime_error(std::string("SparseSolv: cannot open GPU ") + std::to_string(device) + ": " +
                                 ssgpu_last_error());
    r.ctx[device] = c;
    return c;
}

template <typename D> constexpr int dtype_code();
template <> constexpr int dtype_code<double>() { return SSGPU_R64; }
template <> constexpr int dtype_code<dcomplex>() { return SSGPU_C128; }

template <typename D> ssgpu_matrix device_matrix(const SparseMatTMPL<D>& A, int device, int abmc) {
    auto& c = A.deviceCache();
    if (c && c->h && c->version == A.version() && c->device == device && c->abmc == abmc) return c->h;
    if (A.getSize() != A.getColSize())
        throw std::runtime_error("SparseSolv: the solvers need a square matrix");
    auto nc = std::make_shared<DeviceMatrixCache>();
    const int st = abmc > 0 ? ssgpu_matrix_create_ordered(context_for(device), dtype_code<D>(), A.getSize(),
                                                          A.getRowPtr(), A.getColPtr(), (const void*)A.getValuePtr(),
                                                          abmc, &nc->h, nullptr)
                            : ssgpu_matrix_create(context_for(device), dtype_code<D>(), A.getSize(), A.getRowPtr(),
                                                  A.getColPtr(), (const void*)A.getValuePtr(), &nc->h);
    if (st != SSGPU_OK)
        throw std::runtime_error(std::string("SparseSolv: matrix upload failed: ") + ssgpu_status_string(st) + ": " +
                                 ssgpu_last_error());
    nc->version = A.version();
    nc->device = device;
    nc->abmc = abmc;
    c = nc;
    return c->h;
}

}  // namespace

// ============================================================ SparseMat
template <typename DType>
SparseMatTMPL<DType>::SparseMatTMPL(const SparseMatTMPL& o)
    : rows_(o.rows_), cols_(o.cols_), rp_(o.rp_), ci_(o.ci_), v_(o.v_), version_(o.version_) {}

template <typename DType> SparseMatTMPL<DType>& SparseMatTMPL<DType>::operator=(const SparseMatTMPL& o) {
    if (this != &o) {
        rows_ = o.rows_;
        cols_ = o.cols_;
        rp_ = o.rp_;
        ci_ = o.ci_;
        v_ = o.v_;
        version_ = o.version_ + 1;
        dev_.reset();
    }
    return *this;
}

template <typename DType> SparseMatTMPL<DType>::~SparseMatTMPL() = default;

// Eigen::SparseMatrix::setFromTriplets: rows sorted by column, duplicates
// summed in the order they appear.
template <typename DType>
void SparseMatTMPL<DType>::from_triplets(slv_int nrows, slv_int ncols, std::vector<Triplet<DType>>& t) {
    std::stable_sort(t.begin(), t.end(), [](const Triplet<DType>& a, const Triplet<DType>& b) {
        return a.row != b.row ? a.row < b.row : a.col < b.col;
    });
    rows_ = nrows;
    cols_ = ncols;
    rp_.assign((size_t)nrows + 1, 0);
    ci_.clear();
    v_.clear();
    for (size_t k = 0; k < t.size();) {
        const slv_int r = t[k].row, c = t[k].col;
        DType s = t[k].value;
        size_t e = k + 1;
        while (e < t.size() && t[e].row == r && t[e].col == c) s += t[e++].value;
        ci_.push_back(c);
        v_.push_back(s);
        rp_[r + 1]++;
        k = e;
    }
    for (slv_int i = 0; i < nrows; ++i) rp_[i + 1] += rp_[i];
    touch();
    dev_.reset();
}

template <typename DType>
SparseMatTMPL<DType>::SparseMatTMPL(const std::vector<Triplet<DType>>& sparse_data) {
    slv_int max1 = 0, max2 = 0;
    for (const auto& t : sparse_data) {
        max1 = std::max(max1, t.row);
        max2 = std::max(max2, t.col);
    }
    std::vector<Triplet<DType>> t(sparse_data);
    from_triplets(max1 + 1, max2 + 1, t);
}

// SparseMat.cpp:79-104
template <typename DType>
SparseMatTMPL<DType>::SparseMatTMPL(slv_int n_nonZero, const std::vector<slv_int>& rows,
                                    const std::vector<slv_int>& cols, const std::vector<DType>& vals) {
    slv_int max_gyo = 0, max_retu = 0;
    std::vector<Triplet<DType>> t;
    t.reserve(n_nonZero);
    for (slv_int i = 0; i < n_nonZero; i++) {
        t.push_back(Triplet<DType>{rows[i], cols[i], vals[i]});
        max_gyo = std::max(max_gyo, rows[i]);
        max_retu = std::max(max_retu, cols[i]);
    }
    from_triplets(max_gyo + 1, max_retu + 1, t);
}

template <typename DType>
SparseMatTMPL<DType>::SparseMatTMPL(slv_int nrows, slv_int ncols, std::vector<slv_int> row_ptr,
                                    std::vector<slv_int> col_idx, std::vector<DType> values)
    : rows_(nrows), cols_(ncols), rp_(std::move(row_ptr)), ci_(std::move(col_idx)), v_(std::move(values)) {
    if ((slv_int)rp_.size() != nrows + 1 || ci_.size() != v_.size() || rp_.back() != (slv_int)ci_.size())
        throw std::invalid_argument("SparseMat: inconsistent CSR arrays");
    touch();
}

// SparseMat.cpp:113-122
template <typename DType> void SparseMatTMPL<DType>::build(SparseBuilderTMPL<DType>& builder, bool to_square) {
    const slv_int size = builder.getSize();
    std::vector<Triplet<DType>> t;
    const slv_int max_retu = builder.build(t, to_square);
    from_triplets(size, max_retu, t);
}

template <typename DType> slv_int SparseMatTMPL<DType>::getMaxCol() const {
    slv_int m = 0;
    for (slv_int c : ci_) m = std::max(m, c);
    return m;
}

template <typename DType> void SparseMatTMPL<DType>::add(slv_int gyo, slv_int retu, DType val) {
    const auto b = ci_.begin() + rp_[gyo], e = ci_.begin() + rp_[gyo + 1];
    auto it = std::lower_bound(b, e, retu);
    const size_t pos = (size_t)(it - ci_.begin());
    if (it != e && *it == retu) {
        v_[pos] += val;
    } else {
        ci_.insert(ci_.begin() + pos, retu);
        v_.insert(v_.begin() + pos, val);
        for (slv_int i = gyo + 1; i <= rows_; ++i) rp_[i]++;
        if (retu >= cols_) cols_ = retu + 1;
    }
    touch();
}

// SparseMat.cpp:159-179
template <typename DType> void SparseMatTMPL<DType>::getCols(slv_int* c1, slv_int* c2) const {
    for (slv_int i = 0; i < rows_; i++) {
        c1[i] = rp_[i];
        c2[i] = rp_[i + 1];
    }
}
template <typename DType>
void SparseMatTMPL<DType>::getCols(std::vector<slv_int>& c1, std::vector<slv_int>& c2) const {
    c1.resize(rows_);
    c2.resize(rows_);
    getCols(c1.data(), c2.data());
}

template <typename DType> void SparseMatTMPL<DType>::resetMat() {
    std::fill(v_.begin(), v_.end(), DType(0.0));
    touch();
}

template <typename DType> bool SparseMatTMPL<DType>::isInclude(slv_int gyo, slv_int target_r, slv_int& res) const {
    if (gyo >= 0 && gyo < rows_)
        for (slv_int k = rp_[gyo]; k < rp_[gyo + 1]; ++k)
            if (ci_[k] == target_r) {
                res = k - rp_[gyo];
                return true;
            }
    res = 0;
    return false;
}

template <typename DType>
void SparseMatTMPL<DType>::getTargetRowVal(slv_int target, std::vector<slv_int>& row_pos,
                                           std::vector<DType>& row_val) const {
    row_pos.clear();
    row_val.clear();
    for (slv_int i = 0; i < rows_; ++i)
        for (slv_int k = rp_[i]; k < rp_[i + 1]; ++k)
            if (ci_[k] == target) {
                row_pos.push_back(i);
                row_val.push_back(v_[k]);
            }
}

template <typename DType>
void SparseMatTMPL<DType>::getTargetColVal(slv_int target, std::vector<slv_int>& col_pos,
                                           std::vector<DType>& col_val) const {
    col_pos.clear();
    col_val.clear();
    for (slv_int k = rp_[target]; k < rp_[target + 1]; ++k) {
        col_pos.push_back(ci_[k]);
        col_val.push_back(v_[k]);
    }
}

// Eigen pruned(reference = thres): keeps |v| > thres * 1e-12
template <typename DType> void SparseMatTMPL<DType>::prune(double thres) {
    const double lim = thres * 1e-12;
    std::vector<slv_int> rp(rows_ + 1, 0), ci;
    std::vector<DType> v;
    for (slv_int i = 0; i < rows_; ++i) {
        for (slv_int k = rp_[i]; k < rp_[i + 1]; ++k)
            if (std::abs(v_[k]) > lim) {
                ci.push_back(ci_[k]);
                v.push_back(v_[k]);
            }
        rp[i + 1] = (slv_int)ci.size();
    }
    rp_.swap(rp);
    ci_.swap(ci);
    v_.swap(v);
    touch();
    dev_.reset();
}

// text format of SparseMat.cpp:302-344
template <typename DType> void SparseMatTMPL<DType>::printMat(const std::string& str) const {
    std::fstream fp(str, std::ios::out);
    fp << "Column size : " << rows_ << std::endl;
    for (slv_int i = 0; i < rows_; i++) fp << rp_[i + 1] - rp_[i] << std::endl;
    fp << "Column : " << std::endl;
    for (slv_int i = 0; i < rows_; i++) {
        if (rp_[i + 1] == rp_[i]) {
            fp << "-1" << std::endl;
        } else {
            for (slv_int k = rp_[i]; k < rp_[i + 1]; k++) fp << ci_[k] << " ";
            fp << std::endl;
        }
    }
    fp << "Mat value : " << std::endl;
    for (slv_int i = 0; i < rows_; i++) {
        for (slv_int k = rp_[i]; k < rp_[i + 1]; k++) fp << std::scientific << v_[k] << " ";
        fp << std::endl;
    }
}

template <typename DType> void SparseMatTMPL<DType>::print() const {
    for (slv_int i = 0; i < rows_; ++i)
        for (slv_int k = rp_[i]; k < rp_[i + 1]; ++k) std::cout << "(" << i << "," << ci_[k] << ") " << v_[k] << "\n";
}

// SparseMat.cpp:352-411 (without the per-entry echo to stdout)
template <typename DType> void SparseMatTMPL<DType>::readMat(const std::string& filename) {
    std::fstream fp(filename, std::ios::in);
    if (!fp) throw std::runtime_error("SparseMat::readMat: cannot open " + filename);
    std::string s1, s2, s3;
    slv_int size = 0;
    fp >> s1 >> s2 >> s3 >> size;
    SparseBuilderTMPL<DType> builder(size);
    std::vector<slv_int> sizes(size);
    std::vector<std::vector<slv_int>> cols(size);
    for (slv_int i = 0; i < size; i++) fp >> sizes[i];
    fp >> s1 >> s2;
    for (slv_int i = 0; i < size; i++) {
        if (sizes[i] == 0) {
            int dummy;
            fp >> dummy;
        } else {
            cols[i].resize(sizes[i]);
            for (slv_int j = 0; j < sizes[i]; j++) fp >> cols[i][j];
        }
    }
    fp >> s1 >> s2 >> s3;
    for (slv_int i = 0; i < size; i++)
        for (slv_int j = 0; j < sizes[i]; j++) {
            DType v;
            fp >> v;
            builder.add(i, cols[i][j], v);
        }
    build(builder, false);
}

template class SparseMatTMPL<double>;
template class SparseMatTMPL<dcomplex>;

// ============================================================ norms
double norm_naive(const double* b, slv_int n) {
    double s = 0;
    for (slv_int i = 0; i < n; i++) s += b[i] * b[i];
    return std::sqrt(s);
}

// Eigen's vectorized sum of squares (Redux.h:275-325, SSE2 Packet2d): two
// packet accumulators, then the tail.
double norm_eigen_order(const double* a, slv_int n) {
    if (n <= 0) return 0.0;
    const long long aS2 = (n / 4) * 4, aS = (n / 2) * 2;
    double r;
    auto e = [&](long long k) { return a[k] * a[k]; };
    if (aS) {
        double p0[2] = {e(0), e(1)};
        if (aS > 2) {
            double p1[2] = {e(2), e(3)};
            for (long long k = 4; k < aS2; k += 4) {
                p0[0] = p0[0] + e(k);
                p0[1] = p0[1] + e(k + 1);
                p1[0] = p1[0] + e(k + 2);
                p1[1] = p1[1] + e(k + 3);
            }
            p0[0] = p0[0] + p1[0];
            p0[1] = p0[1] + p1[1];
            if (aS > aS2) {
                p0[0] = p0[0] + e(aS2);
                p0[1] = p0[1] + e(aS2 + 1);
            }
        }
        r = p0[0] + p0[1];
        for (long long k = aS; k < n; ++k) r = r + e(k);
    } else {
        r = e(0);
    }
    return std::sqrt(r);
}

// complex: (re^2, im^2) summed with two Packet1cd accumulators, then re + im (Dot.h:21-27)
double norm_eigen_order(const dcomplex* a, slv_int n) {
    if (n <= 0) return 0.0;
    auto e = [&](long long k) { return dcomplex(a[k].real() * a[k].real(), a[k].imag() * a[k].imag()); };
    dcomplex p0 = e(0);
    if (n > 1) {
        const long long aE2 = (n / 2) * 2;
        dcomplex p1 = e(1);
        for (long long k = 2; k < aE2; k += 2) {
            const dcomplex q0 = e(k), q1 = e(k + 1);
            p0 = dcomplex(p0.real() + q0.real(), p0.imag() + q0.imag());
            p1 = dcomplex(p1.real() + q1.real(), p1.imag() + q1.imag());
        }
        p0 = dcomplex(p0.real() + p1.real(), p0.imag() + p1.imag());
        if (n > aE2) {
            const dcomplex q = e(aE2);
            p0 = dcomplex(p0.real() + q.real(), p0.imag() + q.imag());
        }
    }
    return std::sqrt(p0.real() + p0.imag());
}

// ============================================================ solvers
void MatSolverBase::getResidualLog(std::vector<double>& log) {
    log.clear();
    log.resize(residual_log.size());
    if (is_save_residual_log) std::copy(residual_log.begin(), residual_log.end(), log.begin());
}

bool MatSolverBase::run(int kind, double accel, int mrtr_fix, const double conv_cri, const int max_ite,
                        const double normB, const SparseMat* A, const SparseMatC* Ac, const void* b, void* x) {
    ssgpu_matrix h = A ? device_matrix(*A, device, abmc_block) : device_matrix(*Ac, device, abmc_block);
    ssgpu_options o;
    ssgpu_options_default(&o);
    o.is_init = is_init;
    o.is_diag_scale = is_diag_scale;
    o.is_save_best = is_save_best;
    o.is_save_log = is_save_residual_log;
    o.diverge_judge_type = diverge_judge_type;
    o.bad_div_val = bad_div_val;
    o.bad_div_count_thres = bad_div_count_thres;
    o.conv_normalize_type = conv_normalize_type;
    o.conv_normalize_const = conv_normalize_const;
    o.accel_factor = accel;
    o.factor_precision = factor_precision;
    o.mrtr_direction_fix = mrtr_fix;
    const int cap = is_save_residual_log ? (int)std::min<long long>((long long)std::max(max_ite, 0) + 2, 1 << 22) : 0;
    std::vector<double> log(std::max(cap, 1));
    ssgpu_result res{};
    const int st = ssgpu_solve(h, kind, &o, conv_cri, max_ite, normB, b, x, log.data(), cap, &res);
    if (st == SSGPU_NO_DIAGONAL) {
        std::cout << "there is not diagonal element! " << std::endl;
        if (strict_exit) std::exit(1);
        throw std::runtime_error("SparseSolv: a row of the matrix has no diagonal entry");
    }
    if (st != SSGPU_OK && st != SSGPU_NOT_CONVERGED)
        throw std::runtime_error(std::string("SparseSolv: GPU solve failed: ") + ssgpu_status_string(st) + ": " +
                                 ssgpu_last_error());
    if (is_save_residual_log) residual_log.assign(log.begin(), log.begin() + std::min(res.log_len, cap));
    last_iterations = res.iterations;
    last_setup_ms = res.setup_ms;
    last_iterate_ms = res.iterate_ms;
    if (!res.converged) std::cout << "not Convergence!!! " << std::endl;
    return res.converged != 0;
}

bool MatSolverBase::solve(const double conv_cri, const int max_ite, const SparseMat& A, const std::vector<double>& b,
                          std::vector<double>& x) {
    return solve(conv_cri, max_ite, norm_eigen_order(b.data(), (slv_int)b.size()), A, b.data(), x.data());
}
bool MatSolverBase::solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& A,
                          const std::vector<double>& b, std::vector<double>& x) {
    return solve(conv_cri, max_ite, normB, A, b.data(), x.data());
}
bool MatSolverBase::solve(const double conv_cri, const int max_ite, const SparseMatC& A,
                          const std::vector<dcomplex>& b, std::vector<dcomplex>& x) {
    return solve(conv_cri, max_ite, norm_eigen_order(b.data(), (slv_int)b.size()), A, b.data(), x.data());
}
bool MatSolverBase::solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& A,
                          const std::vector<dcomplex>& b, std::vector<dcomplex>& x) {
    return solve(conv_cri, max_ite, normB, A, b.data(), x.data());
}

// MatSolverICCG.cpp:44-58 (naive normB for the real pointer overload), :348-360 (Eigen norm for complex)
bool MatSolverICCG::solve(const double conv_cri, const int max_ite, const SparseMat& A, const double* b, double* x) {
    return solve(conv_cri, max_ite, norm_naive(b, A.getSize()), A, b, x);
}
bool MatSolverICCG::solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& A,
                          const double* b, double* x) {
    return run(SSGPU_ICCG, accel_factor, 0, conv_cri, max_ite, normB, &A, nullptr, b, x);
}
bool MatSolverICCG::solve(const double conv_cri, const int max_ite, const SparseMatC& A, const dcomplex* b,
                          dcomplex* x) {
    return solve(conv_cri, max_ite, norm_eigen_order(b, A.getSize()), A, b, x);
}
bool MatSolverICCG::solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& A,
                          const dcomplex* b, dcomplex* x) {
    return run(SSGPU_ICCG, accel_factor, 0, conv_cri, max_ite, normB, nullptr, &A, b, x);
}

// MatSolverMRTR_ICver.cpp:44-58, :99-104 (type 1 = SGS)
bool MatSolverMRTR::solve(const double conv_cri, const int max_ite, const SparseMat& A, const double* b, double* x) {
    return solve(conv_cri, max_ite, norm_naive(b, A.getSize()), A, b, x);
}
bool MatSolverMRTR::solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& A,
                          const double* b, double* x) {
    return run(precondition_type == 1 ? SSGPU_MRTR_SGS : SSGPU_MRTR_IC, accel_factor, direction_fix ? 1 : 0,
               conv_cri, max_ite, normB, &A, nullptr, b, x);
}
bool MatSolverMRTR::solve(const double conv_cri, const int max_ite, const SparseMatC& A, const dcomplex* b,
                          dcomplex* x) {
    return solve(conv_cri, max_ite, norm_eigen_order(b, A.getSize()), A, b, x);
}
bool MatSolverMRTR::solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& A,
                          const dcomplex* b, dcomplex* x) {
    return run(precondition_type == 1 ? SSGPU_MRTR_SGS : SSGPU_MRTR_IC, accel_factor, direction_fix ? 1 : 0,
               conv_cri, max_ite, normB, nullptr, &A, b, x);
}

}  // namespace SparseSolv
}  // namespace JPMRspace
