// Driver around the compiled reference (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/Makefile from the reference's own sources into
// oracle/_ref/libssref.so; exposes the reference's public C++ API
// (SparseMat, SprsPreconditioner, MatSolverICCG, MatSolverMRTR) through a
// C-ABI so tests can check the oracle restatement against the real thing.
// Nothing here is part of the product.
#include <SparseSolv/MatSolverICCG.hpp>
#include <SparseSolv/SparseBuilder.hpp>
#include <SparseSolv/MatSolverMRTR.hpp>
#include <SparseSolv/SprsPreconditioner.hpp>
#include <SparseSolv/SprsTrans.hpp>

#include <complex>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cpu_ref.h"

using namespace JPMRspace::SparseSolv;
using JPMRspace::dcomplex;

namespace {

template <typename T, typename M>
M make_mat(int n, const int* rp, const int* ci, const double* v) {
    std::vector<int> rows, cols;
    std::vector<T> vals;
    const T* tv = reinterpret_cast<const T*>(v);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            rows.push_back(i);
            cols.push_back(ci[k]);
            vals.push_back(tv[k]);
        }
    return M((int)rows.size(), rows, cols, vals);
}

void apply_opts(MatSolverBase& s, const ref_options* o) {
    s.setInitSolution(o->is_init != 0);
    s.setDiagScale(o->is_diag_scale != 0);
    s.setSaveBest(o->is_save_best != 0);
    s.setSaveLog(o->is_save_log != 0);
    s.setDirvegeType(o->diverge_judge_type);
    s.setBadDivVal(o->bad_div_val);
    s.setBadDivCount(o->bad_div_count_thres);
    s.setConvNormalizeType(o->conv_normalize_type);
    s.setConvNormalizeConst(o->conv_normalize_const);
}

template <typename T, typename M, typename V>
int solve_t(int solver, const ref_options* o, double conv_cri, int max_ite, int normb_mode, double normB, int n,
            const int* rp, const int* ci, const double* v, const double* b, double* x, double* log, int log_cap,
            int* log_len) {
    M A = make_mat<T, M>(n, rp, ci, v);
    MatSolverICCG iccg;
    MatSolverMRTR mrtr;
    MatSolverBase* base;
    if (solver == 0) {
        iccg.setAccelFactor(o->accel_factor);
        base = &iccg;
    } else {
        mrtr.setAccelFactor(o->accel_factor);
        mrtr.setPreconditionType(o->precondition_type);
        base = &mrtr;
    }
    apply_opts(*base, o);
    T* tx = reinterpret_cast<T*>(x);
    const T* tb = reinterpret_cast<const T*>(b);
    bool ok;
    if (normb_mode == 0) {            // pointer overload, explicit normB
        ok = base->solve(conv_cri, max_ite, normB, A, tb, tx);
    } else if (normb_mode == 1) {     // pointer overload, normB computed inside
        ok = base->solve(conv_cri, max_ite, A, tb, tx);
    } else {                          // Eigen overload, normB = b.norm()
        V eb(n), ex(n);
        for (int i = 0; i < n; ++i) { eb(i) = tb[i]; ex(i) = tx[i]; }
        ok = base->solve(conv_cri, max_ite, A, eb, ex);
        for (int i = 0; i < n; ++i) tx[i] = ex(i);
    }
    std::vector<double> lg;
    base->getResidualLog(lg);
    if (log_len) *log_len = (int)lg.size();
    for (int i = 0; i < (int)lg.size() && i < log_cap; ++i) log[i] = lg[i];
    return ok ? 1 : 0;
}

template <typename T, typename M>
void csr_out(const M& m, int* rp, int* ci, double* v) {
    const auto& e = m.mat();
    const int n = (int)e.rows();
    for (int i = 0; i <= n; ++i) rp[i] = e.outerIndexPtr()[i];
    std::memcpy(ci, e.innerIndexPtr(), sizeof(int) * e.nonZeros());
    std::memcpy(v, e.valuePtr(), sizeof(T) * e.nonZeros());
}

}  // namespace

extern "C" {

int refdrv_solve(int solver, int cplx, const ref_options* o, double conv_cri, int max_ite, int normb_mode,
                 double normB, int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                 double* log, int log_cap, int* log_len) {
    if (cplx)
        return solve_t<dcomplex, SparseMatC, Eigen::VectorXcd>(solver, o, conv_cri, max_ite, normb_mode, normB, n, rp,
                                                               ci, v, b, x, log, log_cap, log_len);
    return solve_t<double, SparseMat, Eigen::VectorXd>(solver, o, conv_cri, max_ite, normb_mode, normB, n, rp, ci, v,
                                                       b, x, log, log_cap, log_len);
}

void refdrv_ic_decomp(int cplx, int n, const int* rp, const int* ci, const double* v, double accel, int* Lrp,
                      int* Lci, double* Lv, double* diagD) {
    if (cplx) {
        SparseMatC A = make_mat<dcomplex, SparseMatC>(n, rp, ci, v);
        SparseMatC L = SprsPreconditioner::IC_decomp(A, reinterpret_cast<dcomplex*>(diagD), accel);
        csr_out<dcomplex>(L, Lrp, Lci, Lv);
    } else {
        SparseMat A = make_mat<double, SparseMat>(n, rp, ci, v);
        SparseMat L = SprsPreconditioner::IC_decomp(A, diagD, accel);
        csr_out<double>(L, Lrp, Lci, Lv);
    }
}

void refdrv_spmv(int cplx, int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    if (cplx) {
        SparseMatC A = make_mat<dcomplex, SparseMatC>(n, rp, ci, v);
        Eigen::VectorXcd ex(n), ey;
        for (int i = 0; i < n; ++i) ex(i) = reinterpret_cast<const dcomplex*>(x)[i];
        ey = A.mat() * ex;
        for (int i = 0; i < n; ++i) reinterpret_cast<dcomplex*>(y)[i] = ey(i);
    } else {
        SparseMat A = make_mat<double, SparseMat>(n, rp, ci, v);
        Eigen::VectorXd ex(n), ey;
        for (int i = 0; i < n; ++i) ex(i) = x[i];
        ey = A.mat() * ex;
        for (int i = 0; i < n; ++i) y[i] = ey(i);
    }
}

void refdrv_dot(int cplx, int n, const double* a, const double* b, double* out) {
    if (cplx) {
        Eigen::VectorXcd ea(n), eb(n);
        for (int i = 0; i < n; ++i) {
            ea(i) = reinterpret_cast<const dcomplex*>(a)[i];
            eb(i) = reinterpret_cast<const dcomplex*>(b)[i];
        }
        const dcomplex r = (ea.array() * eb.array()).sum();
        out[0] = r.real();
        out[1] = r.imag();
    } else {
        Eigen::VectorXd ea(n), eb(n);
        for (int i = 0; i < n; ++i) { ea(i) = a[i]; eb(i) = b[i]; }
        out[0] = ea.dot(eb);
    }
}

// The reference's own assembly: the COO constructor (mode 0), or
// SparseBuilder::add in the order given and SparseMat(builder, to_square)
// (mode 1: to_square false, mode 2: true).  Returns nnz.
long long refdrv_assemble(int cplx, int mode, int n, long long ntrip, const int* rows, const int* cols,
                          const double* vals, int* rp, int* ci, double* v) {
    auto run = [&](auto tag, auto* mat_tag) -> long long {
        using T = decltype(tag);
        using M = typename std::remove_pointer<decltype(mat_tag)>::type;
        const T* tv = reinterpret_cast<const T*>(vals);
        if (mode == 0) {
            std::vector<int> r(rows, rows + ntrip), c(cols, cols + ntrip);
            std::vector<T> x(tv, tv + ntrip);
            M A((int)ntrip, r, c, x);
            csr_out<T>(A, rp, ci, v);
            return (long long)A.mat().nonZeros();
        }
        SparseBuilderTMPL<T> b(n);
        for (long long i = 0; i < ntrip; ++i) b.add(rows[i], cols[i], tv[i]);
        M A(b, mode == 2);
        csr_out<T>(A, rp, ci, v);
        return (long long)A.mat().nonZeros();
    };
    if (cplx) return run(dcomplex{}, (SparseMatC*)nullptr);
    return run(double{}, (SparseMat*)nullptr);
}

double refdrv_norm(int cplx, int n, const double* a) {
    if (cplx) {
        Eigen::VectorXcd ea(n);
        for (int i = 0; i < n; ++i) ea(i) = reinterpret_cast<const dcomplex*>(a)[i];
        return ea.norm();
    }
    Eigen::VectorXd ea(n);
    for (int i = 0; i < n; ++i) ea(i) = a[i];
    return ea.norm();
}

}  // extern "C"
