/*
 * MatSolverBase.hpp -- base of the GPU solver classes.
 *
 * Same settings, defaults and residual log as the reference's MatSolverBase
 * (include/SparseSolv/MatSolverBase.hpp:42-152, src/MatSolverBase.cpp:32-87).
 * The solve overloads take raw pointers (the reference's pointer overloads),
 * std::vector, or any vector type with data() and norm() -- Eigen::VectorXd /
 * VectorXcd among them, so the reference's Eigen overloads compile unchanged
 * while this library does not depend on Eigen.  Work runs on one MI355X
 * through the C-ABI in sparsesolv_amd.h.
 */
#ifndef SPARSESOLV_AMD_MAT_SOLVER_BASE_HPP
#define SPARSESOLV_AMD_MAT_SOLVER_BASE_HPP

#include <type_traits>
#include <utility>
#include <vector>

#include "SparseMat.hpp"

namespace JPMRspace {
namespace SparseSolv {

class MatSolverBase {
  protected:
    bool is_init = true;
    bool is_diag_scale = false;
    bool is_save_best = false;
    bool is_save_residual_log = false;
    std::vector<double> residual_log;
    int diverge_judge_type = 0;
    double bad_div_val = 1000.0;
    int bad_div_count_thres = 1000;
    int conv_normalize_type = 0;
    double conv_normalize_const = 1.0;
    /* additions of this library */
    int device = 0;
    int factor_precision = 0;     /* 0 FP64 factor, 1 FP32 factor (FP64 pivots / recurrences) */
    bool strict_exit = false;     /* missing diagonal: false throws, true prints and exit(1) like the reference */
    int abmc_block = 0;
    int last_iterations = 0;
    double last_setup_ms = 0.0, last_iterate_ms = 0.0;

    /* common driver: solver kind of sparsesolv_amd.h, values as interleaved doubles */
    bool run(int kind, double accel, int mrtr_fix, const double conv_cri, const int max_ite, const double normB,
             const SparseMat* A, const SparseMatC* Ac, const void* b, void* x);

  public:
    MatSolverBase() = default;
    virtual ~MatSolverBase() = default;
    void setInitSolution(bool bl) { is_init = bl; }
    void setDiagScale(bool bl) { is_diag_scale = bl; }
    void setSaveBest(bool bl) { is_save_best = bl; }
    void setSaveLog(bool bl) { is_save_residual_log = bl; }
    void getResidualLog(std::vector<double>& log);
    void setDirvegeType(int x) { diverge_judge_type = x; }
    void setBadDivVal(double x) { bad_div_val = x; }
    void setBadDivCount(int x) { bad_div_count_thres = x; }
    void setConvNormalizeType(int x) { conv_normalize_type = x; }
    void setConvNormalizeConst(double x) { conv_normalize_const = x; }
    /* additions */
    void setDevice(int d) { device = d; }
    void setFactorPrecision(int p) { factor_precision = p; }
    void setStrictExit(bool b) { strict_exit = b; }
    /* 0: the matrix's own row order (parity with the reference); b > 0: solve
       on the algebraic block multi-color reordering with blocks of b rows
       (fewer dependent sweep levels, different iteration counts) */
    void setABMC(int block_size) { abmc_block = block_size > 0 ? block_size : 0; }
    int getIterations() const { return last_iterations; }
    double getSetupMs() const { return last_setup_ms; }
    double getIterateMs() const { return last_iterate_ms; }

    virtual bool solve(const double conv_cri, const int max_ite, const SparseMat& matA, const double* vecB,
                       double* results) = 0;
    virtual bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& matA,
                       const double* vecB, double* results) = 0;
    virtual bool solve(const double conv_cri, const int max_ite, const SparseMatC& matA, const dcomplex* vecB,
                       dcomplex* results) = 0;
    virtual bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& matA,
                       const dcomplex* vecB, dcomplex* results) = 0;
    /* std::vector forms of the reference's Eigen overloads (normB = Eigen-order ||b||) */
    bool solve(const double conv_cri, const int max_ite, const SparseMat& matA, const std::vector<double>& vecB,
               std::vector<double>& results);
    bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& matA,
               const std::vector<double>& vecB, std::vector<double>& results);
    bool solve(const double conv_cri, const int max_ite, const SparseMatC& matA, const std::vector<dcomplex>& vecB,
               std::vector<dcomplex>& results);
    bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& matA,
               const std::vector<dcomplex>& vecB, std::vector<dcomplex>& results);
    /* the reference's Eigen overloads (normB = vecB.norm()) for any vector
       type with data() and norm(), e.g. Eigen::VectorXd / VectorXcd */
    template <typename Vec, typename = decltype(std::declval<const Vec&>().norm()),
              typename = decltype(std::declval<Vec&>().data())>
    bool solve(const double conv_cri, const int max_ite, const SparseMat& matA, const Vec& vecB, Vec& results) {
        return solve(conv_cri, max_ite, (double)vecB.norm(), matA, vecB.data(), results.data());
    }
    template <typename Vec, typename = decltype(std::declval<const Vec&>().norm()),
              typename = decltype(std::declval<Vec&>().data())>
    bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& matA, const Vec& vecB,
               Vec& results) {
        return solve(conv_cri, max_ite, normB, matA, vecB.data(), results.data());
    }
    template <typename Vec, typename = decltype(std::declval<const Vec&>().norm()),
              typename = decltype(std::declval<Vec&>().data())>
    bool solve(const double conv_cri, const int max_ite, const SparseMatC& matA, const Vec& vecB, Vec& results) {
        return solve(conv_cri, max_ite, (double)vecB.norm(), matA, vecB.data(), results.data());
    }
    template <typename Vec, typename = decltype(std::declval<const Vec&>().norm()),
              typename = decltype(std::declval<Vec&>().data())>
    bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& matA, const Vec& vecB,
               Vec& results) {
        return solve(conv_cri, max_ite, normB, matA, vecB.data(), results.data());
    }
};

/* ||b|| in the reference's orders: the sequential loop of its real pointer
   overloads (MatSolverICCG.cpp:48-54) and Eigen's vectorized norm()
   (complex pointer and all Eigen overloads). */
double norm_naive(const double* b, slv_int n);
double norm_eigen_order(const double* b, slv_int n);
double norm_eigen_order(const dcomplex* b, slv_int n);

}  // namespace SparseSolv
}  // namespace JPMRspace

#endif
