/*
 * MatSolverICCG.hpp -- shifted ICCG (real) / ICCOCG (complex symmetric) on the GPU.
 * Interface of the reference's MatSolverICCG (include/SparseSolv/
 * MatSolverICCG.hpp:36-76); default shift 1.05, negative = automatic search.
 */
#ifndef SPARSESOLV_AMD_MAT_SOLVER_ICCG_HPP
#define SPARSESOLV_AMD_MAT_SOLVER_ICCG_HPP

#include "MatSolverBase.hpp"

namespace JPMRspace {
namespace SparseSolv {

class MatSolverICCG : public MatSolverBase {
  private:
    double accel_factor = 1.05;

  public:
    MatSolverICCG() = default;
    ~MatSolverICCG() override = default;
    void setAccelFactor(double accel_factor0) { accel_factor = accel_factor0; }
    using MatSolverBase::solve;
    bool solve(const double conv_cri, const int max_ite, const SparseMat& matA, const double* vecB,
               double* results) override;
    bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMat& matA,
               const double* vecB, double* results) override;
    bool solve(const double conv_cri, const int max_ite, const SparseMatC& matA, const dcomplex* vecB,
               dcomplex* results) override;
    bool solve(const double conv_cri, const int max_ite, const double normB, const SparseMatC& matA,
               const dcomplex* vecB, dcomplex* results) override;
};

}  // namespace SparseSolv
}  // namespace JPMRspace

#endif
