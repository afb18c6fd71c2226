/*
 * MatSolverMRTR.hpp -- MRTR with shifted-IC (type 0) or Eisenstat SGS (type 1,
 * default) preconditioning on the GPU.  Interface of the reference's
 * MatSolverMRTR (include/SparseSolv/MatSolverMRTR.hpp:36-88).
 * setDirectionFix(true) selects the standard search direction p = u + t*p
 * instead of the reference's p = u + t*u (MatSolverMRTR_ICver.cpp:302-304).
 */
#ifndef SPARSESOLV_AMD_MAT_SOLVER_MRTR_HPP
#define SPARSESOLV_AMD_MAT_SOLVER_MRTR_HPP

#include "MatSolverBase.hpp"

namespace JPMRspace {
namespace SparseSolv {

class MatSolverMRTR : public MatSolverBase {
  private:
    double accel_factor = 1.05;
    int precondition_type = 1;
    bool direction_fix = false;

  public:
    MatSolverMRTR() = default;
    ~MatSolverMRTR() override = default;
    void setAccelFactor(double accel_factor0) { accel_factor = accel_factor0; }
    void setPreconditionType(int precondition_type0) { precondition_type = precondition_type0; }
    void setDirectionFix(bool b) { direction_fix = b; }
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
