/*
 * SparseTypeDefine.hpp -- integer and complex types of the solver API.
 * Same names as the reference (include/SparseSolv/SparseTypeDefine.hpp:41,
 * BasicDefinesC.hpp:10): slv_int = int, dcomplex = std::complex<double>.
 */
#ifndef SPARSESOLV_AMD_SPARSE_TYPE_DEFINE_HPP
#define SPARSESOLV_AMD_SPARSE_TYPE_DEFINE_HPP

#include <complex>

namespace JPMRspace {
using dcomplex = std::complex<double>;
namespace SparseSolv {
using slv_int = int;
}  // namespace SparseSolv
}  // namespace JPMRspace

#endif
