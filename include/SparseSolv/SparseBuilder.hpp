/*
 * SparseBuilder.hpp -- FEM-style sparse assembly (row maps, duplicates summed).
 *
 * Interface of the reference's SparseBuilderTMPL (include/SparseSolv/
 * SparseBuilder.hpp:53-80, src/SparseBuilder.cpp:37-141): add() accumulates
 * into per-row ordered maps; build() turns them into (row, col, value)
 * triplets and releases the maps.  Triplets are a plain struct here instead
 * of Eigen::Triplet.
 */
#ifndef SPARSESOLV_AMD_SPARSE_BUILDER_HPP
#define SPARSESOLV_AMD_SPARSE_BUILDER_HPP

#include <map>
#include <vector>

#include "SparseTypeDefine.hpp"

namespace JPMRspace {
namespace SparseSolv {

template <typename DType> struct Triplet {
    slv_int row, col;
    DType value;
};

template <typename DType> class SparseBuilderTMPL {
  private:
    slv_int size = 0;
    std::vector<std::map<slv_int, DType>> tempMat;

  public:
    SparseBuilderTMPL() = default;
    explicit SparseBuilderTMPL(slv_int size0) { this->initialize(size0); }
    slv_int getSize() const { return size; }
    void initialize(slv_int ss) { size = ss; initialize(); }
    void initialize() { tempMat.assign(size, std::map<slv_int, DType>()); }
    void reInitialize(slv_int new_size) { this->initialize(new_size); }
    void resetMat();
    void add(slv_int gyo, slv_int retu, DType val) { tempMat[gyo][retu] += val; }
    /* triplets row by row in ascending column order; returns max column + 1
       (padded to a square size when to_square and the matrix is narrower) */
    slv_int build(std::vector<Triplet<DType>>& tripletList, bool to_square);
    void printTempMat() const;
};

using SparseBuilder = SparseBuilderTMPL<double>;
using SparseBuilderC = SparseBuilderTMPL<dcomplex>;

extern template class SparseBuilderTMPL<double>;
extern template class SparseBuilderTMPL<dcomplex>;

}  // namespace SparseSolv
}  // namespace JPMRspace

#endif
