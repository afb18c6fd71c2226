/*
 * SprsTrans.hpp -- structural transforms of SparseMat (host side).
 *
 * Interface of the reference's SprsTrans namespace (include/SparseSolv/
 * SprsTrans.hpp:26-72, src/SprsTrans.cpp:252-425): the same functions with
 * the same results, on the plain-vector CSR of this library's SparseMat.
 * DelFlagPosition / DiagFlagPosition are the Dirichlet eliminations a finite
 * element caller applies before solving.
 */
#ifndef SPARSESOLV_AMD_SPRS_TRANS_HPP
#define SPARSESOLV_AMD_SPRS_TRANS_HPP

#include "SparseMat.hpp"

namespace JPMRspace {
namespace SparseSolv {
namespace SprsTrans {

/* round every value to the nearest integer, then drop the zeros (SprsTrans.cpp:252-260) */
void round(SparseMat& matA);
/* matA = a1 * matA (SprsTrans.cpp:268-282) */
void SelfDot(SparseMat& matA, double a1);
void SelfDot(SparseMatC& matA, double a1);
/* drop entries with |v| <= thres * 1e-12 (SprsTrans.cpp:290-293) */
void prune(SparseMat& matA, double thres);
/* transpose (SprsTrans.cpp:300-307) */
SparseMat trans(const SparseMat& matA);
SparseMatC trans(const SparseMatC& matA);
/* rows [r1a, r1b] x columns [r2a, r2b] as a new matrix (SprsTrans.cpp:36-61, 317-326) */
SparseMat makeSubMat(const SparseMat& matA, slv_int range1a, slv_int range1b, slv_int range2a, slv_int range2b);
SparseMatC makeSubMat(const SparseMatC& matA, slv_int range1a, slv_int range1b, slv_int range2a, slv_int range2b);
/* rows 0..rangeA split at column rangeB: K11 = columns <= rangeB, K12 = the rest (SprsTrans.cpp:76-113) */
void MatDiv(SparseMat& matK11, SparseMat& matK12, const SparseMat& matA, slv_int rangeA, slv_int rangeB);
void MatDiv(SparseMatC& matK11, SparseMatC& matK12, const SparseMatC& matA, slv_int rangeA, slv_int rangeB);
/* entries with column <= row / column >= row (SprsTrans.cpp:358-385) */
SparseMat MatLower(const SparseMat& matA);
SparseMatC MatLower(const SparseMatC& matA);
SparseMat MatUpper(const SparseMat& matA);
SparseMatC MatUpper(const SparseMatC& matA);
/* delete the rows and columns whose flag is true, renumbering the rest (SprsTrans.cpp:125-177) */
void DelFlagPosition(SparseMat& matA, const bool* flag);
void DelFlagPosition(SparseMatC& matA, const bool* flag);
/* zero the flagged rows and columns, 1 on a flagged row's stored diagonal (SprsTrans.cpp:188-237) */
void DiagFlagPosition(SparseMat& matA, const bool* flag);
void DiagFlagPosition(SparseMatC& matA, const bool* flag);

}  // namespace SprsTrans
}  // namespace SparseSolv
}  // namespace JPMRspace

#endif
