/*
 * cpu_ref.h -- C-ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A plain C++ restatement of the reference's preconditioned Krylov path:
 * shifted IC(0) (SprsPreconditioner.cpp:42-111), IC forward/backward apply and
 * plain sweeps (MatSolverBase.cpp:177-327), diagonal scaling
 * (SprsPreconditioner.cpp:150-231, MatSolverBase.cpp:99-141), automatic shift
 * (MatSolverBase.cpp:344-391), shifted ICCG/ICCOCG (MatSolverICCG.cpp), IC-MRTR
 * (MatSolverMRTR_ICver.cpp) and Eisenstat SGS-MRTR (MatSolverMRTR_SGSver.cpp),
 * with the arithmetic order of the vendored Eigen 5.0.0 for the row-major
 * SpMV (SparseCore/SparseDenseProduct.h:66-80) and for the SSE2 vectorized
 * reductions behind dot / squaredNorm (Core/Redux.h:275-325, Core/Dot.h:21-27).
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may load
 * this library, and only as the checker or the timed CPU baseline; the product
 * path never calls it.
 *
 * Complex data is passed as interleaved (re, im) doubles.
 */
#ifndef SPARSESOLV_CPU_REF_H
#define SPARSESOLV_CPU_REF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int is_init;               /* MatSolverBase.cpp:33 default 1 */
    int is_diag_scale;         /* default 0 */
    int is_save_best;          /* default 0 */
    int is_save_log;           /* default 0 */
    int diverge_judge_type;    /* default 0 */
    double bad_div_val;        /* default 1000 */
    int bad_div_count_thres;   /* default 1000 */
    int conv_normalize_type;   /* default 0 */
    double conv_normalize_const; /* default 1 */
    double accel_factor;       /* MatSolverICCG.hpp:40 default 1.05 */
    int precondition_type;     /* MRTR only: 0 IC, 1 SGS (MatSolverMRTR.hpp:43 default 1) */
    int mrtr_direction_fix;    /* IC-MRTR only: 0 reference p = u + t*u, 1 standard p = u + t*p */
} ref_options;

void ref_options_default(ref_options* o);

/* Reduction order of dots and norms: 0 = the reference's (Eigen packet order,
 * default), 1 = the GPU library's fixed tree.  Mode 1 exists so tests can show
 * that the GPU path differs from the reference only through this order. */
void ref_set_reduction_order(int mode);

/* Number of stored entries of lower(A) (col <= row). */
long long ref_lower_nnz(int n, const int* rp, const int* ci);

/* y = A x, Eigen row-major two-accumulator order; OpenMP over rows when
 * nnz > 20000 (the reference's only parallel region). */
void ref_spmv(int cplx, int n, const int* rp, const int* ci, const double* v,
              const double* x, double* y);

/* Eigen-order reductions: real dot / complex non-conjugated sum(a.*b), and
 * the 2-norm (Hermitian for complex). */
void ref_dot(int cplx, int n, const double* a, const double* b, double* out /* 1 or 2 */);
double ref_norm(int cplx, int n, const double* a);
/* Naive sequential sqrt(sum b_i^2) of the real pointer overloads
 * (MatSolverICCG.cpp:48-54). */
double ref_norm_naive(int n, const double* b);

/* Shifted IC(0): writes lower(A) factor (Lrp n+1, Lci/Lv nnzL) and diagD (n).
 * Returns 0, or 2 when a row of lower(A) does not end with its diagonal (the
 * reference exits the process there, SprsPreconditioner.cpp:91-94). */
int ref_ic_decomp(int cplx, int n, const int* rp, const int* ci, const double* v, double accel,
                  int* Lrp, int* Lci, double* Lv, double* diagD);

/* Automatic shift (MatSolverBase.cpp:344-391); returns the shift used. */
double ref_auto_accel(int cplx, int n, const int* rp, const int* ci, const double* v, double accel_ini,
                      int* Lrp, int* Lci, double* Lv, double* diagD);

/* Transpose of a square CSR (Eigen transpose into row-major). */
void ref_transpose(int cplx, int n, const int* rp, const int* ci, const double* v,
                   int* Trp, int* Tci, double* Tv);

/* z = (L D L^T)^{-1} r with L and U = L^T as produced above. */
void ref_ic_apply(int cplx, int n, const int* Lrp, const int* Lci, const double* Lv,
                  const int* Urp, const int* Uci, const double* Uv, const double* diagD,
                  const double* r, double* z);
/* Plain lower sweep dividing by each row's last entry / upper sweep dividing
 * by each row's first entry (MatSolverBase.cpp:243-327). */
void ref_fr(int cplx, int n, const int* Lrp, const int* Lci, const double* Lv, const double* r, double* z);
void ref_bc(int cplx, int n, const int* Urp, const int* Uci, const double* Uv, const double* r, double* z);

/* Diagonal scaling factors s_i = 1/sqrt|a_ii| and b~ = s.b (returns 0). */
int ref_diag_scaling(int cplx, int n, const int* rp, const int* ci, const double* v,
                     const double* b, double* s, double* bs);

/* Full solve, mirroring MatSolver*::solve(conv_cri, max_ite, normB, A, b, x).
 * solver: 0 = MatSolverICCG, 1 = MatSolverMRTR (precondition_type selects IC
 * or SGS).  x is in/out.  The residual log (if is_save_log) is copied into
 * log[0..log_cap) and its full length stored in *log_len.  Returns 1 when
 * converged, 0 when not, -2 on a missing diagonal. */
int ref_solve(int solver, int cplx, const ref_options* opt, double conv_cri, int max_ite, double normB,
              int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
              double* log, int log_cap, int* log_len);

/* CSR of n rows from ntrip (row, col, value) triplets taken in the order
 * given.  mode 0: SparseMat(n_nonZero, rows, cols, vals) (SparseMat.cpp:79-104,
 * Eigen's setFromTriplets: duplicates summed in order from the first value);
 * mode 1: SparseBuilder::add then SparseMat::build (SparseBuilder.hpp:75,
 * SparseBuilder.cpp:107-141, SparseMat.cpp:113-122: each entry summed from
 * zero in order); mode 2: the same with to_square (an explicit zero at
 * (n-1, n-1) when no column reaches n-1).  rp: n+1; ci, v: ntrip+1 entries.
 * Returns nnz, or -1 when an index is outside [0, n). */
long long ref_assemble(int cplx, int mode, int n, long long ntrip, const int* rows, const int* cols,
                       const double* vals, int* rp, int* ci, double* v);

#ifdef __cplusplus
}
#endif
#endif
