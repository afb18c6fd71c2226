/*
 * sparsesolv_amd.h -- C-ABI of the MI355X (gfx950) shifted-IC Krylov solvers.
 *
 * This is the drop-in boundary below the reference's C++ solver classes
 * (include/SparseSolv/MatSolverBase.hpp:42-152, MatSolverICCG.hpp:36-76,
 * MatSolverMRTR.hpp:36-88) and its pybind module (pybind/pybind.cpp:16-141).
 * Plain pointers and sizes only; every entry point returns an ssgpu_status.
 * No exception or process exit crosses this boundary: the reference's
 * exit(1) on a missing diagonal (SprsPreconditioner.cpp:91-94) is returned
 * as SSGPU_NO_DIAGONAL.
 *
 * Data: square CSR with int32 row offsets / column indices, both triangles
 * stored, column indices ascending within each row (the reference's
 * Eigen::SparseMatrix<T, RowMajor> storage, SparseMat.hpp:52).  Values are
 * double (SSGPU_R64) or interleaved (re, im) double pairs (SSGPU_C128).
 * Host arrays passed in are borrowed and copied; handles are owned by the
 * caller and released with the matching destroy call.
 */
#ifndef SPARSESOLV_AMD_H
#define SPARSESOLV_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SSGPU_OK = 0,
    SSGPU_NOT_CONVERGED = 1,   /* solve finished without meeting the tolerance */
    SSGPU_NO_DIAGONAL = 2,     /* a row of lower(A) does not end with its diagonal */
    SSGPU_BAD_ARG = 3,
    SSGPU_OOM = 4,
    SSGPU_HIP_ERROR = 5,       /* a HIP call failed, or a sync-free sweep exceeded its spin limit
                                  (a dependency was never produced); ssgpu_last_error() says which.
                                  After a spin-limit error the context and the matrix stay usable. */
    SSGPU_NOT_SQUARE = 6,
    SSGPU_UNSUPPORTED = 7
} ssgpu_status;

typedef enum { SSGPU_R64 = 0, SSGPU_C128 = 1 } ssgpu_dtype;

/* Solver kinds: MatSolverICCG (MatSolverICCG.cpp:178-328 / 477-626),
 * MatSolverMRTR with precondition type 0 (MatSolverMRTR_ICver.cpp:183-373)
 * and type 1 (MatSolverMRTR_SGSver.cpp:42-282). */
typedef enum { SSGPU_ICCG = 0, SSGPU_MRTR_IC = 1, SSGPU_MRTR_SGS = 2 } ssgpu_solver;

/* Storage precision of the IC factor and of the triangular sweeps. */
typedef enum { SSGPU_FACTOR_F64 = 0, SSGPU_FACTOR_F32 = 1 } ssgpu_precision;

typedef struct ssgpu_context_s* ssgpu_context;
typedef struct ssgpu_matrix_s* ssgpu_matrix;

/* Solver settings.  Field meaning and defaults follow MatSolverBase
 * (MatSolverBase.hpp:58-90, defaults MatSolverBase.cpp:32-47) and the
 * solver classes (accel 1.05: MatSolverICCG.hpp:40; MRTR precondition type
 * 1: MatSolverMRTR.hpp:43).  Fields after `accel_factor` are additions of
 * this library. */
typedef struct {
    int is_init;                 /* zero x before solving (1) or use it as start (0) */
    int is_diag_scale;           /* symmetric diagonal scaling before IC (ICCG / MRTR-IC) */
    int is_save_best;            /* restore best iterate on failure */
    int is_save_log;             /* record the normalized residual history */
    int diverge_judge_type;      /* 0 none, 1 best*bad_div_val for bad_div_count_thres its */
    double bad_div_val;
    int bad_div_count_thres;
    int conv_normalize_type;     /* 0 ||b||, 1 ||r0||, 2 conv_normalize_const */
    double conv_normalize_const;
    double accel_factor;         /* IC shift; < 0 selects the automatic shift search */
    int factor_precision;        /* ssgpu_precision: FP32 stores the IC factor (SGS-MRTR: the triangles
                                    of S A S) in FP32 for the sweeps; pivots, SpMV, dots and recurrences
                                    stay FP64 */
    int mrtr_direction_fix;      /* 0: reference p = u + t*u (MatSolverMRTR_ICver.cpp:302-304);
                                    1: standard MRTR p = u + t*p */
} ssgpu_options;

typedef struct {
    int status;                  /* SSGPU_OK or SSGPU_NOT_CONVERGED on a finished solve */
    int converged;
    int iterations;              /* Krylov iterations run */
    int log_len;                 /* entries of the residual log (0 unless is_save_log) */
    double final_residual;       /* last normalized residual */
    double accel_used;           /* IC shift actually used */
    int pivot_fixes;             /* pivots with |s| < 1e-12 replaced (SprsPreconditioner.cpp:105-106) */
    double setup_ms;             /* factorization + setup time on the device */
    double iterate_ms;           /* Krylov loop time on the device */
    double true_residual;        /* ||b - A x|| / normB of the returned x, recomputed in FP64 */
    int refinements;             /* FP32 factor: outer refinement passes on the FP64 residual */
    int refine_iterations;       /* Krylov iterations spent in those passes */
} ssgpu_result;

void ssgpu_options_default(ssgpu_options* o);
const char* ssgpu_status_string(int status);
/* Thread-local text of the last error raised in this thread. */
const char* ssgpu_last_error(void);

/* Context: one device and one HIP stream (default priority). */
int ssgpu_context_create(int device, ssgpu_context* out);
int ssgpu_context_destroy(ssgpu_context ctx);
int ssgpu_context_synchronize(ssgpu_context ctx);
/* Bytes of device memory currently held by the context: the matrices
 * created on it (with their preconditioners) and its work pool. */
int ssgpu_context_resident_bytes(ssgpu_context ctx, int64_t* bytes);
/* Free the context's work pool (Krylov vectors, reduction partials and block
 * scratch kept between solves); the next solve grows it again.  Waits for a
 * call running on the context to return.  Matrices are not touched.  (No
 * counterpart in the reference, whose work vectors live per solve call:
 * MatSolverICCG.cpp:191-200.  For callers that cap device memory.) */
int ssgpu_context_trim(ssgpu_context ctx);
/* Device memory held by the library in this process, all contexts: bytes
 * now, and the high-water mark since the process started or since the last
 * call with reset_peak != 0 (which then restarts it from the bytes now). */
int ssgpu_memory(int64_t* resident, int64_t* peak, int reset_peak);

/* Upload a square CSR matrix and analyse its structure (lower/upper split,
 * transpose map, level schedules).  Values are copied to the device. */
int ssgpu_matrix_create(ssgpu_context ctx, int dtype, int32_t n, const int32_t* row_ptr,
                        const int32_t* col_idx, const void* values, ssgpu_matrix* out);
/* Same matrix in the algebraic block multi-color (ABMC) order: rows
 * aggregated into blocks of at most block_size, blocks colored so that a
 * color's blocks share no entry (Iwashita et al., IPDPS 2012; the reference
 * names this only as a planned solver, pybind/MatSolvers_pybnd_add_memo.txt:13).
 * The library holds P A P^T and every solve on the handle gathers b and x
 * into that order and returns x in the caller's order; the kernel-level
 * entry points below work in the library's order.  Not a parity mode: the
 * IC(0) factor of P A P^T is a different preconditioner (other iteration
 * counts), in exchange for far fewer dependent sweep levels (at most
 * colors x block_size).  *ncolors (may be NULL) receives the color count. */
/* The ABMC order ssgpu_matrix_create_ordered uses: perm[new] = old row
 * (host only; needs no device).  No counterpart in the reference. */
int ssgpu_abmc_order(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int block_size,
                     int32_t* perm, int32_t* ncolors);

int ssgpu_matrix_create_ordered(ssgpu_context ctx, int dtype, int32_t n, const int32_t* row_ptr,
                                const int32_t* col_idx, const void* values, int block_size,
                                ssgpu_matrix* out, int32_t* ncolors);

/* Assemble a matrix of n rows on the device from ntrip (row, col, value)
 * triplets, taken in the order given, and create it as ssgpu_matrix_create
 * does.  Replaces the reference's host assembly: mode SSGPU_ASSEMBLE_COO is
 * the COO constructor SparseMat(n_nonZero, rows, cols, vals)
 * (SparseMat.cpp:79-104: Eigen's setFromTriplets, duplicates summed in the
 * order given starting from the first value); SSGPU_ASSEMBLE_BUILDER is
 * SparseBuilder(n) with one add(row, col, val) per triplet in order, then
 * SparseMat(builder) (SparseBuilder.hpp:75, SparseBuilder.cpp:107-141,
 * SparseMat.cpp:113-122: every entry summed from zero in the order given);
 * SSGPU_ASSEMBLE_BUILDER_SQUARE the same with to_square (an explicit zero at
 * (n-1, n-1) when no column reaches n-1).  The values are bit-identical to
 * the reference's.  rows, cols (int32) and values (dtype) are device
 * pointers when on_device != 0, host arrays otherwise; indices must lie in
 * [0, n) (SSGPU_BAD_ARG otherwise) and ntrip < 2^31 - 1.  The sort needs
 * about 40 bytes of device memory per triplet while it runs. */
typedef enum { SSGPU_ASSEMBLE_COO = 0, SSGPU_ASSEMBLE_BUILDER = 1, SSGPU_ASSEMBLE_BUILDER_SQUARE = 2 } ssgpu_assembly;
int ssgpu_matrix_create_triplets(ssgpu_context ctx, int dtype, int32_t n, int64_t ntrip, const int32_t* rows,
                                 const int32_t* cols, const void* values, int mode, int on_device,
                                 ssgpu_matrix* out);
/* The matrix as the library stores it (an ABMC-ordered handle holds
 * P A P^T): row_ptr (n+1), col_idx and values (nnz from ssgpu_matrix_info)
 * into host arrays; any pointer may be NULL.  The reference's accessors of
 * the stored matrix (SparseMat.hpp:71-78, getCols SparseMat.cpp:159-179). */
int ssgpu_matrix_get_csr(ssgpu_matrix m, int32_t* row_ptr, int32_t* col_idx, void* values);

/* Replace the values (same pattern) from host memory.  A factor built from
 * the old values is dropped: cached solves need ssgpu_factorize again. */
int ssgpu_matrix_update_values(ssgpu_matrix m, const void* values);
int ssgpu_matrix_destroy(ssgpu_matrix m);
int ssgpu_matrix_info(ssgpu_matrix m, int32_t* n, int64_t* nnz, int* dtype,
                      int32_t* fwd_levels, int32_t* bwd_levels);
/* The IC sweeps' launch shape, as chosen by the tuner at the first
 * factorization (or fixed by SSGPU_SWEEP_MODE): *mode 0 sync-free 64-row
 * chunks, 1 one launch per level, 2 tiles (level bands cut into row-index
 * intervals, in-tile inputs handed over in LDS); *grid its persistent
 * workgroups; *tiles the forward sweep's tile count (0 when none were
 * built).  Every shape gives the same bits.  No counterpart in the reference
 * (its sweeps are sequential loops, MatSolverBase.cpp:177-231). */
int ssgpu_matrix_sweep_info(ssgpu_matrix m, int* mode, int* grid, int32_t* tiles);

/* Solve A x = b like MatSolver*::solve(conv_cri, max_ite, normB, A, b, x)
 * (MatSolverBase.hpp:144-150): the factor is rebuilt on every call, as the
 * reference does.  b and x are host arrays of n values (x in/out).
 * normB is the right-hand-side norm the caller computed (the reference's
 * overloads compute it from b on the host).  log (may be NULL) receives up to
 * log_cap residual-log entries; res->log_len holds the full length. */
int ssgpu_solve(ssgpu_matrix m, int solver, const ssgpu_options* opt, double conv_cri, int max_ite,
                double normB, const void* b, void* x, double* log, int log_cap, ssgpu_result* res);

/* Same, with b and x already resident in device memory (ordered on the
 * context's stream). */
int ssgpu_solve_device(ssgpu_matrix m, int solver, const ssgpu_options* opt, double conv_cri,
                       int max_ite, double normB, const void* d_b, void* d_x, double* log,
                       int log_cap, ssgpu_result* res);

/* Cached preconditioner for repeated-matrix solves (serving, batched RHS).
 * The reference rebuilds the factor on every solve (MatSolverICCG.cpp:109-118,
 * MatSolverMRTR_ICver.cpp:113-123); these two calls split that: build the
 * shifted IC(0) factor once (same kernels and result as inside ssgpu_solve;
 * accel < 0 runs the automatic shift search of MatSolverBase.cpp:344-391),
 * or for SSGPU_MRTR_SGS the scaled matrix S A S and its triangles
 * (MatSolverMRTR_SGSver.cpp:42-69), then solve any number of right-hand
 * sides against it.  ssgpu_solve / ssgpu_solve_device on the same matrix
 * overwrite the cached preconditioner. */
int ssgpu_factorize(ssgpu_matrix m, int solver, double accel, int factor_precision, double* accel_used,
                    int* pivot_fixes);

/* One solve against the preconditioner of ssgpu_factorize, on ctx's stream
 * and with ctx's work space; b and x are device pointers.  Nothing of the
 * matrix is written: threads may run solves on one matrix at once through
 * different contexts (of the matrix's device), and their work overlaps on
 * the GPU.  Results are the same as ssgpu_solve_device's with that factor
 * (setup_ms is 0).  Diagonal scaling (is_diag_scale) is not available here. */
int ssgpu_solve_factored(ssgpu_context ctx, ssgpu_matrix m, int solver, const ssgpu_options* opt,
                         double conv_cri, int max_ite, double normB, const void* d_b, void* d_x,
                         double* log, int log_cap, ssgpu_result* res);

/* Allocation-free form of ssgpu_solve_factored, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and the same results, bit for bit; the solve's work vectors,
 * reduction partials, scalar block and sweep tickets come from a fixed device
 * array the library holds on each GPU (256 MiB, enough for about 4 million
 * real or 2 million complex unknowns in ICCG, fewer for the MRTR solvers),
 * and the call makes no device allocation (it needs a factor built by
 * ssgpu_factorize and a context created beforehand).  A system whose work
 * space does not fit returns SSGPU_BAD_ARG with the sizes in
 * ssgpu_last_error(); ABMC-ordered matrices return SSGPU_UNSUPPORTED.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from different threads or contexts) race and must be serialized by the
 * caller.  iterate_ms is host wall time. */
int ssgpu_solve_factored_noalloc(ssgpu_context ctx, ssgpu_matrix m, int solver, const ssgpu_options* opt,
                                 double conv_cri, int max_ite, double normB, const void* d_b, void* d_x,
                                 double* log, int log_cap, ssgpu_result* res);

/* Block solve: nrhs (1..32) right-hand sides of one system against the
 * preparation of ssgpu_factorize for `solver` (the IC factor for SSGPU_ICCG
 * and SSGPU_MRTR_IC, S A S and its triangles for SSGPU_MRTR_SGS), in one
 * pass of kernels per iteration: the sweeps carry every column through the
 * dependency chain at once and the SpMV reads the matrix once for all
 * columns.  d_B and d_X hold nrhs vectors of n entries one after the other
 * (entry i of vector j at j * n + i); normB[j] as for ssgpu_solve; res[j]
 * per column (no residual log).  Column j returns the same x, iteration
 * count, residuals and refinement passes as ssgpu_solve_factored with the
 * same solver on that column alone.  No counterpart in the reference, which
 * solves one right-hand side per call (MatSolverBase.hpp:142-150); the
 * iterations are those of MatSolverICCG.cpp:257-318,
 * MatSolverMRTR_ICver.cpp:270-360 and MatSolverMRTR_SGSver.cpp:182-262. */
int ssgpu_solve_block(ssgpu_context ctx, ssgpu_matrix m, int solver, const ssgpu_options* opt, double conv_cri,
                      int max_ite, int nrhs, const double* normB, const void* d_B, void* d_X,
                      ssgpu_result* res);

/* Batched ICCG / COCG, IC-MRTR or SGS-MRTR over many independent systems (the caller's own
 * matrices, of any sizes; one matrix may appear several times with different
 * right-hand sides).  No counterpart in the reference, which solves one system
 * per call with its Krylov loop on the host (MatSolverICCG.cpp:257-318): here
 * every phase of the iteration is one launch for all systems of the batch,
 * the scalar recurrences and the convergence bookkeeping run on the device,
 * and the iteration is replayed from a hipGraph.  Each system returns the same
 * x, iteration count and residuals as ssgpu_solve_factored on it alone (same
 * factor, same options).
 *
 * A batch is a session on one context for one solver (SSGPU_ICCG,
 * SSGPU_MRTR_IC: MatSolverMRTR_ICver.cpp:183-373 with its quirks, or the
 * standard direction with mrtr_direction_fix; or SSGPU_MRTR_SGS:
 * MatSolverMRTR_SGSver.cpp:42-282 with S A S from ssgpu_factorize(...,
 * SSGPU_MRTR_SGS, ...)): systems join with ssgpu_batch_add
 * (the matrix prepared by ssgpu_factorize for that solver, and not
 * refactorized, updated or destroyed while any of its systems is in the
 * session; no diagonal scaling for the IC solvers, SGS-MRTR always scales;
 * d_b and d_x device vectors of the matrix's n entries that stay valid until
 * the system's result is polled; d_x is the start vector when
 * opt->is_init == 0 and receives the solution) and enter the loop at the
 * next ssgpu_batch_step, which runs `iterations` iterations of every system
 * still iterating and then finishes those that left the loop (converged,
 * diverged or at max_ite).  ssgpu_batch_poll returns one finished system at a
 * time (*got = 0 when none is waiting), with the tag given at add and its
 * result (no residual log).  Systems may be added between any two steps while
 * others run (continuous batching).  All matrices of a batch share the value
 * type (dtype) and the factor precision.  A batch is used by one thread at a
 * time; it runs on its context's stream.  After a sweep error (SSGPU_HIP_ERROR
 * from a step) every system of the batch has been dropped with that status;
 * the batch and the matrices stay usable. */
typedef struct ssgpu_batch_s* ssgpu_batch;
int ssgpu_batch_create(ssgpu_context ctx, int solver, int dtype, int factor_precision, ssgpu_batch* out);
int ssgpu_batch_add(ssgpu_batch b, ssgpu_matrix m, const ssgpu_options* opt, double conv_cri, int max_ite,
                    double normB, const void* d_b, void* d_x, int64_t tag);
/* Same with b and x in host memory (the reference's pointer overloads,
 * MatSolverBase.hpp:144-146): both are copied in at the call; x must stay
 * valid until the result is polled, and receives the solution when the
 * system finishes. */
int ssgpu_batch_add_host(ssgpu_batch b, ssgpu_matrix m, const ssgpu_options* opt, double conv_cri, int max_ite,
                         double normB, const void* b_host, void* x_host, int64_t tag);
int ssgpu_batch_step(ssgpu_batch b, int iterations, int* running, int* finished);
int ssgpu_batch_poll(ssgpu_batch b, int* got, int64_t* tag, ssgpu_result* res);
/* Systems in the batch, iterations run so far, and the rebuilds whose graph
 * capture the runtime refused (those steps ran the same kernels eagerly). */
int ssgpu_batch_info(ssgpu_batch b, int* systems, int64_t* iterations_run, int* graph_fallbacks);
int ssgpu_batch_destroy(ssgpu_batch b);
/* Free the device vectors of the session's idle slots (kept for reuse by the
 * next system added); systems in the session are not touched.  For callers
 * that cap device memory between bursts of work. */
int ssgpu_batch_trim(ssgpu_batch b);
/* One-shot form: nsys systems with shared options, tolerance and iteration
 * limit; res[i] for system i.  Returns SSGPU_OK when every system converged,
 * SSGPU_NOT_CONVERGED otherwise. */
int ssgpu_solve_batch(ssgpu_context ctx, int solver, int nsys, const ssgpu_matrix* mats, const ssgpu_options* opt,
                      double conv_cri, int max_ite, const double* normB, const void* const* d_b, void* const* d_x,
                      ssgpu_result* res);

/* Kernel-level entry points (device pointers, context stream), used by the
 * parity tests to check each piece against the reference's arithmetic. */
int ssgpu_spmv_device(ssgpu_matrix m, const void* d_x, void* d_y);
/* Shifted IC(0) on the device; copies L (lower(A) layout, n+1 / nnzL /
 * nnzL) and diagD back to the host arrays given. */
int ssgpu_ic_factor_host(ssgpu_matrix m, double accel, int32_t* L_row_ptr, int32_t* L_col_idx,
                         void* L_values, void* diagD, int* pivot_fixes);
/* z = (L D L^T)^{-1} r with the factor of the last ssgpu_ic_factor_host call. */
int ssgpu_ic_apply_device(ssgpu_matrix m, const void* d_r, void* d_z);
/* Plain sweeps on lower(A) / upper(A) (MatSolverBase.cpp:243-327). */
int ssgpu_lower_sweep_device(ssgpu_matrix m, const void* d_r, void* d_z);
int ssgpu_upper_sweep_device(ssgpu_matrix m, const void* d_r, void* d_z);
/* Dot products: dot (non-conjugated for complex) and 2-norm. */
int ssgpu_dot_device(ssgpu_context ctx, int dtype, int32_t n, const void* d_a, const void* d_b,
                     double* out /* 1 or 2 doubles */);
int ssgpu_norm_device(ssgpu_context ctx, int dtype, int32_t n, const void* d_a, double* out);

/* Measurement hook: average device time (HIP events on the context stream)
 * of `reps` back-to-back runs of one piece of the path.
 *   which = 0 SpMV y = A x          (d_in = x, d_out = y)
 *           1 IC forward sweep      (d_in = r, d_out = t)   needs a factor
 *           2 IC backward sweep     (d_in = t, d_out = z)   needs a factor
 *           3 IC apply (both)       (d_in = r, d_out = z)   needs a factor
 *           4 shifted IC(0) factorization at shift `accel` (d_in, d_out unused)
 * factor_precision selects the FP64 or FP32 factor copy for 1-3. */
int ssgpu_time_kernel(ssgpu_matrix m, int which, int reps, double accel, int factor_precision,
                      const void* d_in, void* d_out, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
