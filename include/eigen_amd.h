/*
 * eigen_amd.h -- C ABI of the MI355X (gfx950) dense factor-and-solve library.
 *
 * Plain C: pointers, sizes and status codes only.  Matrices are column-major
 * with a leading dimension (Eigen's default storage, Map<Matrix,0,
 * OuterStride<>>).  Sizes are int64_t; pivots are int32 like Eigen's
 * default PermutationMatrix StorageIndex.  Unless a function is named
 * `eamd_host_*`, every matrix/vector pointer is a DEVICE pointer and the work
 * is queued on the handle's HIP stream.
 *
 * Status: 0 = success; -i = argument i (1-based) is invalid, the reference's
 * xerbla convention (lapack/lu.cpp:17-24, blas/level3_impl.h:47-57);
 * > 0 = numerical outcome documented per function; <= -1000 = library error
 * (eamd_last_error() has the message).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the Eigen 3.3.90 snapshot).
 */
#ifndef EIGEN_AMD_H
#define EIGEN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EAMD_VERSION 10000

#define EAMD_OK 0
#define EAMD_ERR_HIP (-1000)
#define EAMD_ERR_ALLOC (-1001)
#define EAMD_ERR_UNSUPPORTED (-1002)
#define EAMD_ERR_NOT_CONVERGED (-1003)

typedef struct eamd_handle* eamd_handle_t;

/* ------------------------------------------------------------ lifetime -- */
int eamd_version(void);
/* Creates a handle on the current HIP device with its own stream.  Handles
 * are thread-confined (one per host thread), as Eigen objects are not
 * internally synchronised (Parallelizer.h:20,47-54). */
int eamd_create(eamd_handle_t* handle);
int eamd_destroy(eamd_handle_t handle);
/* Queue the handle's work on a caller-owned hipStream_t (passed as void*);
 * NULL selects the default stream.  Pending work on the previous stream is
 * synchronised first. */
int eamd_set_stream(eamd_handle_t handle, void* hip_stream);
void* eamd_get_stream(eamd_handle_t handle);
int eamd_synchronize(eamd_handle_t handle);
/* Thread-local message for the last failure. */
const char* eamd_last_error(void);
/* Frees the handle's cached device workspace. */
int eamd_release_workspace(eamd_handle_t handle);

/* Library instrumentation (no reference counterpart; the reference has
 * only offline timers, bench/BenchTimer.h): device time of the library's
 * own kernels per phase, from hipEvent pairs on the handle's stream, with
 * the algorithmic bytes and flops of those launches. */
#define EAMD_PHASE_PANEL 0     /* LU panel factorization */
#define EAMD_PHASE_LASWP 1     /* row interchanges */
#define EAMD_PHASE_TRSM 2      /* U12 = L11^{-1} A12 */
#define EAMD_PHASE_UPDATE 3    /* LU trailing update A22 -= A21 A12 */
#define EAMD_PHASE_PACK 4      /* FP16/BF16 operand packing of the mixed update */
#define EAMD_PHASE_TRSV 5      /* triangular solves */
#define EAMD_PHASE_RESIDUAL 6  /* FP64 residuals */
#define EAMD_PHASE_GEMM 7      /* eamd_dgemm */
#define EAMD_PHASE_POTF2 8     /* Cholesky diagonal blocks */
#define EAMD_PHASE_BATCHED 9   /* batched solves */
#define EAMD_PHASE_NEXT 10     /* look-ahead LU: the next panel's FP32 update */
int eamd_profile_enable(eamd_handle_t handle, int on);
int eamd_profile_read(eamd_handle_t handle, int phase, double* ms, double* bytes, double* flops,
                      int64_t* count);

/* --------------------------------------------------------------- level 3 -- */
/* C = alpha*op(A)*op(B) + beta*C ('N' or 'T').
 * Replaces general_matrix_matrix_product<...,ColMajor>::run
 * (Eigen/src/Core/products/GeneralMatrixMatrix.h:49-199, res += alpha*lhs*rhs)
 * and dgemm_ (blas/level3_impl.h:12-76). */
int eamd_dgemm(eamd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
               double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
               double beta, double* C, int64_t ldc);

/* B = alpha*op(A)^{-1}*B (side 'L') or alpha*B*op(A)^{-1} (side 'R');
 * uplo 'L'/'U', transa 'N'/'T', diag 'N'/'U'.
 * Replaces triangular_solve_matrix<...,OnTheLeft|OnTheRight,...>::run
 * (Eigen/src/Core/products/TriangularSolverMatrix.h:38-329; dispatch
 * SolveTriangular.h:81-104) and dtrsm_ (blas/level3_impl.h:78-179). */
int eamd_dtrsm(eamd_handle_t h, char side, char uplo, char transa, char diag, int64_t m,
               int64_t n, double alpha, const double* A, int64_t lda, double* B, int64_t ldb);

/* C(uplo) = alpha*op(A)*op(A)^T + beta*C(uplo), trans 'N' (A n x k) or 'T'.
 * Replaces SelfAdjointView::rankUpdate -> general_matrix_matrix_triangular_
 * product (SelfadjointProduct.h:82-129, GeneralMatrixMatrixTriangular.h:56-125)
 * and dsyrk_ (blas/level3_impl.h:357). */
int eamd_dsyrk(eamd_handle_t h, char uplo, char trans, int64_t n, int64_t k, double alpha,
               const double* A, int64_t lda, double beta, double* C, int64_t ldc);

/* ------------------------------------------------------------------- LU -- */
/* In-place LU with partial pivoting of an m x n matrix.  As in the
 * reference's blocked loop, a wide matrix (m < n) with min(m, n) > 16 has
 * only its first m columns factored; columns m..n-1 keep their values
 * (PartialPivLU.h:450-465).  transp[min(m,n)] receives 0-based row
 * transpositions (device int32), *nb_transp (host) their non-trivial count
 * and *first_zero_pivot (host) the first exactly-zero pivot or -1; the
 * factorization continues past it, as the reference does.
 * Replaces partial_lu_impl<double,ColMajor,int>::blocked_lu
 * (Eigen/src/LU/PartialPivLU.h:426-496), i.e. PartialPivLU::compute
 * (:514-536), and dgetrf_ (lapack/lu.cpp:14-41, which reports 1-based ipiv
 * and info = first_zero+1). */
int eamd_dgetrf(eamd_handle_t h, int64_t m, int64_t n, double* A, int64_t lda, int32_t* transp,
                int32_t* nb_transp, int64_t* first_zero_pivot);

/* Solves A X = B ('N') or A^T X = B ('T') with the factors of eamd_dgetrf.
 * Replaces PartialPivLU::_solve_impl / _solve_impl_transposed
 * (PartialPivLU.h:226-271) and dgetrs_ (lapack/lu.cpp:46-89). */
int eamd_dgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const double* LU,
                int64_t lda, const int32_t* transp, double* B, int64_t ldb);

/* PermutationMatrix indices from a transposition sequence (device arrays):
 * identity, then swap(perm[k], perm[t_k]) for k = n-1..0, i.e.
 * PartialPivLU::permutationP() (PermutationMatrix.h:82-88). */
int eamd_transpositions_to_permutation(eamd_handle_t h, int64_t n, const int32_t* transp,
                                       int32_t* perm);

/* ------------------------------------------------------------------ LLT -- */
/* In-place Cholesky; uplo 'L' (A = L L^T, lower triangle read/written) or
 * 'U' (A = U^T U).  *fail_col (host) = first column whose pivot is <= 0, or
 * -1 (LLT::info() == NumericalIssue when >= 0).
 * Replaces llt_inplace<double,Lower|Upper>::blocked (Eigen/src/Cholesky/
 * LLT.h:330-391), LLT::compute (:422-452) and dpotrf_ (lapack/cholesky.cpp:14-36,
 * info = fail_col+1). */
int eamd_dpotrf(eamd_handle_t h, char uplo, int64_t n, double* A, int64_t lda,
                int64_t* fail_col);

/* Solves A X = B with the factor of eamd_dpotrf.
 * Replaces LLT::_solve_impl (LLT.h:475-505) and dpotrs_ (lapack/cholesky.cpp:41-72). */
int eamd_dpotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const double* A,
                int64_t lda, double* B, int64_t ldb);

/* ----------------------------------------------- FP32 single matrices -- */
/* FP32 storage, FP64 arithmetic: the operands are widened on the device, the
 * FP64 routine above runs, and the results are rounded back to FP32.
 * Arguments, transpositions, first-zero-pivot and failing-column rules are
 * those of the FP64 entry points.  The pivots are those of the FP64
 * factorization of the FP32 values (Eigen's FP32 arithmetic can choose
 * differently between candidates within FP32 rounding of each other), and
 * the factors are that factorization rounded once.
 * eamd_sgetrf/eamd_sgetrs replace partial_lu_impl<float,ColMajor,int>::
 * blocked_lu as PartialPivLU_LAPACKE.h:75 binds it (sgetrf) and the FP32
 * PartialPivLU::_solve_impl (PartialPivLU.h:226-271); eamd_spotrf/
 * eamd_spotrs replace llt_inplace<float,Lower|Upper>::blocked as
 * LLT_LAPACKE.h:91 binds it (spotrf) and the FP32 LLT::_solve_impl
 * (LLT.h:475-505). */
int eamd_sgetrf(eamd_handle_t h, int64_t m, int64_t n, float* A, int64_t lda, int32_t* transp,
                int32_t* nb_transp, int64_t* first_zero_pivot);
int eamd_sgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const float* LU,
                int64_t lda, const int32_t* transp, float* B, int64_t ldb);
int eamd_spotrf(eamd_handle_t h, char uplo, int64_t n, float* A, int64_t lda,
                int64_t* fail_col);
int eamd_spotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const float* A,
                int64_t lda, float* B, int64_t ldb);

/* --------------------------------------------------- mixed precision IR -- */
/* No reference counterpart (Eigen has no iterative refinement): a new
 * capability behind PartialPivLU's solve() surface. */
#define EAMD_PREC_FP64 0      /* plain FP64 LU (the drop-in) */
#define EAMD_PREC_FP32 1      /* FP32 LU, FP32 trailing update */
#define EAMD_PREC_FP16 2      /* FP32 panel, FP16 MFMA trailing update */
#define EAMD_PREC_BF16 3      /* FP32 panel, BF16 MFMA trailing update */
#define EAMD_PREC_FP16X3 4    /* FP32 panel, FP16 hi/lo split update (3 MFMA products) */

typedef struct eamd_ir_opts {
  int factor_precision;   /* EAMD_PREC_* */
  int max_iter;           /* refinement steps (classic IR), default 30 */
  double tol;             /* target normwise backward error, default 1e-14 */
  int gmres;              /* 0: never, 1: fall back when IR stalls, 2: always */
  int gmres_restart;      /* Krylov subspace size, default 30 */
  int gmres_max_iter;     /* total GMRES iterations, default 300 */
  int nb;                 /* LU block width, 0 = automatic */
} eamd_ir_opts;

typedef struct eamd_ir_report {
  int ir_iterations;
  int gmres_iterations;
  int converged;          /* 1 when backward_error <= tol */
  int used_gmres;
  double backward_error;  /* ||b-Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf), max over rhs */
  double factor_ms;
  double refine_ms;
  int64_t first_zero_pivot;
} eamd_ir_report;

void eamd_ir_opts_default(eamd_ir_opts* opts);

/* Solves A X = B: low-precision LU of A, FP64 residuals on the device,
 * classic refinement and LU-preconditioned FP64 GMRES when it stalls.  A is
 * not modified.  Returns EAMD_ERR_NOT_CONVERGED (report filled) when the
 * tolerance is not met. */
int eamd_dgesv_ir(eamd_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double* X, int64_t ldx, const eamd_ir_opts* opts,
                  eamd_ir_report* report);

/* Normwise backward error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf
 * + ||b_j||_inf), residual in FP64 on the device; result on the host. */
int eamd_dbackward_error(eamd_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                         const double* X, int64_t ldx, const double* B, int64_t ldb,
                         double* berr);

/* ------------------------------------------------------ LLT rank update -- */
/* Replaces LLT<MatrixXd,UpLo>::rankUpdate(v, sigma) (LLT.h:459-472, through
 * internal::llt_rank_update_lower, LLT.h:231-299) for many factors in one
 * launch: factor i is n[i] x n[i], column-major, at L + offL[i] (the lower
 * triangle, or with upper != 0 the upper one holding L^T), v_i at v + offv[i],
 * sigma_i = sigma[i] (sigma may be NULL: 1).  On return L_i L_i^T = A_i +
 * sigma_i v_i v_i^T and info[i] = -1, or the column where a downdate (sigma
 * <= 0) stopped (the reference's NumericalIssue), or -2 for n[i] outside
 * 0..256.  All arrays are device arrays; asynchronous on the handle's stream. */
int eamd_dllt_rank_update_batched(eamd_handle_t h, int64_t count, const int64_t* offL,
                                  const int32_t* n, double* L, const int64_t* offv,
                                  const double* v, const double* sigma, int32_t upper,
                                  int32_t* info);

/* ----------------------------------------------------------------- LDLT -- */
/* Replaces LDLT<MatrixXd, Lower>::compute (LDLT.h:491-522, through
 * internal::ldlt_inplace<Lower>::unblocked, LDLT.h:290-395) for many small
 * symmetric matrices in one launch: matrix i is n[i] x n[i], column-major, at
 * A + offA[i]; its lower triangle is read and replaced by matrixLDLT()'s (L
 * strictly below the diagonal, D on it; the strict upper part is left as it
 * was).  transpositions + offT[i] receives transpositionsP().indices()
 * (n[i] entries, 0-based), sign[i] the reference's internal::SignMatrix
 * (0 PositiveSemiDef, 1 NegativeSemiDef, 2 ZeroSign, 3 Indefinite: isPositive()
 * is 0 or 2, isNegative() 1 or 2), info[i] 0 (Success), 1 (NumericalIssue)
 * or -2 for n[i] outside 0..max_n.  max_n (at most 64) bounds every n[i] and
 * sizes the launch's LDS.  Device arrays; asynchronous on the handle's
 * stream.  Returns -5 for max_n outside 0..64. */
int eamd_dldlt_batched(eamd_handle_t h, int64_t count, const int64_t* offA, const int32_t* n,
                       int32_t max_n, double* A, const int64_t* offT, int32_t* transpositions,
                       int32_t* sign, int32_t* info);

/* Replaces LDLT<MatrixXd, Lower>::rankUpdate(w, sigma) (LDLT.h:529-555,
 * through internal::ldlt_inplace<Lower>::update and updateInPlace,
 * LDLT.h:403-449) for many factors of eamd_dldlt_batched in one launch:
 * factor i (n[i] x n[i] at A + offA[i]: L strictly below the diagonal, D on
 * it) with its transpositions (transpositions + offT[i]) becomes the factor
 * of P^T L D L^T P + sigma[i] w w^T, w at w + offw[i] (n[i] values, read
 * only); sigma null means 1 for every factor.  The transpositions and the
 * sign are unchanged, as in the reference.  info[i] 0, or -2 for n[i]
 * outside 0..256.  Device arrays; asynchronous on the handle's stream. */
int eamd_dldlt_rank_update_batched(eamd_handle_t h, int64_t count, const int64_t* offA,
                                   const int32_t* n, double* A, const int64_t* offT,
                                   const int32_t* transpositions, const int64_t* offw,
                                   const double* w, const double* sigma, int32_t* info);

/* Replaces LDLT<MatrixXd, Lower>::solve(b) (LDLT.h:559-593) for each system
 * factored by eamd_dldlt_batched: b + offb[i] (n[i] values) is overwritten
 * with P^T L^-T D^+ L^-1 P b, D^+ zeroing the rows whose |d| is not above
 * 1 / DBL_MAX.  max_n as for the factorization (returns -5 outside 0..64);
 * systems with n[i] outside 0..max_n are left as they are. */
int eamd_dldlt_solve_batched(eamd_handle_t h, int64_t count, const int64_t* offA,
                             const int32_t* n, int32_t max_n, const double* A,
                             const int64_t* offT, const int32_t* transpositions,
                             const int64_t* offb, double* b);

/* -------------------------------------------------------------- batched -- */
/* No reference counterpart as one launch; per system it is
 * LLT<MatrixXf>(A_i).solve(b_i) (LLT.h:422-505).  Systems are packed
 * column-major n_i x n_i at A + offA[i], right-hand sides at b + offb[i]
 * (overwritten by x_i).  info[i] = failing column or -1.  All arrays are
 * device arrays.  Sizes up to 32. */
int eamd_spotrf_solve_batched(eamd_handle_t h, int64_t count, const int64_t* offA,
                              const int32_t* n, const float* A, const int64_t* offb, float* b,
                              int32_t* info);

/* No counterpart in the reference: eamd_spotrf_solve_batched for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and size rule (n[i] in 1..32); at most 1,048,576 systems per call (larger
 * counts return EAMD_ERR_UNSUPPORTED).  Its plan buffers are a static
 * device array of about 4 MiB per GPU, held until the process ends and
 * shared by every call on that GPU, so calls that overlap on one GPU race.
 * It synchronises the handle's stream once (the size histogram is read on
 * the host), uses the per-size widths tuned by an earlier planned call and
 * otherwise the narrowest, and calls no allocator. */
int eamd_spotrf_solve_batched_noalloc(eamd_handle_t h, int64_t count, const int64_t* offA,
                                      const int32_t* n, const float* A, const int64_t* offb,
                                      float* b, int32_t* info);

/* The same split into a plan (size bucketing of n[], one host
 * synchronisation, autotuned lanes-per-system per size) and an execution
 * that only queues kernels -- no host synchronisation, no memset -- so it can
 * be captured in a HIP graph and replayed for new A/b of the same layout. */
typedef struct eamd_batched_plan* eamd_batched_plan_t;
int eamd_batched_plan_create(eamd_handle_t h, int64_t count, const int32_t* n, int32_t max_n,
                             eamd_batched_plan_t* plan);
int eamd_batched_plan_destroy(eamd_batched_plan_t plan);
/* lanes of a wavefront per system (batched LLT plan) or threads per
 * workgroup (batched LU plan) chosen for size n (0: n not in the plan) */
int eamd_batched_plan_lanes(eamd_batched_plan_t plan, int32_t n);
int eamd_spotrf_solve_batched_planned(eamd_handle_t h, eamd_batched_plan_t plan,
                                      const int64_t* offA, const float* A, const int64_t* offb,
                                      float* b, int32_t* info);

/* Per system PartialPivLU<MatrixXd>(A_i).solve(B_i) (PartialPivLU.h:128,
 * 226-244, 426-496), computed as an FP32 LU on matrix cores with FP64
 * iterative refinement to a normwise backward error <= 1e-14 per right-hand
 * side; a system the FP32 factor cannot refine is refactored in FP64.  A_i
 * is n_i x n_i column-major at A + offA[i]; B_i and X_i are n_i x nrhs
 * column-major at offB[i] (1 <= nrhs <= 8).  info[i] = 0 refined from the
 * FP32 factor, 1 refined after the FP64 refactor, 2 not converged (e.g.
 * singular); iters[i] (nullable) = refinement steps of the final factor.
 * Sizes 1..256.  All arrays are device arrays. */
int eamd_dgesv_ir_batched(eamd_handle_t h, int64_t count, const int64_t* offA, const int32_t* n,
                          const double* A, const int64_t* offB, int64_t nrhs, const double* B,
                          double* X, int32_t* info, int32_t* iters);

/* The same split into a plan (size sort of n[], one host synchronisation,
 * autotuned workgroup size per 16-wide size bucket, working-copy layout) and
 * an execution that only queues kernels (capturable).  Destroy the plan with
 * eamd_batched_plan_destroy. */
int eamd_batched_lu_plan_create(eamd_handle_t h, int64_t count, const int32_t* n, int64_t nrhs,
                                eamd_batched_plan_t* plan);
int eamd_dgesv_ir_batched_planned(eamd_handle_t h, eamd_batched_plan_t plan,
                                  const int64_t* offA, const int32_t* n, const double* A,
                                  const int64_t* offB, const double* B, double* X,
                                  int32_t* info, int32_t* iters);

/* ------------------------------------------------- host-buffer variants -- */
/* Same operations on host (pageable) buffers: upload, compute, download.
 * These back the Eigen plug-point specialisations, whose arguments are host
 * pointers (PartialPivLU_LAPACKE.h:42-77, LLT_LAPACKE.h:42-93,
 * GeneralMatrixMatrix_BLAS.h:49-116, TriangularSolverMatrix_BLAS.h:41-157,
 * GeneralMatrixMatrixTriangular_BLAS.h:79-98 for eamd_host_dsyrk), and the
 * Fortran ABI of eigen_amd_lapack.h. */
int eamd_host_dgetrf(eamd_handle_t h, int64_t m, int64_t n, double* A, int64_t lda,
                     int32_t* transp, int32_t* nb_transp, int64_t* first_zero_pivot);
int eamd_host_dgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const double* LU,
                     int64_t lda, const int32_t* transp, double* B, int64_t ldb);
int eamd_host_dpotrf(eamd_handle_t h, char uplo, int64_t n, double* A, int64_t lda,
                     int64_t* fail_col);
int eamd_host_dpotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const double* A,
                     int64_t lda, double* B, int64_t ldb);
int eamd_host_sgetrf(eamd_handle_t h, int64_t m, int64_t n, float* A, int64_t lda,
                     int32_t* transp, int32_t* nb_transp, int64_t* first_zero_pivot);
int eamd_host_sgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const float* LU,
                     int64_t lda, const int32_t* transp, float* B, int64_t ldb);
int eamd_host_spotrf(eamd_handle_t h, char uplo, int64_t n, float* A, int64_t lda,
                     int64_t* fail_col);
int eamd_host_spotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const float* A,
                     int64_t lda, float* B, int64_t ldb);
int eamd_host_dgemm(eamd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                    double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                    double beta, double* C, int64_t ldc);
int eamd_host_dsyrk(eamd_handle_t h, char uplo, char trans, int64_t n, int64_t k, double alpha,
                    const double* A, int64_t lda, double beta, double* C, int64_t ldc);
int eamd_host_dtrsm(eamd_handle_t h, char side, char uplo, char transa, char diag, int64_t m,
                    int64_t n, double alpha, const double* A, int64_t lda, double* B,
                    int64_t ldb);
int eamd_host_dgesv_ir(eamd_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                       const double* B, int64_t ldb, double* X, int64_t ldx,
                       const eamd_ir_opts* opts, eamd_ir_report* report);

/* ------------------------------------------------- device memory helpers -- */
/* Thin wrappers so that C callers (and the adapter) need no HIP headers. */
int eamd_malloc(void** ptr, int64_t bytes);
int eamd_free(void* ptr);
int eamd_memcpy_htod(eamd_handle_t h, void* dst, const void* src, int64_t bytes);
int eamd_memcpy_dtoh(eamd_handle_t h, void* dst, const void* src, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* EIGEN_AMD_H */
