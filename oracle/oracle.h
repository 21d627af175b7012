/*
 * oracle.h -- CPU restatement of Eigen's dense factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the eigen_amd package,
 * its HIP kernels or its C-ABI) may include, link or call this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker / CPU baseline.
 *
 * Every function follows a named reference routine (file:line relative to the
 * Eigen 3.3.90 snapshot).  Storage is column-major with a leading dimension,
 * indices are 0-based unless stated, pivots are int32 like Eigen's default
 * PermutationMatrix StorageIndex.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   orc_lu_*        golden vectors from compiled Eigen (oracle/_ref) for the
 *                   pivot sequence, L/U and det sign; PA=LU reconstruction;
 *                   closed-form cases (upper-triangular input -> P=I, exact
 *                   ties -> lowest row, zero column -> first_zero_pivot).
 *   orc_llt_*       golden L from compiled Eigen; uniqueness of the Cholesky
 *                   factor; diagonal matrices (closed form sqrt); the failing
 *                   column of a diagonal matrix with one negative entry.
 *   orc_gemm_d      identity / permutation / rank-1 closed forms, brute force.
 *   orc_trsm_d      round trip T*X == B, identity and diagonal cases.
 *   orc_syrk_d      equals the lower triangle of a brute-force A*A^T.
 *   batched         per-system equality with the single-system routines and
 *                   with compiled Eigen on a sample.
 */
#ifndef EAMD_ORACLE_H
#define EAMD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PartialPivLU (Eigen/src/LU/PartialPivLU.h) ---- */
/* unblocked_lu, PartialPivLU.h:365-409.  margin (nullable) receives, per
 * column, (|best|-|runner-up|)/|best| of the pivot candidates. */
long orc_lu_unblocked_d(long rows, long cols, double* a, long lda, int* transp,
                        int* nb_transp, double* margin);
/* blocked_lu, PartialPivLU.h:426-496 (recursive panel with maxBlockSize 16). */
long orc_lu_blocked_d(long rows, long cols, double* a, long lda, int* transp,
                      int* nb_transp, long max_block, double* margin);
long orc_lu_blocked_f(long rows, long cols, float* a, long lda, int* transp,
                      int* nb_transp, long max_block);
/* PermutationMatrix = Transpositions, PermutationMatrix.h:82-88 */
void orc_transpositions_to_perm(long n, const int* transp, int* perm);
/* PartialPivLU::_solve_impl, PartialPivLU.h:226-244: dst = P*b, L\, U\ */
void orc_lu_solve_d(long n, long nrhs, const double* lu, long lda,
                    const int* transp, double* b, long ldb);
void orc_lu_solve_f(long n, long nrhs, const float* lu, long lda,
                    const int* transp, float* b, long ldb);

/* ---- LLT (Eigen/src/Cholesky/LLT.h) ---- */
/* llt_inplace<.,Lower>::unblocked, LLT.h:304-328; returns failing col or -1 */
long orc_llt_unblocked_d(long n, double* a, long lda);
long orc_llt_unblocked_f(long n, float* a, long lda);
/* llt_inplace<.,Lower>::blocked, LLT.h:330-360 */
long orc_llt_blocked_d(long n, double* a, long lda);
long orc_llt_blocked_f(long n, float* a, long lda);
/* LLT::_solve_impl / solveInPlace, LLT.h:475-505 */
void orc_llt_solve_d(long n, long nrhs, const double* l, long ldl, double* b, long ldb);
void orc_llt_solve_f(long n, long nrhs, const float* l, long ldl, float* b, long ldb);
/* LLT rank-one update of a lower factor (LLT.h:231-299); -1 or failing column */
long orc_llt_rank_update_d(long n, double* L, long ld, const double* v, double sigma);
long orc_ldlt_d(long n, double* A, long ld, int* trans, int* sign);
void orc_ldlt_solve_d(long n, long nrhs, const double* A, long ld, const int* trans, double* B,
                      long ldb);

/* ---- level-3 kernels ---- */
/* C = alpha*op(A)*op(B) + beta*C  (general_matrix_matrix_product,
 * GeneralMatrixMatrix.h:49-199 computes res += alpha*lhs*rhs) */
void orc_gemm_d(char transa, char transb, long m, long n, long k, double alpha,
                const double* A, long lda, const double* B, long ldb,
                double beta, double* C, long ldc);
/* triangular_solve_matrix, TriangularSolverMatrix.h:38-329 (BLAS dtrsm
 * argument convention, blas/level3_impl.h:78-179) */
void orc_trsm_d(char side, char uplo, char transa, char diag, long m, long n,
                double alpha, const double* A, long lda, double* B, long ldb);
/* rankUpdate -> general_matrix_matrix_triangular_product,
 * SelfadjointProduct.h:82-129; C(uplo) = alpha*op(A)*op(A)^T + beta*C */
void orc_syrk_d(char uplo, char trans, long n, long k, double alpha,
                const double* A, long lda, double beta, double* C, long ldc);

/* ---- diagnostics ---- */
/* max over rhs columns of ||A x - b||_inf / (||A||_inf ||x||_inf + ||b||_inf),
 * residual accumulated in long double. */
double orc_backward_error_d(long n, long nrhs, const double* A, long lda,
                            const double* X, long ldx, const double* B, long ldb);

/* ---- batched (one Eigen decomposition per system) ---- */
/* LLT<MatrixXf>(A_i).solve(b_i) per system; packed col-major n_i x n_i at
 * A+offA[i], b at b+offb[i] overwritten with x; info[i] = failing col or -1 */
void orc_batched_llt_solve_f(long count, const int64_t* offA, const int32_t* n,
                             const float* A, const int64_t* offb, float* b,
                             int32_t* info);
/* PartialPivLU<MatrixXd>(A_i).solve(B_i), B_i n_i x nrhs (ld n_i) */
void orc_batched_lu_solve_d(long count, const int64_t* offA, const int32_t* n,
                            const double* A, const int64_t* offb, long nrhs,
                            double* B, int32_t* info);

#ifdef __cplusplus
}
#endif
#endif
