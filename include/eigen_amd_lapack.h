/* eigen_amd_lapack.h -- the reference's Fortran/C ABI for this path
 * (libeigen_lapack / libeigen_blas), exported by libeigen_amd_lapack.so over
 * the host-buffer entry points of eigen_amd.h.  Column-major host arrays,
 * every argument by pointer, 1-based ipiv, info as LAPACK defines it:
 * < 0 the index of an illegal argument, > 0 a numerical failure.  A device
 * failure sets info = -1000 and prints the library's message. */
#ifndef EIGEN_AMD_LAPACK_H
#define EIGEN_AMD_LAPACK_H

#ifdef __cplusplus
extern "C" {
#endif

/* LU with partial pivoting; info > 0: U(info,info) is exactly zero
 * (lapack/lu.cpp:14-41) */
int dgetrf_(int* m, int* n, double* a, int* lda, int* ipiv, int* info);
/* solve with the LU of dgetrf_, trans 'N' or 'T' (lapack/lu.cpp:46-89) */
int dgetrs_(char* trans, int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b, int* ldb,
            int* info);
/* Cholesky; info > 0: the leading minor of that order is not positive
 * (lapack/cholesky.cpp:14-36) */
int dpotrf_(char* uplo, int* n, double* a, int* lda, int* info);
/* solve with the factor of dpotrf_ (lapack/cholesky.cpp:41-72) */
int dpotrs_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb, int* info);
/* The same four for FP32 arrays (lapack/lu.cpp and lapack/cholesky.cpp are
 * compiled for float too); FP32 storage, FP64 arithmetic inside */
int sgetrf_(int* m, int* n, float* a, int* lda, int* ipiv, int* info);
int sgetrs_(char* trans, int* n, int* nrhs, float* a, int* lda, int* ipiv, float* b, int* ldb,
            int* info);
int spotrf_(char* uplo, int* n, float* a, int* lda, int* info);
int spotrs_(char* uplo, int* n, int* nrhs, float* a, int* lda, float* b, int* ldb, int* info);
/* C = alpha op(A) op(B) + beta C (blas/level3_impl.h:12-76) */
int dgemm_(const char* opa, const char* opb, const int* m, const int* n, const int* k,
           const double* alpha, const double* a, const int* lda, const double* b, const int* ldb,
           const double* beta, double* c, const int* ldc);
/* B = alpha op(A)^-1 B or alpha B op(A)^-1 (blas/level3_impl.h:78-179) */
int dtrsm_(const char* side, const char* uplo, const char* opa, const char* diag, const int* m,
           const int* n, const double* alpha, const double* a, const int* lda, double* b,
           const int* ldb);

/* C(uplo) = alpha op(A) op(A)^T + beta C(uplo), op 'N' (A n x k) or 'T'/'C'
 * (A k x n) (blas/level3_impl.h:357-391) */
int dsyrk_(const char* uplo, const char* op, const int* n, const int* k, const double* alpha,
           const double* a, const int* lda, const double* beta, double* c, const int* ldc);

#ifdef __cplusplus
}
#endif

#endif /* EIGEN_AMD_LAPACK_H */
