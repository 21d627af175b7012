// kernels.h -- internal launchers shared between the translation units of
// the eigen_amd library.  All pointers are device pointers; all launches go
// to the given stream and return a status code (0 ok).
#pragma once
#include "common.h"

// A size-bucketing plan of a batched solve (opaque to C callers).  Buckets
// cover sizes 1..256; the batched LLT uses 1..32.
struct eamd_batched_plan {
  int64_t count = 0;
  int max_n = 0;
  int* order = nullptr;     // device: system indices sorted by size (descending)
  int bucket_start[258];    // host: first position of each size in `order`
  int bucket_count[258];    // host: systems of each size
  int lanes[258];           // host: lanes (or threads) per system chosen per size
};

namespace eamd {

// ---- level 3 (gemm_f64.hip) ----
// C = alpha*op(A)*op(B) + beta*C, column-major.  tri: 0 full C, 1 only
// i >= j (lower SYRK), 2 only i <= j (upper).  abort_flag (nullable, device):
// the kernel returns at once when *abort_flag >= 0.
int dgemm_device(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc, int tri, const int* abort_flag = nullptr);

// ---- LU pieces (lu.hip) ----
// Panel factorization with partial pivoting of a row-major m x nb panel
// (element (i,j) at P[i*ldp + j]).  piv[j] receives the panel-relative row
// exchanged with row j.  Returns the first exactly-zero pivot (panel
// relative) through *zero_col_dev (device int, -1 if none, min-reduced).
template <typename T>
int panel_lu_rowmajor(eamd_handle* h, T* P, int64_t ldp, int m, int nb, int* piv,
                      int* zero_col_dev);
// Apply transpositions piv[0..np) (relative to row r0) to columns
// [c0, c1) of a row-major matrix.
template <typename T>
int laswp_rowmajor(hipStream_t s, T* W, int64_t ldw, int r0, const int* piv, int np,
                   int c0, int c1);
// X <- L^{-1} X with L unit lower (nb x nb, row-major), X nb x ncols
// row-major.
template <typename T>
int trsm_llu_rowmajor(hipStream_t s, const T* L, int64_t ldl, int nb, T* X, int64_t ldx,
                      int ncols);

// row-major <-> column-major copies with optional precision change
template <typename Ti, typename To>
int transpose_copy(hipStream_t s, const Ti* A, int64_t lda, int rows, int cols, To* B,
                   int64_t ldb);

// FP64 LU driver on a row-major working matrix (n x n, ld ldw), pivots
// (0-based transpositions, absolute) into piv_dev.  first_zero on the host.
int dgetrf_rowmajor(eamd_handle* h, double* W, int64_t ldw, int n, int* piv_dev,
                    int64_t* first_zero, int nb);
int sgetrf_rowmajor(eamd_handle* h, float* W, int64_t ldw, int n, int* piv_dev,
                    int64_t* first_zero, int nb, int update_precision);

// ---- triangular solves (trsv.hip) ----
// op(T) X = X in place; T n x n in row- or column-major layout, X
// column-major n x nrhs.
template <typename T>
int trsv_device(eamd_handle* h, int n, const T* A, int64_t lda, bool rowmajor, bool lower,
                bool unit, T* X, int64_t ldx, int nrhs);
int trsv_check_error(eamd_handle* h);
// X <- P X (forward) or P^T X with P given as a transposition sequence
template <typename T>
int apply_transpositions(eamd_handle* h, int n, const int* t, bool forward, T* X, int64_t ldx,
                         int nrhs);

// ---- Cholesky (llt.hip) ----
// Lower-triangular in-place factorization of a column-major n x n matrix;
// *fail_col = failing column or -1.
int dpotrf_lower(eamd_handle* h, int n, double* A, int64_t lda, int64_t* fail_col, int nb);

}  // namespace eamd
