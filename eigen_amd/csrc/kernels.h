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
  int bucket_start[258] = {0};  // host: first position of each size in `order`
  int bucket_count[258] = {0};  // host: systems of each size
  int lanes[258] = {0};         // host: threads per system chosen per size (batched LU)
  unsigned char llt_path[33] = {0};  // host: code path chosen per size (batched LLT)
  // batched LU only (kind 1)
  int kind = 0;             // 0 batched LLT, 1 batched LU
  int nrhs = 0;
  int64_t total_elems = 0;  // sum of n_i^2: the working copies, laid out by size
  int64_t* wbase = nullptr; // device [257]: first working-copy element of each size
  int* bstart = nullptr;    // device [257]: bucket_start on the device
  double lu_flops = 0.0;    // algorithmic flops of one execution
  double lu_bytes = 0.0;    // algorithmic HBM bytes of one execution
};

namespace eamd {

// integer environment setting, dflt when unset or empty
int env_int(const char* name, int dflt);

// LLT rank-one update of packed factors (llt.hip)
int llt_rank_update_batched(eamd_handle* h, int64_t count, const int64_t* offL, const int32_t* n,
                            double* L, const int64_t* offv, const double* v, const double* sigma,
                            int upper, int32_t* info);
int ldlt_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n, int max_n,
                 double* A, const int64_t* offT, int32_t* T, int32_t* sign, int32_t* info);
int ldlt_solve_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                       int max_n, const double* A, const int64_t* offT, const int32_t* T,
                       const int64_t* offb, double* b);
int ldlt_rank_update_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                             double* A, const int64_t* offT, const int32_t* T,
                             const int64_t* offw, const double* w, const double* sigma,
                             int32_t* info);

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

// FP64 LU driver on a row-major working matrix (m x n, ld ldw), min(m, n)
// pivots (0-based transpositions, absolute) into piv_dev.  first_zero on the
// host.
int dgetrf_rowmajor(eamd_handle* h, double* W, int64_t ldw, int m, int n, int* piv_dev,
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

// ---- batched plans (batched.hip) ----
// Size sort of n[] (sizes 1..max_n <= 256) into a plan; the solvers then
// fill in their per-size choices.
// The handle's second stream (look-ahead LU) and its events (capi.hip).
int ensure_side_stream(eamd_handle* h);

// ---- mixed-precision trailing update (mixed.hip) ----
// A packed operand: FP16/BF16 halves (hi, and lo for FP16x3) of K values per
// row or column, padded to 128-row blocks, with one power-of-two scale per
// block.
struct MixedOperand {
  unsigned short* hi = nullptr;
  unsigned short* lo = nullptr;
  float* scale = nullptr;
};
size_t mixed_operand_bytes(int len, int k, int precision);
int mixed_pack(eamd_handle* h, bool rows, const float* S, int64_t ld, int len, int k,
               int precision, void* buf, MixedOperand* out);
// FP32 update A22 -= A21 A12 (row-major, ld), timed under `phase`.
int f32_update_rowmajor(eamd_handle* h, const float* A12, const float* A21, float* A22,
                        int64_t ld, int mrows, int ncols, int k, int phase);
// max_wgs > 0 caps the grid (rounded down to a multiple of 8; the
// workgroups loop over the tiles).
int mixed_update_packed(eamd_handle* h, const MixedOperand& a, int M, const MixedOperand& b, int N,
                        int k, int precision, float* A22, int64_t ld, int max_wgs = 0);

int batched_plan_sort(eamd_handle* h, int64_t count, const int32_t* n, int max_n,
                      eamd_batched_plan** out);
void batched_plan_destroy(eamd_batched_plan* p);

}  // namespace eamd
