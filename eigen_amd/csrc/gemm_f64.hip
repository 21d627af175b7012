// gemm_f64.hip -- FP64 GEMM / SYRK on CDNA4 matrix cores.
//
// C = alpha * op(A) * op(B) + beta * C, column-major with leading dimensions.
// This is the MI355X replacement for general_matrix_matrix_product::run
// (Eigen/src/Core/products/GeneralMatrixMatrix.h:49-199) and, with the
// lower-triangle filter, for general_matrix_matrix_triangular_product
// (GeneralMatrixMatrixTriangular.h:56-125).
//
// Design: a 128x128 C tile per 256-thread workgroup (4 waves in 2x2, each
// wave 64x64 = 4x4 tiles of v_mfma_f64_16x16x4f64), K staged through LDS in
// slabs of 16 with double buffering.  The MFMA is fed the transposed product
// (op(B)^T op(A)^T) so that each lane's four accumulator rows map to four
// *columns* of C and the sixteen lanes of a row group cover sixteen
// consecutive rows: the epilogue writes 128-byte runs of a column.
// Workgroups are remapped so that consecutive tiles share an XCD (the
// dispatcher deals workgroups round-robin over the 8 XCDs, each with its own
// L2) and swizzled in groups of 8 tile-rows for L2 reuse of B.
#include <algorithm>
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int BM = 128, BN = 128, BK = 16, PAD = 2;
constexpr int LDS_LD = BM + PAD;  // row length of the LDS slabs (doubles)

__device__ __forceinline__ void tile_coords(int bid, int nblocks, int tiles_m, int tiles_n,
                                            int& tm, int& tn) {
  // XCD-aware remap: workgroup bid runs on XCD bid % 8.
  int xcd = bid & 7, loc = bid >> 3;
  int per = nblocks >> 3, rem = nblocks & 7;
  int logical = (xcd < rem) ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
  const int G = 8;
  int group = G * tiles_n;
  int g = logical / group;
  int first_m = g * G;
  int gm = min(tiles_m - first_m, G);
  int r = logical - g * group;
  tm = first_m + r % gm;
  tn = r / gm;
}

// Tile `idx` of the lower trapezoid {(tr, tc): tc <= tr < tr_n, tc < tc_n} of
// a tr_n x tc_n tile grid, row by row: the triangle tr < T = min(tr_n, tc_n)
// first, then full rows of tc_n tiles.
__device__ __forceinline__ void tri_tile(int idx, int tr_n, int tc_n, int& tr, int& tc) {
  const int T = min(tr_n, tc_n), tri = T * (T + 1) / 2;
  if (idx < tri) {
    int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    if ((r + 1) * (r + 2) / 2 <= idx) ++r;
    if (r * (r + 1) / 2 > idx) --r;
    tr = r;
    tc = idx - r * (r + 1) / 2;
  } else {
    const int q = idx - tri;
    tr = T + q / tc_n;
    tc = q - (q / tc_n) * tc_n;
  }
}

// Number of tiles of that trapezoid.
inline int tri_tiles(int tr_n, int tc_n) {
  const int T = std::min(tr_n, tc_n);
  return T * (T + 1) / 2 + (tr_n > T ? (tr_n - T) * tc_n : 0);
}

template <bool TA, bool TB>
__device__ __forceinline__ void load_slab(const double* __restrict__ A, int64_t lda,
                                          const double* __restrict__ B, int64_t ldb, int m0,
                                          int n0, int k0, int m, int n, int k, double ra[8],
                                          double rb[8]) {
  const int t = threadIdx.x;
  // ---- op(A): BM x BK
  if (!TA) {
    int kk = t >> 4, mm = (t & 15) * 8;
    int gk = k0 + kk;
    const double* p = A + (int64_t)gk * lda + (m0 + mm);
#pragma unroll
    for (int q = 0; q < 8; ++q) ra[q] = (gk < k && m0 + mm + q < m) ? p[q] : 0.0;
  } else {
    int mm = t >> 1, kk = (t & 1) * 8;
    int gm = m0 + mm;
    const double* p = A + (int64_t)gm * lda + (k0 + kk);
#pragma unroll
    for (int q = 0; q < 8; ++q) ra[q] = (gm < m && k0 + kk + q < k) ? p[q] : 0.0;
  }
  // ---- op(B): BK x BN
  if (!TB) {
    int nn = t >> 1, kk = (t & 1) * 8;
    int gn = n0 + nn;
    const double* p = B + (int64_t)gn * ldb + (k0 + kk);
#pragma unroll
    for (int q = 0; q < 8; ++q) rb[q] = (gn < n && k0 + kk + q < k) ? p[q] : 0.0;
  } else {
    int kk = t >> 4, nn = (t & 15) * 8;
    int gk = k0 + kk;
    const double* p = B + (int64_t)gk * ldb + (n0 + nn);
#pragma unroll
    for (int q = 0; q < 8; ++q) rb[q] = (gk < k && n0 + nn + q < n) ? p[q] : 0.0;
  }
}

template <bool TA, bool TB>
__device__ __forceinline__ void store_slab(double* As, double* Bs, const double ra[8],
                                           const double rb[8]) {
  const int t = threadIdx.x;
  if (!TA) {
    int kk = t >> 4, mm = (t & 15) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) As[kk * LDS_LD + mm + q] = ra[q];
  } else {
    int mm = t >> 1, kk = (t & 1) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) As[(kk + q) * LDS_LD + mm] = ra[q];
  }
  if (!TB) {
    int nn = t >> 1, kk = (t & 1) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) Bs[(kk + q) * LDS_LD + nn] = rb[q];
  } else {
    int kk = t >> 4, nn = (t & 15) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) Bs[kk * LDS_LD + nn + q] = rb[q];
  }
}

// The products of one BK slab, staged in as / bs, into the 4 x 4 tiles.
__device__ __forceinline__ void slab_products(const double* as, const double* bs, int wm, int wn,
                                              int l16, int lk, v4d (&acc)[4][4]) {
#pragma unroll
  for (int kk = 0; kk < BK / 4; ++kk) {
    const int krow = (kk * 4 + lk) * LDS_LD;
    double af[4], bf[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) af[i] = as[krow + wm + i * 16 + l16];
#pragma unroll
    for (int j = 0; j < 4; ++j) bf[j] = bs[krow + wn + j * 16 + l16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
  }
}

// The K loop of C += alpha op(A) op(B) (alpha = +-1) with the C tile read in
// eight parts of eight words per lane, one part beside each of the first
// eight slabs' operand loads and added into the accumulators one slab
// later.  Loading the whole tile before the first slab made every
// workgroup of a launch wait for 128 KB of C at once: with a short inner
// dimension (the rank-128 and rank-256 trailing updates of the blocked
// factorizations) that burst was a large share of the tile's time (LLT
// n = 8192: 13.35 -> 12.18 ms; FP64 LU n = 16384: 195 -> 178 ms).  Part p
// holds tiles (i, j) = (p / 2, 2 (p % 2) + {0, 1}).
template <bool TA, bool TB>
__device__ __forceinline__ void dgemm_body_cin(int m, int n, int k, double alpha,
                                                const double* __restrict__ A, int64_t lda,
                                                const double* __restrict__ B, int64_t ldb,
                                                const double* __restrict__ C, int64_t ldc, int m0,
                                                int n0, double (*As)[BK * LDS_LD],
                                                double (*Bs)[BK * LDS_LD], v4d (&acc)[4][4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int l16 = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  double cb[8];
  auto cload = [&](auto P) {
    constexpr int p = decltype(P)::value, i = p >> 1, j0 = (p & 1) * 2;
    const int gi = m0 + wm + i * 16 + l16;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn + (j0 + jj) * 16 + lk + 4 * r;
        cb[jj * 4 + r] = (gi < m && gj < n) ? C[(int64_t)gj * ldc + gi] : 0.0;
      }
  };
  auto cadd = [&](auto P) {
    constexpr int p = decltype(P)::value, i = p >> 1, j0 = (p & 1) * 2;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j0 + jj][r] += alpha * cb[jj * 4 + r];
  };
  const int ktiles = (k + BK - 1) / BK;
  double ra[8], rb[8];
  if (ktiles > 0) {
    load_slab<TA, TB>(A, lda, B, ldb, m0, n0, 0, m, n, k, ra, rb);
    store_slab<TA, TB>(As[0], Bs[0], ra, rb);
  }
  __syncthreads();
  // slab p with C part p (a slab that does not exist only moves C)
  auto step = [&](auto P) {
    constexpr int p = decltype(P)::value;
    if (p < ktiles) {
      if (p + 1 < ktiles)
        load_slab<TA, TB>(A, lda, B, ldb, m0, n0, (p + 1) * BK, m, n, k, ra, rb);
      if constexpr (p > 0) cadd(std::integral_constant<int, p - 1>{});
      cload(P);
      slab_products(As[p & 1], Bs[p & 1], wm, wn, l16, lk, acc);
      if (p + 1 < ktiles) store_slab<TA, TB>(As[(p + 1) & 1], Bs[(p + 1) & 1], ra, rb);
      __syncthreads();
    } else {
      if constexpr (p > 0) cadd(std::integral_constant<int, p - 1>{});
      cload(P);
    }
  };
  step(std::integral_constant<int, 0>{});
  step(std::integral_constant<int, 1>{});
  step(std::integral_constant<int, 2>{});
  step(std::integral_constant<int, 3>{});
  step(std::integral_constant<int, 4>{});
  step(std::integral_constant<int, 5>{});
  step(std::integral_constant<int, 6>{});
  step(std::integral_constant<int, 7>{});
  cadd(std::integral_constant<int, 7>{});
  // the rest of K: the plain loop (written out, not through slab_products,
  // whose per-slab LDS address arithmetic cost the 8192^3 product 5 %)
  for (int kt = 8; kt < ktiles; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < ktiles) load_slab<TA, TB>(A, lda, B, ldb, m0, n0, (kt + 1) * BK, m, n, k, ra, rb);
    const double* as = As[buf];
    const double* bs = Bs[buf];
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int krow = (kk * 4 + lk) * LDS_LD;
      double af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = as[krow + wm + i * 16 + l16];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = bs[krow + wn + j * 16 + l16];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < ktiles) store_slab<TA, TB>(As[buf ^ 1], Bs[buf ^ 1], ra, rb);
    __syncthreads();
  }
}

// MODE 0: full C.  MODE 1: lower triangle only (i >= j), tiles strictly above
// the diagonal are skipped.  MODE 2: upper triangle only (i <= j).
// CIN: beta == 1 and alpha == +-1 (the trailing updates of the blocked
// factorizations): C is read during the first K slabs and added into the
// accumulators (dgemm_body_cin); the epilogue then only stores alpha * acc
// (alpha^2 = 1, both scalings exact).  Otherwise C is read in the epilogue.
// The two K loops are separate instantiations, so neither constrains the
// other's registers or code layout.
template <bool TA, bool TB, int MODE, bool CIN>
__global__ __launch_bounds__(256, 2) void dgemm_kernel(int m, int n, int k, double alpha,
                                                    const double* __restrict__ A, int64_t lda,
                                                    const double* __restrict__ B, int64_t ldb,
                                                    double beta, double* __restrict__ C,
                                                    int64_t ldc, const int* __restrict__ abort_flag) {
  __shared__ double As[2][BK * LDS_LD];
  if (abort_flag && *abort_flag >= 0) return;
  __shared__ double Bs[2][BK * LDS_LD];

  const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
  int tm, tn;
  if (MODE == 0) {
    tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  } else {
    // the grid holds only the tiles that meet the triangle; each XCD takes
    // a contiguous run of them (equal shares of the work)
    const int bid = blockIdx.x, nblocks = gridDim.x;
    const int xcd = bid & 7, loc = bid >> 3, per = nblocks >> 3, rem = nblocks & 7;
    const int logical =
        (xcd < rem) ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
    if (MODE == 1) tri_tile(logical, tiles_m, tiles_n, tm, tn);
    else tri_tile(logical, tiles_n, tiles_m, tn, tm);
  }
  const int m0 = tm * BM, n0 = tn * BN;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int l16 = lane & 15, lk = lane >> 4;

  v4d acc[4][4];
  if constexpr (CIN) {
    dgemm_body_cin<TA, TB>(m, n, k, alpha, A, lda, B, ldb, C, ldc, m0, n0, As, Bs, acc);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    const int ktiles = (k + BK - 1) / BK;
    double ra[8], rb[8];
    if (ktiles > 0) {
      load_slab<TA, TB>(A, lda, B, ldb, m0, n0, 0, m, n, k, ra, rb);
      store_slab<TA, TB>(As[0], Bs[0], ra, rb);
    }
    __syncthreads();
    for (int kt = 0; kt < ktiles; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < ktiles)
        load_slab<TA, TB>(A, lda, B, ldb, m0, n0, (kt + 1) * BK, m, n, k, ra, rb);
      const double* as = As[buf];
      const double* bs = Bs[buf];
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int krow = (kk * 4 + lk) * LDS_LD;
        double af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) af[i] = as[krow + wm + i * 16 + l16];
#pragma unroll
        for (int j = 0; j < 4; ++j) bf[j] = bs[krow + wn + j * 16 + l16];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
      }
      if (kt + 1 < ktiles) store_slab<TA, TB>(As[buf ^ 1], Bs[buf ^ 1], ra, rb);
      __syncthreads();
    }
  }

  // epilogue: v_mfma_f64_16x16x4f64 leaves result row lk + 4r of the 16x16
  // tile in register r of lane (lk, l16): acc[i][j][r] = C(m0+wm+16i+l16,
  // n0+wn+16j+lk+4r)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm + i * 16 + l16;
    if (gi >= m) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn + j * 16 + lk + 4 * r;
        if (gj >= n) continue;
        if (MODE == 1 && gi < gj) continue;
        if (MODE == 2 && gi > gj) continue;
        double* c = C + (int64_t)gj * ldc + gi;
        double v = alpha * acc[i][j][r];
        if (!CIN && beta != 0.0) v += beta * *c;
        *c = v;
      }
  }
}

template <bool TA, bool TB, int MODE>
void launch_t(hipStream_t s, int m, int n, int k, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
              const int* abort_flag) {
  const int tm = ceil_div(m, BM), tn = ceil_div(n, BN);
  const int tiles = MODE == 0 ? tm * tn : (MODE == 1 ? tri_tiles(tm, tn) : tri_tiles(tn, tm));
  if (beta == 1.0 && (alpha == 1.0 || alpha == -1.0))
    hipLaunchKernelGGL((dgemm_kernel<TA, TB, MODE, true>), dim3(tiles), dim3(256), 0, s, m, n, k,
                       alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  else
    hipLaunchKernelGGL((dgemm_kernel<TA, TB, MODE, false>), dim3(tiles), dim3(256), 0, s, m, n,
                       k, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
}

// C = beta*C for k == 0 or alpha == 0 (BLAS semantics: beta == 0 clears C,
// NaNs included)
__global__ void dscale_kernel(int m, int n, double beta, double* C, int64_t ldc, int tri) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)m * n;
  if (idx >= total) return;
  int i = (int)(idx % m), j = (int)(idx / m);
  if ((tri == 1 && i < j) || (tri == 2 && i > j)) return;
  double* c = C + (int64_t)j * ldc + i;
  *c = beta == 0.0 ? 0.0 : beta * *c;
}

}  // namespace

int dgemm_device(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc, int tri, const int* abort_flag) {
  if (m <= 0 || n <= 0) return kOk;
  if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) {
    set_error("dgemm: dimensions beyond int32 are not supported");
    return kErrUnsupported;
  }
  if (k == 0 || alpha == 0.0) {
    if (beta == 1.0) return kOk;
    int64_t total = m * n;
    hipLaunchKernelGGL(dscale_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (int)m, (int)n, beta, C, ldc, tri);
    EAMD_HIP(hipGetLastError());
    return kOk;
  }
  int M = (int)m, N = (int)n, K = (int)k;
#define EAMD_GEMM_DISPATCH(MODE)                                                              \
  if (!ta && !tb) launch_t<false, false, MODE>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag); \
  else if (!ta && tb) launch_t<false, true, MODE>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag); \
  else if (ta && !tb) launch_t<true, false, MODE>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag); \
  else launch_t<true, true, MODE>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, abort_flag);
  if (tri == 1) {
    EAMD_GEMM_DISPATCH(1)
  } else if (tri == 2) {
    EAMD_GEMM_DISPATCH(2)
  } else {
    EAMD_GEMM_DISPATCH(0)
  }
#undef EAMD_GEMM_DISPATCH
  EAMD_HIP(hipGetLastError());
  return kOk;
}

}  // namespace eamd
