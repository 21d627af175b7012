// lu.hip -- blocked LU with partial pivoting on a row-major working matrix.
//
// Replaces partial_lu_impl<Scalar,ColMajor,int>::blocked_lu
// (Eigen/src/LU/PartialPivLU.h:426-496).  The reference factors a block
// column (panel) by recursing on it with a small block size (:470-471), then
// applies the panel's interchanges left and right (:477-488), solves
// A12 <- L11^{-1} A12 (:490) and updates A22 -= A21 A12 (:492).  Here:
//
//   * the caller's column-major matrix is transposed into a row-major working
//     copy so that row interchanges move contiguous memory;
//   * a panel of width nb (<= 256) is factored as leaves of 32 columns
//     (leaf_lu_kernel, leaf.hip) with, between leaves, the leaf's
//     interchanges on the other panel columns, the leaf's U rows
//     (trsm_llu_kernel) and the update of the remaining panel columns;
//   * after the panel: interchanges on the columns outside it
//     (laswp_perm_kernel, from the panel's composed row permutation), the
//     triangular solve for U12 and the trailing update (FP64 MFMA GEMM, or
//     the mixed-precision update for the FP32 working copy).
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "../../include/eigen_amd.h"

#include <cmath>
#include <vector>

namespace eamd {

namespace {

// Apply a row permutation to columns [0,c_split) U [c_split+skip, ...) of the
// rows r0 + pos[*]: row pos[s] receives the content of row pos[src[s]].
// Each workgroup stages CW columns of every touched row in LDS, then writes
// them back permuted.
template <typename T, int CW>
__global__ __launch_bounds__(256) void laswp_perm_kernel(T* __restrict__ W, int64_t ldw, int r0,
                                                         const int* __restrict__ perm_n,
                                                         const int* __restrict__ pos,
                                                         const int* __restrict__ src,
                                                         int c_split, int skip, int ncols_eff) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* buf = reinterpret_cast<T*>(smem_raw);
  __shared__ int s_pos[512], s_src[512];
  const int nS = *perm_n;
  for (int e = threadIdx.x; e < nS; e += 256) {
    s_pos[e] = pos[e];
    s_src[e] = src[e];
  }
  __syncthreads();
  const int cbase = blockIdx.x * CW;
  const int tc = threadIdx.x % CW, tr = threadIdx.x / CW;
  constexpr int RS = 256 / CW, U = 8;
  const int ce = cbase + tc;
  const bool ok = ce < ncols_eff;
  const int c = ce < c_split ? ce : ce + skip;
  // U rows in flight per thread before their values are staged
  for (int s0 = tr; s0 < nS; s0 += RS * U) {
    T v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int s2 = s0 + RS * j;
      v[j] = (ok && s2 < nS) ? W[(int64_t)(r0 + s_pos[s2]) * ldw + c] : T(0);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int s2 = s0 + RS * j;
      if (s2 < nS) buf[s2 * CW + tc] = v[j];
    }
  }
  __syncthreads();
  for (int s2 = tr; s2 < nS; s2 += RS)
    if (ok) W[(int64_t)(r0 + s_pos[s2]) * ldw + c] = buf[s_src[s2] * CW + tc];
}

// X <- L^{-1} X, L unit lower nb x nb (row-major, ld ldl), X nb x ncols
// (row-major).  One workgroup per CB columns, the X block resident in LDS.
// Per 32-row diagonal block: one thread per column solves the block from
// registers (compile-time indices, L broadcast from LDS, no barrier inside),
// then all threads update the rows below with 32-long dot products.
template <typename T, int CB>
__global__ __launch_bounds__(256) void trsm_llu_kernel(const T* __restrict__ L, int64_t ldl,
                                                       int nb, T* __restrict__ X, int64_t ldx,
                                                       int ncols) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int XP = CB + 1;
  T* Xs = reinterpret_cast<T*>(smem_raw);  // nb x XP
  __shared__ T Ls[32][33];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CB;
  const int cw = min(CB, ncols - c0);
  const int nbp = (nb + 31) & ~31;  // rows padded to whole blocks, padding zero
  for (int e = tid; e < nbp * CB; e += 256) {
    const int i = e / CB, c = e - i * CB;
    Xs[i * XP + c] = (c < cw && i < nb) ? X[(int64_t)i * ldx + c0 + c] : T(0);
  }
  for (int b = 0; b < nb; b += 32) {
    const int bs = min(32, nb - b);
    for (int e = tid; e < 32 * 32; e += 256) {
      const int i = e >> 5, q = e & 31;
      Ls[i][q] = (i < bs && q < bs) ? L[(int64_t)(b + i) * ldl + b + q] : T(0);
    }
    __syncthreads();
    if (tid < CB) {
      T x[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) x[i] = Xs[(b + i) * XP + tid];
#pragma unroll
      for (int i = 1; i < 32; ++i) {
        T acc = x[i];
#pragma unroll
        for (int q = 0; q < i; ++q) acc -= Ls[i][q] * x[q];
        x[i] = acc;
      }
#pragma unroll
      for (int i = 0; i < 32; ++i) Xs[(b + i) * XP + tid] = x[i];
    }
    __syncthreads();
    // rows below the block, 32 at a time: X[r] -= L[r, b:b+32] X[b:b+32],
    // the L rows staged in LDS (zero padded)
    for (int rb = b + 32; rb < nbp; rb += 32) {
      for (int e = tid; e < 32 * 32; e += 256) {
        const int i = e >> 5, q = e & 31;
        Ls[i][q] = (rb + i < nb && q < bs) ? L[(int64_t)(rb + i) * ldl + b + q] : T(0);
      }
      __syncthreads();
      for (int e = tid; e < 32 * CB; e += 256) {
        const int i = e / CB, c = e - i * CB;
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < 32; ++q) acc += Ls[i][q] * Xs[(b + q) * XP + c];
        Xs[(rb + i) * XP + c] -= acc;
      }
      __syncthreads();
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * CB; e += 256) {
    const int i = e / CB, c = e - i * CB;
    if (c < cw) X[(int64_t)i * ldx + c0 + c] = Xs[i * XP + c];
  }
}

// The panel's row permutation (row pos[s] receives the content of slot
// src[s]; slots 0..nb-1 are rows 0..nb-1) on CB columns of X, then
// X[0..nb) <- L^{-1} X[0..nb) for those columns (FP32, nb <= 128).  The whole
// L (unit lower, nb x nb) and every touched row are staged in LDS first, with
// their loads in flight together; the rows below the top nb are written back
// permuted and the top nb are solved in LDS, 32-row blocks: one thread per
// column solves the diagonal block from registers, then all threads update
// the rows below.
// LBLK: L kept as its lower 32 x 32 blocks only (under 70 KB of LDS with CB
// = 16, so the launch fits on a CU beside a trailing-update workgroup).
template <int CB, bool LBLK>
__global__ __launch_bounds__(256) void laswp_trsm_kernel(const float* __restrict__ L, int64_t ldl,
                                                         int nb, float* __restrict__ X,
                                                         int64_t ldx, int ncols,
                                                         const int* __restrict__ perm_n,
                                                         const int* __restrict__ pos,
                                                         const int* __restrict__ src) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // (LBLK: the next panel's interchanges and U rows, on the critical path
  // beside the trailing update's waves: issue priority)
  if (LBLK) __builtin_amdgcn_s_setprio(3);
  constexpr int XP = CB + 1;
  const int nbp = (nb + 31) & ~31, LP = nbp + 1, nbb = nbp / 32;
  // full: nbp x LP; blocks: (bi, bj), bi >= bj, 32 x 33 each at bi(bi+1)/2 + bj
  float* Lf = reinterpret_cast<float*>(smem_raw);
  float* Xs = Lf + (LBLK ? nbb * (nbb + 1) / 2 * 32 * 33 : nbp * LP);  // nbp x XP
  auto lat = [&](int r, int q) -> float& {
    if (LBLK) {
      const int bi = r >> 5, bj = q >> 5;
      return Lf[((bi * (bi + 1) / 2 + bj) * 32 + (r & 31)) * 33 + (q & 31)];
    }
    return Lf[r * LP + q];
  };
  float* buf = Xs + nbp * XP;                      // touched rows x CB
  int* s_pos = reinterpret_cast<int*>(buf + 2 * nb * CB);
  int* s_src = s_pos + 2 * nb;
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CB;
  const int cw = min(CB, ncols - c0);
  // L (strictly lower part; zero elsewhere) with the permutation lists, then
  // the touched rows: every load in flight before the first LDS store (nb <=
  // 128: at most 64 L words and 2 nb CB / 256 row words per thread)
  constexpr int NL = 128 * 128 / 256, NR = 2 * 128 * CB / 256;
  float lv[NL], rv[NR];
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int e = tid + 256 * j, r = e / nbp, q = e - r * nbp;
    lv[j] = (e < nbp * nbp && q < r && r < nb) ? L[(int64_t)r * ldl + q] : 0.f;
  }
  const int nS = *perm_n;
  for (int e = tid; e < nS; e += 256) {
    s_pos[e] = pos[e];
    s_src[e] = src[e];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int e = tid + 256 * j, sl = e / CB, c = e - sl * CB;
    rv[j] = (e < nS * CB && c < cw) ? X[(int64_t)s_pos[sl] * ldx + c0 + c] : 0.f;
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int e = tid + 256 * j, r = e / nbp, q = e - r * nbp;
    if (e < nbp * nbp && (!LBLK || (q >> 5) <= (r >> 5))) lat(r, q) = lv[j];
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int e = tid + 256 * j;
    if (e < nS * CB) buf[e] = rv[j];
  }
  __syncthreads();
  for (int e = tid; e < nbp * CB; e += 256) {
    const int i = e / CB, c = e - i * CB;
    Xs[i * XP + c] = i < nb ? buf[s_src[i] * CB + c] : 0.f;
  }
  for (int e = nb * CB + tid; e < nS * CB; e += 256) {
    const int sl = e / CB, c = e - sl * CB;
    if (c < cw) X[(int64_t)s_pos[sl] * ldx + c0 + c] = buf[s_src[sl] * CB + c];
  }
  __syncthreads();
  for (int b = 0; b < nb; b += 32) {
    if constexpr (LBLK) {
      // (the critical-path instance) the diagonal block by substitution with
      // lanes over rows: a wave takes two columns (lanes 0..31 and 32..63,
      // row i = lane % 32), each step's solved value broadcast by readlane
      // (the same operations in the same order as the per-column loop below,
      // without its dependent chain of 496 FMAs)
      for (int cp = 2 * (tid >> 6); cp < CB; cp += 2 * 4) {
        const int lane = tid & 63, i = lane & 31, c = cp + (lane >> 5);
        float x = c < CB ? Xs[(b + i) * XP + c] : 0.f;
#pragma unroll
        for (int t = 0; t < 31; ++t) {
          const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), t));
          const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 32 + t));
          const float xt = lane < 32 ? lo : hi;
          if (i > t) x -= lat(b + i, b + t) * xt;
        }
        if (c < CB) Xs[(b + i) * XP + c] = x;
      }
    } else if (tid < CB) {
      float x[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) x[i] = Xs[(b + i) * XP + tid];
#pragma unroll
      for (int i = 1; i < 32; ++i) {
        float acc = x[i];
#pragma unroll
        for (int q = 0; q < i; ++q) acc -= lat(b + i, b + q) * x[q];
        x[i] = acc;
      }
#pragma unroll
      for (int i = 0; i < 32; ++i) Xs[(b + i) * XP + tid] = x[i];
    }
    __syncthreads();
    for (int e = tid; e < (nbp - b - 32) * CB; e += 256) {
      const int i = b + 32 + e / CB, c = e % CB;
      float acc = 0.f;
#pragma unroll
      for (int q = 0; q < 32; ++q) acc += lat(i, b + q) * Xs[(b + q) * XP + c];
      Xs[i * XP + c] -= acc;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * CB; e += 256) {
    const int i = e / CB, c = e - i * CB;
    if (c < cw) X[(int64_t)i * ldx + c0 + c] = Xs[i * XP + c];
  }
}

// B[i*ldb + j] = A[i + j*lda] for i < rows, j < cols: a column-major matrix
// into the row-major layout (or back: a row-major matrix is the
// column-major storage of its transpose).
template <typename Ti, typename To>
__global__ __launch_bounds__(256) void transpose_kernel(const Ti* __restrict__ A, int64_t lda,
                                                        int rows, int cols, To* __restrict__ B,
                                                        int64_t ldb) {
  __shared__ Ti tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int jj = ty; jj < 32; jj += 8) {
    int i = i0 + tx, j = j0 + jj;
    if (i < rows && j < cols) tile[jj][tx] = A[i + (int64_t)j * lda];
  }
  __syncthreads();
  for (int ii = ty; ii < 32; ii += 8) {
    int i = i0 + ii, j = j0 + tx;
    if (i < rows && j < cols) B[(int64_t)i * ldb + j] = (To)tile[tx][ii];
  }
}


}  // namespace

// ---------------------------------------------------------------------------
template <typename T>
int laswp_perm_rowmajor(hipStream_t s, T* W, int64_t ldw, int r0, const int* perm_n,
                        const int* pos, const int* src, int nb, int c_split, int skip,
                        int ncols_eff) {
  if (ncols_eff <= 0) return kOk;
  constexpr int CW = 32;
  size_t lds = sizeof(T) * (size_t)2 * nb * CW;
  auto kern = laswp_perm_kernel<T, CW>;
  static bool attr_set = false;
  if (!attr_set) {
    EAMD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(ceil_div(ncols_eff, CW)), dim3(256), lds, s, W, ldw, r0, perm_n,
                     pos, src, c_split, skip, ncols_eff);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int trsm_llu_rowmajor(hipStream_t s, const T* L, int64_t ldl, int nb, T* X, int64_t ldx,
                      int ncols) {
  if (ncols <= 0 || nb <= 0) return kOk;
  constexpr int CB = 32;
  size_t lds = sizeof(T) * (size_t)((nb + 31) & ~31) * (CB + 1);
  auto kern = trsm_llu_kernel<T, CB>;
  static bool attr_set = false;
  if (!attr_set) {
    EAMD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(ceil_div(ncols, CB)), dim3(256), lds, s, L, ldl, nb, X, ldx,
                     ncols);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// The panel's row permutation on ncols columns starting at X (row 0 = the
// panel's first row), then X[0..nb) <- L^{-1} X[0..nb): the interchanges and
// the U rows of those columns in one launch (FP32, nb <= 128).
int laswp_trsm_rowmajor(hipStream_t s, const float* L, int64_t ldl, int nb, float* X, int64_t ldx,
                        int ncols, const int* perm_n, const int* pos, const int* src,
                        bool compact) {
  if (ncols <= 0 || nb <= 0) return kOk;
  if (nb > 128) {
    set_error("laswp_trsm: panel width %d > 128", nb);
    return kErrUnsupported;
  }
  const size_t nbp = (size_t)((nb + 31) & ~31), nbb = nbp / 32;
  auto lds_of = [&](int cb, bool lblk) {
    return sizeof(float) * ((lblk ? nbb * (nbb + 1) / 2 * 32 * 33 : nbp * (nbp + 1)) +
                            nbp * (cb + 1) + 2 * (size_t)nb * cb) +
           sizeof(int) * 4 * (size_t)nb;
  };
  static bool attr_set = false;
  if (!attr_set) {
    EAMD_HIP(hipFuncSetAttribute((const void*)laswp_trsm_kernel<32, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    EAMD_HIP(hipFuncSetAttribute((const void*)laswp_trsm_kernel<4, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr_set = true;
  }
  if (compact)
    hipLaunchKernelGGL((laswp_trsm_kernel<4, true>), dim3(ceil_div(ncols, 4)), dim3(256),
                       lds_of(4, true), s, L, ldl, nb, X, ldx, ncols, perm_n, pos, src);
  else
    hipLaunchKernelGGL((laswp_trsm_kernel<32, false>), dim3(ceil_div(ncols, 32)), dim3(256),
                       lds_of(32, false), s, L, ldl, nb, X, ldx, ncols, perm_n, pos, src);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

template <typename Ti, typename To>
int transpose_copy(hipStream_t s, const Ti* A, int64_t lda, int rows, int cols, To* B,
                   int64_t ldb) {
  if (rows <= 0 || cols <= 0) return kOk;
  dim3 grid(ceil_div(rows, 32), ceil_div(cols, 32));
  hipLaunchKernelGGL((transpose_kernel<Ti, To>), grid, dim3(256), 0, s, A, lda, rows, cols, B,
                     ldb);
  EAMD_HIP(hipGetLastError());
  return kOk;
}
template int transpose_copy<double, double>(hipStream_t, const double*, int64_t, int, int,
                                            double*, int64_t);
template int transpose_copy<double, float>(hipStream_t, const double*, int64_t, int, int, float*,
                                           int64_t);
template int transpose_copy<float, double>(hipStream_t, const float*, int64_t, int, int, double*,
                                           int64_t);
template int transpose_copy<float, float>(hipStream_t, const float*, int64_t, int, int, float*,
                                          int64_t);

// ---------------------------------------------------------------------------
int update_row_order(hipStream_t s, int* ord, int k, const int* perm_n, const int* pos,
                     const int* src, int nb);
int init_row_order(hipStream_t s, int* ord, int n);
template <typename T>
int leaf_lu(eamd_handle* h, T* P, int64_t ldp, int m, int w, int* piv, int* lperm_n, int* lpos,
            int* lsrc, int* pperm_n, int* ppos, int* psrc, int row0, int pnb, int* zero_col,
            int piv_offset, int* fz_global);

int leaf_lu_tournament(eamd_handle* h, float* P, int64_t ldp, int m, int w, int* piv,
                       int* lperm_n, int* lpos, int* lsrc, int* pperm_n, int* ppos, int* psrc,
                       int row0, int pnb, int* zero_col, int piv_offset, int* fz_global,
                       bool* fused);
int env_int(const char* name, int dflt);

constexpr int kLeaf = 32;

// FP32 leaves use tournament pivoting (the mixed-precision solver judges
// its factor by the refined backward error); EAMD_TOURNAMENT=0 selects
// partial pivoting, as the FP64 factorization always uses.  *fused (set to
// false unless the tournament leaf also did them) reports whether the
// leaf's interchanges on the other panel columns, the U rows right of it
// and their update are done.
template <typename T>
int panel_leaf(eamd_handle* h, T* P, int64_t ldp, int m, int w, int* piv, int* lperm_n,
               int* lpos, int* lsrc, int* pperm_n, int* ppos, int* psrc, int row0, int pnb,
               int* zero_col, int piv_offset, int* fz_global, bool* fused) {
  *fused = false;
  if constexpr (std::is_same<T, float>::value) {
    static const bool tournament = env_int("EAMD_TOURNAMENT", 1) != 0;
    static const bool fuse = env_int("EAMD_FUSE_PANEL", 1) != 0;
    if (tournament)
      return leaf_lu_tournament(h, P, ldp, m, w, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc,
                                row0, pnb, zero_col, piv_offset, fz_global,
                                fuse ? fused : nullptr);
  }
  return leaf_lu<T>(h, P, ldp, m, w, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc, row0, pnb,
                    zero_col, piv_offset, fz_global);
}

// FP32 LU trailing update in the requested precision (defined in mixed.hip);
// kPrecPanelFp32 marks the FP32 updates inside a panel.
constexpr int kPrecPanelFp32 = -1;
int sgemm_update_rowmajor(eamd_handle* h, const float* A12, const float* A21, float* A22,
                          int64_t ld, int mrows, int ncols, int k, int precision);

// Factors the panel of columns [k, k+bs) over rows [k, k+m) of the
// row-major W: 32-column leaves, each followed by its interchanges on the
// other panel columns, the leaf's U rows and the update of the remaining
// panel columns (`update`, within_panel = true).  The panel's composed row
// permutation is left in pperm_n/ppos/psrc.
template <typename T, typename Update>
int factor_panel(eamd_handle* h, T* W, int64_t ldw, int k, int bs, int m, int* piv_dev,
                 int* lperm_n, int* lpos, int* lsrc, int* pperm_n, int* ppos, int* psrc,
                 int* zero_col, int* fz_dev, Update& update) {
  hipStream_t s = h->stream;
  T* Pk = W + (int64_t)k * ldw + k;
  PhaseScope ps(h, kPhPanel, 2.0 * sizeof(T) * (double)m * bs, (double)m * bs * bs);
  // (the panel's first leaf starts the panel permutation list)
  for (int j0 = 0; j0 < bs; j0 += kLeaf) {
    const int w = min(kLeaf, bs - j0);
    T* Lk = Pk + (int64_t)j0 * ldw + j0;
    // transpositions stored with the panel offset, zero pivot recorded
    bool fused = false;
    EAMD_CHECK(panel_leaf<T>(h, Lk, ldw, m - j0, w, piv_dev + k + j0, lperm_n, lpos, lsrc,
                             pperm_n, ppos, psrc, j0, bs, zero_col, k + j0, fz_dev, &fused));
    if (fused) continue;
    // the leaf's interchanges on the other columns of the panel
    EAMD_CHECK(laswp_perm_rowmajor<T>(s, Pk + (int64_t)j0 * ldw, ldw, 0, lperm_n, lpos, lsrc, w,
                                      j0, w, bs - w));
    const int rest = bs - j0 - w;
    if (rest > 0) {
      T* X = Lk + w;  // leaf rows x remaining panel columns
      EAMD_CHECK(trsm_llu_rowmajor<T>(s, Lk, ldw, w, X, ldw, rest));
      if (m - j0 - w > 0)
        EAMD_CHECK(update(X, Lk + (int64_t)w * ldw, Lk + (int64_t)w * ldw + w, m - j0 - w, rest,
                          w, true));
    }
  }
  return kOk;
}

// Reads back the first zero pivot and the panel exchange error flag.
int finish_getrf(eamd_handle* h, const int* fz_dev, int64_t* first_zero) {
  hipStream_t s = h->stream;
  int host_fz = -1, host_err = 0;
  EAMD_HIP(hipMemcpyAsync(&host_fz, fz_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  if (h->ws.ptr[kWsPanelSync])
    EAMD_HIP(hipMemcpyAsync(&host_err, h->ws.ptr[kWsPanelSync], sizeof(int),
                            hipMemcpyDeviceToHost, s));
  EAMD_HIP(hipStreamSynchronize(s));
  *first_zero = host_fz;
  if (host_err) {
    EAMD_HIP(hipMemset(h->ws.ptr[kWsPanelSync], 0, sizeof(int)));
    set_error("panel factorization: grid exchange timed out (workgroups not co-resident?)");
    return kErrHip;
  }
  return kOk;
}

// Row-major blocked LU driver of an mr x n matrix (min(mr, n) pivots).
// `update(A12, A21, A22, mrows, ncols, k, within_panel)` performs
// A22 -= A21 * A12 (row-major, ld ldw).  ord (nullable, device, mr ints)
// receives the final row order: row i of the factors is row ord[i] of the
// input, i.e. (P b)[i] = b[ord[i]].
template <typename T, typename Update>
int getrf_rowmajor_t(eamd_handle* h, T* W, int64_t ldw, int mr, int n, int* piv_dev, int* ord,
                     int64_t* first_zero, int nb, Update&& update) {
  hipStream_t s = h->stream;
  if (nb > 256) nb = 256;
  if (nb < kLeaf) nb = kLeaf;
  void* aux = nullptr;
  // aux ints: lperm_n | zero_col | first_zero | pperm_n | pad | lpos[2*32] |
  //           lsrc[2*32] | ppos[2nb] | psrc[2nb]
  size_t need = sizeof(int) * (8 + 4 * kLeaf + 4 * (size_t)nb);
  EAMD_CHECK(h->ws.get(kWsPiv, need, &aux));
  int* ai = static_cast<int*>(aux);
  int* lperm_n = ai;
  int* zero_col = ai + 1;
  int* fz_dev = ai + 2;
  int* pperm_n = ai + 3;
  int* lpos = ai + 8;
  int* lsrc = lpos + 2 * kLeaf;
  int* ppos = lsrc + 2 * kLeaf;
  int* psrc = ppos + 2 * nb;
  EAMD_HIP(hipMemsetAsync(fz_dev, 0xff, sizeof(int), s));  // -1
  if (ord) EAMD_CHECK(init_row_order(s, ord, mr));
  const int size = min(mr, n);
  for (int k = 0; k < size; k += nb) {
    const int bs = min(nb, size - k);
    const int m = mr - k;
    T* Pk = W + (int64_t)k * ldw + k;
    EAMD_CHECK(factor_panel<T>(h, W, ldw, k, bs, m, piv_dev, lperm_n, lpos, lsrc, pperm_n, ppos,
                               psrc, zero_col, fz_dev, update));
    {
      // the panel's interchanges on the columns outside it: [0,k) and [k+bs,n)
      PhaseScope ps(h, kPhLaswp, 4.0 * sizeof(T) * bs * (double)(n - bs), 0.0);
      EAMD_CHECK(laswp_perm_rowmajor<T>(s, W + (int64_t)k * ldw, ldw, 0, pperm_n, ppos, psrc, bs,
                                        k, bs, n - bs));
    }
    if (ord) EAMD_CHECK(update_row_order(s, ord, k, pperm_n, ppos, psrc, bs));
    const int tn = n - k - bs;
    if (tn > 0) {
      T* A12 = W + (int64_t)k * ldw + (k + bs);
      {
        PhaseScope ps(h, kPhTrsm, sizeof(T) * (2.0 * bs * tn + 0.5 * bs * bs),
                      (double)bs * bs * tn);
        EAMD_CHECK(trsm_llu_rowmajor<T>(s, Pk, ldw, bs, A12, ldw, tn));
      }
      T* A21 = W + (int64_t)(k + bs) * ldw + k;
      T* A22 = W + (int64_t)(k + bs) * ldw + (k + bs);
      if (m - bs > 0) EAMD_CHECK(update(A12, A21, A22, m - bs, tn, bs, false));
    }
  }
  return finish_getrf(h, fz_dev, first_zero);
}

// Sets the handle's stream for the lifetime of the object (launchers and
// phase timers use h->stream).
struct StreamSwap {
  eamd_handle* h;
  hipStream_t saved;
  StreamSwap(eamd_handle* h_, hipStream_t s) : h(h_), saved(h_->stream) { h->stream = s; }
  ~StreamSwap() { h->stream = saved; }
};

// FP32 LU with a mixed-precision trailing update and a look-ahead of one
// panel.  The handle's stream factors panel k, then (after the rest of panel
// k-1's update) applies panel k's interchanges, U rows and update to the
// next panel's columns only, and goes on to factor that panel; a second
// stream applies panel k's interchanges to the columns left of it and
// beyond the next panel and does their U rows and update, which so overlap
// the next panel's factorization.  Packed panel columns and permutation
// lists alternate between two buffers (panel k's are read by the second
// stream while panel k+1 is factored).
int getrf_lookahead_body(eamd_handle* h, hipStream_t sa, hipStream_t sb, float* W, int64_t ldw,
                         int n, int* piv_dev, int* ord, int nb, int precision);
int getrf_lookahead_mixed(eamd_handle* h, float* W, int64_t ldw, int n, int* piv_dev, int* ord,
                          int64_t* first_zero, int nb, int precision) {
  if (nb > 128) nb = 128;  // the next panel's interchanges and U rows keep L in LDS
  if (nb < kLeaf) nb = kLeaf;
  EAMD_CHECK(ensure_side_stream(h));
  // the panels run on the reserved-CU stream when there is one
  const hipStream_t user = h->stream;
  const hipStream_t sa = h->pstream ? h->pstream : user, sb = h->side;
  if (sa != user) {
    EAMD_HIP(hipEventRecord(h->la_ev[4], user));
    EAMD_HIP(hipStreamWaitEvent(sa, h->la_ev[4], 0));
  }
  int st = getrf_lookahead_body(h, sa, sb, W, ldw, n, piv_dev, ord, nb, precision);
  if (sa != user) {
    // join, also after a failure
    EAMD_HIP(hipEventRecord(h->la_ev[5], sa));
    EAMD_HIP(hipStreamWaitEvent(user, h->la_ev[5], 0));
  }
  EAMD_CHECK(st);
  return finish_getrf(h, static_cast<int*>(h->ws.ptr[kWsPiv]) + 2, first_zero);
}

int getrf_lookahead_body(eamd_handle* h, hipStream_t sa, hipStream_t sb, float* W, int64_t ldw,
                         int n, int* piv_dev, int* ord, int nb, int precision) {
  StreamSwap swa(h, sa);
  void* aux = nullptr;
  // aux ints: lperm_n | zero_col | first_zero | pperm_n[2] | pad | lpos[2*32] |
  //           lsrc[2*32] | (ppos[2nb], psrc[2nb]) x 2
  EAMD_CHECK(h->ws.get(kWsPiv, sizeof(int) * (8 + 4 * kLeaf + 8 * (size_t)nb), &aux));
  int* ai = static_cast<int*>(aux);
  int* lperm_n = ai;
  int* zero_col = ai + 1;
  int* fz_dev = ai + 2;
  int* lpos = ai + 8;
  int* lsrc = lpos + 2 * kLeaf;
  int* pn[2] = {ai + 3, ai + 4};
  int* pp[2] = {lsrc + 2 * kLeaf, lsrc + 2 * kLeaf + 4 * nb};
  int* psl[2] = {pp[0] + 2 * nb, pp[1] + 2 * nb};
  // all buffers sized before the second stream has work (a growing
  // workspace synchronizes the device)
  const size_t opL = mixed_operand_bytes(n, nb, precision);
  void* pl = nullptr;
  void* pu = nullptr;
  EAMD_CHECK(h->ws.get(kWsPackL, 2 * opL, &pl));
  EAMD_CHECK(h->ws.get(kWsPackU, 2 * opL, &pu));
  char* packL[2] = {static_cast<char*>(pl), static_cast<char*>(pl) + opL};
  char* packU[2] = {static_cast<char*>(pu), static_cast<char*>(pu) + opL};
  auto in_panel = [&](float* A12, float* A21, float* A22, int mrows, int ncols, int k,
                      bool) -> int {
    return sgemm_update_rowmajor(h, A12, A21, A22, ldw, mrows, ncols, k, kPrecPanelFp32);
  };
  hipEvent_t evP = h->la_ev[0], evB[2] = {h->la_ev[1], h->la_ev[2]};
  EAMD_HIP(hipMemsetAsync(fz_dev, 0xff, sizeof(int), sa));  // -1
  if (ord) EAMD_CHECK(init_row_order(sa, ord, n));
  bool side_used = false;
  for (int k = 0, it = 0; k < n; k += nb, ++it) {
    const int par = it & 1;
    const int bs = min(nb, n - k), m = n - k;
    const int tn = n - k - bs, n1 = min(nb, tn), n2 = tn - n1;
    float* Pk = W + (int64_t)k * ldw + k;
    EAMD_CHECK(factor_panel<float>(h, W, ldw, k, bs, m, piv_dev, lperm_n, lpos, lsrc, pn[par],
                                   pp[par], psl[par], zero_col, fz_dev, in_panel));
    if (tn > 0) EAMD_HIP(hipEventRecord(evP, sa));  // the panel is factored
    // the next panel's columns were last written by the previous panel's
    // second-stream update
    if (it > 0) EAMD_HIP(hipStreamWaitEvent(sa, evB[par ^ 1], 0));
    if (tn == 0) {  // last panel: its interchanges on the columns to its left
      if (ord) EAMD_CHECK(update_row_order(sa, ord, k, pn[par], pp[par], psl[par], bs));
      PhaseScope ps(h, kPhLaswp, 4.0 * sizeof(float) * bs * (double)k, 0.0);
      EAMD_CHECK(laswp_perm_rowmajor<float>(sa, W + (int64_t)k * ldw, ldw, 0, pn[par], pp[par],
                                            psl[par], bs, k, bs, k));
      break;
    }
    float* A12n = W + (int64_t)k * ldw + (k + bs);
    float* A21 = W + (int64_t)(k + bs) * ldw + k;
    {
      // the next panel's columns: interchanges and U rows in one launch,
      // then their update in FP32 (no packing on this path)
      PhaseScope ps(h, kPhTrsm, 4.0 * (4.0 * bs * n1 + 0.5 * bs * bs), (double)bs * bs * n1);
      EAMD_CHECK(laswp_trsm_rowmajor(sa, Pk, ldw, bs, A12n, ldw, n1, pn[par], pp[par],
                                            psl[par], true));
    }
    EAMD_CHECK(f32_update_rowmajor(h, A12n, A21, W + (int64_t)(k + bs) * ldw + (k + bs), ldw,
                                   m - bs, n1, bs, kPhNext));
    // second stream: the row order, the columns left of the panel and
    // beyond the next panel
    EAMD_HIP(hipStreamWaitEvent(sb, evP, 0));
    side_used = true;
    {
      StreamSwap sw(h, sb);
      if (ord) EAMD_CHECK(update_row_order(sb, ord, k, pn[par], pp[par], psl[par], bs));
      if (k > 0) {
        PhaseScope ps(h, kPhLaswp, 4.0 * sizeof(float) * bs * (double)k, 0.0);
        EAMD_CHECK(laswp_perm_rowmajor<float>(sb, W + (int64_t)k * ldw, ldw, 0, pn[par], pp[par],
                                              psl[par], bs, k, 0, k));
      }
      if (n2 > 0) {
        float* A12r = A12n + n1;
        {
          PhaseScope ps(h, kPhTrsm, 4.0 * (4.0 * bs * n2 + 0.5 * bs * bs), (double)bs * bs * n2);
          EAMD_CHECK(laswp_trsm_rowmajor(sb, Pk, ldw, bs, A12r, ldw, n2, pn[par], pp[par],
                                                psl[par], false));
        }
        MixedOperand L, Ur;
        EAMD_CHECK(mixed_pack(h, true, A21, ldw, m - bs, bs, precision, packL[par], &L));
        EAMD_CHECK(mixed_pack(h, false, A12r, ldw, n2, bs, precision, packU[1], &Ur));
        EAMD_CHECK(mixed_update_packed(h, Ur, n2, L, m - bs, bs, precision,
                                       W + (int64_t)(k + bs) * ldw + (k + bs + n1), ldw));
      }
      EAMD_HIP(hipEventRecord(evB[par], sb));
    }
  }
  if (side_used) {
    EAMD_HIP(hipEventRecord(h->la_ev[3], sb));
    EAMD_HIP(hipStreamWaitEvent(sa, h->la_ev[3], 0));
  }
  return kOk;
}

namespace {
// row-major A22 -= A21*A12  <=>  column-major A22^T -= A12^T * A21^T
int dgemm_rowmajor_update(eamd_handle* h, int64_t ldw, double* A12, double* A21, double* A22,
                          int mrows, int ncols, int k, bool within_panel) {
  PhaseScope ps(h, within_panel ? kPhPanel : kPhUpdate,
                within_panel ? 0.0 : 8.0 * (2.0 * mrows * ncols + (double)k * (mrows + ncols)),
                within_panel ? 0.0 : 2.0 * mrows * ncols * k);
  return dgemm_device(h->stream, false, false, ncols, mrows, k, -1.0, A12, ldw, A21, ldw, 1.0,
                      A22, ldw, 0);
}
}  // namespace

int dgetrf_rowmajor(eamd_handle* h, double* W, int64_t ldw, int m, int n, int* piv_dev,
                    int64_t* first_zero, int nb) {
  auto upd = [&](double* A12, double* A21, double* A22, int mrows, int ncols, int k,
                 bool within) -> int {
    return dgemm_rowmajor_update(h, ldw, A12, A21, A22, mrows, ncols, k, within);
  };
  return getrf_rowmajor_t<double>(h, W, ldw, m, n, piv_dev, nullptr, first_zero, nb, upd);
}

int sgetrf_rowmajor(eamd_handle* h, float* W, int64_t ldw, int n, int* piv_dev,
                    int64_t* first_zero, int nb, int update_precision) {
  auto upd = [&](float* A12, float* A21, float* A22, int mrows, int ncols, int k,
                 bool within) -> int {
    // inside the panel the update stays in FP32 (the panel is factored in FP32)
    return sgemm_update_rowmajor(h, A12, A21, A22, ldw, mrows, ncols, k,
                                 within ? kPrecPanelFp32 : update_precision);
  };
  return getrf_rowmajor_t<float>(h, W, ldw, n, n, piv_dev, nullptr, first_zero, nb, upd);
}

// The refinement solver's factorization: FP64 (plain) or FP32 with the
// chosen trailing-update precision, recording the final row order.
int getrf_rowmajor_order(eamd_handle* h, void* W, bool fp64, int64_t ldw, int n, int* piv,
                         int* ord, int64_t* first_zero, int nb, int precision) {
  if (fp64) {
    double* Wd = static_cast<double*>(W);
    auto upd = [&](double* A12, double* A21, double* A22, int mrows, int ncols, int k,
                   bool within) -> int {
      return dgemm_rowmajor_update(h, ldw, A12, A21, A22, mrows, ncols, k, within);
    };
    return getrf_rowmajor_t<double>(h, Wd, ldw, n, n, piv, ord, first_zero, nb, upd);
  }
  float* Wf = static_cast<float*>(W);
  static const bool lookahead = env_int("EAMD_LOOKAHEAD", 1) != 0;
  if (lookahead && (precision == EAMD_PREC_FP16X3 || precision == EAMD_PREC_FP16 ||
                    precision == EAMD_PREC_BF16))
    return getrf_lookahead_mixed(h, Wf, ldw, n, piv, ord, first_zero, nb, precision);
  auto upd = [&](float* A12, float* A21, float* A22, int mrows, int ncols, int k,
                 bool within) -> int {
    return sgemm_update_rowmajor(h, A12, A21, A22, ldw, mrows, ncols, k,
                                 within ? kPrecPanelFp32 : precision);
  };
  return getrf_rowmajor_t<float>(h, Wf, ldw, n, n, piv, ord, first_zero, nb, upd);
}

}  // namespace eamd
