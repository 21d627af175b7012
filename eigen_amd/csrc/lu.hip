// lu.hip -- blocked LU with partial pivoting on a row-major working matrix.
//
// Replaces partial_lu_impl<Scalar,ColMajor,int>::blocked_lu
// (Eigen/src/LU/PartialPivLU.h:426-496) and unblocked_lu (:365-409).
//
// The caller's column-major matrix is transposed into a row-major working
// copy so that row interchanges touch contiguous memory (a column-major
// row swap reads one 8-byte element per cache line).  For each block column
// of width nb:
//   1. panel_lu_kernel: the (n-k) x nb panel is spread over G workgroups,
//      each holding R rows in LDS.  Per column: local |a| argmax (strict '>'
//      scan in row order; ties resolve to the lowest row; a NaN wins only
//      from the diagonal position, as the reference's visitor scan does),
//      one grid-wide exchange of (score, row, candidate row), swap, scale
//      by the pivot (true division, as `col /= pivot` in the reference) and
//      the rank-1 update of the panel in LDS.
//   2. laswp_perm_kernel: the panel's row permutation (a permutation of at
//      most 2nb rows, tracked by the panel kernel) is applied to the columns
//      left and right of the panel.
//   3. trsm_llu_kernel: A12 <- L11^{-1} A12.
//   4. FP64 MFMA GEMM: A22 -= A21 * A12.
#include "common.h"
#include "kernels.h"

#include <cmath>
#include <vector>

namespace eamd {

namespace {

constexpr int kSpinLimit = 1 << 22;

// ---- grid exchange words ---------------------------------------------------
// Every exchanged quantity travels in 8-byte words whose high half is the
// epoch of the column that produced it and whose low half is 32 bits of
// payload.  An aligned 8-byte store is single-copy atomic, so a reader that
// sees the expected epoch also sees the payload written with it: no release
// or acquire fence (and none of their L2 write-backs) is needed per column.
typedef unsigned long long u64;

__device__ __forceinline__ void ll_store(u64* p, unsigned payload, unsigned epoch) {
  __hip_atomic_store(p, ((u64)epoch << 32) | payload, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
struct LLWords;  // number of exchange words per value
template <>
struct LLWords<double> {
  static constexpr int n = 2;
};
template <>
struct LLWords<float> {
  static constexpr int n = 1;
};

template <typename T>
__device__ __forceinline__ void ll_put(u64* p, T v, unsigned epoch) {
  if (sizeof(T) == 8) {
    u64 bits = __double_as_longlong((double)v);
    ll_store(p, (unsigned)bits, epoch);
    ll_store(p + 1, (unsigned)(bits >> 32), epoch);
  } else {
    ll_store(p, __float_as_uint((float)v), epoch);
  }
}

// Polls K consecutive words together (their loads are independent, so one
// poll costs one memory round trip) until all carry `epoch`.
template <int K>
__device__ __forceinline__ bool ll_load_n(const u64* p, unsigned epoch, unsigned (&payload)[K],
                                          const int* err) {
  for (int spins = 0;; ++spins) {
    u64 w[K];
#pragma unroll
    for (int q = 0; q < K; ++q)
      w[q] = __hip_atomic_load(const_cast<u64*>(p + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
#pragma unroll
    for (int q = 0; q < K; ++q) ok &= (unsigned)(w[q] >> 32) == epoch;
    if (ok) {
#pragma unroll
      for (int q = 0; q < K; ++q) payload[q] = (unsigned)w[q];
      return true;
    }
    if ((spins & 255) == 255) {
      if (spins > kSpinLimit ||
          __hip_atomic_load(const_cast<int*>(err), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return false;
    }
  }
}

template <typename T>
__device__ __forceinline__ bool ll_get(const u64* p, unsigned epoch, T& v, const int* err) {
  if (sizeof(T) == 8) {
    unsigned pl[2];
    if (!ll_load_n<2>(p, epoch, pl, err)) return false;
    v = (T)__longlong_as_double((long long)(((u64)pl[1] << 32) | pl[0]));
  } else {
    unsigned pl[1];
    if (!ll_load_n<1>(p, epoch, pl, err)) return false;
    v = (T)__uint_as_float(pl[0]);
  }
  return true;
}

__device__ __forceinline__ bool beats(double s, int i, double bs, int bi) {
  return s > bs || (s == bs && i < bi);
}

// |a| as the pivot score with the reference's NaN rule: a NaN ranks below
// every number except in the diagonal (first scanned) position.
__device__ __forceinline__ double pivot_score(double v, bool first) {
  double s = fabs(v);
  if (s != s) s = first ? INFINITY : -1.0;
  return s;
}

template <int BLOCK>
__device__ __forceinline__ void block_argmax(double& best, int& bi, double* red_s, int* red_i) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double os = __shfl_xor(best, off);
    int oi = __shfl_xor(bi, off);
    if (beats(os, oi, best, bi)) { best = os; bi = oi; }
  }
  if (lane == 0) { red_s[wave] = best; red_i[wave] = bi; }
  __syncthreads();
  best = red_s[0];
  bi = red_i[0];
#pragma unroll
  for (int w = 1; w < BLOCK / 64; ++w)
    if (beats(red_s[w], red_i[w], best, bi)) { best = red_s[w]; bi = red_i[w]; }
  __syncthreads();
}

// Exchange buffers of one parity: per workgroup a slot of 4 words (pivot
// value, row index) and its candidate row (nb values); one row-j buffer.
template <typename T>
struct PanelXchg {
  u64* slots;    // [2][G][4]
  u64* cand;     // [2][G][nb * LLWords<T>::n]
  u64* rowj;     // [2][nb * LLWords<T>::n]
};

// Outputs besides the factored panel:
//   piv[j]            panel-relative transposition of column j
//   perm_n, perm_pos  the rows touched by the interchanges (slot s -> row
//   perm_src          perm_pos[s]); after the panel, row perm_pos[s] holds the
//                     original content of row perm_pos[perm_src[s]]
//   zero_col          first column with an exactly zero pivot (or -1)
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void panel_lu_kernel(
    T* __restrict__ P, int64_t ldp, int m, int nb, int R, int G, int* __restrict__ piv,
    int* __restrict__ perm_n, int* __restrict__ perm_pos, int* __restrict__ perm_src,
    int* __restrict__ zero_col, PanelXchg<T> xg, int* __restrict__ err, unsigned epoch_base) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int W = LLWords<T>::n;
  const int nbp = nb + 1;
  T* rows = reinterpret_cast<T*>(smem_raw);
  T* urow = rows + (size_t)R * nbp;
  int* sPos = reinterpret_cast<int*>(urow + nb + 1);
  int* sCont = sPos + 2 * nb;

  __shared__ double red_s[BLOCK / 64];
  __shared__ int red_i[BLOCK / 64];
  __shared__ int s_win, s_abort, s_nS;
  __shared__ T s_pv;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x;
  const int r0 = g * R;
  const int r1 = min(m, r0 + R);
  const int nr = max(0, r1 - r0);

  for (int e = tid; e < nr * nb; e += BLOCK) {
    int i = e / nb, c = e - i * nb;
    rows[i * nbp + c] = P[(int64_t)(r0 + i) * ldp + c];
  }
  if (g == 0) {
    for (int s = tid; s < nb; s += BLOCK) { sPos[s] = s; sCont[s] = s; }
    if (tid == 0) { s_nS = nb; }
  }
  if (tid == 0) s_abort = 0;
  __syncthreads();

  int first_zero = -1;
  for (int j = 0; j < nb; ++j) {
    const int par = j & 1;
    const unsigned epoch = epoch_base + (unsigned)j + 1u;
    // ---- 1. local candidate: first maximum of |a_ij| over local rows i >= j
    double best = -2.0;
    int bi = 0x7fffffff;
    for (int i = max(r0, j) + tid; i < r1; i += BLOCK) {
      double s = pivot_score((double)rows[(i - r0) * nbp + j], i == j);
      if (s > best) { best = s; bi = i; }
    }
    block_argmax<BLOCK>(best, bi, red_s, red_i);

    int p;
    if (G == 1) {
      p = bi;
      for (int c = tid; c < nb; c += BLOCK) urow[c] = rows[(p - r0) * nbp + c];
      __syncthreads();
      if (p != j) {
        for (int c = tid; c < nb; c += BLOCK) {
          T t = rows[(j - r0) * nbp + c];
          rows[(j - r0) * nbp + c] = urow[c];
          rows[(p - r0) * nbp + c] = t;
        }
      }
    } else {
      // ---- 2. publish: slot (pivot value, row), candidate row, row j
      u64* slot = xg.slots + ((size_t)par * G + g) * 4;
      u64* crow = xg.cand + ((size_t)par * G + g) * (size_t)nb * W;
      const bool have = bi != 0x7fffffff;
      if (tid == 0) {
        T v = have ? rows[(bi - r0) * nbp + j] : T(0);
        ll_put<double>(slot, (double)v, epoch);
        ll_store(slot + 2, have ? (unsigned)bi : 0xffffffffu, epoch);
      }
      if (have)
        for (int c = tid; c < nb; c += BLOCK) ll_put<T>(crow + (size_t)c * W, rows[(bi - r0) * nbp + c], epoch);
      if (j >= r0 && j < r1) {
        u64* rj = xg.rowj + (size_t)par * nb * W;
        for (int c = tid; c < nb; c += BLOCK) ll_put<T>(rj + (size_t)c * W, rows[(j - r0) * nbp + c], epoch);
      }
      // ---- 3. every workgroup's slot (one lane per workgroup), reduce
      if (wave == 0) {
        double wb = -2.0;
        int wi = 0x7fffffff;
        double wv = 0.0;
        int abort = 0;
        for (int g2 = lane; g2 < G; g2 += 64) {
          const u64* sl = xg.slots + ((size_t)par * G + g2) * 4;
          unsigned pl[3];
          if (!ll_load_n<3>(sl, epoch, pl, err)) {
            abort = 1;
            break;
          }
          const double v = __longlong_as_double((long long)(((u64)pl[1] << 32) | pl[0]));
          const unsigned idx = pl[2];
          if (idx != 0xffffffffu) {
            double s2 = pivot_score(v, (int)idx == j);
            if (beats(s2, (int)idx, wb, wi)) { wb = s2; wi = (int)idx; wv = v; }
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          double os = __shfl_xor(wb, off);
          int oi = __shfl_xor(wi, off);
          double ov = __shfl_xor(wv, off);
          if (beats(os, oi, wb, wi)) { wb = os; wi = oi; wv = ov; }
          abort |= __shfl_xor(abort, off);
        }
        if (lane == 0) {
          s_win = wi;
          s_abort = abort;
          if (abort) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      __syncthreads();
      if (s_abort) return;
      p = s_win;
      // ---- 4. pivot row and interchange
      const u64* wrow = xg.cand + ((size_t)par * G + p / R) * (size_t)nb * W;
      int bad = 0;
      for (int c = tid; c < nb; c += BLOCK) {
        T v;
        if (!ll_get<T>(wrow + (size_t)c * W, epoch, v, err)) bad = 1;
        urow[c] = v;
      }
      if (p != j && p >= r0 && p < r1) {
        const u64* rj = xg.rowj + (size_t)par * nb * W;
        for (int c = tid; c < nb; c += BLOCK) {
          T v;
          if (!ll_get<T>(rj + (size_t)c * W, epoch, v, err)) bad = 1;
          rows[(p - r0) * nbp + c] = v;
        }
      }
      if (bad) s_abort = 1;
      __syncthreads();
      if (s_abort) {
        if (tid == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      if (p != j && j >= r0 && j < r1)
        for (int c = tid; c < nb; c += BLOCK) rows[(j - r0) * nbp + c] = urow[c];
    }
    // ---- bookkeeping of the row permutation (workgroup 0)
    if (g == 0) {
      if (tid == 0) piv[j] = p;
      if (p != j && wave == 0) {
        int slotb = p;
        if (p >= nb) {
          int nS = s_nS;
          int found = -1;
          for (int s = nb + lane; s < nS; s += 64)
            if (sPos[s] == p) found = s;
          unsigned long long bal = __ballot(found >= 0);
          if (bal) {
            int src = __ffsll((long long)bal) - 1;
            slotb = __shfl(found, src);
          } else {
            slotb = nS;
          }
          if (lane == 0 && !bal) {
            sPos[nS] = p;
            sCont[nS] = nS;
            s_nS = nS + 1;
          }
        }
        if (lane == 0) {
          int t = sCont[j];
          sCont[j] = sCont[slotb];
          sCont[slotb] = t;
        }
      }
    }
    __syncthreads();
    // ---- 5. scale and rank-1 update
    const T pv = urow[j];
    if (pv != T(0)) {
      for (int i = max(r0, j + 1) + tid; i < r1; i += BLOCK) rows[(i - r0) * nbp + j] /= pv;
      __syncthreads();
      const int c0 = j + 1;
      if (c0 < nb) {
        const int tc = tid & 31, tr = tid >> 5;
        for (int i = max(r0, j + 1) + tr; i < r1; i += BLOCK / 32) {
          T* row = rows + (i - r0) * nbp;
          const T l = row[j];
          for (int c = c0 + tc; c < nb; c += 32) row[c] -= l * urow[c];
        }
      }
    } else if (first_zero < 0) {
      first_zero = j;
    }
    __syncthreads();
  }

  for (int e = tid; e < nr * nb; e += BLOCK) {
    int i = e / nb, c = e - i * nb;
    P[(int64_t)(r0 + i) * ldp + c] = rows[i * nbp + c];
  }
  if (g == 0) {
    const int nS = s_nS;
    for (int s = tid; s < nS; s += BLOCK) {
      perm_pos[s] = sPos[s];
      perm_src[s] = sCont[s];
    }
    if (tid == 0) {
      *perm_n = nS;
      *zero_col = first_zero;
    }
  }
}

// Apply the panel permutation to columns [0,c_split) U [c_split+skip, ncols)
// of rows r0 + perm_pos[*].  Each workgroup stages CW columns of every
// touched row in LDS, then writes them back permuted.
template <typename T, int CW>
__global__ __launch_bounds__(256) void laswp_perm_kernel(T* __restrict__ W, int64_t ldw, int r0,
                                                         const int* __restrict__ perm_n,
                                                         const int* __restrict__ pos,
                                                         const int* __restrict__ src,
                                                         int c_split, int skip, int ncols_eff) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* buf = reinterpret_cast<T*>(smem_raw);
  const int nS = *perm_n;
  const int cbase = blockIdx.x * CW;
  const int tc = threadIdx.x % CW, tr = threadIdx.x / CW;
  const int RS = 256 / CW;
  int ce = cbase + tc;
  bool ok = ce < ncols_eff;
  int c = ce < c_split ? ce : ce + skip;
  for (int s = tr; s < nS; s += RS)
    if (ok) buf[s * CW + tc] = W[(int64_t)(r0 + pos[s]) * ldw + c];
  __syncthreads();
  for (int s = tr; s < nS; s += RS)
    if (ok) W[(int64_t)(r0 + pos[s]) * ldw + c] = buf[src[s] * CW + tc];
}

// X <- L^{-1} X, L unit lower nb x nb (row-major, ld ldl), X nb x ncols
// (row-major).  One workgroup per CB columns; X block resident in LDS.
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
  for (int e = tid; e < nb * CB; e += 256) {
    int i = e / CB, c = e - i * CB;
    Xs[i * XP + c] = c < cw ? X[(int64_t)i * ldx + c0 + c] : T(0);
  }
  for (int b = 0; b < nb; b += 32) {
    const int bs = min(32, nb - b);
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += 256) {
      int i = e >> 5, s = e & 31;
      Ls[i][s] = (i < bs && s < bs) ? L[(int64_t)(b + i) * ldl + b + s] : T(0);
    }
    __syncthreads();
    // diagonal block, column-oriented substitution
    for (int s = 0; s < bs - 1; ++s) {
      for (int e = tid; e < 32 * CB; e += 256) {
        int i = e / CB, c = e - i * CB;
        if (i > s && i < bs) Xs[(b + i) * XP + c] -= Ls[i][s] * Xs[(b + s) * XP + c];
      }
      __syncthreads();
    }
    // rows below the block: X[r] -= L[r, b:b+bs] * X[b:b+bs]
    for (int rb = b + bs; rb < nb; rb += 32) {
      const int rs = min(32, nb - rb);
      __syncthreads();
      for (int e = tid; e < 32 * 32; e += 256) {
        int i = e >> 5, s = e & 31;
        Ls[i][s] = (i < rs && s < bs) ? L[(int64_t)(rb + i) * ldl + b + s] : T(0);
      }
      __syncthreads();
      for (int e = tid; e < 32 * CB; e += 256) {
        int i = e / CB, c = e - i * CB;
        if (i < rs) {
          T acc = T(0);
          for (int s = 0; s < bs; ++s) acc += Ls[i][s] * Xs[(b + s) * XP + c];
          Xs[(rb + i) * XP + c] -= acc;
        }
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < nb * CB; e += 256) {
    int i = e / CB, c = e - i * CB;
    if (c < cw) X[(int64_t)i * ldx + c0 + c] = Xs[i * XP + c];
  }
}

// B (row-major rows x cols, ld ldb) = A^T where A is column-major
// (rows x cols)... i.e. B[i*ldb + j] = A[i + j*lda].  The same kernel
// converts in both directions because a column-major matrix is the
// row-major storage of its transpose.
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

__global__ void count_swaps_kernel(const int* piv, int n, int base, int* out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int local = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) local += (piv[i] != i + base) ? 1 : 0;
  atomicAdd(&cnt, local);
  __syncthreads();
  if (threadIdx.x == 0) *out = cnt;
}

__global__ void add_offset_kernel(int* v, int n, int off) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] += off;
}

__global__ void record_zero_kernel(const int* zc, int k, int* first_zero) {
  if (threadIdx.x == 0 && *zc >= 0 && *first_zero < 0) *first_zero = k + *zc;
}

template <typename T>
size_t panel_lds_bytes(int R, int nb) {
  size_t b = (size_t)R * (nb + 1) * sizeof(T) + (size_t)(nb + 1) * sizeof(T);
  b = (b + 15) & ~(size_t)15;
  b += (size_t)4 * nb * sizeof(int);
  return b;
}

constexpr size_t kLdsBudget = 150 * 1024;

}  // namespace

// ---------------------------------------------------------------------------
template <typename T>
int panel_lu_rowmajor_impl(eamd_handle* h, T* P, int64_t ldp, int m, int nb, int* piv,
                           int* perm_n, int* perm_pos, int* perm_src, int* zero_col) {
  constexpr int BLOCK = 256;
  // rows per workgroup bounded by LDS; spread rows over at most num_cus WGs
  size_t per_row = (size_t)(nb + 1) * sizeof(T);
  size_t fixed = panel_lds_bytes<T>(0, nb);
  int rmax = (int)((kLdsBudget - fixed) / per_row);
  if (rmax < 1) {
    set_error("panel width %d too large for LDS", nb);
    return kErrUnsupported;
  }
  int G = ceil_div(m, rmax);
  if (G > h->num_cus) {
    set_error("panel of %d rows needs %d workgroups (> %d CUs)", m, G, h->num_cus);
    return kErrUnsupported;
  }
  int R = ceil_div(m, G);
  size_t lds = panel_lds_bytes<T>(R, nb);

  // sync area: err flag | slots[2][G][4] | cand[2][G][nb*W] | rowj[2][nb*W]
  constexpr int W = LLWords<T>::n;
  size_t off_slots = 256;
  size_t off_cand = off_slots + sizeof(u64) * 2 * (size_t)G * 4;
  size_t off_rowj = off_cand + sizeof(u64) * 2 * (size_t)G * nb * W;
  size_t need = off_rowj + sizeof(u64) * 2 * (size_t)nb * W;
  void* base = nullptr;
  bool fresh = h->ws.bytes[kWsPanelSync] < need;
  EAMD_CHECK(h->ws.get(kWsPanelSync, need, &base));
  char* b = static_cast<char*>(base);
  if (fresh || h->panel_epoch > (1u << 31)) {
    EAMD_HIP(hipMemsetAsync(base, 0, h->ws.bytes[kWsPanelSync], h->stream));
    h->panel_epoch = 0;
  }
  unsigned epoch_base = h->panel_epoch;
  h->panel_epoch += (unsigned)nb + 1u;
  PanelXchg<T> xg;
  xg.slots = reinterpret_cast<u64*>(b + off_slots);
  xg.cand = reinterpret_cast<u64*>(b + off_cand);
  xg.rowj = reinterpret_cast<u64*>(b + off_rowj);
  auto kern = panel_lu_kernel<T, BLOCK>;
  static bool attr_set = false;
  if (!attr_set) {
    EAMD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kLdsBudget + 8192));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(G), dim3(BLOCK), lds, h->stream, P, ldp, m, nb, R, G, piv, perm_n,
                     perm_pos, perm_src, zero_col, xg, reinterpret_cast<int*>(b), epoch_base);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

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
                                 160 * 1024));
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
  size_t lds = sizeof(T) * (size_t)nb * (CB + 1);
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

// Row-major blocked LU driver (any T with a trailing-update callback).  ord
// (nullable, device, n ints) receives the final row order: row i of the
// factors is row ord[i] of the input, i.e. (P b)[i] = b[ord[i]].
template <typename T, typename Update>
int getrf_rowmajor_t(eamd_handle* h, T* W, int64_t ldw, int n, int* piv_dev, int* ord,
                     int64_t* first_zero, int nb, Update&& update) {
  hipStream_t s = h->stream;
  void* aux = nullptr;
  // aux ints: perm_n | zero_col | first_zero | nswaps | perm_pos[2nb] | perm_src[2nb]
  size_t need = sizeof(int) * (8 + 4 * (size_t)nb);
  EAMD_CHECK(h->ws.get(kWsPiv, need, &aux));
  int* ai = static_cast<int*>(aux);
  int* perm_n = ai;
  int* zero_col = ai + 1;
  int* fz_dev = ai + 2;
  int* pos = ai + 8;
  int* src = pos + 2 * nb;
  EAMD_HIP(hipMemsetAsync(fz_dev, 0xff, sizeof(int), s));  // -1
  if (ord) EAMD_CHECK(init_row_order(s, ord, n));
  for (int k = 0; k < n; k += nb) {
    const int bs = min(nb, n - k);
    const int m = n - k;
    T* Pk = W + (int64_t)k * ldw + k;
    {
      PhaseScope ps(h, kPhPanel, 2.0 * sizeof(T) * (double)m * bs, (double)m * bs * bs);
      EAMD_CHECK(panel_lu_rowmajor_impl<T>(h, Pk, ldw, m, bs, piv_dev + k, perm_n, pos, src,
                                           zero_col));
    }
    hipLaunchKernelGGL(record_zero_kernel, dim3(1), dim3(64), 0, s, zero_col, k, fz_dev);
    {
      // interchanges on the columns outside the panel: [0,k) and [k+bs,n)
      PhaseScope ps(h, kPhLaswp, 4.0 * sizeof(T) * bs * (double)(n - bs), 0.0);
      EAMD_CHECK(laswp_perm_rowmajor<T>(s, W + (int64_t)k * ldw, ldw, 0, perm_n, pos, src, bs, k,
                                        bs, n - bs));
    }
    if (ord) EAMD_CHECK(update_row_order(s, ord, k, perm_n, pos, src, bs));
    hipLaunchKernelGGL(add_offset_kernel, dim3(ceil_div(bs, 256)), dim3(256), 0, s, piv_dev + k,
                       bs, k);
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
      EAMD_CHECK(update(A12, A21, A22, m - bs, tn, bs));
    }
  }
  // read back the first zero pivot and the panel exchange error flag
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

int dgetrf_rowmajor(eamd_handle* h, double* W, int64_t ldw, int n, int* piv_dev,
                    int64_t* first_zero, int nb) {
  hipStream_t s = h->stream;
  auto upd = [&](double* A12, double* A21, double* A22, int mrows, int ncols, int k) -> int {
    // row-major A22 -= A21*A12  <=>  column-major A22^T -= A12^T * A21^T
    PhaseScope ps(h, kPhUpdate, 8.0 * (2.0 * mrows * ncols + (double)k * (mrows + ncols)),
                  2.0 * mrows * ncols * k);
    return dgemm_device(s, false, false, ncols, mrows, k, -1.0, A12, ldw, A21, ldw, 1.0, A22,
                        ldw, 0);
  };
  return getrf_rowmajor_t<double>(h, W, ldw, n, piv_dev, nullptr, first_zero, nb, upd);
}

// FP32 LU trailing update in the requested precision (defined in mixed.hip)
int sgemm_update_rowmajor(eamd_handle* h, const float* A12, const float* A21, float* A22,
                          int64_t ld, int mrows, int ncols, int k, int precision);

int sgetrf_rowmajor(eamd_handle* h, float* W, int64_t ldw, int n, int* piv_dev,
                    int64_t* first_zero, int nb, int update_precision) {
  auto upd = [&](float* A12, float* A21, float* A22, int mrows, int ncols, int k) -> int {
    return sgemm_update_rowmajor(h, A12, A21, A22, ldw, mrows, ncols, k, update_precision);
  };
  return getrf_rowmajor_t<float>(h, W, ldw, n, piv_dev, nullptr, first_zero, nb, upd);
}

// The refinement solver's factorization: FP64 (plain) or FP32 with the
// chosen trailing-update precision, recording the final row order.
int getrf_rowmajor_order(eamd_handle* h, void* W, bool fp64, int64_t ldw, int n, int* piv,
                         int* ord, int64_t* first_zero, int nb, int precision) {
  hipStream_t s = h->stream;
  if (fp64) {
    double* Wd = static_cast<double*>(W);
    auto upd = [&](double* A12, double* A21, double* A22, int mrows, int ncols, int k) -> int {
      PhaseScope ps(h, kPhUpdate, 8.0 * (2.0 * mrows * ncols + (double)k * (mrows + ncols)),
                    2.0 * mrows * ncols * k);
      return dgemm_device(s, false, false, ncols, mrows, k, -1.0, A12, ldw, A21, ldw, 1.0, A22,
                          ldw, 0);
    };
    return getrf_rowmajor_t<double>(h, Wd, ldw, n, piv, ord, first_zero, nb, upd);
  }
  float* Wf = static_cast<float*>(W);
  auto upd = [&](float* A12, float* A21, float* A22, int mrows, int ncols, int k) -> int {
    return sgemm_update_rowmajor(h, A12, A21, A22, ldw, mrows, ncols, k, precision);
  };
  return getrf_rowmajor_t<float>(h, Wf, ldw, n, piv, ord, first_zero, nb, upd);
}

}  // namespace eamd
