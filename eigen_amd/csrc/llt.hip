// llt.hip -- blocked right-looking Cholesky (lower) in FP64.
//
// Replaces llt_inplace<Scalar,Lower>::blocked (Eigen/src/Cholesky/LLT.h:330-360)
// and ::unblocked (:304-328).  Per block column of width nb (<= 128):
//   potf2_inv_kernel   the nb x nb diagonal block is factored by one
//                      workgroup in LDS (x <= 0 fails at that column, NaN
//                      passes, as in the reference's
//                      `if (x <= RealScalar(0)) return k`), which also forms
//                      L11^-1;
//   skinny_nt_kernel<0> A21 <- A21 * L11^{-T} as the MFMA product A21 L11^-T
//                      (16 rows per workgroup);
//   skinny_nt_kernel<1> the next block column's A22 -= A21 A21^T (the look-
//                      ahead's critical path), and
//   dgemm (lower)      the rest of A22 -= A21 * A21^T on MFMA, only the tiles
//                      that meet the lower triangle launched (the SYRK
//                      "bottleneck" of LLT.h:357).
// A failure is recorded on the device; later kernels of the same call see
// it and return early, so no host synchronisation per block is needed.
#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kMaxNb = 128;

// 1 / d from the hardware reciprocal and two Newton steps (within an ulp
// of the division for finite d > 0; d = inf gives 0 and NaN gives NaN, as
// the division; d = 0 only occurs after the column has already failed).
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return __builtin_isinf(d) ? 0.0 : r;
}

__device__ __forceinline__ double readlane_f64(double x, int l) {
  const long long bits = __double_as_longlong(x);
  return __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(bits >> 32), l) << 32) |
                              (unsigned)__builtin_amdgcn_readlane((int)bits, l));
}

// One 16 x 16 f64 MFMA tile product over LDS operands, accumulated into
// acc: lane (lk = lane / 16, l16 = lane % 16) supplies F(l16, k) and
// G(k, l16) for k = 4 q + lk, and register r of acc holds D(lk + 4 r, l16)
// (the v_mfma_f64_16x16x4f64 layout).
// kdim is a multiple of 8; two accumulator chains halve the dependent
// MFMA latency.
template <typename FF, typename GG>
__device__ __forceinline__ void tile16(v4d& acc, int kdim, int l16, int lk, FF F, GG G) {
  v4d acc2 = {0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < kdim; q += 8) {
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(F(l16, q + lk), G(q + lk, l16), acc, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(F(l16, q + 4 + lk), G(q + 4 + lk, l16), acc2, 0,
                                                0, 0);
  }
  acc += acc2;
}

// Factor the bs x bs (bs <= 128) lower block at A (column-major, lda) in
// LDS and also form L^-1 there; L goes back to A, L^-1 (bsp x bsp, zero
// above the diagonal, identity on the padding) to Li (column-major, ld
// kMaxNb).  Factorization in 32-column sub-blocks: wave 0 factors the 32 x 32
// diagonal sub-block with one row per lane in registers (each column is
// broadcast through LDS, one division per column: the column is scaled by
// 1 / L_cc), one thread per row below solves against it, and the trailing
// lower part takes the rank-32 update as 16 x 16 MFMA tiles.  The inverse:
// the four 32 x 32 diagonal blocks are inverted by one wave each (lane j
// forms column j), then block row i of L^-1 is X_ij = -X_ii (sum over
// k = j..i-1 of L_ik X_kj), block rows in order, each product on MFMA.
// fail[0] keeps the first failing absolute column (k0 + local) or -1: a
// diagonal x <= 0 fails, a NaN passes, as in the reference's
// `if (x <= RealScalar(0)) return k` (LLT.h:316).
__global__ __launch_bounds__(256) void potf2_inv_kernel(double* __restrict__ A, int64_t lda,
                                                        int bs, int k0, int* __restrict__ fail,
                                                        double* __restrict__ Li) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // column-major, leading dimension a compile-time constant so every LDS
  // offset of the unrolled loops is an immediate
  constexpr int ld = kMaxNb + 1;
  double* S = reinterpret_cast<double*>(smem_raw);  // kMaxNb x ld
  const int bsp = (bs + 31) & ~31;
  double* Tm = S + kMaxNb * ld;  // up to three 32 x 32 temporaries, ld 33
  __shared__ int s_fail;
  __shared__ double s_dinv[kMaxNb];  // 1 / L_cc
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lk = lane >> 4;
  if (*fail >= 0) return;
  auto Sx = [&](int r, int c) -> double& { return S[r + c * ld]; };
  {
    // thread: row tid % 128 of columns tid / 128 + 2 q; every load is
    // requested before the first store (one memory round trip)
    const int i = tid & (kMaxNb - 1), jb = tid >> 7;
    double v[kMaxNb / 2];
#pragma unroll
    for (int q = 0; q < kMaxNb / 2; ++q) {
      const int j = jb + 2 * q;
      v[q] = (i < bs && j < bs) ? (i >= j ? A[i + (int64_t)j * lda] : 0.0) : (i == j ? 1.0 : 0.0);
    }
#pragma unroll
    for (int q = 0; q < kMaxNb / 2; ++q) Sx(i, jb + 2 * q) = v[q];
  }
  if (tid == 0) s_fail = -1;
  __syncthreads();
  for (int kb = 0; kb < bsp; kb += 32) {
    if (wave == 0) {
      const int li = lane & 31, i = kb + li;
      double x[32];
#pragma unroll
      for (int c = 0; c < 32; ++c) x[c] = Sx(i, kb + c);
      int f = -1;
      double myr = 0.0;  // 1 / L_cc of this lane's row c = li
      double diag = readlane_f64(x[0], 0);
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        f = (f < 0 && diag <= 0.0) ? c : f;
        const double d = sqrt(diag), r = recip(d);
        x[c] = li == c ? d : (li > c ? x[c] * r : x[c]);
        myr = li == c ? r : myr;
        if (c < 31) {
          // the next column first (its diagonal is the next step's input),
          // then the rest of the rank-1 update, column c read lane by lane
          // (lanes above the diagonal carry values that are never stored)
          x[c + 1] -= x[c] * readlane_f64(x[c], c + 1);
          diag = readlane_f64(x[c + 1], c + 1);
#pragma unroll
          for (int j = c + 2; j < 32; ++j) x[j] -= x[c] * readlane_f64(x[c], j);
        }
      }
      if (lane < 32) {
        s_dinv[kb + li] = myr;
#pragma unroll
        for (int c = 0; c < 32; ++c)
          if (c <= li) Sx(i, kb + c) = x[c];
      }
      if (lane == 0 && f >= 0) s_fail = kb + f;
    }
    __syncthreads();
    if (s_fail >= 0) break;
    const int below = bsp - kb - 32;
    if (below <= 0) break;
    if (tid < below) {
      const int i = kb + 32 + tid;
      double y[32];
#pragma unroll
      for (int c = 0; c < 32; ++c) y[c] = Sx(i, kb + c);
      // row c of L is read one step ahead (the LDS reads of a step are in
      // flight while the previous step computes); four partial sums, the
      // newest term (q = c - 1) last, so only it waits for y[c - 1]
      double lr[2][32];
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        if (c == 0) {
#pragma unroll
          for (int q = 0; q < 1; ++q) lr[1][q] = Sx(kb + 1, kb + q);
        } else if (c < 31) {
#pragma unroll
          for (int q = 0; q <= c; ++q) lr[(c + 1) & 1][q] = Sx(kb + c + 1, kb + q);
        }
        double v[4] = {y[c], 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < c; ++q) v[(c - 1 - q) & 3] -= y[q] * lr[c & 1][q];
        y[c] = ((v[0] + v[1]) + (v[2] + v[3])) * s_dinv[kb + c];
      }
#pragma unroll
      for (int c = 0; c < 32; ++c) Sx(i, kb + c) = y[c];
    }
    __syncthreads();
    // trailing lower update on 16 x 16 tiles (ti >= tj), one wave per tile
    const int t0 = kb + 32, T16 = below / 16;
    for (int t = wave; t < T16 * (T16 + 1) / 2; t += 4) {
      int ti = 0;
      while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
      const int tj = t - ti * (ti + 1) / 2;
      const int i0 = t0 + 16 * ti, j0 = t0 + 16 * tj;
      v4d acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = Sx(i0 + lk + 4 * r, j0 + l16);
      tile16(acc, 32, l16, lk, [&](int a, int k) { return -Sx(i0 + a, kb + k); },
             [&](int k, int b) { return Sx(j0 + b, kb + k); });
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (ti > tj || lk + 4 * r >= l16) Sx(i0 + lk + 4 * r, j0 + l16) = acc[r];
    }
    __syncthreads();
  }
  if (s_fail >= 0) {
    if (tid == 0) *fail = k0 + s_fail;
    return;
  }
  {
    const int i = tid & (kMaxNb - 1), jb = tid >> 7;
    double v[kMaxNb / 2];
#pragma unroll
    for (int q = 0; q < kMaxNb / 2; ++q) v[q] = Sx(i, jb + 2 * q);
#pragma unroll
    for (int q = 0; q < kMaxNb / 2; ++q) {
      const int j = jb + 2 * q;
      if (i < bs && j < bs && i >= j) A[i + (int64_t)j * lda] = v[q];
    }
  }
  // ---- L^-1: diagonal 32 x 32 blocks, lane j of wave b forms column j of
  // block b (entries above the block's diagonal become zero)
  const int nbk = bsp / 32;
  if (wave < nbk) {
    const int b0 = wave * 32, j = lane & 31;
    double x[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      double sp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < i; ++k) sp[(i - 1 - k) & 3] += Sx(b0 + i, b0 + k) * x[k];
      const double sacc = (sp[0] + sp[1]) + (sp[2] + sp[3]);
      const double di = s_dinv[b0 + i];
      x[i] = i < j ? 0.0 : (i == j ? di : -sacc * di);
      // keeps the next row's LDS reads from being hoisted (register bound)
      asm volatile("" ::: "memory");
    }
    asm volatile("" ::: "memory");
    if (lane < 32) {
#pragma unroll
      for (int i = 0; i < 32; ++i) Sx(b0 + i, b0 + j) = x[i];
    }
  }
  __syncthreads();
  // ---- off-diagonal blocks, block row by block row
  for (int bi = 1; bi < nbk; ++bi) {
    // T_j = sum_{k=j}^{bi-1} L_{bi,k} X_{k,j}, j < bi: 4 tiles per block
    for (int t = wave; t < 4 * bi; t += 4) {
      const int bj = t >> 2, ti = (t >> 1) & 1, tj = t & 1;
      const int r0 = 32 * bi + 16 * ti, c0 = 32 * bj + 16 * tj, kk0 = 32 * bj;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      tile16(acc, 32 * bi - kk0, l16, lk, [&](int a, int k) { return Sx(r0 + a, kk0 + k); },
             [&](int k, int b) { return Sx(kk0 + k, c0 + b); });
      double* T = Tm + bj * 32 * 33;
#pragma unroll
      for (int r = 0; r < 4; ++r) T[(16 * ti + lk + 4 * r) + (16 * tj + l16) * 33] = acc[r];
    }
    __syncthreads();
    // X_{bi,j} = -X_{bi,bi} T_j
    for (int t = wave; t < 4 * bi; t += 4) {
      const int bj = t >> 2, ti = (t >> 1) & 1, tj = t & 1;
      const double* T = Tm + bj * 32 * 33;
      const int d0 = 32 * bi;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      tile16(acc, 32, l16, lk, [&](int a, int k) { return -Sx(d0 + 16 * ti + a, d0 + k); },
             [&](int k, int b) { return T[k + (16 * tj + b) * 33]; });
#pragma unroll
      for (int r = 0; r < 4; ++r) Sx(d0 + 16 * ti + lk + 4 * r, 32 * bj + 16 * tj + l16) = acc[r];
    }
    __syncthreads();
  }
  {
    const int i = tid & (kMaxNb - 1), jb = tid >> 7;
    double v[kMaxNb / 2];
#pragma unroll
    for (int q = 0; q < kMaxNb / 2; ++q) v[q] = i >= jb + 2 * q ? Sx(i, jb + 2 * q) : 0.0;
#pragma unroll
    for (int q = 0; q < kMaxNb / 2; ++q) Li[i + (jb + 2 * q) * kMaxNb] = v[q];
  }
}

// Products of a 16-row block of an m-row panel with the transpose of a
// K x nc block B (K, nc <= 128), one workgroup per 16 rows:
//   MODE 0, the panel TRSM X <- X L^{-T}: out(r, c) = sum_k X(r, k) Li(c, k)
//     with Li = L^{-1} from potf2_inv_kernel (ld kMaxNb, zero above the
//     diagonal), written over X;
//   MODE 1, the next block column's update A22 -= A21 A21^T:
//     C(r, c) -= sum_k A(r, k) A(c, k) on the lower part (r >= c) only.
// The workgroup's 16 x K rows are staged in LDS in one round trip; wave w
// computes columns [32 w, 32 w + 32) as two MFMA tiles with B read from
// memory (L2), consecutive lanes on consecutive rows.  17 KB of LDS and
// under 100 VGPRs, so a workgroup fits on a CU beside two GEMM workgroups
// of the concurrent trailing update.
template <int MODE>
__global__ __launch_bounds__(256) void skinny_nt_kernel(const double* __restrict__ B, int64_t ldb,
                                                        int K, const double* __restrict__ Ain,
                                                        int64_t lda, double* __restrict__ C,
                                                        int64_t ldc, int m, int nc,
                                                        const int* __restrict__ fail,
                                                        double* __restrict__ Bc) {
  // the factorization's chain, beside the SYRK's waves: issue priority
  __builtin_amdgcn_s_setprio(3);
  constexpr int RB = 16, LDA_S = RB + 1;
  __shared__ double As[kMaxNb * LDA_S];  // As[k * 17 + r]
  if (*fail >= 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * RB, c0 = wave * 32;
  {
    // thread: row tid % 16 of columns tid / 16 + 16 q
    const int r = tid & (RB - 1), kb = tid >> 4;
    const bool rin = r0 + r < m;
    double v[kMaxNb / 16];
#pragma unroll
    for (int q = 0; q < kMaxNb / 16; ++q) {
      const int k = kb + 16 * q;
      v[q] = (rin && k < K) ? Ain[(r0 + r) + (int64_t)k * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kMaxNb / 16; ++q) As[(kb + 16 * q) * LDA_S + r] = v[q];
  }
  const bool active = c0 < nc;
  const int row = r0 + l16;
  v4d acc[2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      // MODE 1 starts from -C (the result is -acc)
      const int col = c0 + 16 * u + lk + 4 * e;
      acc[u][e] = (MODE == 1 && active && row < m && col < nc) ? -C[row + (int64_t)col * ldc]
                                                                 : 0.0;
    }
  __syncthreads();
  if (!active) return;
  const double* b0p = B + min(c0 + l16, nc - 1);
  const double* b1p = B + min(c0 + 16 + l16, nc - 1);
  if constexpr (MODE == 0) {
    // every B value of the wave (2 x K / 4 per lane, from the 128 KB L^-1)
    // requested at once: one memory round trip (14 vs 19 us per launch)
    double bv[2][kMaxNb / 4];
#pragma unroll
    for (int s = 0; s < kMaxNb / 4; ++s) {
      const int k = 4 * s + lk;
      const bool kin = k < K;
      bv[0][s] = kin ? b0p[(int64_t)k * ldb] : 0.0;
      bv[1][s] = kin ? b1p[(int64_t)k * ldb] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kMaxNb / 4; ++s) {
      const double a0 = As[(4 * s + lk) * LDA_S + l16];
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[0][s], a0, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[1][s], a0, acc[1], 0, 0, 0);
    }
  } else {
    // K in chunks of 32 (eight MFMA steps, their 16 loads issued together):
    // fewer registers, so the workgroup still fits beside the concurrent
    // SYRK's (the all-at-once form ran slower here, 52 vs 45 us)
    for (int q0 = 0; q0 < K; q0 += 32) {
#pragma unroll
      for (int q = q0; q < q0 + 32; q += 4) {
        const int k = q + lk;
        const bool kin = k < K;
        const double b0 = kin ? b0p[(int64_t)k * ldb] : 0.0;
        const double b1 = kin ? b1p[(int64_t)k * ldb] : 0.0;
        const double a0 = As[k * LDA_S + l16];
        acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[1], 0, 0, 0);
      }
    }
  }
  // acc[u][e] = out(r0 + l16, c0 + 16 u + lk + 4 e)
  if (row >= m) return;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int col = c0 + 16 * u + lk + 4 * e;
      if (col >= nc || (MODE == 1 && row < col)) continue;
      C[row + (int64_t)col * ldc] = MODE == 1 ? -acc[u][e] : acc[u][e];
      // MODE 0: the first kMaxNb rows of the result also go to the compact
      // Bc (ld kMaxNb), the B operand of the next block column's update (a
      // 128 KB block instead of 128 columns of the matrix)
      if (MODE == 0 && row < kMaxNb) Bc[row + col * kMaxNb] = acc[u][e];
    }
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

}  // namespace

namespace {
// LLT rank-one update, L L^T + sigma v v^T = L' L'^T, in place: one wave per
// system, row i of the factor on lane i % 64 (up to 4 rows per lane, n <=
// 256).  Restates internal::llt_rank_update_lower (LLT.h:231-299): for
// sigma > 0 a Givens rotation per column (JacobiRotation::makeGivens for
// reals, Jacobi.h:226-262, applied as x' = c x + s y, y' = -s x + c y,
// Jacobi.h:315-330); otherwise the scalar downdate, stopping at the first
// column whose new diagonal would not be positive.  Column j's rows are read
// and written by their own lanes in step j only, and the diagonal element
// (j, j) is read by all lanes before lane j % 64 writes it, so the steps
// need no other synchronisation.  info: -1, the failing column, or -2 for a
// size outside 0..256.
constexpr int kRankUpdMaxN = 256;
__global__ __launch_bounds__(256) void llt_rank_update_kernel(
    int64_t count, const int64_t* __restrict__ offL, const int32_t* __restrict__ nv,
    double* __restrict__ Lall, const int64_t* __restrict__ offv, const double* __restrict__ V,
    const double* __restrict__ sigmas, int upper, int32_t* __restrict__ info) {
  constexpr int K = kRankUpdMaxN / 64;
  const int lane = threadIdx.x & 63;
  const int64_t s = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  if (s >= count) return;  // whole waves
  const int n = nv[s];
  if (n < 0 || n > kRankUpdMaxN) {
    if (lane == 0) info[s] = -2;
    return;
  }
  double* L = Lall + offL[s];
  const double* v = V + offv[s];
  const double sigma = sigmas ? sigmas[s] : 1.0;
  // element (i, j) of the lower factor (the upper one stores its transpose)
  auto at = [&](int i, int j) -> double& {
    return upper ? L[j + (int64_t)i * n] : L[i + (int64_t)j * n];
  };
  const double ss = sigma > 0.0 ? sqrt(sigma) : 1.0;
  double t[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = lane + 64 * k;
    t[k] = i < n ? (sigma > 0.0 ? ss * v[i] : v[i]) : 0.0;
  }
  int ret = -1;
  double beta = 1.0;
  for (int j = 0; j < n; ++j) {
    const int jk = j >> 6, jl = j & 63;
    const double tsel = jk == 0 ? t[0] : (jk == 1 ? t[1] : (jk == 2 ? t[2] : t[3]));
    const long long tb = __double_as_longlong(tsel);
    const double tj = __longlong_as_double(
        ((long long)__builtin_amdgcn_readlane((int)(tb >> 32), jl) << 32) |
        (unsigned)__builtin_amdgcn_readlane((int)tb, jl));
    const double p = at(j, j);
    if (sigma > 0.0) {
      const double q = -tj;
      double c, sn, r;
      if (q == 0.0) {
        c = p < 0.0 ? -1.0 : 1.0;
        sn = 0.0;
        r = fabs(p);
      } else if (p == 0.0) {
        c = 0.0;
        sn = q < 0.0 ? 1.0 : -1.0;
        r = fabs(q);
      } else if (fabs(p) > fabs(q)) {
        const double tt = q / p;
        double u = sqrt(1.0 + tt * tt);
        if (p < 0.0) u = -u;
        c = 1.0 / u;
        sn = -tt * c;
        r = p * u;
      } else {
        const double tt = p / q;
        double u = sqrt(1.0 + tt * tt);
        if (q < 0.0) u = -u;
        sn = -1.0 / u;
        c = -tt * sn;
        r = q * u;
      }
      if (lane == jl) at(j, j) = r;
      // the identity rotation leaves the column and the work vector as they
      // are (Jacobi.h:460-461), also where they hold non-finite values
      if (c == 1.0 && sn == 0.0) continue;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int i = lane + 64 * k;
        if (i > j && i < n) {
          const double x = at(i, j), y = t[k];
          at(i, j) = c * x + sn * y;
          t[k] = -sn * x + c * y;
        }
      }
    } else {
      const double Ljj = p, dj = Ljj * Ljj, wj = tj;
      const double swj2 = sigma * wj * wj, gamma = dj * beta + swj2;
      const double x = dj + swj2 / beta;
      if (x <= 0.0) {
        ret = j;
        break;
      }
      const double nLjj = sqrt(x);
      if (lane == jl) at(j, j) = nLjj;
      beta += swj2 / dj;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int i = lane + 64 * k;
        if (i > j && i < n) {
          const double lij = at(i, j);
          t[k] -= (wj / Ljj) * lij;
          if (gamma != 0.0) at(i, j) = (nLjj / Ljj) * lij + (nLjj * sigma * wj / gamma) * t[k];
        }
      }
    }
  }
  if (lane == 0) info[s] = ret;
}
}  // namespace

int llt_rank_update_batched(eamd_handle* h, int64_t count, const int64_t* offL, const int32_t* n,
                            double* L, const int64_t* offv, const double* v, const double* sigma,
                            int upper, int32_t* info) {
  if (count <= 0) return kOk;
  const int64_t blocks = (count + 3) / 4;  // four waves (systems) per workgroup
  if (blocks > INT32_MAX) {
    set_error("rank update: %lld systems exceed one launch", (long long)count);
    return kErrUnsupported;
  }
  hipLaunchKernelGGL(llt_rank_update_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                     count, offL, n, L, offv, v, sigma, upper, info);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int dpotrf_lower(eamd_handle* h, int n, double* A, int64_t lda, int64_t* fail_col, int nb) {
  hipStream_t s = h->stream;
  if (nb > kMaxNb) nb = kMaxNb;
  if (nb < 8) nb = 8;
  void* aux = nullptr;
  EAMD_CHECK(h->ws.get(kWsAux, 256, &aux));
  int* fail = static_cast<int*>(aux);
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, s, fail, -1);
  static bool attr = false;
  if (!attr) {
    EAMD_HIP(hipFuncSetAttribute((const void*)potf2_inv_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024));
    attr = true;
  }
  void* lip = nullptr;
  EAMD_CHECK(h->ws.get(kWsInv, 2 * sizeof(double) * kMaxNb * kMaxNb, &lip));
  double* Li = static_cast<double*>(lip);       // L11^-1 of the current block column
  double* Bc = Li + kMaxNb * kMaxNb;            // its first 128 rows of A21 L11^-T
  // One-panel look-ahead on two streams (no pivoting, so nothing waits on
  // a grid-wide exchange): the handle's stream factors panel k (POTF2 of the
  // diagonal block, then A21 L11^-T), updates the next panel's columns only
  // (rows below the panel, lower part of their diagonal block) and goes on
  // to panel k+1; the second stream, started by an event after the next
  // panel's columns are updated, does the SYRK of the rest of the trailing
  // matrix.  The handle's stream waits for it only before updating the next
  // panel's columns again.  The SYRK starts after that column update (not
  // straight after panel k) so that panel k+1's POTF2, whose 157 KB of LDS
  // needs a CU of its own, is dispatched before the SYRK's workgroups (two
  // of 65 KB per CU) occupy every CU; started earlier, the SYRK made each
  // POTF2 wait for its tail (13.9 vs 13.4 ms for n = 8192).
  static const bool lookahead = env_int("EAMD_LOOKAHEAD", 1) != 0;
  hipStream_t sb = nullptr;
  if (lookahead && n > nb) {
    EAMD_CHECK(ensure_side_stream(h));
    sb = h->side;
  }
  hipEvent_t evP = h->la_ev[0], evB[2] = {h->la_ev[1], h->la_ev[2]};
  bool side_used = false;
  for (int k = 0, it = 0; k < n; k += nb, ++it) {
    const int bs = min(nb, n - k);
    const int rs = n - k - bs;
    double* A11 = A + k + (int64_t)k * lda;
    const int bsp = (bs + 31) & ~31;
    const size_t lds = sizeof(double) * (kMaxNb * (kMaxNb + 1) + 3 * 32 * 33);
    hipLaunchKernelGGL(potf2_inv_kernel, dim3(1), dim3(256), lds, s, A11, lda, bs, k, fail, Li);
    if (rs > 0) {
      double* A21 = A11 + bs;
      hipLaunchKernelGGL(skinny_nt_kernel<0>, dim3(ceil_div(rs, 16)), dim3(256), 0, s,
                         (const double*)Li, (int64_t)kMaxNb, bs, (const double*)A21, lda, A21,
                         lda, rs, bs, (const int*)fail, Bc);
      double* A22 = A21 + (int64_t)bs * lda;
      if (!sb) {
        EAMD_CHECK(dgemm_device(s, false, true, rs, rs, bs, -1.0, A21, lda, A21, lda, 1.0, A22,
                                lda, 1, fail));
        continue;
      }
      const int par = it & 1, n1 = min(nb, rs), n2 = rs - n1;
      // the next panel's columns were last written by the previous panel's
      // second-stream update
      if (it > 0) EAMD_HIP(hipStreamWaitEvent(s, evB[par ^ 1], 0));
      hipLaunchKernelGGL(skinny_nt_kernel<1>, dim3(ceil_div(rs, 16)), dim3(256), 0, s,
                         (const double*)Bc, (int64_t)kMaxNb, bs, (const double*)A21, lda, A22,
                         lda, rs, n1, (const int*)fail, (double*)nullptr);
      EAMD_HIP(hipEventRecord(evP, s));  // panel factored, next panel's columns updated
      if (n2 > 0) {
        EAMD_HIP(hipStreamWaitEvent(sb, evP, 0));
        side_used = true;
        EAMD_CHECK(dgemm_device(sb, false, true, n2, n2, bs, -1.0, A21 + n1, lda, A21 + n1, lda,
                                1.0, A22 + n1 + (int64_t)n1 * lda, lda, 1, fail));
        EAMD_HIP(hipEventRecord(evB[par], sb));
      } else {
        // nothing for the second stream: the next panel has no later wait
        EAMD_HIP(hipEventRecord(evB[par], s));
      }
    }
  }
  if (side_used) {
    EAMD_HIP(hipEventRecord(h->la_ev[3], sb));
    EAMD_HIP(hipStreamWaitEvent(s, h->la_ev[3], 0));
  }
  EAMD_HIP(hipGetLastError());
  int hf = -1;
  EAMD_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, s));
  EAMD_HIP(hipStreamSynchronize(s));
  *fail_col = hf;
  return kOk;
}

}  // namespace eamd
