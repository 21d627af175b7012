// llt.hip -- blocked right-looking Cholesky (lower) in FP64.
//
// Replaces llt_inplace<Scalar,Lower>::blocked (Eigen/src/Cholesky/LLT.h:330-360)
// and ::unblocked (:304-328).  Per block column of width nb:
//   potf2_kernel   the nb x nb diagonal block is factored by one workgroup in
//                  LDS (x <= 0 fails at that column, NaN passes, as in the
//                  reference's `if (x <= RealScalar(0)) return k`);
//   trsm_rlt_kernel A21 <- A21 * L11^{-T}, one workgroup per 32 rows;
//   dgemm (lower)  A22 -= A21 * A21^T on MFMA, tiles above the diagonal
//                  skipped (the SYRK "bottleneck" of LLT.h:357).
// A failure is recorded on the device; later kernels of the same call see
// it and return early, so no host synchronisation per block is needed.
#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kMaxNb = 128;

// Factor the bs x bs (bs <= 128) lower block at A (column-major, lda) in
// LDS, in 32-column sub-blocks: wave 0 factors the 32 x 32 diagonal
// sub-block with one row per lane in registers (broadcasts by readlane, no
// barrier), the rows below solve against it one row per thread, and all
// threads update the trailing lower part.  fail[0] keeps the first failing
// absolute column (k0 + local) or -1: a diagonal x <= 0 fails, a NaN passes,
// as in the reference's `if (x <= RealScalar(0)) return k` (LLT.h:316).
__global__ __launch_bounds__(256) void potf2_kernel(double* __restrict__ A, int64_t lda, int bs,
                                                    int k0, int* __restrict__ fail) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* S = reinterpret_cast<double*>(smem_raw);  // bsp x (bsp+1), column-major, padded
  const int bsp = (bs + 31) & ~31;
  const int ld = bsp + 1;
  __shared__ int s_fail;
  __shared__ double s_dinv[32];  // 1 / L_cc of the current diagonal sub-block
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (*fail >= 0) return;
  // padding: identity on the diagonal, zero elsewhere (factors to itself)
  for (int e = tid; e < bsp * bsp; e += 256) {
    const int j = e / bsp, i = e - j * bsp;
    double v = 0.0;
    if (i < bs && j < bs) v = (i >= j) ? A[i + (int64_t)j * lda] : 0.0;
    else if (i == j) v = 1.0;
    S[i + j * ld] = v;
  }
  if (tid == 0) s_fail = -1;
  __syncthreads();
  for (int kb = 0; kb < bsp; kb += 32) {
    // ---- 32 x 32 diagonal sub-block on wave 0, row kb+lane in x[]
    if (wave == 0) {
      double x[32];
      const int i = kb + (lane & 31);
#pragma unroll
      for (int c = 0; c < 32; ++c) x[c] = S[i + (kb + c) * ld];
      int f = -1;
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        const double diag = __longlong_as_double(
            ((long long)__builtin_amdgcn_readlane((int)(__double_as_longlong(x[c]) >> 32), c) << 32) |
            (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(x[c]), c));
        if (f < 0 && diag <= 0.0) f = c;
        const double d = sqrt(diag);
        x[c] = (lane & 31) == c ? d : ((lane & 31) > c ? x[c] / d : x[c]);
#pragma unroll
        for (int j = c + 1; j < 32; ++j) {
          const long long bits = __double_as_longlong(x[c]);
          const double ljc = __longlong_as_double(
              ((long long)__builtin_amdgcn_readlane((int)(bits >> 32), j) << 32) |
              (unsigned)__builtin_amdgcn_readlane((int)bits, j));
          if ((lane & 31) >= j) x[j] -= x[c] * ljc;
        }
      }
      if (lane < 32) {
#pragma unroll
        for (int c = 0; c < 32; ++c)
          if (c <= lane) S[i + (kb + c) * ld] = x[c];
#pragma unroll
        for (int c = 0; c < 32; ++c)
          if (c == lane) s_dinv[c] = 1.0 / x[c];
      }
      if (lane == 0 && f >= 0) s_fail = kb + f;
    }
    __syncthreads();
    if (s_fail >= 0) break;
    const int below = bsp - kb - 32;
    if (below <= 0) break;
    // ---- rows below: y L11^T = s, one row per thread
    if (tid < below) {
      const int i = kb + 32 + tid;
      double y[32];
#pragma unroll
      for (int c = 0; c < 32; ++c) y[c] = S[i + (kb + c) * ld];
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        double v = y[c];
#pragma unroll
        for (int q = 0; q < c; ++q) v -= y[q] * S[kb + c + (kb + q) * ld];
        y[c] = v * s_dinv[c];
      }
#pragma unroll
      for (int c = 0; c < 32; ++c) S[i + (kb + c) * ld] = y[c];
    }
    __syncthreads();
    // ---- trailing lower update: S[i][j] -= sum_q S[i][kb+q] S[j][kb+q], in
    // 4 x 4 blocks (bi >= bj) so each loaded value serves four products
    const int t0 = kb + 32, nb4 = below / 4;
    for (int e = tid; e < nb4 * (nb4 + 1) / 2; e += 256) {
      // e -> (bi, bj) with bi >= bj: bi(bi+1)/2 <= e < (bi+1)(bi+2)/2
      int bi = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
      if ((bi + 1) * (bi + 2) / 2 <= e) ++bi;
      if (bi * (bi + 1) / 2 > e) --bi;
      const int bj = e - bi * (bi + 1) / 2;
      const int i0 = t0 + 4 * bi, j0 = t0 + 4 * bj;
      double acc[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
#pragma unroll 4
      for (int q = 0; q < 32; ++q) {
        const double* col = S + (kb + q) * ld;
        double a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          a[r] = col[i0 + r];
          b[r] = col[j0 + r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (bi > bj || r >= c) S[i0 + r + (j0 + c) * ld] -= acc[r][c];
    }
    __syncthreads();
  }
  if (s_fail >= 0) {
    if (tid == 0) *fail = k0 + s_fail;
    return;
  }
  for (int e = tid; e < bs * bs; e += 256) {
    const int j = e / bs, i = e - j * bs;
    if (i >= j) A[i + (int64_t)j * lda] = S[i + j * ld];
  }
}

// X (m x bs, column-major) <- X * L^{-T}, L lower bs x bs (column-major):
// row r of X solves y L^T = x_r.  One workgroup per 32 rows, X block in LDS;
// per 32-column block of L: one thread per row solves the block from
// registers (L broadcast from LDS), then 8 threads per row update the later
// columns.
__global__ __launch_bounds__(256) void trsm_rlt_kernel(const double* __restrict__ L, int64_t ldl,
                                                       int bs, double* __restrict__ X,
                                                       int64_t ldx, int m,
                                                       const int* __restrict__ fail) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int RB = 32;
  const int bsp = (bs + 31) & ~31;
  double* Xs = reinterpret_cast<double*>(smem_raw);  // [col][row], bsp x (RB+1)
  double* Lp = Xs + (size_t)bsp * (RB + 1);          // [row j][q], bsp x 33
  const int tid = threadIdx.x;
  if (*fail >= 0) return;
  const int r0 = blockIdx.x * RB;
  const int rr = min(RB, m - r0);
  const int r = tid & (RB - 1), qd = tid >> 5;  // 8 threads per row
  for (int e = tid; e < RB * bsp; e += 256) {
    const int c = e / RB, rw = e - c * RB;
    Xs[rw + c * (RB + 1)] = (rw < rr && c < bs) ? X[(r0 + rw) + (int64_t)c * ldx] : 0.0;
  }
  for (int b = 0; b < bsp; b += 32) {
    const int w = min(32, bs - b);
    const int rest = bsp - b;  // rows j of the L panel (b .. bsp-1)
    __syncthreads();
    for (int e = tid; e < rest * 32; e += 256) {
      const int j = e >> 5, q = e & 31;  // Lp[j][q] = L(b + j, b + q); padding: identity
      double v = 0.0;
      if (b + j < bs && q < w) v = L[(b + j) + (int64_t)(b + q) * ldl];
      else if (j == q) v = 1.0;
      Lp[j * 33 + q] = v;
    }
    __syncthreads();
    if (qd == 0) {
      double y[32];
#pragma unroll
      for (int c = 0; c < 32; ++c) y[c] = Xs[r + (b + c) * (RB + 1)];
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        double v = y[c];
#pragma unroll
        for (int q = 0; q < c; ++q) v -= y[q] * Lp[c * 33 + q];
        y[c] = v / Lp[c * 33 + c];
      }
#pragma unroll
      for (int c = 0; c < 32; ++c) Xs[r + (b + c) * (RB + 1)] = y[c];
    }
    __syncthreads();
    for (int j = 32 + qd; j < rest; j += 8) {
      double acc = 0.0;
#pragma unroll 8
      for (int q = 0; q < 32; ++q) acc += Xs[r + (b + q) * (RB + 1)] * Lp[j * 33 + q];
      Xs[r + (b + j) * (RB + 1)] -= acc;
    }
  }
  __syncthreads();
  for (int e = tid; e < RB * bs; e += 256) {
    const int c = e / RB, rw = e - c * RB;
    if (rw < rr) X[(r0 + rw) + (int64_t)c * ldx] = Xs[rw + c * (RB + 1)];
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
    EAMD_HIP(hipFuncSetAttribute((const void*)potf2_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    EAMD_HIP(hipFuncSetAttribute((const void*)trsm_rlt_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr = true;
  }
  // One-panel look-ahead on two streams (no pivoting, so nothing waits on
  // a grid-wide exchange): the handle's stream factors panel k (POTF2 of the
  // diagonal block, then A21 L11^-T), updates the next panel's columns only
  // (rows below the panel, lower part of their diagonal block) and goes on
  // to panel k+1; the second stream, started by an event after panel k,
  // does the SYRK of the rest of the trailing matrix.  The handle's stream
  // waits for it only before updating the next panel's columns again.
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
    size_t lds = sizeof(double) * (size_t)bsp * (bsp + 1);
    hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(256), lds, s, A11, lda, bs, k, fail);
    if (rs > 0) {
      double* A21 = A11 + bs;
      size_t lds2 = sizeof(double) * (size_t)bsp * (32 + 1) * 2;
      hipLaunchKernelGGL(trsm_rlt_kernel, dim3(ceil_div(rs, 32)), dim3(256), lds2, s, A11, lda,
                         bs, A21, lda, rs, fail);
      double* A22 = A21 + (int64_t)bs * lda;
      if (!sb) {
        EAMD_CHECK(dgemm_device(s, false, true, rs, rs, bs, -1.0, A21, lda, A21, lda, 1.0, A22,
                                lda, 1, fail));
        continue;
      }
      const int par = it & 1, n1 = min(nb, rs), n2 = rs - n1;
      EAMD_HIP(hipEventRecord(evP, s));  // the panel is factored
      // the next panel's columns were last written by the previous panel's
      // second-stream update
      if (it > 0) EAMD_HIP(hipStreamWaitEvent(s, evB[par ^ 1], 0));
      EAMD_CHECK(dgemm_device(s, false, true, rs, n1, bs, -1.0, A21, lda, A21, lda, 1.0, A22,
                              lda, 1, fail));
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
