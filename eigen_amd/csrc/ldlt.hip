// ldlt.hip -- batched LDLT (symmetric diagonal pivoting) of small FP64
// matrices, and the solve with its factors.
//
// Per system this is LDLT<MatrixXd, Lower>::compute (Eigen/src/Cholesky/
// LDLT.h:491-522, through internal::ldlt_inplace<Lower>::unblocked,
// :290-395) and ::solve (:559-593).  The factorization is the reference's
// left-looking column loop, step k:
//   p = first index of the largest |a_ii| over i >= k (maxCoeff's strict '>'
//       scan: a NaN at i = k wins, NaNs elsewhere never do);
//   the lower-only symmetric exchange of k and p;
//   temp_j = d_j l_kj (j < k); a_ik -= sum_j l_ij temp_j (i >= k);
//   l_ik = a_ik / a_kk (i > k) when a_kk != 0, else the column must be zero;
//   the sign state machine and the NumericalIssue rules of the reference.
//
// MI355X mapping.  A group of W lanes per system (W = 16, 32 or 64, the
// smallest power of two >= the batch's largest n; 64 / W systems per
// wavefront): the lower triangle is staged in LDS (column stride n + 1),
// lane i of the group owns row i, so the pivot search is one butterfly
// max inside the group plus a ballot for the first lane holding it, the
// exchange is a set of disjoint element swaps done by all lanes at once,
// and the column update is one length-k dot product per lane against temp
// held in LDS.  The solve uses the same groups with x_i in lane i: each
// permutation is one gather (every lane follows its position through the
// transpositions), the substitutions exchange values by lane shuffles
// inside the group, and each lane's operands are loaded before the sweeps
// (W <= 32: row i and column i of L in registers).
// No host synchronisation: both launches can be captured in a HIP graph.
#include "../../include/eigen_amd.h"
#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kLdltMaxN = 64;

// info: 0 Success, 1 NumericalIssue (ComputationInfo's values), -2 for a
// size outside 0..max_n; sign: 0 PositiveSemiDef, 1 NegativeSemiDef,
// 2 ZeroSign, 3 Indefinite (internal::SignMatrix, LDLT.h:22).
// W lanes per system (W >= max_n, a power of two), 64 / W systems per
// wavefront; the step loop runs to the largest size in the wavefront so
// every barrier is reached by all lanes, steps past a system's own size
// doing nothing for it.
template <int W>
__global__ __launch_bounds__(64) void ldlt_batch_kernel(int64_t count,
                                                        const int64_t* __restrict__ offA,
                                                        const int32_t* __restrict__ nv, int max_n,
                                                        double* __restrict__ Aall,
                                                        const int64_t* __restrict__ offT,
                                                        int32_t* __restrict__ Tall,
                                                        int32_t* __restrict__ sign_out,
                                                        int32_t* __restrict__ info) {
  constexpr int G = 64 / W;  // systems per wavefront
  // LDS per system sized by the batch's largest n, not by W: the factor is
  // latency-bound, and LDS is what limits the workgroups resident per CU
  extern __shared__ double s_lds[];
  const int stride = max_n * (max_n + 1) + max_n;  // matrix, then temp
  const int lane = threadIdx.x;
  const int sub = lane / W, sl = lane % W;
  const int64_t s = (int64_t)blockIdx.x * G + sub;
  const bool act = s < count;
  int n = act ? nv[s] : -1;
  if (act && (n < 0 || n > max_n || n > W)) {
    if (sl == 0) info[s] = -2;
    n = -1;
  }
  double* a = n >= 0 ? Aall + offA[s] : nullptr;
  int32_t* tr = n >= 0 ? Tall + offT[s] : nullptr;
  if ((n == 0 || n == 1) && sl == 0) {
    int sg = 2;
    if (n == 1) {
      tr[0] = 0;
      const double d = a[0];
      sg = d > 0.0 ? 0 : (d < 0.0 ? 1 : 2);
    }
    sign_out[s] = sg;
    info[s] = 0;
  }
  const int nl = n > 1 ? n : 0;  // steps of this system
  int nmax = nl;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) nmax = max(nmax, __shfl_xor(nmax, off));
  const int ld = nl + 1;
  double* M = s_lds + sub * stride;  // nl x nl, column stride ld
  double* temp = M + max_n * (max_n + 1);
  const unsigned long long segmask =
      W == 64 ? ~0ull : (((1ull << (W & 63)) - 1ull) << (sub * W));
  // the matrix as one contiguous run of nl * nl values, eight loads per
  // lane in flight before any is stored (the lower triangle is kept)
  const int ne = nl * nl;
  for (int e0 = 0; e0 < ne; e0 += 8 * W) {
    double vals[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * W + sl;
      vals[u] = e < ne ? a[e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * W + sl;
      if (e < ne) {
        const int j = e / nl, i = e - j * nl;
        if (i >= j) M[i + j * ld] = vals[u];
      }
    }
  }
  __syncthreads();

  int sg = 2;
  bool ret = true, found_zero = false, stop = false, zero_exit = false;
  for (int k = 0; k < nmax; ++k) {
    const bool on = !stop && k < nl;  // the same for the W lanes of a system
    // ---- pivot: the largest |a_ii| over the system's lanes, lower index on ties
    const int ci = k + sl;
    double v = -1.0;
    if (on && ci < nl) {
      v = fabs(M[ci + ci * ld]);
      if (v != v) v = sl == 0 ? __builtin_inf() : -1.0;
    }
    // the group's largest value by a butterfly max, then the lowest lane
    // holding it from one ballot (invalid lanes hold -1, below any |a_ii|)
    double vmax = v;
#pragma unroll
    for (int off = W / 2; off >= 1; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off));
    const unsigned long long hit = __ballot(v == vmax) & segmask;
    const int p = k + (hit ? __builtin_ctzll(hit) - sub * W : 0);
    if (on && sl == 0) tr[k] = p;
    if (on && p != k) {
      // the lower-only exchange (LDLT.h:321-337); the four sets of
      // elements are disjoint, so every lane swaps its own at once
      if (sl < k) {
        const double t = M[k + sl * ld];
        M[k + sl * ld] = M[p + sl * ld];
        M[p + sl * ld] = t;
      }
      const int it = p + 1 + sl;
      if (it < nl) {
        const double t = M[it + k * ld];
        M[it + k * ld] = M[it + p * ld];
        M[it + p * ld] = t;
      }
      if (sl == 0) {
        const double t = M[k + k * ld];
        M[k + k * ld] = M[p + p * ld];
        M[p + p * ld] = t;
      }
      const int im = k + 1 + sl;
      if (im < p) {
        const double t = M[im + k * ld];
        M[im + k * ld] = M[p + im * ld];
        M[p + im * ld] = t;
      }
    }
    __syncthreads();
    // ---- column k: a_ik -= sum_j l_ij d_j l_kj (LDLT.h:348-354)
    if (k > 0) {
      if (on && sl < k) temp[sl] = M[sl + sl * ld] * M[k + sl * ld];
      __syncthreads();
      if (on && ci < nl) {
        double r = 0.0;
#pragma unroll 4
        for (int j = 0; j < k; ++j) r += M[ci + j * ld] * temp[j];  // loads run ahead, order kept
        M[ci + k * ld] -= r;
      }
      __syncthreads();
    }
    const double akk = on ? M[k + k * ld] : 1.0;
    const bool valid = fabs(akk) > 0.0;
    const bool zx = on && k == 0 && !valid;  // the whole diagonal is zero
    bool nz = false;
    if (zx) {
      for (int j = 0; j < nl; ++j)
        if (sl > j && sl < nl && !(M[sl + j * ld] == 0.0)) nz = true;
    } else if (on) {
      const int ib = k + 1 + sl;
      if (ib < nl) {
        if (valid)
          M[ib + k * ld] /= akk;
        else
          nz = !(M[ib + k * ld] == 0.0);
      }
    }
    const bool anynz = (__ballot(nz) & segmask) != 0ull;
    if (zx) {
      // LDLT.h:363-374: identity transpositions, success iff L is zero
      ret = !anynz;
      if (sl < nl) tr[sl] = sl;
      sg = 2;
      stop = true;
      zero_exit = true;
    } else if (on) {
      if (anynz) ret = false;
      if (found_zero && valid)
        ret = false;
      else if (!valid)
        found_zero = true;
      if (sg == 0) {
        if (akk < 0.0) sg = 3;
      } else if (sg == 1) {
        if (akk > 0.0) sg = 3;
      } else if (sg == 2) {
        if (akk > 0.0)
          sg = 0;
        else if (akk < 0.0)
          sg = 1;
      }
    }
    __syncthreads();
  }
  if (nl > 0) {
    if (!zero_exit)
      for (int j = 0; j < nl; ++j)
        if (sl >= j && sl < nl) a[sl + (int64_t)j * nl] = M[sl + j * ld];
    if (sl == 0) {
      sign_out[s] = sg;
      info[s] = ret ? 0 : 1;
    }
  }
}

// x <- A^{-1} b for each system from ldlt_batch_kernel's factors; W lanes
// per system (x_i in lane i of its group), 64 / W systems per wavefront,
// four wavefronts per workgroup.  Systems of a size outside 0..max_n are
// left as they are (their factorization reported -2 and wrote no
// transpositions).
template <int W>
__global__ __launch_bounds__(256) void ldlt_solve_batch_kernel(
    int64_t count, const int64_t* __restrict__ offA, const int32_t* __restrict__ nv, int max_n,
    const double* __restrict__ Aall, const int64_t* __restrict__ offT,
    const int32_t* __restrict__ Tall, const int64_t* __restrict__ offb, double* __restrict__ B) {
  constexpr int G = 64 / W;
  const int lane = threadIdx.x & 63;
  const int sub = lane / W, sl = lane % W, base = sub * W;
  const int64_t s = (((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6) * G + sub;
  if (s >= count) return;  // whole groups: every shuffle below stays in the group
  const int n = nv[s];
  if (n <= 0 || n > max_n || n > W) return;
  const double* a = Aall + offA[s];
  double* xp = B + offb[s];
  // everything a lane needs, loaded up front so the loads overlap: its
  // transposition, right-hand side, d_i, and (W <= 32) row i and column i
  // of L in registers
  const int trv = sl < n ? Tall[offT[s] + sl] : 0;
  double x = sl < n ? xp[sl] : 0.0;
  const double d = sl < n ? a[sl + (int64_t)sl * n] : 1.0;
  constexpr bool kRegs = W <= 32;
  double lrow[kRegs ? W : 1], lcol[kRegs ? W : 1];
  if constexpr (kRegs) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      lrow[k] = (k < sl && sl < n) ? a[sl + (int64_t)k * n] : 0.0;  // L(i, k)
      lcol[k] = (k > sl && k < n) ? a[k + (int64_t)sl * n] : 0.0;   // L(k, i)
    }
  }
  // P and P^T as gathers: every lane follows its own position through the
  // transpositions (all W of them fetched by independent shuffles), so each
  // permutation costs one shuffle of x instead of a chain of swaps
  int pk[W];
#pragma unroll
  for (int k = 0; k < W; ++k) pk[k] = __shfl(trv, base + k);
  int src_p = sl, src_pt = sl;  // x (P b)_i = b_src_p; (P^T y)_i = y_src_pt
#pragma unroll
  for (int k = W - 1; k >= 0; --k)
    if (k < n) {
      if (src_p == k)
        src_p = pk[k];
      else if (src_p == pk[k])
        src_p = k;
    }
#pragma unroll
  for (int k = 0; k < W; ++k)
    if (k < n) {
      if (src_pt == k)
        src_pt = pk[k];
      else if (src_pt == pk[k])
        src_pt = k;
    }
  x = __shfl(x, base + src_p);
  // L y = x (unit lower), column k of L per step
  if constexpr (kRegs) {
#pragma unroll
    for (int k = 0; k < W - 1; ++k)
      if (k < n - 1) {
        const double xk = __shfl(x, base + k);
        if (sl > k) x -= lrow[k] * xk;
      }
  } else {
    for (int k = 0; k < n - 1; ++k) {
      const double xk = __shfl(x, base + k);
      if (sl > k && sl < n) x -= a[sl + (int64_t)k * n] * xk;
    }
  }
  // D^+ y: rows whose |d| is not above 1 / DBL_MAX become zero
  if (sl < n) x = fabs(d) > 1.0 / 1.7976931348623157e308 ? x / d : 0.0;
  // L^T x = y, row i of L per step
  if constexpr (kRegs) {
#pragma unroll
    for (int i = W - 1; i > 0; --i)
      if (i < n) {
        const double xi = __shfl(x, base + i);
        if (sl < i) x -= lcol[i] * xi;
      }
  } else {
    for (int i = n - 1; i > 0; --i) {
      const double xi = __shfl(x, base + i);
      if (sl < i) x -= a[i + (int64_t)sl * n] * xi;
    }
  }
  // x = P^T x
  x = __shfl(x, base + src_pt);
  if (sl < n) xp[sl] = x;
}

int group_width(int max_n) { return max_n <= 16 ? 16 : (max_n <= 32 ? 32 : 64); }

// LDLT rank-one update in place: one wave per factor, row i on lane i % 64
// (up to four rows per lane, n <= 256).  Restates
// internal::ldlt_inplace<Lower>::update (LDLT.h:440-449): w is permuted by
// the factor's transpositions in order (tmp = P w), then updateInPlace
// (LDLT.h:404-437) runs column by column: d_j += sigma w_j^2 / alpha,
// alpha += sigma w_j^2 / d_j, w_i -= w_j l_ij, and, if gamma = d_j alpha +
// sigma w_j^2 is not zero, l_ij += (sigma w_j / gamma) w_i below the
// diagonal, stopping once alpha is not finite.  Column j's entries are read
// and written by their own lanes in step j only, and d_j is read by all
// lanes before lane j % 64 writes it.  info: 0, or -2 for a size outside
// 0..256.
constexpr int kLdltRankMaxN = 256;

// element idx of a vector kept as four register slots per lane (slot idx /
// 64 of lane idx % 64), as a wave-uniform value
__device__ __forceinline__ double lane_get(const double (&t)[4], int idx) {
  const int k = idx >> 6;
  const double sel = k == 0 ? t[0] : (k == 1 ? t[1] : (k == 2 ? t[2] : t[3]));
  const long long b = __double_as_longlong(sel);
  return __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(b >> 32), idx & 63) << 32) |
                              (unsigned)__builtin_amdgcn_readlane((int)b, idx & 63));
}

__global__ __launch_bounds__(256) void ldlt_rank_update_kernel(
    int64_t count, const int64_t* __restrict__ offA, const int32_t* __restrict__ nv,
    double* __restrict__ Aall, const int64_t* __restrict__ offT, const int32_t* __restrict__ Tall,
    const int64_t* __restrict__ offw, const double* __restrict__ Wall,
    const double* __restrict__ sigmas, int32_t* __restrict__ info) {
  constexpr int K = kLdltRankMaxN / 64;
  const int lane = threadIdx.x & 63;
  const int64_t s = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  if (s >= count) return;  // whole waves
  const int n = nv[s];
  if (n < 0 || n > kLdltRankMaxN) {
    if (lane == 0) info[s] = -2;
    return;
  }
  double* A = Aall + offA[s];
  const int32_t* tr = Tall + offT[s];
  const double* w = Wall + offw[s];
  const double sigma = sigmas ? sigmas[s] : 1.0;
  double t[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = lane + 64 * k;
    t[k] = i < n ? w[i] : 0.0;
  }
  // tmp = P w: transposition k swaps entries k and tr[k], k = 0, 1, ...
  for (int k = 0; k < n; ++k) {
    const int j = tr[k];
    if (j == k || j < 0 || j >= n) continue;
    const double vk = lane_get(t, k), vj = lane_get(t, j);
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = lane + 64 * q;
      t[q] = i == k ? vj : (i == j ? vk : t[q]);
    }
  }
  double alpha = 1.0;
  for (int j = 0; j < n; ++j) {
    if (!__builtin_isfinite(alpha)) break;
    const double dj = A[j + (int64_t)j * n];
    const double wj = lane_get(t, j);
    const double swj2 = sigma * wj * wj;
    const double gamma = dj * alpha + swj2;
    if (lane == (j & 63)) A[j + (int64_t)j * n] = dj + swj2 / alpha;
    alpha += swj2 / dj;
    const double c = sigma * wj / gamma;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = lane + 64 * q;
      if (i > j && i < n) {
        const double lij = A[i + (int64_t)j * n];
        t[q] -= wj * lij;
        if (gamma != 0.0) A[i + (int64_t)j * n] = lij + c * t[q];
      }
    }
  }
  if (lane == 0) info[s] = 0;
}

}  // namespace

int ldlt_rank_update_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                             double* A, const int64_t* offT, const int32_t* T,
                             const int64_t* offw, const double* w, const double* sigma,
                             int32_t* info) {
  if (count <= 0) return kOk;
  const int64_t blocks = (count + 3) / 4;  // four waves (factors) per workgroup
  if (blocks > INT32_MAX) {
    set_error("ldlt rank update: %lld factors exceed one launch", (long long)count);
    return kErrUnsupported;
  }
  hipLaunchKernelGGL(ldlt_rank_update_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                     count, offA, n, A, offT, T, offw, w, sigma, info);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int ldlt_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n, int max_n,
                 double* A, const int64_t* offT, int32_t* T, int32_t* sign, int32_t* info) {
  if (count == 0) return kOk;
  if (max_n < 0 || max_n > kLdltMaxN) {
    set_error("ldlt: max_n %d outside 0..%d", max_n, kLdltMaxN);
    return kErrUnsupported;
  }
  const int W = group_width(max_n);
  const int64_t blocks = (count + 64 / W - 1) / (64 / W);
  if (blocks > INT32_MAX) {
    set_error("ldlt: %lld systems exceed one launch", (long long)count);
    return kErrUnsupported;
  }
  const dim3 grid((unsigned)blocks), block(64);
  const size_t lds = sizeof(double) * (64 / W) * ((size_t)max_n * (max_n + 1) + max_n);
  if (W == 16)
    hipLaunchKernelGGL(ldlt_batch_kernel<16>, grid, block, lds, h->stream, count, offA, n, max_n,
                       A, offT, T, sign, info);
  else if (W == 32)
    hipLaunchKernelGGL(ldlt_batch_kernel<32>, grid, block, lds, h->stream, count, offA, n, max_n,
                       A, offT, T, sign, info);
  else
    hipLaunchKernelGGL(ldlt_batch_kernel<64>, grid, block, lds, h->stream, count, offA, n, max_n,
                       A, offT, T, sign, info);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int ldlt_solve_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                       int max_n, const double* A, const int64_t* offT, const int32_t* T,
                       const int64_t* offb, double* b) {
  if (count == 0) return kOk;
  if (max_n < 0 || max_n > kLdltMaxN) {
    set_error("ldlt solve: max_n %d outside 0..%d", max_n, kLdltMaxN);
    return kErrUnsupported;
  }
  const int W = group_width(max_n);
  const int64_t per_block = 4 * (64 / W);  // four wavefronts per workgroup
  const int64_t blocks = (count + per_block - 1) / per_block;
  if (blocks > INT32_MAX) {
    set_error("ldlt solve: %lld systems exceed one launch", (long long)count);
    return kErrUnsupported;
  }
  const dim3 grid((unsigned)blocks), block(256);
  if (W == 16)
    hipLaunchKernelGGL(ldlt_solve_batch_kernel<16>, grid, block, 0, h->stream, count, offA, n,
                       max_n, A, offT, T, offb, b);
  else if (W == 32)
    hipLaunchKernelGGL(ldlt_solve_batch_kernel<32>, grid, block, 0, h->stream, count, offA, n,
                       max_n, A, offT, T, offb, b);
  else
    hipLaunchKernelGGL(ldlt_solve_batch_kernel<64>, grid, block, 0, h->stream, count, offA, n,
                       max_n, A, offT, T, offb, b);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

}  // namespace eamd
