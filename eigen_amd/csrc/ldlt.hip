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
// MI355X mapping.  One wavefront per system (n <= 64): the lower triangle is
// staged in LDS (column stride n + 1), lane i owns row i, so the pivot
// search is one butterfly arg-max over the wave, the exchange is a set of
// disjoint element swaps done by all lanes at once, and the column update is
// one length-k dot product per lane against temp held in LDS.  The solve is
// one wavefront per system with x_i in lane i: the transpositions and the
// substitutions exchange values by lane shuffles, and L is read straight
// from HBM (column k for the forward sweep, row i for the backward one).
// No host synchronisation: both launches can be captured in a HIP graph.
#include "../../include/eigen_amd.h"
#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kLdltMaxN = 64;

// info: 0 Success, 1 NumericalIssue (ComputationInfo's values), -2 for a
// size outside 0..min(max_n, 64); sign: 0 PositiveSemiDef, 1 NegativeSemiDef,
// 2 ZeroSign, 3 Indefinite (internal::SignMatrix, LDLT.h:22)
__global__ __launch_bounds__(64) void ldlt_batch_kernel(const int64_t* __restrict__ offA,
                                                        const int32_t* __restrict__ nv, int max_n,
                                                        double* __restrict__ Aall,
                                                        const int64_t* __restrict__ offT,
                                                        int32_t* __restrict__ Tall,
                                                        int32_t* __restrict__ sign_out,
                                                        int32_t* __restrict__ info) {
  extern __shared__ double s_lds[];
  const int64_t s = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = nv[s];
  if (n < 0 || n > max_n || n > kLdltMaxN) {
    if (lane == 0) info[s] = -2;
    return;
  }
  double* a = Aall + offA[s];
  int32_t* tr = Tall + offT[s];
  const int ld = n + 1;
  double* M = s_lds;                       // n x n, column stride ld
  double* temp = s_lds + max_n * (max_n + 1);  // max_n
  if (n <= 1) {
    if (lane == 0) {
      int sg = 2;
      if (n == 1) {
        tr[0] = 0;
        const double d = a[0];
        sg = d > 0.0 ? 0 : (d < 0.0 ? 1 : 2);
      }
      sign_out[s] = sg;
      info[s] = 0;
    }
    return;
  }
  // the lower triangle, one coalesced column at a time (lane = row)
#pragma unroll 8
  for (int j = 0; j < n; ++j)
    if (lane >= j && lane < n) M[lane + j * ld] = a[lane + (int64_t)j * n];
  __syncthreads();

  int sg = 2;
  bool ret = true, found_zero = false, zero_exit = false;
  for (int k = 0; k < n; ++k) {
    // ---- pivot: (|a_ii|, i) arg-max, lower index on ties
    const int ci = k + lane;
    double v = -1.0;
    if (ci < n) {
      v = fabs(M[ci + ci * ld]);
      if (v != v) v = lane == 0 ? __builtin_inf() : -1.0;
    }
    int idx = ci;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(v, off);
      const int oi = __shfl_xor(idx, off);
      if (ov > v || (ov == v && oi < idx)) {
        v = ov;
        idx = oi;
      }
    }
    const int p = __builtin_amdgcn_readfirstlane(idx);
    if (lane == 0) tr[k] = p;
    if (p != k) {
      // the lower-only exchange (LDLT.h:321-337); the four sets of
      // elements are disjoint, so every lane swaps its own at once
      if (lane < k) {
        const double t = M[k + lane * ld];
        M[k + lane * ld] = M[p + lane * ld];
        M[p + lane * ld] = t;
      }
      const int it = p + 1 + lane;
      if (it < n) {
        const double t = M[it + k * ld];
        M[it + k * ld] = M[it + p * ld];
        M[it + p * ld] = t;
      }
      if (lane == 0) {
        const double t = M[k + k * ld];
        M[k + k * ld] = M[p + p * ld];
        M[p + p * ld] = t;
      }
      const int im = k + 1 + lane;
      if (im < p) {
        const double t = M[im + k * ld];
        M[im + k * ld] = M[p + im * ld];
        M[p + im * ld] = t;
      }
      __syncthreads();
    }
    // ---- column k: a_ik -= sum_j l_ij d_j l_kj (LDLT.h:348-354)
    if (k > 0) {
      if (lane < k) temp[lane] = M[lane + lane * ld] * M[k + lane * ld];
      __syncthreads();
      if (ci < n) {
        double r = 0.0;
        for (int j = 0; j < k; ++j) r += M[ci + j * ld] * temp[j];
        M[ci + k * ld] -= r;
      }
      __syncthreads();
    }
    const double akk = M[k + k * ld];
    const bool valid = fabs(akk) > 0.0;
    if (k == 0 && !valid) {
      // the whole diagonal is zero (LDLT.h:363-374)
      bool nz = false;
      for (int j = 0; j < n; ++j)
        if (lane > j && lane < n && !(M[lane + j * ld] == 0.0)) nz = true;
      ret = !__any(nz);
      for (int j = lane; j < n; j += 64) tr[j] = j;
      sg = 2;
      zero_exit = true;
      break;
    }
    const int ib = k + 1 + lane;
    bool nz = false;
    if (ib < n) {
      if (valid)
        M[ib + k * ld] /= akk;
      else
        nz = !(M[ib + k * ld] == 0.0);
    }
    if (__any(nz)) ret = false;
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
    __syncthreads();
  }
  if (!zero_exit) {
#pragma unroll 8
    for (int j = 0; j < n; ++j)
      if (lane >= j && lane < n) a[lane + (int64_t)j * n] = M[lane + j * ld];
  }
  if (lane == 0) {
    sign_out[s] = sg;
    info[s] = ret ? 0 : 1;
  }
}

// x <- A^{-1} b for each system from ldlt_batch_kernel's factors; four
// systems (wavefronts) per workgroup, x_i in lane i.  Systems of a size
// outside 0..64 are left as they are (their factorization reported -2).
__global__ __launch_bounds__(256) void ldlt_solve_batch_kernel(
    int64_t count, const int64_t* __restrict__ offA, const int32_t* __restrict__ nv,
    const double* __restrict__ Aall, const int64_t* __restrict__ offT,
    const int32_t* __restrict__ Tall, const int64_t* __restrict__ offb, double* __restrict__ B) {
  const int lane = threadIdx.x & 63;
  const int64_t s = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  if (s >= count) return;  // whole waves
  const int n = nv[s];
  if (n <= 0 || n > kLdltMaxN) return;
  const double* a = Aall + offA[s];
  const int32_t* tr = Tall + offT[s];
  double* xp = B + offb[s];
  double x = lane < n ? xp[lane] : 0.0;
  // x = P b: the transpositions in order
  for (int k = 0; k < n; ++k) {
    const int p = tr[k];
    if (p != k) {
      const double xk = __shfl(x, k), xq = __shfl(x, p);
      if (lane == k) x = xq;
      if (lane == p) x = xk;
    }
  }
  // L y = x (unit lower), column k of L per step
  for (int k = 0; k < n - 1; ++k) {
    const double xk = __shfl(x, k);
    if (lane > k && lane < n) x -= a[lane + (int64_t)k * n] * xk;
  }
  // D^+ y: rows whose |d| is not above 1 / DBL_MAX become zero
  if (lane < n) {
    const double d = a[lane + (int64_t)lane * n];
    x = fabs(d) > 1.0 / 1.7976931348623157e308 ? x / d : 0.0;
  }
  // L^T x = y, row i of L per step
  for (int i = n - 1; i > 0; --i) {
    const double xi = __shfl(x, i);
    if (lane < i) x -= a[i + (int64_t)lane * n] * xi;
  }
  // x = P^T x: the transpositions in reverse order
  for (int k = n - 1; k >= 0; --k) {
    const int p = tr[k];
    if (p != k) {
      const double xk = __shfl(x, k), xq = __shfl(x, p);
      if (lane == k) x = xq;
      if (lane == p) x = xk;
    }
  }
  if (lane < n) xp[lane] = x;
}

}  // namespace

int ldlt_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n, int max_n,
                 double* A, const int64_t* offT, int32_t* T, int32_t* sign, int32_t* info) {
  if (count == 0) return kOk;
  if (max_n < 0 || max_n > kLdltMaxN) {
    set_error("ldlt: max_n %d outside 0..%d", max_n, kLdltMaxN);
    return kErrUnsupported;
  }
  if (count > INT32_MAX) {
    set_error("ldlt: %lld systems exceed one launch", (long long)count);
    return kErrUnsupported;
  }
  const size_t lds = sizeof(double) * ((size_t)max_n * (max_n + 1) + max_n);
  hipLaunchKernelGGL(ldlt_batch_kernel, dim3((unsigned)count), dim3(64), lds, h->stream, offA, n,
                     max_n, A, offT, T, sign, info);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int ldlt_solve_batched(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                       const double* A, const int64_t* offT, const int32_t* T,
                       const int64_t* offb, double* b) {
  if (count == 0) return kOk;
  const int64_t blocks = (count + 3) / 4;  // four waves (systems) per workgroup
  if (blocks > INT32_MAX) {
    set_error("ldlt solve: %lld systems exceed one launch", (long long)count);
    return kErrUnsupported;
  }
  hipLaunchKernelGGL(ldlt_solve_batch_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                     count, offA, n, A, offT, T, offb, b);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

}  // namespace eamd
