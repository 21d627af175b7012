// llt.hip -- blocked right-looking Cholesky (lower) in FP64.
//
// Replaces llt_inplace<Scalar,Lower>::blocked (Eigen/src/Cholesky/LLT.h:330-360)
// and ::unblocked (:304-328).  Per block column of width nb:
//   potf2_kernel   the nb x nb diagonal block is factored by one workgroup in
//                  LDS (x <= 0 fails at that column, NaN passes, as in the
//                  reference's `if (x <= RealScalar(0)) return k`);
//   trsm_rlt_kernel A21 <- A21 * L11^{-T}, one workgroup per 64 rows;
//   dgemm (lower)  A22 -= A21 * A21^T on MFMA, tiles above the diagonal
//                  skipped (the SYRK "bottleneck" of LLT.h:357).
// A failure is recorded on the device; later kernels of the same call see
// it and return early, so no host synchronisation per block is needed.
#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kMaxNb = 128;

// Factor the bs x bs lower block at A (column-major, lda).  fail[0] keeps
// the first failing absolute column (k0 + local) or -1.
__global__ __launch_bounds__(256) void potf2_kernel(double* __restrict__ A, int64_t lda, int bs,
                                                    int k0, int* __restrict__ fail) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* S = reinterpret_cast<double*>(smem_raw);  // bs x (bs+1), column-major
  const int ld = bs + 1;
  __shared__ int s_fail;
  __shared__ double s_d;
  const int tid = threadIdx.x;
  if (*fail >= 0) return;
  for (int e = tid; e < bs * bs; e += 256) {
    int j = e / bs, i = e - j * bs;
    S[i + j * ld] = (i >= j) ? A[i + (int64_t)j * lda] : 0.0;
  }
  if (tid == 0) s_fail = -1;
  __syncthreads();
  for (int k = 0; k < bs; ++k) {
    if (tid == 0) {
      double x = S[k + k * ld];
      if (x <= 0.0) {
        s_fail = k;
      } else {
        x = sqrt(x);
        S[k + k * ld] = x;
        s_d = x;
      }
    }
    __syncthreads();
    if (s_fail >= 0) break;
    const double d = s_d;
    for (int i = k + 1 + tid; i < bs; i += 256) S[i + k * ld] /= d;
    __syncthreads();
    // trailing lower update: S[i][j] -= S[i][k] * S[j][k], k < j <= i
    const int rs = bs - k - 1;
    const int tri = rs * (rs + 1) / 2;
    for (int e = tid; e < rs * rs; e += 256) {
      int jj = e / rs, ii = e - jj * rs;
      if (ii >= jj) {
        int i = k + 1 + ii, j = k + 1 + jj;
        S[i + j * ld] -= S[i + k * ld] * S[j + k * ld];
      }
    }
    (void)tri;
    __syncthreads();
  }
  if (s_fail >= 0) {
    if (tid == 0) *fail = k0 + s_fail;
    return;
  }
  for (int e = tid; e < bs * bs; e += 256) {
    int j = e / bs, i = e - j * bs;
    if (i >= j) A[i + (int64_t)j * lda] = S[i + j * ld];
  }
}

// X (m x bs, column-major) <- X * L^{-T}, L lower bs x bs (column-major).
// Row r of X solves L y = x_r^T.  One workgroup per RB rows; X rows in LDS,
// L staged in 32x32 tiles.
template <int RB>
__global__ __launch_bounds__(256) void trsm_rlt_kernel(const double* __restrict__ L, int64_t ldl,
                                                       int bs, double* __restrict__ X,
                                                       int64_t ldx, int m,
                                                       const int* __restrict__ fail) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* Xs = reinterpret_cast<double*>(smem_raw);  // RB x (bs+1)... stored [col][row]
  __shared__ double Ls[32][33];
  const int tid = threadIdx.x;
  if (*fail >= 0) return;
  const int r0 = blockIdx.x * RB;
  const int rr = min(RB, m - r0);
  const int ld = RB + 1;
  for (int e = tid; e < RB * bs; e += 256) {
    int c = e / RB, r = e - c * RB;
    Xs[r + c * ld] = r < rr ? X[(r0 + r) + (int64_t)c * ldx] : 0.0;
  }
  for (int b = 0; b < bs; b += 32) {
    const int w = min(32, bs - b);
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += 256) {
      int i = e & 31, j = e >> 5;  // Ls[i][j] = L(b+i, b+j)
      Ls[i][j] = (i < w && j < w) ? L[(b + i) + (int64_t)(b + j) * ldl] : 0.0;
    }
    __syncthreads();
    // diagonal block: y_c = (x_c - sum_{q<c} y_q L(c,q)) / L(c,c), column by column
    for (int c = 0; c < w; ++c) {
      for (int r = tid; r < RB; r += 256) {
        double v = Xs[r + (b + c) * ld];
        for (int q = 0; q < c; ++q) v -= Xs[r + (b + q) * ld] * Ls[c][q];
        Xs[r + (b + c) * ld] = v / Ls[c][c];
      }
    }
    // later columns: x_j -= sum_{q in block} y_q L(j, q)
    for (int jb = b + w; jb < bs; jb += 32) {
      const int w2 = min(32, bs - jb);
      __syncthreads();
      for (int e = tid; e < 32 * 32; e += 256) {
        int i = e & 31, j = e >> 5;  // Ls[i][j] = L(jb+i, b+j)
        Ls[i][j] = (i < w2 && j < w) ? L[(jb + i) + (int64_t)(b + j) * ldl] : 0.0;
      }
      __syncthreads();
      for (int e = tid; e < RB * 32; e += 256) {
        int r = e % RB, i = e / RB;
        if (i < w2) {
          double acc = 0.0;
          for (int q = 0; q < w; ++q) acc += Xs[r + (b + q) * ld] * Ls[i][q];
          Xs[r + (jb + i) * ld] -= acc;
        }
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < RB * bs; e += 256) {
    int c = e / RB, r = e - c * RB;
    if (r < rr) X[(r0 + r) + (int64_t)c * ldx] = Xs[r + c * ld];
  }
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

}  // namespace

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
    EAMD_HIP(hipFuncSetAttribute((const void*)trsm_rlt_kernel<64>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr = true;
  }
  for (int k = 0; k < n; k += nb) {
    const int bs = min(nb, n - k);
    const int rs = n - k - bs;
    double* A11 = A + k + (int64_t)k * lda;
    size_t lds = sizeof(double) * (size_t)bs * (bs + 1);
    hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(256), lds, s, A11, lda, bs, k, fail);
    if (rs > 0) {
      double* A21 = A11 + bs;
      size_t lds2 = sizeof(double) * (size_t)(64 + 1) * bs;
      hipLaunchKernelGGL(trsm_rlt_kernel<64>, dim3(ceil_div(rs, 64)), dim3(256), lds2, s, A11,
                         lda, bs, A21, lda, rs, fail);
      double* A22 = A21 + (int64_t)bs * lda;
      EAMD_CHECK(dgemm_device(s, false, true, rs, rs, bs, -1.0, A21, lda, A21, lda, 1.0, A22, lda,
                              1, fail));
    }
  }
  EAMD_HIP(hipGetLastError());
  int hf = -1;
  EAMD_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, s));
  EAMD_HIP(hipStreamSynchronize(s));
  *fail_col = hf;
  return kOk;
}

}  // namespace eamd
