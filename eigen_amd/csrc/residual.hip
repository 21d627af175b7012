// residual.hip -- FP64 residuals r = b - A x, norms and the normwise
// backward error on the device.
//
// The product A x is accumulated in double-double (Knuth TwoSum + FMA
// TwoProd), so the computed residual is accurate to about u^2 |A||x| rather
// than n u |A||x|: at n = 16384 a plain FP64 dot product would carry
// rounding noise of the same order as the 1e-14 backward-error target.
// Work is split over (row block, column chunk) tiles so that a few thousand
// workgroups stream A (column-major, coalesced down each column) from HBM;
// a second pass reduces the chunk partials.
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace eamd {

namespace {

struct dd {
  double hi, lo;
};

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void dd_add_prod(double& hi, double& lo, double a, double b) {
  double p = a * b;
  double pe = fma(a, b, -p);
  double s, e;
  two_sum(hi, p, s, e);
  hi = s;
  lo += e + pe;
}

__device__ __forceinline__ void dd_add(double& hi, double& lo, double h2, double l2) {
  double s, e;
  two_sum(hi, h2, s, e);
  hi = s;
  lo += e + l2;
}

constexpr int RB = 256;     // rows per workgroup
constexpr int CCH = 256;    // columns per chunk

// partial[chunk][rhs][row] (hi, lo) of sum_{j in chunk} A(i,j) x(j,rhs)
__global__ __launch_bounds__(RB) void gemv_dd_partial_kernel(int n, int ncols, const double* __restrict__ A,
                                                             int64_t lda, const double* __restrict__ X,
                                                             int64_t ldx, int nrhs, dd* __restrict__ part) {
  const int i = blockIdx.x * RB + threadIdx.x;
  const int chunk = blockIdx.y;
  const int j0 = chunk * CCH, j1 = min(ncols, j0 + CCH);
  __shared__ double xs[CCH * 4];
  for (int c = 0; c < nrhs; c += 4) {
    const int nc = min(4, nrhs - c);
    __syncthreads();
    for (int e = threadIdx.x; e < CCH * 4; e += RB) {
      int jj = e % CCH, q = e / CCH;
      xs[e] = (j0 + jj < j1 && q < nc) ? X[(int64_t)(c + q) * ldx + j0 + jj] : 0.0;
    }
    __syncthreads();
    if (i < n) {
      double hi[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0};
      const double* a = A + (int64_t)j0 * lda + i;
      for (int j = j0; j < j1; ++j, a += lda) {
        const double v = *a;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < nc) dd_add_prod(hi[q], lo[q], v, xs[q * CCH + (j - j0)]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < nc) part[((size_t)chunk * nrhs + c + q) * n + i] = dd{hi[q], lo[q]};
    }
  }
}

// r = b - sum_chunks partial (double-double, rounded once); also |r| and
// (optionally) the row abs-sum of A when x = nullptr is not used here.
__global__ __launch_bounds__(256) void residual_finish_kernel(int n, int nchunks, int nrhs,
                                                              const dd* __restrict__ part,
                                                              const double* __restrict__ B,
                                                              int64_t ldb, double* __restrict__ R,
                                                              int64_t ldr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || c >= nrhs) return;
  double hi = B ? B[(int64_t)c * ldb + i] : 0.0, lo = 0.0;  // B == nullptr: R = -A X
  for (int k = 0; k < nchunks; ++k) {
    dd p = part[((size_t)k * nrhs + c) * n + i];
    dd_add(hi, lo, -p.hi, -p.lo);
  }
  R[(int64_t)c * ldr + i] = hi + lo;
}

// row abs sums of A (for ||A||_inf), chunked like the GEMV
__global__ __launch_bounds__(RB) void rowabs_partial_kernel(int n, const double* __restrict__ A,
                                                            int64_t lda, double* __restrict__ part,
                                                            double* __restrict__ out) {
  const int i = blockIdx.x * RB + threadIdx.x;
  const int chunk = blockIdx.y;
  const int j0 = chunk * CCH, j1 = min(n, j0 + CCH);
  if (i == 0 && chunk == 0) *out = 0.0;  // rowabs_finish_kernel's maximum starts here
  if (i >= n) return;
  double s = 0.0;
  const double* a = A + (int64_t)j0 * lda + i;
  for (int j = j0; j < j1; ++j, a += lda) s += fabs(*a);
  part[(size_t)chunk * n + i] = s;
}

// max over i of sum over chunks -> out[0]: one thread per row, the
// workgroup's maximum merged by an atomic maximum on the bit pattern (for
// non-negative doubles, NaN included, the unsigned order of the bits is the
// numeric order); out[0] was cleared by rowabs_partial_kernel
__global__ __launch_bounds__(256) void rowabs_finish_kernel(int n, int nchunks,
                                                            const double* __restrict__ part,
                                                            double* __restrict__ out) {
  __shared__ double red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (i < n)
    for (int k = 0; k < nchunks; ++k) s += part[(size_t)k * n + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double a = red[threadIdx.x], b = red[threadIdx.x + w];
      red[threadIdx.x] = (b > a || b != b) ? b : a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax(reinterpret_cast<unsigned long long*>(out),
              (unsigned long long)__double_as_longlong(red[0]));
}

// per rhs: max|r|, max|x|, max|b| -> out[3*c + {0,1,2}]; NaN propagates
__global__ __launch_bounds__(256) void inf_norms_kernel(int n, int nrhs, const double* __restrict__ R,
                                                        int64_t ldr, const double* __restrict__ X,
                                                        int64_t ldx, const double* __restrict__ B,
                                                        int64_t ldb, double* __restrict__ out) {
  __shared__ double red[3][256];
  const int c = blockIdx.x;
  double mr = 0, mx = 0, mb = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double r = fabs(R[(int64_t)c * ldr + i]);
    double x = fabs(X[(int64_t)c * ldx + i]);
    double b = fabs(B[(int64_t)c * ldb + i]);
    mr = (r > mr || r != r) ? r : mr;
    mx = (x > mx || x != x) ? x : mx;
    mb = fmax(mb, b);
  }
  red[0][threadIdx.x] = mr;
  red[1][threadIdx.x] = mx;
  red[2][threadIdx.x] = mb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int q = 0; q < 3; ++q) {
        double a = red[q][threadIdx.x], b = red[q][threadIdx.x + w];
        red[q][threadIdx.x] = (b > a || b != b) ? b : a;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) out[3 * c + q] = red[q][0];
}

}  // namespace

int residual_dd(eamd_handle* h, int n, int nrhs, const double* A, int64_t lda, const double* X,
                int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr) {
  const int nchunks = ceil_div(n, CCH);
  void* p = nullptr;
  EAMD_CHECK(h->ws.get(kWsAux2, sizeof(dd) * (size_t)nchunks * nrhs * n + 256, &p));
  dd* part = static_cast<dd*>(p);
  PhaseScope ps(h, kPhResidual, 8.0 * ((double)n * n + 3.0 * (double)n * nrhs),
                2.0 * (double)n * n * nrhs);
  dim3 g1(ceil_div(n, RB), nchunks);
  hipLaunchKernelGGL(gemv_dd_partial_kernel, g1, dim3(RB), 0, h->stream, n, n, A, lda, X, ldx,
                     nrhs, part);
  dim3 g2(ceil_div(n, 256), nrhs);
  hipLaunchKernelGGL(residual_finish_kernel, g2, dim3(256), 0, h->stream, n, nchunks, nrhs, part,
                     B, ldb, R, ldr);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int matrix_inf_norm(eamd_handle* h, int n, const double* A, int64_t lda, double* out_dev) {
  const int nchunks = ceil_div(n, CCH);
  void* p = nullptr;
  EAMD_CHECK(h->ws.get(kWsAux2, sizeof(double) * (size_t)nchunks * n + 256, &p));
  double* part = static_cast<double*>(p);
  dim3 g1(ceil_div(n, RB), nchunks);
  hipLaunchKernelGGL(rowabs_partial_kernel, g1, dim3(RB), 0, h->stream, n, A, lda, part,
                     out_dev);
  hipLaunchKernelGGL(rowabs_finish_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, n,
                     nchunks, part, out_dev);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int inf_norms(eamd_handle* h, int n, int nrhs, const double* R, int64_t ldr, const double* X,
              int64_t ldx, const double* B, int64_t ldb, double* out_dev) {
  hipLaunchKernelGGL(inf_norms_kernel, dim3(nrhs), dim3(256), 0, h->stream, n, nrhs, R, ldr, X,
                     ldx, B, ldb, out_dev);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// Backward error: returns max over rhs of ||r||/(||A|| ||x|| + ||b||).
int backward_error_host(eamd_handle* h, int n, int nrhs, const double* A, int64_t lda,
                        const double* X, int64_t ldx, const double* B, int64_t ldb, double* R,
                        double anorm, double* berr_out, double* per_rhs) {
  EAMD_CHECK(residual_dd(h, n, nrhs, A, lda, X, ldx, B, ldb, R, n));
  void* p = nullptr;
  EAMD_CHECK(h->ws.get(kWsAux, sizeof(double) * (3 * (size_t)nrhs + 8), &p));
  double* nd = static_cast<double*>(p);
  EAMD_CHECK(inf_norms(h, n, nrhs, R, n, X, ldx, B, ldb, nd));
  double hn[3 * 64];
  int cnt = nrhs < 64 ? nrhs : 64;
  EAMD_HIP(hipMemcpyAsync(hn, nd, sizeof(double) * 3 * cnt, hipMemcpyDeviceToHost, h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  double worst = 0.0;
  for (int c = 0; c < cnt; ++c) {
    double den = anorm * hn[3 * c + 1] + hn[3 * c + 2];
    double e = den > 0 ? hn[3 * c] / den : hn[3 * c];
    if (per_rhs) per_rhs[c] = e;
    if (e > worst || e != e) worst = e;
  }
  *berr_out = worst;
  return kOk;
}

int dbackward_error_impl(eamd_handle* h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                         const double* X, int64_t ldx, const double* B, int64_t ldb,
                         double* berr) {
  void* p = nullptr;
  EAMD_CHECK(h->ws.get(kWsVec, sizeof(double) * ((size_t)n * nrhs + 64), &p));
  double* R = static_cast<double*>(p);
  double* anorm_dev = R + (size_t)n * nrhs;
  EAMD_CHECK(matrix_inf_norm(h, (int)n, A, lda, anorm_dev));
  double anorm = 0;
  EAMD_HIP(hipMemcpyAsync(&anorm, anorm_dev, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  double worst = 0;
  for (int64_t c0 = 0; c0 < nrhs; c0 += 64) {
    int nc = (int)std::min<int64_t>(64, nrhs - c0);
    double e = 0;
    EAMD_CHECK(backward_error_host(h, (int)n, nc, A, lda, X + c0 * ldx, ldx, B + c0 * ldb, ldb,
                                   R, anorm, &e, nullptr));
    if (e > worst || e != e) worst = e;
  }
  *berr = worst;
  return kOk;
}

}  // namespace eamd
