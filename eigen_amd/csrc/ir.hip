// ir.hip -- mixed-precision LU solver with iterative refinement.
//
// No counterpart in the reference (Eigen has no iterative refinement); it
// sits behind PartialPivLU's solve() surface as a new capability:
//   1. W = fp32(A) in a row-major working copy; LU with partial pivoting,
//      FP32 panels and the trailing update in the requested precision
//      (FP32 / FP16 / BF16 / FP16 hi-lo split on the matrix cores);
//   2. x0 = U^{-1} L^{-1} P b in FP32, promoted to FP64;
//   3. classic refinement: r = b - A x in FP64 (double-double accumulation),
//      d = U^{-1} L^{-1} P r in FP32, x += d in FP64, until the normwise
//      backward error ||r||/(||A|| ||x|| + ||b||) reaches the tolerance;
//   4. when the classic iteration stalls (contraction worse than 1/2 twice in
//      a row, or divergence), GMRES-IR: restarted FP64 GMRES on A d = r,
//      right-preconditioned by the same FP32 LU (the preconditioned GMRES of
//      unsupported/Eigen/src/IterativeSolvers/GMRES.h:56-200 is the design
//      reference; here the Arnoldi basis uses classical Gram-Schmidt applied
//      twice, which maps to two dot-product passes on the device).
#include "../../include/eigen_amd.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace eamd {

int residual_dd(eamd_handle* h, int n, int nrhs, const double* A, int64_t lda, const double* X,
                int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr);
int matrix_inf_norm(eamd_handle* h, int n, const double* A, int64_t lda, double* out_dev);
int inf_norms(eamd_handle* h, int n, int nrhs, const double* R, int64_t ldr, const double* X,
              int64_t ldx, const double* B, int64_t ldb, double* out_dev);
int lu_block_width(int64_t n, bool fp64);

namespace {

__global__ void iota_kernel(int* v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// y[i,c] = (To) x[q[i], c]   (row gather + precision change)
template <typename Ti, typename To>
__global__ void gather_rows_kernel(int n, int nrhs, const int* __restrict__ q,
                                   const Ti* __restrict__ x, int64_t ldx, To* __restrict__ y,
                                   int64_t ldy) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (i < n && c < nrhs) y[(int64_t)c * ldy + i] = (To)x[(int64_t)c * ldx + q[i]];
}

// y = (To) x
template <typename Ti, typename To>
__global__ void convert_kernel(int n, int nrhs, const Ti* __restrict__ x, int64_t ldx,
                               To* __restrict__ y, int64_t ldy) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (i < n && c < nrhs) y[(int64_t)c * ldy + i] = (To)x[(int64_t)c * ldx + i];
}

// x += (double) d
template <typename T>
__global__ void add_kernel(int n, int nrhs, const T* __restrict__ d, int64_t ldd,
                           double* __restrict__ x, int64_t ldx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (i < n && c < nrhs) x[(int64_t)c * ldx + i] += (double)d[(int64_t)c * ldd + i];
}

// dots[k] = V[:,k] . w for k < nv (one workgroup per k); V may alias w's
// storage only for k's that are not w itself.
__global__ __launch_bounds__(256) void multidot_kernel(int n, const double* __restrict__ V,
                                                       int64_t ldv, const double* __restrict__ w,
                                                       double* __restrict__ dots) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += V[(int64_t)k * ldv + i] * w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) dots[k] = red[0];
}

// w -= V[:, :nv] * c   (c on the device)
__global__ void vmul_sub_kernel(int n, int nv, const double* __restrict__ V, int64_t ldv,
                                const double* __restrict__ c, double* __restrict__ w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int k = 0; k < nv; ++k) s -= V[(int64_t)k * ldv + i] * c[k];
  w[i] = s;
}

// y = alpha * x
__global__ void scale_copy_kernel(int n, double alpha, const double* __restrict__ x,
                                  double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = alpha * x[i];
}

// ord[k + pos[s]] = old ord[k + pos[src[s]]]  (row order after a panel)
__global__ __launch_bounds__(256) void update_order_kernel(int* __restrict__ ord, int k,
                                                           const int* __restrict__ perm_n,
                                                           const int* __restrict__ pos,
                                                           const int* __restrict__ src) {
  extern __shared__ int old[];
  const int nS = *perm_n;
  for (int s = threadIdx.x; s < nS; s += 256) old[s] = ord[k + pos[s]];
  __syncthreads();
  for (int s = threadIdx.x; s < nS; s += 256) ord[k + pos[s]] = old[src[s]];
}

}  // namespace

int update_row_order(hipStream_t s, int* ord, int k, const int* perm_n, const int* pos,
                     const int* src, int nb) {
  hipLaunchKernelGGL(update_order_kernel, dim3(1), dim3(256), sizeof(int) * 2 * (size_t)nb, s,
                     ord, k, perm_n, pos, src);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int init_row_order(hipStream_t s, int* ord, int n) {
  hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, ord, n);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// ---------------------------------------------------------------------------
namespace {

struct IrContext {
  eamd_handle* h;
  int n;
  const double* A;
  int64_t lda;
  const void* W;  // factors, row-major, ld ldw
  int64_t ldw;
  bool fp64_factors;
  const int* q;  // row order of the factorization: (P b)[i] = b[q[i]]
  float* Y32;    // n x ncap scratch
  double* Y64;
  int ncap;
};

// D = U^{-1} L^{-1} P R  (column-major n x nrhs, nrhs <= ncap)
int lu_apply_inverse(IrContext& c, const double* R, int64_t ldr, double* D, int64_t ldd,
                     int nrhs) {
  hipStream_t s = c.h->stream;
  dim3 g(ceil_div(c.n, 256), nrhs);
  if (c.fp64_factors) {
    const double* W = static_cast<const double*>(c.W);
    hipLaunchKernelGGL((gather_rows_kernel<double, double>), g, dim3(256), 0, s, c.n, nrhs, c.q,
                       R, ldr, D, ldd);
    EAMD_CHECK(trsv_device<double>(c.h, c.n, W, c.ldw, true, true, true, D, ldd, nrhs));
    EAMD_CHECK(trsv_device<double>(c.h, c.n, W, c.ldw, true, false, false, D, ldd, nrhs));
  } else {
    const float* W = static_cast<const float*>(c.W);
    hipLaunchKernelGGL((gather_rows_kernel<double, float>), g, dim3(256), 0, s, c.n, nrhs, c.q,
                       R, ldr, c.Y32, (int64_t)c.n);
    EAMD_CHECK(trsv_device<float>(c.h, c.n, W, c.ldw, true, true, true, c.Y32, c.n, nrhs));
    EAMD_CHECK(trsv_device<float>(c.h, c.n, W, c.ldw, true, false, false, c.Y32, c.n, nrhs));
    hipLaunchKernelGGL((convert_kernel<float, double>), g, dim3(256), 0, s, c.n, nrhs, c.Y32,
                       (int64_t)c.n, D, ldd);
  }
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// per-rhs backward errors (host) of the residual R for X, B
int backward_errors(eamd_handle* h, int n, int nrhs, const double* R, const double* X,
                    int64_t ldx, const double* B, int64_t ldb, double anorm, double* nd,
                    std::vector<double>& out) {
  EAMD_CHECK(inf_norms(h, n, nrhs, R, n, X, ldx, B, ldb, nd));
  std::vector<double> hn(3 * (size_t)nrhs);
  EAMD_HIP(hipMemcpyAsync(hn.data(), nd, sizeof(double) * 3 * nrhs, hipMemcpyDeviceToHost,
                          h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  out.resize(nrhs);
  for (int c = 0; c < nrhs; ++c) {
    double den = anorm * hn[3 * c + 1] + hn[3 * c + 2];
    out[c] = den > 0 ? hn[3 * c] / den : hn[3 * c];
  }
  return kOk;
}

// One GMRES(m) cycle on A d = r (single right-hand side, r overwritten as
// scratch), right preconditioned by the LU: d = Z y with Z = M^{-1} V.
int gmres_cycle(IrContext& c, const double* r, double* d, int m, double rtol, int* iters,
                double* Vbuf, double* Zbuf, double* dev_small) {
  eamd_handle* h = c.h;
  hipStream_t s = h->stream;
  const int n = c.n;
  const int64_t ldv = n;
  const dim3 g1(ceil_div(n, 256));
  // beta = ||r||_2
  hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(256), 0, s, n, r, ldv, r, dev_small);
  double beta2 = 0;
  EAMD_HIP(hipMemcpyAsync(&beta2, dev_small, sizeof(double), hipMemcpyDeviceToHost, s));
  EAMD_HIP(hipStreamSynchronize(s));
  const double beta = std::sqrt(beta2);
  if (!(beta > 0)) {
    EAMD_HIP(hipMemsetAsync(d, 0, sizeof(double) * n, s));
    *iters = 0;
    return kOk;
  }
  hipLaunchKernelGGL(scale_copy_kernel, g1, dim3(256), 0, s, n, 1.0 / beta, r, Vbuf);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1, 0.0);
  gv[0] = beta;
  int k = 0;
  std::vector<double> hcol(m + 1), h2(m + 1);
  for (; k < m; ++k) {
    double* vk = Vbuf + (size_t)k * ldv;
    double* zk = Zbuf + (size_t)k * ldv;
    double* w = Vbuf + (size_t)(k + 1) * ldv;
    EAMD_CHECK(lu_apply_inverse(c, vk, ldv, zk, ldv, 1));
    // w = A z  (as -(0 - A z))
    EAMD_CHECK(residual_dd(h, n, 1, c.A, c.lda, zk, ldv, nullptr, 0, w, ldv));
    hipLaunchKernelGGL(scale_copy_kernel, g1, dim3(256), 0, s, n, -1.0, w, w);
    // classical Gram-Schmidt, twice
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(multidot_kernel, dim3(k + 1), dim3(256), 0, s, n, Vbuf, ldv, w,
                         dev_small);
      hipLaunchKernelGGL(vmul_sub_kernel, g1, dim3(256), 0, s, n, k + 1, Vbuf, ldv, dev_small,
                         w);
      EAMD_HIP(hipMemcpyAsync(pass ? h2.data() : hcol.data(), dev_small,
                              sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
    }
    hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(256), 0, s, n, w, ldv, w, dev_small + 300);
    double nw2 = 0;
    EAMD_HIP(hipMemcpyAsync(&nw2, dev_small + 300, sizeof(double), hipMemcpyDeviceToHost, s));
    EAMD_HIP(hipStreamSynchronize(s));
    for (int i = 0; i <= k; ++i) hcol[i] += h2[i];
    double hk1 = std::sqrt(nw2);
    hcol[k + 1] = hk1;
    if (hk1 > 0)
      hipLaunchKernelGGL(scale_copy_kernel, g1, dim3(256), 0, s, n, 1.0 / hk1, w, w);
    // Givens rotations
    for (int i = 0; i < k; ++i) {
      double t = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
      hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
      hcol[i] = t;
    }
    double den = std::hypot(hcol[k], hcol[k + 1]);
    cs[k] = den > 0 ? hcol[k] / den : 1.0;
    sn[k] = den > 0 ? hcol[k + 1] / den : 0.0;
    hcol[k] = den;
    hcol[k + 1] = 0.0;
    gv[k + 1] = -sn[k] * gv[k];
    gv[k] = cs[k] * gv[k];
    for (int i = 0; i <= k; ++i) H[(size_t)k * (m + 1) + i] = hcol[i];
    if (std::fabs(gv[k + 1]) <= rtol * beta || hk1 == 0.0) {
      ++k;
      break;
    }
  }
  // back substitution H y = g, then d = Z y
  std::vector<double> y(k, 0.0);
  for (int i = k - 1; i >= 0; --i) {
    double sacc = gv[i];
    for (int j = i + 1; j < k; ++j) sacc -= H[(size_t)j * (m + 1) + i] * y[j];
    double hii = H[(size_t)i * (m + 1) + i];
    y[i] = hii != 0 ? sacc / hii : 0.0;
  }
  EAMD_HIP(hipMemsetAsync(d, 0, sizeof(double) * n, s));
  // d = -( 0 - Z y ): reuse vmul_sub with negated coefficients
  for (int i = 0; i < k; ++i) y[i] = -y[i];
  EAMD_HIP(hipMemcpyAsync(dev_small, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(vmul_sub_kernel, g1, dim3(256), 0, s, n, k, Zbuf, ldv, dev_small, d);
  EAMD_HIP(hipGetLastError());
  EAMD_HIP(hipStreamSynchronize(s));
  *iters = k;
  return kOk;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int getrf_rowmajor_order(eamd_handle* h, void* W, bool fp64, int64_t ldw, int n, int* piv,
                         int* ord, int64_t* first_zero, int nb, int precision);

int dgesv_ir_impl(eamd_handle* h, int64_t n64, int64_t nrhs64, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double* X, int64_t ldx, const eamd_ir_opts* o,
                  eamd_ir_report* rep) {
  if (n64 > (1 << 20) || nrhs64 > (1 << 16)) return kErrUnsupported;
  const int n = (int)n64, nrhs = (int)nrhs64;
  hipStream_t s = h->stream;
  const bool fp64 = o->factor_precision == EAMD_PREC_FP64;
  const int64_t ldw = ((n + 63) / 64) * 64;
  const int nb = o->nb > 0 ? o->nb : lu_block_width(n, fp64);
  auto t0 = std::chrono::steady_clock::now();
  EAMD_HIP(hipStreamSynchronize(s));
  t0 = std::chrono::steady_clock::now();

  // ---- workspace
  void* wp = nullptr;
  EAMD_CHECK(h->ws.get(kWsFactor32, (fp64 ? sizeof(double) : sizeof(float)) * (size_t)ldw * n,
                       &wp));
  void* op = nullptr;
  EAMD_CHECK(h->ws.get(kWsTmp, sizeof(int) * 2 * (size_t)n + 256, &op));
  int* ord = static_cast<int*>(op);
  int* piv = ord + n;
  const int m = std::max(1, std::min(o->gmres_restart > 0 ? o->gmres_restart : 30, 200));  // <= 255 dots
  const int ncap = std::max(nrhs, 1);
  size_t vec_need = sizeof(double) * ((size_t)n * (3 * ncap) + (size_t)n * 2 * (m + 1) + 1024) +
                    sizeof(float) * (size_t)n * ncap;
  void* vp = nullptr;
  EAMD_CHECK(h->ws.get(kWsGmres, vec_need, &vp));
  double* R = static_cast<double*>(vp);            // n x ncap residual
  double* D = R + (size_t)n * ncap;                // n x ncap correction
  double* V = D + (size_t)n * ncap;                // n x (m+1) Krylov basis
  double* Z = V + (size_t)n * (m + 1);             // n x (m+1) preconditioned basis
  double* small = Z + (size_t)n * (m + 1);         // 1024 doubles
  float* Y32 = reinterpret_cast<float*>(small + 1024);

  // ---- factorization
  if (fp64) {
    EAMD_CHECK(transpose_copy<double, double>(s, A, lda, n, n, static_cast<double*>(wp), ldw));
  } else {
    EAMD_CHECK(transpose_copy<double, float>(s, A, lda, n, n, static_cast<float*>(wp), ldw));
  }
  int64_t fz = -1;
  EAMD_CHECK(getrf_rowmajor_order(h, wp, fp64, ldw, n, piv, ord, &fz, nb, o->factor_precision));
  rep->first_zero_pivot = fz;
  EAMD_CHECK(matrix_inf_norm(h, n, A, lda, small + 512));
  double anorm = 0;
  EAMD_HIP(hipMemcpyAsync(&anorm, small + 512, sizeof(double), hipMemcpyDeviceToHost, s));
  EAMD_HIP(hipStreamSynchronize(s));
  rep->factor_ms = ms_since(t0);
  auto t1 = std::chrono::steady_clock::now();

  IrContext c{h, n, A, lda, wp, ldw, fp64, ord, Y32, nullptr, ncap};
  // ---- x0 = LU^{-1} P b
  for (int c0 = 0; c0 < nrhs; ++c0) {
    EAMD_CHECK(lu_apply_inverse(c, B + (int64_t)c0 * ldb, ldb, X + (int64_t)c0 * ldx, ldx, 1));
  }
  // ---- refinement, one right-hand side at a time (columns converge independently)
  double worst = 0.0;
  int total_ir = 0, total_gm = 0, used_gmres = 0;
  const double tol = o->tol > 0 ? o->tol : 1e-14;
  const int max_iter = o->max_iter > 0 ? o->max_iter : 30;
  const int gmax = o->gmres_max_iter > 0 ? o->gmres_max_iter : 300;
  std::vector<double> be;
  for (int col = 0; col < nrhs; ++col) {
    const double* b = B + (int64_t)col * ldb;
    double* x = X + (int64_t)col * ldx;
    double prev = INFINITY, eta = INFINITY;
    int slow = 0, it = 0, git = 0;
    bool use_gmres = o->gmres == 2;
    while (true) {
      EAMD_CHECK(residual_dd(h, n, 1, A, lda, x, ldx, b, ldb, R, n));
      EAMD_CHECK(backward_errors(h, n, 1, R, x, ldx, b, ldb, anorm, small + 600, be));
      eta = be[0];
      if (eta <= tol || !(eta == eta)) break;
      if (it + git >= max_iter + gmax) break;
      if (!use_gmres) {
        if (eta > 0.5 * prev) ++slow; else slow = 0;
        if ((slow >= 2 || eta > 2.0 * prev) && o->gmres >= 1) use_gmres = true;
        if (!use_gmres && it >= max_iter) break;
      }
      prev = std::min(prev, eta);
      if (!use_gmres) {
        EAMD_CHECK(lu_apply_inverse(c, R, n, D, n, 1));
        dim3 g(ceil_div(n, 256), 1);
        hipLaunchKernelGGL((add_kernel<double>), g, dim3(256), 0, s, n, 1, D, (int64_t)n, x, ldx);
        ++it;
      } else {
        int k = 0;
        EAMD_CHECK(gmres_cycle(c, R, D, m, 1e-10, &k, V, Z, small));
        dim3 g(ceil_div(n, 256), 1);
        hipLaunchKernelGGL((add_kernel<double>), g, dim3(256), 0, s, n, 1, D, (int64_t)n, x, ldx);
        git += std::max(k, 1);
        used_gmres = 1;
        if (git >= gmax) {
          EAMD_CHECK(residual_dd(h, n, 1, A, lda, x, ldx, b, ldb, R, n));
          EAMD_CHECK(backward_errors(h, n, 1, R, x, ldx, b, ldb, anorm, small + 600, be));
          eta = be[0];
          break;
        }
      }
    }
    total_ir += it;
    total_gm += git;
    if (eta > worst || eta != eta) worst = eta;
  }
  EAMD_CHECK(trsv_check_error(h));
  EAMD_HIP(hipStreamSynchronize(s));
  rep->refine_ms = ms_since(t1);
  rep->ir_iterations = total_ir;
  rep->gmres_iterations = total_gm;
  rep->used_gmres = used_gmres;
  rep->backward_error = worst;
  rep->converged = worst <= tol ? 1 : 0;
  if (!rep->converged) {
    set_error("iterative refinement did not reach %.3g (backward error %.3g)", tol, worst);
    return kErrNotConverged;
  }
  return kOk;
}

}  // namespace eamd
