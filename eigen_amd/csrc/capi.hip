// capi.hip -- extern "C" entry points of include/eigen_amd.h: argument
// checks in the reference's xerbla convention, handle/workspace lifetime,
// dispatch to the kernels, and the host-buffer variants used by the Eigen
// plug points.
#include "../../include/eigen_amd.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <mutex>

namespace eamd {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return e == hipErrorOutOfMemory ? kErrAlloc : kErrHip;
}

int Workspace::get(int slot, size_t need, void** out) {
  if (need == 0) need = 1;
  if (bytes[slot] < need) {
    if (ptr[slot]) {
      EAMD_HIP(hipDeviceSynchronize());
      EAMD_HIP(hipFree(ptr[slot]));
      ptr[slot] = nullptr;
      bytes[slot] = 0;
    }
    size_t sz = (need + 4095) & ~(size_t)4095;
    hipError_t e = hipMalloc(&ptr[slot], sz);
    if (e != hipSuccess) {
      ptr[slot] = nullptr;
      set_error("workspace allocation of %zu bytes failed: %s", sz, hipGetErrorString(e));
      return kErrAlloc;
    }
    bytes[slot] = sz;
  }
  *out = ptr[slot];
  return kOk;
}

// The look-ahead streams and their ordering events, created on first use.
// EAMD_PANEL_CUS > 0 asks for that many CUs (every k-th) reserved for the
// panel stream and masked out of the update stream; where the runtime
// refuses CU-masked streams both are plain streams.
int env_int(const char* name, int dflt);
int ensure_side_stream(eamd_handle* h) {
  if (h->side) return kOk;
  const int ncu = h->num_cus > 0 ? h->num_cus : 256;
  const int want = env_int("EAMD_PANEL_CUS", 0);
  if (want > 0 && want < ncu) {
    std::vector<uint32_t> pm((ncu + 31) / 32, 0u), um((ncu + 31) / 32, 0u);
    const int stride = ncu / want;
    int taken = 0;
    for (int c = 0; c < ncu; ++c) {
      const bool panel = c % stride == 0 && taken < want;
      taken += panel;
      (panel ? pm : um)[c / 32] |= 1u << (c % 32);
    }
    if (hipExtStreamCreateWithCUMask(&h->pstream, (uint32_t)pm.size(), pm.data()) != hipSuccess ||
        hipExtStreamCreateWithCUMask(&h->side, (uint32_t)um.size(), um.data()) != hipSuccess) {
      (void)hipGetLastError();
      if (h->pstream) (void)hipStreamDestroy(h->pstream);
      if (h->side) (void)hipStreamDestroy(h->side);
      h->pstream = h->side = nullptr;
    }
  }
  if (!h->side) EAMD_HIP(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
  for (hipEvent_t& e : h->la_ev) EAMD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return kOk;
}

hipEvent_t PhaseTimer::take() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void PhaseTimer::begin(hipStream_t s, int phase, hipEvent_t* a) {
  *a = take();
  if (*a) (void)hipEventRecord(*a, s);
}

void PhaseTimer::end(hipStream_t s, int phase, hipEvent_t a, double by, double fl) {
  hipEvent_t b = take();
  if (!b) return;
  (void)hipEventRecord(b, s);
  open.push_back(Rec{phase, a, b, by, fl});
  if (open.size() > 4096) (void)resolve();
}

int PhaseTimer::resolve() {
  for (auto& r : open) {
    EAMD_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    EAMD_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms[r.phase] += t;
    bytes[r.phase] += r.bytes;
    flops[r.phase] += r.flops;
    count[r.phase] += 1;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  open.clear();
  return kOk;
}

void PhaseTimer::reset() {
  (void)resolve();
  for (int i = 0; i < kPhCount; ++i) ms[i] = bytes[i] = flops[i] = 0, count[i] = 0;
}

PhaseTimer::~PhaseTimer() {
  for (auto& r : open) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : pool) (void)hipEventDestroy(e);
}

void Workspace::release() {
  for (int i = 0; i < kSlots; ++i) {
    if (ptr[i]) (void)hipFree(ptr[i]);
    ptr[i] = nullptr;
    bytes[i] = 0;
  }
}

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

int lu_block_width(int64_t n, bool fp64) {
  int nb = env_int("EAMD_LU_NB", 0);
  if (nb > 0) return nb;
  if (n <= 2048) return 64;
  // the FP64 factorization's trailing updates gain more from K = 256 than
  // its panels lose (n = 16384: 224.7 ms at 256, 244.0 ms at 128); the
  // mixed solver's FP32 panel is fastest at 128 (67.3 vs 72.0 ms at 96)
  if (fp64 && n >= 8192) return 256;
  return 128;
}

// defined in the other translation units
int dgesv_ir_impl(eamd_handle* h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double* X, int64_t ldx, const eamd_ir_opts* opts,
                  eamd_ir_report* rep);
int dbackward_error_impl(eamd_handle* h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                         const double* X, int64_t ldx, const double* B, int64_t ldb,
                         double* berr);
int spotrf_solve_batched_impl(eamd_handle* h, int64_t count, const int64_t* offA,
                              const int32_t* n, const float* A, const int64_t* offb, float* b,
                              int32_t* info);
int dgesv_ir_batched_impl(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                          const double* A, const int64_t* offB, int64_t nrhs, const double* B,
                          double* X, int32_t* info, int32_t* iters);
int batched_plan_create(eamd_handle* h, int64_t count, const int32_t* n, int max_n,
                        eamd_batched_plan** out);
int batched_plan_sort(eamd_handle* h, int64_t count, const int32_t* n, int max_n,
                      eamd_batched_plan** out);
int batched_lu_plan_create(eamd_handle* h, int64_t count, const int32_t* n, int nrhs,
                           eamd_batched_plan** out);
int dgesv_ir_batched_exec(eamd_handle* h, const eamd_batched_plan* p, const int64_t* offA,
                          const int32_t* n, const double* A, const int64_t* offB,
                          const double* B, double* X, int32_t* info, int32_t* iters);
void batched_plan_destroy(eamd_batched_plan* p);
int batched_lanes(const eamd_batched_plan* p, int n);
int spotrf_solve_batched_exec(eamd_handle* h, const eamd_batched_plan* p, const int64_t* offA,
                              const float* A, const int64_t* offb, float* b, int32_t* info);
int batched_noalloc_init(int device);
int spotrf_solve_batched_noalloc(eamd_handle* h, int64_t count, const int64_t* offA,
                                 const int32_t* n, const float* A, const int64_t* offb, float* b,
                                 int32_t* info);

namespace {

__global__ void count_nontrivial_kernel(const int* t, int n, int* out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) c += (t[i] != i);
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) *out = cnt;
}

__global__ void perm_from_transp_kernel(int n, const int* t, int* perm) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int* p = reinterpret_cast<int*>(smem_raw);
  const bool lds = n <= 36000;
  int* v = lds ? p : perm;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = n - 1; k >= 0; --k) {
      int q = t[k];
      int a = v[k];
      v[k] = v[q];
      v[q] = a;
    }
  __syncthreads();
  if (lds)
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = v[i];
}

bool is_op(char c) { return c == 'N' || c == 'n' || c == 'T' || c == 't' || c == 'C' || c == 'c'; }
bool is_trans(char c) { return !(c == 'N' || c == 'n'); }
bool is_uplo(char c) { return c == 'L' || c == 'l' || c == 'U' || c == 'u'; }
bool is_lower(char c) { return c == 'L' || c == 'l'; }
bool is_diag(char c) { return c == 'N' || c == 'n' || c == 'U' || c == 'u'; }
bool is_unit(char c) { return c == 'U' || c == 'u'; }
bool is_side(char c) { return c == 'L' || c == 'l' || c == 'R' || c == 'r'; }

// B = A converted (rows x cols, column-major, leading dimensions lda/ldb)
template <typename Ti, typename To>
__global__ void convert2d_kernel(int64_t rows, int64_t cols, const Ti* __restrict__ A,
                                 int64_t lda, To* __restrict__ B, int64_t ldb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int64_t i = idx % rows, j = idx / rows;
  B[j * ldb + i] = (To)A[j * lda + i];
}

template <typename Ti, typename To>
int convert2d(hipStream_t s, int64_t rows, int64_t cols, const Ti* A, int64_t lda, To* B,
              int64_t ldb) {
  const int64_t blocks = (rows * cols + 255) / 256;
  if (blocks <= 0) return kOk;
  if (blocks > INT32_MAX) {
    set_error("convert: %lld x %lld exceeds one launch", (long long)rows, (long long)cols);
    return kErrUnsupported;
  }
  hipLaunchKernelGGL((convert2d_kernel<Ti, To>), dim3((unsigned)blocks), dim3(256), 0, s, rows,
                     cols, A, lda, B, ldb);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

__global__ void scale_kernel(int m, int n, double alpha, double* B, int64_t ldb) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  int i = (int)(idx % m), j = (int)(idx / m);
  double* p = B + (int64_t)j * ldb + i;
  *p = alpha == 0.0 ? 0.0 : alpha * *p;
}

}  // namespace
}  // namespace eamd

using namespace eamd;

#define CHECK_HANDLE(h)                 \
  do {                                  \
    if (!(h)) {                         \
      set_error("null handle");         \
      return -1;                        \
    }                                   \
  } while (0)

extern "C" {

int eamd_version(void) { return EAMD_VERSION; }

const char* eamd_last_error(void) { return g_err; }

int eamd_create(eamd_handle_t* handle) {
  if (!handle) return -1;
  *handle = nullptr;
  int dev = 0;
  EAMD_HIP(hipGetDevice(&dev));
  eamd_handle* h = new eamd_handle();
  h->device = dev;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipGetDeviceProperties", __FILE__, __LINE__);
  }
  h->num_cus = prop.multiProcessorCount;
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
  }
  h->own_stream = true;
  // the allocation-free entry point's scratch address on this device
  const int st = batched_noalloc_init(dev);
  if (st != kOk) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return st;
  }
  *handle = h;
  return kOk;
}

int eamd_destroy(eamd_handle_t h) {
  if (!h) return kOk;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->side) {
    (void)hipStreamSynchronize(h->side);
    (void)hipStreamDestroy(h->side);
    if (h->pstream) {
      (void)hipStreamSynchronize(h->pstream);
      (void)hipStreamDestroy(h->pstream);
    }
    for (hipEvent_t e : h->la_ev)
      if (e) (void)hipEventDestroy(e);
  }
  h->ws.release();
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return kOk;
}

int eamd_set_stream(eamd_handle_t h, void* s) {
  CHECK_HANDLE(h);
  hipStream_t ns = static_cast<hipStream_t>(s);
  if (ns == h->stream && !h->own_stream) return kOk;
  if (h->stream) EAMD_HIP(hipStreamSynchronize(h->stream));
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  // NULL selects the default (null) stream, as hipblasSetStream does
  h->stream = ns;
  h->own_stream = false;
  return kOk;
}

void* eamd_get_stream(eamd_handle_t h) { return h ? (void*)h->stream : nullptr; }

int eamd_synchronize(eamd_handle_t h) {
  CHECK_HANDLE(h);
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_profile_enable(eamd_handle_t h, int on) {
  CHECK_HANDLE(h);
  h->prof.reset();
  h->prof.enabled = on != 0;
  return kOk;
}

int eamd_profile_read(eamd_handle_t h, int phase, double* ms, double* bytes, double* flops,
                      int64_t* count) {
  CHECK_HANDLE(h);
  if (phase < 0 || phase >= kPhCount) return -2;
  EAMD_CHECK(h->prof.resolve());
  if (ms) *ms = h->prof.ms[phase];
  if (bytes) *bytes = h->prof.bytes[phase];
  if (flops) *flops = h->prof.flops[phase];
  if (count) *count = h->prof.count[phase];
  return kOk;
}

int eamd_release_workspace(eamd_handle_t h) {
  CHECK_HANDLE(h);
  EAMD_HIP(hipStreamSynchronize(h->stream));
  h->ws.release();
  h->panel_epoch = 0;
  h->trsv_epoch = 0;
  return kOk;
}

// ------------------------------------------------------------------ level 3
int eamd_dgemm(eamd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
               double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
               double beta, double* C, int64_t ldc) {
  // argument numbering follows dgemm_ (blas/level3_impl.h:47-57)
  if (!is_op(transa)) return -1;
  if (!is_op(transb)) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (k < 0) return -5;
  const bool ta = is_trans(transa), tb = is_trans(transb);
  if (lda < std::max<int64_t>(1, ta ? k : m)) return -8;
  if (ldb < std::max<int64_t>(1, tb ? n : k)) return -10;
  if (ldc < std::max<int64_t>(1, m)) return -13;
  CHECK_HANDLE(h);
  PhaseScope ps(h, kPhGemm, 8.0 * ((ta ? k * m : m * k) + (double)k * n + 2.0 * m * n),
                2.0 * m * n * k);
  return dgemm_device(h->stream, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

int eamd_dsyrk(eamd_handle_t h, char uplo, char trans, int64_t n, int64_t k, double alpha,
               const double* A, int64_t lda, double beta, double* C, int64_t ldc) {
  if (!is_uplo(uplo)) return -1;
  if (!is_op(trans)) return -2;
  if (n < 0) return -3;
  if (k < 0) return -4;
  const bool t = is_trans(trans);
  if (lda < std::max<int64_t>(1, t ? k : n)) return -7;
  if (ldc < std::max<int64_t>(1, n)) return -10;
  CHECK_HANDLE(h);
  // op(A) op(A)^T: 'N' -> A * A^T, 'T' -> A^T * A
  return dgemm_device(h->stream, t, !t, n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                      is_lower(uplo) ? 1 : 2);
}

int eamd_dtrsm(eamd_handle_t h, char side, char uplo, char transa, char diag, int64_t m,
               int64_t n, double alpha, const double* A, int64_t lda, double* B, int64_t ldb) {
  if (!is_side(side)) return -1;
  if (!is_uplo(uplo)) return -2;
  if (!is_op(transa)) return -3;
  if (!is_diag(diag)) return -4;
  if (m < 0) return -5;
  if (n < 0) return -6;
  const bool left = (side == 'L' || side == 'l');
  const int64_t na = left ? m : n;
  if (lda < std::max<int64_t>(1, na)) return -9;
  if (ldb < std::max<int64_t>(1, m)) return -11;
  CHECK_HANDLE(h);
  if (m == 0 || n == 0) return kOk;
  if (m > INT32_MAX || n > INT32_MAX) return kErrUnsupported;
  hipStream_t s = h->stream;
  if (alpha != 1.0) {
    int64_t tot = m * n;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       (int)m, (int)n, alpha, B, ldb);
    EAMD_HIP(hipGetLastError());
    if (alpha == 0.0) return kOk;
  }
  const bool tr = is_trans(transa);
  // op(A) = A^T reads A's storage as row-major; its triangle flips.
  const bool rowmajor = tr;
  const bool eff_lower = is_lower(uplo) != tr;
  if (left) {
    EAMD_CHECK(trsv_device<double>(h, (int)m, A, lda, rowmajor, eff_lower, is_unit(diag), B, ldb,
                                   (int)n));
  } else {
    // X op(A) = B  <=>  op(A)^T X^T = B^T
    void* t = nullptr;
    EAMD_CHECK(h->ws.get(kWsTmp, sizeof(double) * (size_t)m * n, &t));
    double* Bt = static_cast<double*>(t);
    // Bt (n x m col-major) = B^T
    EAMD_CHECK(transpose_copy<double, double>(s, B, ldb, (int)m, (int)n, Bt, n));
    // op(A)^T: A^T if not transposed (row-major view), A otherwise
    EAMD_CHECK(trsv_device<double>(h, (int)n, A, lda, !rowmajor, !eff_lower, is_unit(diag), Bt,
                                   n, (int)m));
    EAMD_CHECK(transpose_copy<double, double>(s, Bt, n, (int)n, (int)m, B, ldb));
  }
  return trsv_check_error(h);
}

// ----------------------------------------------------------------------- LU
int eamd_dgetrf(eamd_handle_t h, int64_t m, int64_t n, double* A, int64_t lda, int32_t* transp,
                int32_t* nb_transp, int64_t* first_zero_pivot) {
  // argument numbering follows dgetrf_ (lapack/lu.cpp:16-19)
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, m)) return -4;
  CHECK_HANDLE(h);
  if (nb_transp) *nb_transp = 0;
  if (first_zero_pivot) *first_zero_pivot = -1;
  if (m == 0 || n == 0) return kOk;
  if (m > 1000000 || n > 1000000) return kErrUnsupported;
  hipStream_t s = h->stream;
  const int64_t size = std::min(m, n);
  // A wide matrix whose min(m, n) exceeds the reference's unblocked limit of
  // 16 is factored by its blocked loop over the first m columns only: the
  // trailing size there is min(m, n) - k - bs (PartialPivLU.h:450-452,
  // 464-465), so columns m..n-1 keep their input values, neither permuted
  // nor solved.  The same columns are left alone here; up to 16 the
  // unblocked loop (PartialPivLU.h:365-409) updates every column, and so
  // does this driver.
  const int64_t ncols = (m < n && size > 16) ? m : n;
  const int M = (int)m, NC = (int)ncols;
  const int64_t ldw = ((ncols + 15) / 16) * 16;
  void* wp = nullptr;
  EAMD_CHECK(h->ws.get(kWsWork, sizeof(double) * (size_t)ldw * m, &wp));
  double* W = static_cast<double*>(wp);
  // W (row-major, m x ncols) = A
  EAMD_CHECK(transpose_copy<double, double>(s, A, lda, M, NC, W, ldw));
  int64_t fz = -1;
  EAMD_CHECK(dgetrf_rowmajor(h, W, ldw, M, NC, transp, &fz, lu_block_width(size, true)));
  EAMD_CHECK(transpose_copy<double, double>(s, W, ldw, NC, M, A, lda));
  void* cnt = nullptr;
  EAMD_CHECK(h->ws.get(kWsAux2, 256, &cnt));
  hipLaunchKernelGGL(count_nontrivial_kernel, dim3(1), dim3(1024), 0, s, transp, (int)size,
                     static_cast<int*>(cnt));
  int hc = 0;
  EAMD_HIP(hipMemcpyAsync(&hc, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
  EAMD_HIP(hipStreamSynchronize(s));
  if (nb_transp) *nb_transp = hc;
  if (first_zero_pivot) *first_zero_pivot = fz;
  return kOk;
}

int eamd_dgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const double* LU,
                int64_t lda, const int32_t* transp, double* B, int64_t ldb) {
  // argument numbering follows dgetrs_ (lapack/lu.cpp:50-55)
  if (!is_op(trans)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -8;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  const int N = (int)n, R = (int)nrhs;
  if (!is_trans(trans)) {
    EAMD_CHECK(apply_transpositions<double>(h, N, transp, true, B, ldb, R));
    EAMD_CHECK(trsv_device<double>(h, N, LU, lda, false, true, true, B, ldb, R));
    EAMD_CHECK(trsv_device<double>(h, N, LU, lda, false, false, false, B, ldb, R));
  } else {
    // A^T = U^T L^T P: U^T (lower, row-major view), L^T (upper, unit), P^T
    EAMD_CHECK(trsv_device<double>(h, N, LU, lda, true, true, false, B, ldb, R));
    EAMD_CHECK(trsv_device<double>(h, N, LU, lda, true, false, true, B, ldb, R));
    EAMD_CHECK(apply_transpositions<double>(h, N, transp, false, B, ldb, R));
  }
  return trsv_check_error(h);
}

int eamd_transpositions_to_permutation(eamd_handle_t h, int64_t n, const int32_t* transp,
                                       int32_t* perm) {
  if (n < 0) return -2;
  CHECK_HANDLE(h);
  if (n == 0) return kOk;
  size_t lds = n <= 36000 ? sizeof(int) * (size_t)n : 0;
  EAMD_HIP(hipFuncSetAttribute((const void*)perm_from_transp_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
  hipLaunchKernelGGL(perm_from_transp_kernel, dim3(1), dim3(256), lds, h->stream, (int)n, transp,
                     perm);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// ---------------------------------------------------------------------- LLT
int eamd_dpotrf(eamd_handle_t h, char uplo, int64_t n, double* A, int64_t lda,
                int64_t* fail_col) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, n)) return -4;
  CHECK_HANDLE(h);
  if (fail_col) *fail_col = -1;
  if (n == 0) return kOk;
  const int N = (int)n;
  int64_t fc = -1;
  int nb = env_int("EAMD_LLT_NB", 128);
  if (is_lower(uplo)) {
    EAMD_CHECK(dpotrf_lower(h, N, A, lda, &fc, nb));
  } else {
    // Upper: the Lower algorithm on the transpose (LLT.h:369-391)
    void* wp = nullptr;
    EAMD_CHECK(h->ws.get(kWsWork, sizeof(double) * (size_t)n * n, &wp));
    double* W = static_cast<double*>(wp);
    EAMD_CHECK(transpose_copy<double, double>(h->stream, A, lda, N, N, W, n));
    EAMD_CHECK(dpotrf_lower(h, N, W, n, &fc, nb));
    EAMD_CHECK(transpose_copy<double, double>(h->stream, W, n, N, N, A, lda));
  }
  if (fail_col) *fail_col = fc;
  return kOk;
}

int eamd_dpotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const double* A,
                int64_t lda, double* B, int64_t ldb) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -7;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  const int N = (int)n, R = (int)nrhs;
  if (is_lower(uplo)) {
    EAMD_CHECK(trsv_device<double>(h, N, A, lda, false, true, false, B, ldb, R));
    // L^T: the column-major lower factor read as row-major is upper
    EAMD_CHECK(trsv_device<double>(h, N, A, lda, true, false, false, B, ldb, R));
  } else {
    EAMD_CHECK(trsv_device<double>(h, N, A, lda, true, true, false, B, ldb, R));
    EAMD_CHECK(trsv_device<double>(h, N, A, lda, false, false, false, B, ldb, R));
  }
  return trsv_check_error(h);
}

// ------------------------------------------------- FP32 single matrices
// FP32 storage, FP64 arithmetic: widen into the handle's conversion buffer,
// run the FP64 routine, round the results back.  Every entry point returns
// with its outputs complete, as the FP64 ones do.
int eamd_sgetrf(eamd_handle_t h, int64_t m, int64_t n, float* A, int64_t lda, int32_t* transp,
                int32_t* nb_transp, int64_t* first_zero_pivot) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, m)) return -4;
  CHECK_HANDLE(h);
  if (nb_transp) *nb_transp = 0;
  if (first_zero_pivot) *first_zero_pivot = -1;
  if (m == 0 || n == 0) return kOk;
  if (m > 1000000 || n > 1000000) return kErrUnsupported;
  void* cp = nullptr;
  EAMD_CHECK(h->ws.get(kWsConv, sizeof(double) * (size_t)m * n, &cp));
  double* D = static_cast<double*>(cp);
  EAMD_CHECK(convert2d<float, double>(h->stream, m, n, A, lda, D, m));
  EAMD_CHECK(eamd_dgetrf(h, m, n, D, m, transp, nb_transp, first_zero_pivot));
  EAMD_CHECK(convert2d<double, float>(h->stream, m, n, D, m, A, lda));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_sgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const float* LU,
                int64_t lda, const int32_t* transp, float* B, int64_t ldb) {
  if (!is_op(trans)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -8;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  void* cp = nullptr;
  EAMD_CHECK(h->ws.get(kWsConv, sizeof(double) * (size_t)n * (n + nrhs), &cp));
  double* D = static_cast<double*>(cp);
  double* DB = D + (size_t)n * n;
  EAMD_CHECK(convert2d<float, double>(h->stream, n, n, LU, lda, D, n));
  EAMD_CHECK(convert2d<float, double>(h->stream, n, nrhs, B, ldb, DB, n));
  EAMD_CHECK(eamd_dgetrs(h, trans, n, nrhs, D, n, transp, DB, n));
  EAMD_CHECK(convert2d<double, float>(h->stream, n, nrhs, DB, n, B, ldb));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_spotrf(eamd_handle_t h, char uplo, int64_t n, float* A, int64_t lda,
                int64_t* fail_col) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, n)) return -4;
  CHECK_HANDLE(h);
  if (fail_col) *fail_col = -1;
  if (n == 0) return kOk;
  void* cp = nullptr;
  EAMD_CHECK(h->ws.get(kWsConv, sizeof(double) * (size_t)n * n, &cp));
  double* D = static_cast<double*>(cp);
  EAMD_CHECK(convert2d<float, double>(h->stream, n, n, A, lda, D, n));
  EAMD_CHECK(eamd_dpotrf(h, uplo, n, D, n, fail_col));
  // the triangle not referenced keeps its values (FP32 -> FP64 -> FP32 is exact)
  EAMD_CHECK(convert2d<double, float>(h->stream, n, n, D, n, A, lda));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_spotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const float* A,
                int64_t lda, float* B, int64_t ldb) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -7;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  void* cp = nullptr;
  EAMD_CHECK(h->ws.get(kWsConv, sizeof(double) * (size_t)n * (n + nrhs), &cp));
  double* D = static_cast<double*>(cp);
  double* DB = D + (size_t)n * n;
  EAMD_CHECK(convert2d<float, double>(h->stream, n, n, A, lda, D, n));
  EAMD_CHECK(convert2d<float, double>(h->stream, n, nrhs, B, ldb, DB, n));
  EAMD_CHECK(eamd_dpotrs(h, uplo, n, nrhs, D, n, DB, n));
  EAMD_CHECK(convert2d<double, float>(h->stream, n, nrhs, DB, n, B, ldb));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

// ------------------------------------------------------------ mixed / IR
void eamd_ir_opts_default(eamd_ir_opts* o) {
  if (!o) return;
  o->factor_precision = EAMD_PREC_FP16X3;
  o->max_iter = 30;
  o->tol = 1e-14;
  o->gmres = 1;
  o->gmres_restart = 30;
  o->gmres_max_iter = 300;
  o->nb = 0;
}

int eamd_dgesv_ir(eamd_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double* X, int64_t ldx, const eamd_ir_opts* opts,
                  eamd_ir_report* report) {
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -7;
  if (ldx < std::max<int64_t>(1, n)) return -9;
  CHECK_HANDLE(h);
  eamd_ir_opts o;
  if (opts) o = *opts; else eamd_ir_opts_default(&o);
  eamd_ir_report r;
  memset(&r, 0, sizeof(r));
  r.first_zero_pivot = -1;
  int st = kOk;
  if (n > 0 && nrhs > 0) st = dgesv_ir_impl(h, n, nrhs, A, lda, B, ldb, X, ldx, &o, &r);
  if (report) *report = r;
  return st;
}

int eamd_dbackward_error(eamd_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                         const double* X, int64_t ldx, const double* B, int64_t ldb,
                         double* berr) {
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldx < std::max<int64_t>(1, n)) return -7;
  if (ldb < std::max<int64_t>(1, n)) return -9;
  if (!berr) return -10;
  CHECK_HANDLE(h);
  *berr = 0.0;
  if (n == 0 || nrhs == 0) return kOk;
  return dbackward_error_impl(h, n, nrhs, A, lda, X, ldx, B, ldb, berr);
}

// -------------------------------------------------------------- batched
int eamd_dllt_rank_update_batched(eamd_handle_t h, int64_t count, const int64_t* offL,
                                  const int32_t* n, double* L, const int64_t* offv,
                                  const double* v, const double* sigma, int32_t upper,
                                  int32_t* info) {
  if (count < 0) return -2;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offL) return -3;
  if (!n) return -4;
  if (!L) return -5;
  if (!offv) return -6;
  if (!v) return -7;
  if (!info) return -10;
  return llt_rank_update_batched(h, count, offL, n, L, offv, v, sigma, upper ? 1 : 0, info);
}

int eamd_dldlt_batched(eamd_handle_t h, int64_t count, const int64_t* offA, const int32_t* n,
                       int32_t max_n, double* A, const int64_t* offT, int32_t* transpositions,
                       int32_t* sign, int32_t* info) {
  if (count < 0) return -2;
  if (max_n < 0 || max_n > 64) return -5;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -6;
  if (!offT) return -7;
  if (!transpositions) return -8;
  if (!sign) return -9;
  if (!info) return -10;
  return ldlt_batched(h, count, offA, n, max_n, A, offT, transpositions, sign, info);
}

int eamd_dldlt_solve_batched(eamd_handle_t h, int64_t count, const int64_t* offA,
                             const int32_t* n, int32_t max_n, const double* A,
                             const int64_t* offT, const int32_t* transpositions,
                             const int64_t* offb, double* b) {
  if (count < 0) return -2;
  if (max_n < 0 || max_n > 64) return -5;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -6;
  if (!offT) return -7;
  if (!transpositions) return -8;
  if (!offb) return -9;
  if (!b) return -10;
  return ldlt_solve_batched(h, count, offA, n, max_n, A, offT, transpositions, offb, b);
}

int eamd_dldlt_rank_update_batched(eamd_handle_t h, int64_t count, const int64_t* offA,
                                   const int32_t* n, double* A, const int64_t* offT,
                                   const int32_t* transpositions, const int64_t* offw,
                                   const double* w, const double* sigma, int32_t* info) {
  if (count < 0) return -2;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -5;
  if (!offT) return -6;
  if (!transpositions) return -7;
  if (!offw) return -8;
  if (!w) return -9;
  if (!info) return -11;
  return ldlt_rank_update_batched(h, count, offA, n, A, offT, transpositions, offw, w, sigma,
                                  info);
}

int eamd_spotrf_solve_batched(eamd_handle_t h, int64_t count, const int64_t* offA,
                              const int32_t* n, const float* A, const int64_t* offb, float* b,
                              int32_t* info) {
  if (count < 0) return -2;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -5;
  if (!offb) return -6;
  if (!b) return -7;
  if (!info) return -8;
  return spotrf_solve_batched_impl(h, count, offA, n, A, offb, b, info);
}

int eamd_spotrf_solve_batched_noalloc(eamd_handle_t h, int64_t count, const int64_t* offA,
                                      const int32_t* n, const float* A, const int64_t* offb,
                                      float* b, int32_t* info) {
  if (count < 0) return -2;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -5;
  if (!offb) return -6;
  if (!b) return -7;
  if (!info) return -8;
  return spotrf_solve_batched_noalloc(h, count, offA, n, A, offb, b, info);
}

int eamd_batched_plan_create(eamd_handle_t h, int64_t count, const int32_t* n, int32_t max_n,
                             eamd_batched_plan_t* plan) {
  if (count < 0) return -2;
  if (!n && count > 0) return -3;
  if (max_n < 1 || max_n > 256) return -4;
  if (!plan) return -5;
  CHECK_HANDLE(h);
  return batched_plan_create(h, count, n, max_n, plan);
}

int eamd_batched_plan_destroy(eamd_batched_plan_t plan) {
  batched_plan_destroy(plan);
  return kOk;
}

int eamd_batched_plan_lanes(eamd_batched_plan_t plan, int32_t n) {
  return plan ? batched_lanes(plan, n) : 0;
}

int eamd_spotrf_solve_batched_planned(eamd_handle_t h, eamd_batched_plan_t plan,
                                      const int64_t* offA, const float* A, const int64_t* offb,
                                      float* b, int32_t* info) {
  if (!plan) return -2;
  if (plan->count > 0 && (!offA || !A || !offb || !b || !info)) return -3;
  CHECK_HANDLE(h);
  if (plan->max_n > 32) {
    set_error("plan built for sizes above 32: the batched LLT supports n <= 32");
    return kErrUnsupported;
  }
  return spotrf_solve_batched_exec(h, plan, offA, A, offb, b, info);
}

int eamd_dgesv_ir_batched(eamd_handle_t h, int64_t count, const int64_t* offA, const int32_t* n,
                          const double* A, const int64_t* offB, int64_t nrhs, const double* B,
                          double* X, int32_t* info, int32_t* iters) {
  if (count < 0) return -2;
  if (nrhs < 1 || nrhs > 8) return -7;
  CHECK_HANDLE(h);
  if (count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -5;
  if (!offB) return -6;
  if (!B) return -8;
  if (!X) return -9;
  if (!info) return -10;
  return dgesv_ir_batched_impl(h, count, offA, n, A, offB, nrhs, B, X, info, iters);
}

int eamd_batched_lu_plan_create(eamd_handle_t h, int64_t count, const int32_t* n, int64_t nrhs,
                                eamd_batched_plan_t* plan) {
  if (!plan) return -5;
  *plan = nullptr;
  if (count < 0) return -2;
  if (count > 0 && !n) return -3;
  if (nrhs < 1 || nrhs > 8) return -4;
  CHECK_HANDLE(h);
  return batched_lu_plan_create(h, count, n, (int)nrhs, plan);
}

int eamd_dgesv_ir_batched_planned(eamd_handle_t h, eamd_batched_plan_t plan,
                                  const int64_t* offA, const int32_t* n, const double* A,
                                  const int64_t* offB, const double* B, double* X,
                                  int32_t* info, int32_t* iters) {
  if (!plan) return -2;
  CHECK_HANDLE(h);
  if (plan->count == 0) return kOk;
  if (!offA) return -3;
  if (!n) return -4;
  if (!A) return -5;
  if (!offB) return -6;
  if (!B) return -7;
  if (!X) return -8;
  if (!info) return -9;
  return dgesv_ir_batched_exec(h, plan, offA, n, A, offB, B, X, info, iters);
}

// ------------------------------------------------------------ device memory
int eamd_malloc(void** ptr, int64_t bytes) {
  if (!ptr) return -1;
  if (bytes < 0) return -2;
  EAMD_HIP(hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 1));
  return kOk;
}

int eamd_free(void* ptr) {
  if (ptr) EAMD_HIP(hipFree(ptr));
  return kOk;
}

int eamd_memcpy_htod(eamd_handle_t h, void* dst, const void* src, int64_t bytes) {
  CHECK_HANDLE(h);
  EAMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_memcpy_dtoh(eamd_handle_t h, void* dst, const void* src, int64_t bytes) {
  CHECK_HANDLE(h);
  EAMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

// ------------------------------------------------------- host-buffer variants
}  // extern "C"

namespace {
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    EAMD_HIP(hipMalloc(&p, bytes ? bytes : 1));
    return kOk;
  }
};

template <typename T>
int up2d(eamd_handle* h, T* d, int64_t ldd, const T* src, int64_t lds, int64_t rows,
         int64_t cols) {
  EAMD_HIP(hipMemcpy2DAsync(d, ldd * sizeof(T), src, lds * sizeof(T), rows * sizeof(T), cols,
                            hipMemcpyHostToDevice, h->stream));
  return kOk;
}

template <typename T>
int down2d(eamd_handle* h, T* dst, int64_t ldd, const T* d, int64_t lds, int64_t rows,
           int64_t cols) {
  EAMD_HIP(hipMemcpy2DAsync(dst, ldd * sizeof(T), d, lds * sizeof(T), rows * sizeof(T), cols,
                            hipMemcpyDeviceToHost, h->stream));
  return kOk;
}
}  // namespace

extern "C" {

int eamd_host_dgetrf(eamd_handle_t h, int64_t m, int64_t n, double* A, int64_t lda,
                     int32_t* transp, int32_t* nb_transp, int64_t* first_zero_pivot) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, m)) return -4;
  CHECK_HANDLE(h);
  if (m == 0 || n == 0) return kOk;
  DevBuf dA, dt;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)m * n));
  EAMD_CHECK(dt.alloc(sizeof(int32_t) * (size_t)std::min(m, n)));
  EAMD_CHECK(up2d(h, (double*)dA.p, m, A, lda, m, n));
  EAMD_CHECK(eamd_dgetrf(h, m, n, (double*)dA.p, m, (int32_t*)dt.p, nb_transp, first_zero_pivot));
  EAMD_CHECK(down2d(h, A, lda, (double*)dA.p, m, m, n));
  EAMD_HIP(hipMemcpyAsync(transp, dt.p, sizeof(int32_t) * (size_t)std::min(m, n),
                          hipMemcpyDeviceToHost, h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const double* LU,
                     int64_t lda, const int32_t* transp, double* B, int64_t ldb) {
  if (!is_op(trans)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -8;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  DevBuf dA, dB, dt;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)n * n));
  EAMD_CHECK(dB.alloc(sizeof(double) * (size_t)n * nrhs));
  EAMD_CHECK(dt.alloc(sizeof(int32_t) * (size_t)n));
  EAMD_CHECK(up2d(h, (double*)dA.p, n, LU, lda, n, n));
  EAMD_CHECK(up2d(h, (double*)dB.p, n, B, ldb, n, nrhs));
  EAMD_HIP(hipMemcpyAsync(dt.p, transp, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                          h->stream));
  EAMD_CHECK(eamd_dgetrs(h, trans, n, nrhs, (double*)dA.p, n, (int32_t*)dt.p, (double*)dB.p, n));
  EAMD_CHECK(down2d(h, B, ldb, (double*)dB.p, n, n, nrhs));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dpotrf(eamd_handle_t h, char uplo, int64_t n, double* A, int64_t lda,
                     int64_t* fail_col) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, n)) return -4;
  CHECK_HANDLE(h);
  if (n == 0) {
    if (fail_col) *fail_col = -1;
    return kOk;
  }
  DevBuf dA;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)n * n));
  EAMD_CHECK(up2d(h, (double*)dA.p, n, A, lda, n, n));
  EAMD_CHECK(eamd_dpotrf(h, uplo, n, (double*)dA.p, n, fail_col));
  EAMD_CHECK(down2d(h, A, lda, (double*)dA.p, n, n, n));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dpotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const double* A,
                     int64_t lda, double* B, int64_t ldb) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -7;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  DevBuf dA, dB;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)n * n));
  EAMD_CHECK(dB.alloc(sizeof(double) * (size_t)n * nrhs));
  EAMD_CHECK(up2d(h, (double*)dA.p, n, A, lda, n, n));
  EAMD_CHECK(up2d(h, (double*)dB.p, n, B, ldb, n, nrhs));
  EAMD_CHECK(eamd_dpotrs(h, uplo, n, nrhs, (double*)dA.p, n, (double*)dB.p, n));
  EAMD_CHECK(down2d(h, B, ldb, (double*)dB.p, n, n, nrhs));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_sgetrf(eamd_handle_t h, int64_t m, int64_t n, float* A, int64_t lda,
                     int32_t* transp, int32_t* nb_transp, int64_t* first_zero_pivot) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, m)) return -4;
  CHECK_HANDLE(h);
  if (m == 0 || n == 0) return kOk;
  DevBuf dA, dt;
  EAMD_CHECK(dA.alloc(sizeof(float) * (size_t)m * n));
  EAMD_CHECK(dt.alloc(sizeof(int32_t) * (size_t)std::min(m, n)));
  EAMD_CHECK(up2d(h, (float*)dA.p, m, A, lda, m, n));
  EAMD_CHECK(eamd_sgetrf(h, m, n, (float*)dA.p, m, (int32_t*)dt.p, nb_transp, first_zero_pivot));
  EAMD_CHECK(down2d(h, A, lda, (float*)dA.p, m, m, n));
  EAMD_HIP(hipMemcpyAsync(transp, dt.p, sizeof(int32_t) * (size_t)std::min(m, n),
                          hipMemcpyDeviceToHost, h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_sgetrs(eamd_handle_t h, char trans, int64_t n, int64_t nrhs, const float* LU,
                     int64_t lda, const int32_t* transp, float* B, int64_t ldb) {
  if (!is_op(trans)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -8;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  DevBuf dA, dB, dt;
  EAMD_CHECK(dA.alloc(sizeof(float) * (size_t)n * n));
  EAMD_CHECK(dB.alloc(sizeof(float) * (size_t)n * nrhs));
  EAMD_CHECK(dt.alloc(sizeof(int32_t) * (size_t)n));
  EAMD_CHECK(up2d(h, (float*)dA.p, n, LU, lda, n, n));
  EAMD_CHECK(up2d(h, (float*)dB.p, n, B, ldb, n, nrhs));
  EAMD_HIP(hipMemcpyAsync(dt.p, transp, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                          h->stream));
  EAMD_CHECK(eamd_sgetrs(h, trans, n, nrhs, (float*)dA.p, n, (int32_t*)dt.p, (float*)dB.p, n));
  EAMD_CHECK(down2d(h, B, ldb, (float*)dB.p, n, n, nrhs));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_spotrf(eamd_handle_t h, char uplo, int64_t n, float* A, int64_t lda,
                     int64_t* fail_col) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (lda < std::max<int64_t>(1, n)) return -4;
  CHECK_HANDLE(h);
  if (n == 0) {
    if (fail_col) *fail_col = -1;
    return kOk;
  }
  DevBuf dA;
  EAMD_CHECK(dA.alloc(sizeof(float) * (size_t)n * n));
  EAMD_CHECK(up2d(h, (float*)dA.p, n, A, lda, n, n));
  EAMD_CHECK(eamd_spotrf(h, uplo, n, (float*)dA.p, n, fail_col));
  EAMD_CHECK(down2d(h, A, lda, (float*)dA.p, n, n, n));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_spotrs(eamd_handle_t h, char uplo, int64_t n, int64_t nrhs, const float* A,
                     int64_t lda, float* B, int64_t ldb) {
  if (!is_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -7;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  DevBuf dA, dB;
  EAMD_CHECK(dA.alloc(sizeof(float) * (size_t)n * n));
  EAMD_CHECK(dB.alloc(sizeof(float) * (size_t)n * nrhs));
  EAMD_CHECK(up2d(h, (float*)dA.p, n, A, lda, n, n));
  EAMD_CHECK(up2d(h, (float*)dB.p, n, B, ldb, n, nrhs));
  EAMD_CHECK(eamd_spotrs(h, uplo, n, nrhs, (float*)dA.p, n, (float*)dB.p, n));
  EAMD_CHECK(down2d(h, B, ldb, (float*)dB.p, n, n, nrhs));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dgemm(eamd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                    double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                    double beta, double* C, int64_t ldc) {
  if (!is_op(transa)) return -1;
  if (!is_op(transb)) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (k < 0) return -5;
  const bool ta = is_trans(transa), tb = is_trans(transb);
  const int64_t ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
  if (lda < std::max<int64_t>(1, ar)) return -8;
  if (ldb < std::max<int64_t>(1, br)) return -10;
  if (ldc < std::max<int64_t>(1, m)) return -13;
  CHECK_HANDLE(h);
  if (m == 0 || n == 0) return kOk;
  DevBuf dA, dB, dC;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)std::max<int64_t>(1, ar * ac)));
  EAMD_CHECK(dB.alloc(sizeof(double) * (size_t)std::max<int64_t>(1, br * bc)));
  EAMD_CHECK(dC.alloc(sizeof(double) * (size_t)m * n));
  if (ar * ac) EAMD_CHECK(up2d(h, (double*)dA.p, ar, A, lda, ar, ac));
  if (br * bc) EAMD_CHECK(up2d(h, (double*)dB.p, br, B, ldb, br, bc));
  if (beta != 0.0) EAMD_CHECK(up2d(h, (double*)dC.p, m, C, ldc, m, n));
  EAMD_CHECK(eamd_dgemm(h, transa, transb, m, n, k, alpha, (double*)dA.p, std::max<int64_t>(1, ar),
                        (double*)dB.p, std::max<int64_t>(1, br), beta, (double*)dC.p, m));
  EAMD_CHECK(down2d(h, C, ldc, (double*)dC.p, m, m, n));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dsyrk(eamd_handle_t h, char uplo, char trans, int64_t n, int64_t k, double alpha,
                    const double* A, int64_t lda, double beta, double* C, int64_t ldc) {
  if (!is_uplo(uplo)) return -1;
  if (!is_op(trans)) return -2;
  if (n < 0) return -3;
  if (k < 0) return -4;
  const bool t = is_trans(trans);
  const int64_t ar = t ? k : n, ac = t ? n : k;
  if (lda < std::max<int64_t>(1, ar)) return -7;
  if (ldc < std::max<int64_t>(1, n)) return -10;
  CHECK_HANDLE(h);
  if (n == 0) return kOk;
  DevBuf dA, dC;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)std::max<int64_t>(1, ar * ac)));
  EAMD_CHECK(dC.alloc(sizeof(double) * (size_t)n * n));
  if (ar * ac) EAMD_CHECK(up2d(h, (double*)dA.p, ar, A, lda, ar, ac));
  EAMD_CHECK(up2d(h, (double*)dC.p, n, C, ldc, n, n));
  EAMD_CHECK(eamd_dsyrk(h, uplo, trans, n, k, alpha, (double*)dA.p, std::max<int64_t>(1, ar),
                        beta, (double*)dC.p, n));
  EAMD_CHECK(down2d(h, C, ldc, (double*)dC.p, n, n, n));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dtrsm(eamd_handle_t h, char side, char uplo, char transa, char diag, int64_t m,
                    int64_t n, double alpha, const double* A, int64_t lda, double* B,
                    int64_t ldb) {
  if (!is_side(side)) return -1;
  if (m < 0) return -5;
  if (n < 0) return -6;
  const int64_t na = (side == 'L' || side == 'l') ? m : n;
  if (lda < std::max<int64_t>(1, na)) return -9;
  if (ldb < std::max<int64_t>(1, m)) return -11;
  CHECK_HANDLE(h);
  if (m == 0 || n == 0) return kOk;
  DevBuf dA, dB;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)na * na));
  EAMD_CHECK(dB.alloc(sizeof(double) * (size_t)m * n));
  EAMD_CHECK(up2d(h, (double*)dA.p, na, A, lda, na, na));
  EAMD_CHECK(up2d(h, (double*)dB.p, m, B, ldb, m, n));
  EAMD_CHECK(eamd_dtrsm(h, side, uplo, transa, diag, m, n, alpha, (double*)dA.p, na,
                        (double*)dB.p, m));
  EAMD_CHECK(down2d(h, B, ldb, (double*)dB.p, m, m, n));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return kOk;
}

int eamd_host_dgesv_ir(eamd_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                       const double* B, int64_t ldb, double* X, int64_t ldx,
                       const eamd_ir_opts* opts, eamd_ir_report* report) {
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max<int64_t>(1, n)) return -5;
  if (ldb < std::max<int64_t>(1, n)) return -7;
  if (ldx < std::max<int64_t>(1, n)) return -9;
  CHECK_HANDLE(h);
  if (n == 0 || nrhs == 0) return kOk;
  DevBuf dA, dB, dX;
  EAMD_CHECK(dA.alloc(sizeof(double) * (size_t)n * n));
  EAMD_CHECK(dB.alloc(sizeof(double) * (size_t)n * nrhs));
  EAMD_CHECK(dX.alloc(sizeof(double) * (size_t)n * nrhs));
  EAMD_CHECK(up2d(h, (double*)dA.p, n, A, lda, n, n));
  EAMD_CHECK(up2d(h, (double*)dB.p, n, B, ldb, n, nrhs));
  int st = eamd_dgesv_ir(h, n, nrhs, (double*)dA.p, n, (double*)dB.p, n, (double*)dX.p, n, opts,
                         report);
  if (st != kOk && st != kErrNotConverged) return st;
  EAMD_CHECK(down2d(h, X, ldx, (double*)dX.p, n, n, nrhs));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  return st;
}

}  // extern "C"
