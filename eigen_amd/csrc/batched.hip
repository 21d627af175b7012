// batched.hip -- many small SPD solves in one captured launch sequence.
//
// Per system this is LLT<MatrixXf>(A_i).solve(b_i) (Eigen/src/Cholesky/
// LLT.h:304-328 unblocked, :475-505 solve): every size here is below the
// reference's blocking threshold of 32 (LLT.h:335), so the unblocked
// algorithm is the one to match: x = a_kk - ..., fail at x <= 0 (NaN passes),
// L_kk = sqrt(x), column divided by L_kk.
//
// MI355X mapping.  Systems are bucketed by size once (the plan: a histogram,
// one host read of 33 counts, a scatter into a size-sorted index list).
// Execution is one kernel per non-empty size bucket, each a grid-stride loop
// over groups of T lanes per system: lane t of a group owns the columns
// c = t, t+T, t+2T, ... of L in registers (fully unrolled for the compile-time
// size), the column k being factored is broadcast to the group with
// wavefront shuffles, and the forward / backward substitutions reduce
// partial sums with butterfly shuffles inside the group.  T (1, 2 or 4) is
// picked per size by an autotuner that times the candidates on synthetic
// systems of that size.  The execution phase has no host synchronisation,
// so it can be captured in a HIP graph and replayed.
#include "../../include/eigen_amd.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace eamd {

constexpr int kMaxBatchN = 32;
constexpr int kMaxPlanN = 256;

namespace {

template <int T>
__device__ __forceinline__ float group_bcast(float v, int src_lane) {
  if (T == 1) return v;
  return __shfl(v, src_lane);
}

template <int T>
__device__ __forceinline__ float group_sum(float v) {
  if (T >= 2) v += __shfl_xor(v, 1);
  if (T >= 4) v += __shfl_xor(v, 2);
  if (T >= 8) v += __shfl_xor(v, 4);
  return v;
}

// one size N, T lanes per system
template <int N, int T>
__global__ __launch_bounds__(256) void llt_batch_kernel(int count, const int* __restrict__ order,
                                                        const int64_t* __restrict__ offA,
                                                        const float* __restrict__ A,
                                                        const int64_t* __restrict__ offb,
                                                        const float* __restrict__ bin,
                                                        float* __restrict__ xout,
                                                        int32_t* __restrict__ info) {
  constexpr int GS = 64 / T;               // systems per wave
  constexpr int NS = (N + T - 1) / T;      // column slots per lane
  const int lane = threadIdx.x & 63;
  const int t = lane % T;
  const int gbase = lane - t;
  const long gw = ((long)blockIdx.x * 256 + threadIdx.x) / 64;
  const long nw = (long)gridDim.x * 4;
  for (long base = gw * GS; base < count; base += nw * GS) {
    const long si = base + lane / T;
    const bool act = si < count;
    const int sys = order[act ? si : base];
    const float* a = A + offA[sys];
    float col[NS][N];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int c = s * T + t;
      if (N >= 4) {
        // the whole column as 16-byte words (only 4-byte alignment is
        // needed), the last word overlapping so every read stays inside the
        // column: a quarter of the address work of word loads
        typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
        const float* pc = a + (c < N ? c : 0) * N;
        float cv[N];
#pragma unroll
        for (int q = 0; q + 4 <= N; q += 4) {
          const f4u w = *reinterpret_cast<const f4u*>(pc + q);
          cv[q] = w.x;
          cv[q + 1] = w.y;
          cv[q + 2] = w.z;
          cv[q + 3] = w.w;
        }
        if (N % 4) {
          const f4u w = *reinterpret_cast<const f4u*>(pc + N - 4);
          cv[N - 4] = w.x;
          cv[N - 3] = w.y;
          cv[N - 2] = w.z;
          cv[N - 1] = w.w;
        }
#pragma unroll
        for (int i = s * T; i < N; ++i) col[s][i] = (c < N && i >= c) ? cv[i] : 0.0f;
      } else {
#pragma unroll
        for (int i = s * T; i < N; ++i) col[s][i] = (c < N && i >= c) ? a[i + c * N] : 0.0f;
      }
    }
    // ---- factorization (right-looking; same quantities as the reference's
    // left-looking unblocked loop)
    int fail = -1;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      constexpr int dummy = 0;
      (void)dummy;
      const int sk = k / T, owner = k % T;
      float dkk = group_bcast<T>(col[sk][k], gbase + owner);
      if (fail < 0 && dkk <= 0.0f) fail = k;
      const float d = sqrtf(dkk);
      if (t == owner) {
        col[sk][k] = d;
#pragma unroll
        for (int i = k + 1; i < N; ++i) col[sk][i] /= d;
      }
      float lk[N];
#pragma unroll
      for (int i = k + 1; i < N; ++i) lk[i] = group_bcast<T>(col[sk][i], gbase + owner);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (s * T + T - 1 <= k) continue;  // every column of this slot is <= k
        const int c = s * T + t;
        float lc = 0.0f;
#pragma unroll
        for (int q = 0; q < T; ++q)
          if (t == q && s * T + q < N && s * T + q > k) lc = lk[s * T + q];
        if (c > k && c < N) {
#pragma unroll
          for (int i = (s * T > k + 1 ? s * T : k + 1); i < N; ++i)
            if (i >= c) col[s][i] -= lk[i] * lc;
        }
      }
    }
    // ---- forward substitution L y = b (column partial sums, group-reduced)
    const float* bp = bin + offb[sys];
    float y[NS];
    float p[N];
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = 0.0f;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const int sk = k / T, owner = k % T;
      const float tot = group_sum<T>(p[k]);
      if (t == owner) {
        const float yk = (bp[k] - tot) / col[sk][k];
        y[sk] = yk;
#pragma unroll
        for (int i = k + 1; i < N; ++i) p[i] += col[sk][i] * yk;
      }
    }
    // ---- backward substitution L^T x = y
    float x[N];
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
      const int sk = k / T, owner = k % T;
      float xk = 0.0f;
      if (t == owner) {
        float sacc = y[sk];
#pragma unroll
        for (int i = k + 1; i < N; ++i) sacc -= col[sk][i] * x[i];
        xk = sacc / col[sk][k];
      }
      x[k] = group_bcast<T>(xk, gbase + owner);
    }
    if (act) {
      float* xp = xout + offb[sys];
      if (fail < 0) {
#pragma unroll
        for (int k = 0; k < N; ++k)
          if (k % T == t) xp[k] = x[k];
      }
      if (t == 0) info[sys] = fail;
    }
  }
}

// ---- plan kernels
// hist[v] = systems of size v (1..max_n); hist[kMaxPlanN + 1] = sizes outside
__global__ void size_histogram_kernel(int count, int max_n, const int32_t* __restrict__ n,
                                      int* __restrict__ hist) {
  __shared__ int h[kMaxPlanN + 2];
  for (int i = threadIdx.x; i < kMaxPlanN + 2; i += blockDim.x) h[i] = 0;
  __syncthreads();
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x) {
    int v = n[i];
    if (v < 1 || v > max_n) atomicAdd(&h[kMaxPlanN + 1], 1);
    else atomicAdd(&h[v], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMaxPlanN + 2; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void size_scatter_kernel(int count, int max_n, const int32_t* __restrict__ n,
                                    int* __restrict__ cursor, int* __restrict__ order) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x) {
    int v = n[i];
    if (v >= 1 && v <= max_n) order[atomicAdd(&cursor[v], 1)] = (int)i;
  }
}

// synthetic SPD systems for the autotuner: A = I*N + small symmetric noise
__global__ void tune_fill_kernel(int N, int count, float* A, int64_t* offA, float* b,
                                 int64_t* offb, int* order) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  offA[i] = i * N * N;
  offb[i] = i * N;
  order[i] = (int)i;
  for (int c = 0; c < N; ++c) {
    b[i * N + c] = 1.0f + 0.01f * c;
    for (int r = 0; r < N; ++r)
      A[i * N * N + r + c * N] = (r == c) ? (float)N : 0.1f / (1.0f + (float)((r + c) % 7));
  }
}

typedef void (*BatchKernel)(int, const int*, const int64_t*, const float*, const int64_t*,
                            const float*, float*, int32_t*);

template <int N>
BatchKernel pick(int T) {
  if (T == 1) return llt_batch_kernel<N, 1>;
  if (T == 2) return llt_batch_kernel<N, 2>;
  return llt_batch_kernel<N, 4>;
}

BatchKernel kernel_for(int N, int T) {
  switch (N) {
#define EAMD_CASE(NN) \
  case NN:            \
    return pick<NN>(T);
    EAMD_CASE(1) EAMD_CASE(2) EAMD_CASE(3) EAMD_CASE(4) EAMD_CASE(5) EAMD_CASE(6) EAMD_CASE(7)
    EAMD_CASE(8) EAMD_CASE(9) EAMD_CASE(10) EAMD_CASE(11) EAMD_CASE(12) EAMD_CASE(13)
    EAMD_CASE(14) EAMD_CASE(15) EAMD_CASE(16) EAMD_CASE(17) EAMD_CASE(18) EAMD_CASE(19)
    EAMD_CASE(20) EAMD_CASE(21) EAMD_CASE(22) EAMD_CASE(23) EAMD_CASE(24) EAMD_CASE(25)
    EAMD_CASE(26) EAMD_CASE(27) EAMD_CASE(28) EAMD_CASE(29) EAMD_CASE(30) EAMD_CASE(31)
    EAMD_CASE(32)
#undef EAMD_CASE
    default:
      return nullptr;
  }
}

int default_lanes(int N) { return N <= 12 ? 1 : (N <= 20 ? 2 : 4); }

// process-wide autotune table: lanes per system for each size (0 = unknown)
std::mutex g_tune_mu;
int g_tune[kMaxBatchN + 1] = {0};

int grid_for(eamd_handle* h, int count, int T) {
  const long waves = (count + (64 / T) - 1) / (64 / T);
  long wgs = (waves + 3) / 4;
  long cap = (long)h->num_cus * 8;
  return (int)std::max<long>(1, std::min(wgs, cap));
}

int autotune_size(eamd_handle* h, int N) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune[N]) return g_tune[N];
  }
  const char* env = getenv("EAMD_BATCH_AUTOTUNE");
  int best = default_lanes(N);
  if (!(env && env[0] == '0')) {
    const int cnt = 32768;
    size_t need = sizeof(float) * ((size_t)cnt * N * N + 2 * (size_t)cnt * N) +
                  sizeof(int64_t) * 2 * cnt + sizeof(int) * 2 * (size_t)cnt + 1024;
    void* p = nullptr;
    if (h->ws.get(kWsAux, need, &p) == kOk) {
      char* c = static_cast<char*>(p);
      float* A = reinterpret_cast<float*>(c);
      float* b = A + (size_t)cnt * N * N;
      float* x = b + (size_t)cnt * N;
      int64_t* offA = reinterpret_cast<int64_t*>(x + (size_t)cnt * N);
      int64_t* offb = offA + cnt;
      int* order = reinterpret_cast<int*>(offb + cnt);
      int* info = order + cnt;
      hipLaunchKernelGGL(tune_fill_kernel, dim3(ceil_div(cnt, 256)), dim3(256), 0, h->stream, N,
                         cnt, A, offA, b, offb, order);
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      float best_ms = 1e30f;
      for (int T : {1, 2, 4}) {
        BatchKernel k = kernel_for(N, T);
        float ms_min = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
          (void)hipEventRecord(e0, h->stream);
          hipLaunchKernelGGL(k, dim3(grid_for(h, cnt, T)), dim3(256), 0, h->stream, cnt, order,
                             offA, A, offb, b, x, info);
          (void)hipEventRecord(e1, h->stream);
          (void)hipEventSynchronize(e1);
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) ms_min = std::min(ms_min, ms);
        }
        if (ms_min < best_ms) {
          best_ms = ms_min;
          best = T;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (hipGetLastError() != hipSuccess) best = default_lanes(N);
    }
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[N] = best;
  return best;
}

}  // namespace

}  // namespace eamd

namespace eamd {

// Size sort shared by the batched solvers: one histogram of n[] (one host
// synchronisation), descending-size buckets, a scatter of the system
// indices into `order`.
int batched_plan_sort(eamd_handle* h, int64_t count, const int32_t* n, int max_n,
                      eamd_batched_plan** out) {
  *out = nullptr;
  if (count > INT32_MAX) return kErrUnsupported;
  if (max_n < 1 || max_n > kMaxPlanN) return kErrUnsupported;
  auto* p = new eamd_batched_plan();
  p->count = count;
  p->max_n = max_n;
  int* hist = nullptr;
  hipError_t e = hipMalloc(&hist, sizeof(int) * 2 * (kMaxPlanN + 2));
  if (e != hipSuccess) {
    delete p;
    return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  }
  e = hipMalloc(&p->order, sizeof(int) * std::max<int64_t>(1, count));
  if (e != hipSuccess) {
    (void)hipFree(hist);
    delete p;
    return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  }
  int hh[kMaxPlanN + 2] = {0};
  int st = kOk;
  do {
    if ((e = hipMemsetAsync(hist, 0, sizeof(int) * 2 * (kMaxPlanN + 2), h->stream))) break;
    int grid = (int)std::min<int64_t>(1024, (count + 255) / 256);
    hipLaunchKernelGGL(size_histogram_kernel, dim3(std::max(1, grid)), dim3(256), 0, h->stream,
                       (int)count, max_n, n, hist);
    if ((e = hipMemcpyAsync(hh, hist, sizeof(hh), hipMemcpyDeviceToHost, h->stream))) break;
    if ((e = hipStreamSynchronize(h->stream))) break;
  } while (0);
  if (e != hipSuccess) {
    st = hip_fail(e, "batched plan", __FILE__, __LINE__);
  } else if (hh[kMaxPlanN + 1] > 0) {
    set_error("batched solve: %d systems have a size outside [1, %d]", hh[kMaxPlanN + 1], max_n);
    st = -3;
  }
  if (st == kOk) {
    int acc = 0;
    // descending size: the largest (slowest) systems are scheduled first
    int cur[kMaxPlanN + 2] = {0};
    for (int v = kMaxPlanN; v >= 1; --v) {
      p->bucket_start[v] = acc;
      p->bucket_count[v] = hh[v];
      cur[v] = acc;
      acc += hh[v];
    }
    p->bucket_start[0] = acc;
    p->bucket_count[0] = 0;
    int* cursor = hist + (kMaxPlanN + 2);
    e = hipMemcpyAsync(cursor, cur, sizeof(cur), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
      int grid = (int)std::min<int64_t>(1024, (count + 255) / 256);
      hipLaunchKernelGGL(size_scatter_kernel, dim3(std::max(1, grid)), dim3(256), 0, h->stream,
                         (int)count, max_n, n, cursor, p->order);
      e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) st = hip_fail(e, "batched plan scatter", __FILE__, __LINE__);
  }
  (void)hipFree(hist);
  if (st != kOk) {
    (void)hipFree(p->order);
    delete p;
    return st;
  }
  *out = p;
  return kOk;
}

int batched_plan_create(eamd_handle* h, int64_t count, const int32_t* n, int max_n,
                        eamd_batched_plan** out) {
  EAMD_CHECK(batched_plan_sort(h, count, n, max_n, out));
  eamd_batched_plan* p = *out;
  for (int v = 1; v <= kMaxBatchN; ++v)
    p->lanes[v] = p->bucket_count[v] ? autotune_size(h, v) : default_lanes(v);
  return kOk;
}

void batched_plan_destroy(eamd_batched_plan* p) {
  if (!p) return;
  if (p->order) (void)hipFree(p->order);
  if (p->wbase) (void)hipFree(p->wbase);
  if (p->bstart) (void)hipFree(p->bstart);
  delete p;
}

int spotrf_solve_batched_exec(eamd_handle* h, const eamd_batched_plan* p, const int64_t* offA,
                              const float* A, const int64_t* offb, float* b, int32_t* info) {
  PhaseScope ps(h, kPhBatched, 0.0, 0.0);
  for (int v = kMaxBatchN; v >= 1; --v) {
    const int cnt = p->bucket_count[v];
    if (!cnt) continue;
    const int T = p->lanes[v];
    BatchKernel k = kernel_for(v, T);
    hipLaunchKernelGGL(k, dim3(grid_for(h, cnt, T)), dim3(256), 0, h->stream, cnt,
                       p->order + p->bucket_start[v], offA, A, offb, b, b, info);
  }
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int spotrf_solve_batched_impl(eamd_handle* h, int64_t count, const int64_t* offA,
                              const int32_t* n, const float* A, const int64_t* offb, float* b,
                              int32_t* info) {
  eamd_batched_plan* p = nullptr;
  EAMD_CHECK(batched_plan_create(h, count, n, kMaxBatchN, &p));
  int st = spotrf_solve_batched_exec(h, p, offA, A, offb, b, info);
  (void)hipStreamSynchronize(h->stream);
  batched_plan_destroy(p);
  return st;
}

int batched_lanes(const eamd_batched_plan* p, int n) {
  return (n >= 1 && n <= p->max_n && n <= kMaxPlanN) ? p->lanes[n] : 0;
}

}  // namespace eamd
