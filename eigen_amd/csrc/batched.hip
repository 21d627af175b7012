// batched.hip -- many small SPD solves in one persistent launch.
//
// Per system this is LLT<MatrixXf>(A_i).solve(b_i) (Eigen/src/Cholesky/
// LLT.h:304-328 unblocked, :475-505 solve): every size here is below the
// reference's blocking threshold of 32 (LLT.h:335), so the unblocked
// algorithm is the one to match: x = a_kk - ..., fail at x <= 0 (NaN passes),
// L_kk = sqrt(x), column divided by L_kk.
//
// MI355X mapping.  Systems are bucketed by size once (the plan: a histogram,
// one host read of the counts, a scatter into a size-sorted index list).
// Execution is ONE persistent launch: every wavefront walks a static
// round-robin share of the plan's tasks (a task is one wavefront's batch of
// systems of one size, largest sizes first) and runs the size's code path,
// chosen by a wave-uniform branch.  Inside a task a group of T lanes (1, 2
// or 4) owns one system; lane t of the group keeps the lower triangle of the
// columns c = t, t + T, ... in registers (only rows >= the slot's first
// column, the rows above the lane's own diagonal held at zero), so no lane
// holds more than about N^2 / 2T values and nothing is spilled.  Column k of
// L is broadcast inside the group with DPP quad permutes, the forward
// substitution rides along the factorization (b is replicated in the group
// and updated with each column), and the backward substitution updates each
// lane's own entries of y with each new x_k.  T per size is chosen by an
// autotuner that times the admissible widths on synthetic systems of that
// size.  Execution has no host synchronisation and no memset, so it can be
// captured in a HIP graph and replayed.
#include "../../include/eigen_amd.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <vector>

namespace eamd {

constexpr int kMaxBatchN = 32;
constexpr int kMaxPlanN = 256;

namespace {

// compile-time loop: f(integral_constant<int, i>) for i = B .. E-1
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// value of lane O of this lane's T-lane group (T = 1, 2, 4: groups are
// inside a quad, so one DPP quad permute)
template <int T, int O>
__device__ __forceinline__ float gbcast(float v) {
  if constexpr (T == 1) {
    return v;
  } else {
    constexpr int ctrl = T == 2 ? (O * 5 + (2 + O) * 0x50) : O * 0x55;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xf, 0xf, false));
  }
}

typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

// v[R0 .. NM-1] = p[R0 .. NM-1] for rows below `lim` (rows at or above lim
// read as zero); a whole 16-byte word (4-byte alignment suffices) when
// `wide` allows it for the word's last row, single floats otherwise.
template <int NM, int R0>
__device__ __forceinline__ void load_rows(const float* __restrict__ p, int lim, int wide_lim,
                                          float (&v)[NM]) {
  sfor<0, (NM - R0 + 3) / 4>([&](auto qc) {
    constexpr int r = R0 + 4 * decltype(qc)::value;
    constexpr int nr = NM - r < 4 ? NM - r : 4;
    if (nr == 4 && r + 3 < wide_lim) {
      const f4u w = *reinterpret_cast<const f4u*>(p + r);
      v[r] = r < lim ? w.x : 0.0f;
      v[r + 1] = r + 1 < lim ? w.y : 0.0f;
      v[r + 2] = r + 2 < lim ? w.z : 0.0f;
      v[r + 3] = r + 3 < lim ? w.w : 0.0f;
    } else {
      sfor<0, nr>([&](auto ec) {
        constexpr int i = r + decltype(ec)::value;
        v[i] = i < lim ? p[i] : 0.0f;
      });
    }
  });
}

// One task: systems order[first .. first + cnt) (cnt <= 64 / T), all of
// size N <= NM, T lanes each.  Rows and columns N..NM-1 are held at zero
// and the steps k >= N are skipped (wave-uniform branches), so one code
// path serves every size up to NM.
template <int NM, int T>
__device__ __forceinline__ void llt_task(const int N, const int lane,
                                         const int* __restrict__ order, int first, int cnt,
                                         const int64_t* __restrict__ offA,
                                         const float* __restrict__ A,
                                         const int64_t* __restrict__ offb,
                                         const float* __restrict__ bin, float* __restrict__ xout,
                                         int32_t* __restrict__ info) {
  static_assert(NM % T == 0 && NM % 4 == 0, "padded size");
  constexpr int NS = NM / T;  // column slots per lane
  const int t = T == 1 ? 0 : lane % T;
  const int gi = lane / T;
  const bool act = gi < cnt;
  const int sys = order[first + (act ? gi : 0)];
  const float* a = A + offA[sys];
  const float* bp = bin + offb[sys];
  // col[s][i], i >= s*T: row i of this lane's column c = s*T + t (zero for
  // i < c, for i >= N and for c >= N)
  float col[NS][NM];
  sfor<0, NS>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    constexpr int r0 = s * T;
    const int c = r0 + t;
    const bool cv = c < N;
    // a 16-byte word may run past row N - 1 into the next column, but not
    // past the system's last element
    load_rows<NM, r0>(a + (cv ? c : 0) * N, cv ? N : 0, cv ? (N - c) * N : 0, col[s]);
    if constexpr (T > 1) {
      sfor<r0, r0 + T - 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        col[s][i] = i >= c ? col[s][i] : 0.0f;
      });
    }
  });
  float bv[NM];
  load_rows<NM, 0>(bp, N, N, bv);
  // ---- factorization (right-looking: the same quantities as the
  // reference's left-looking loop) with the forward substitution L y = b
  int fail = -1;
  sfor<0, NM>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k >= N) return;
    constexpr int sk = k / T, o = k % T;
    const float akk = gbcast<T, o>(col[sk][k]);
    if (fail < 0 && akk <= 0.0f) fail = k;
    const float d = sqrtf(akk);
    const float r = 1.0f / d;
    const bool own = T == 1 || t == o;
    float l[NM];
    sfor<k + 1, NM>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      l[i] = gbcast<T, o>(col[sk][i]) * r;
      col[sk][i] = own ? l[i] : col[sk][i];
    });
    col[sk][k] = own ? d : col[sk][k];
    const float yk = bv[k] / d;
    bv[k] = yk;
    sfor<k + 1, NM>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      bv[i] = fmaf(-l[i], yk, bv[i]);
    });
    // columns c > k of this lane take -= l l_c (slots before sk hold only
    // columns < k; slots past the last column are skipped)
    sfor<sk, NS>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      if (s * T >= N) return;
      float lc = 0.0f;
      sfor<0, T>([&](auto qc) {
        constexpr int cq = s * T + decltype(qc)::value;
        if constexpr (cq > k) lc = (T == 1 || t == decltype(qc)::value) ? l[cq] : lc;
      });
      constexpr int i0 = s * T > k + 1 ? s * T : k + 1;
      sfor<i0, NM>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const float v = fmaf(-l[i], lc, col[s][i]);
        if constexpr (i - s * T < T - 1)
          col[s][i] = t <= i - s * T ? v : col[s][i];  // rows at or below c only
        else
          col[s][i] = v;
      });
    });
  });
  // ---- backward substitution L^T x = y: each lane keeps y of its own
  // columns; x_j is broadcast and every lane folds row j of its columns in
  float yo[NS];
  sfor<0, NS>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    float y = 0.0f;
    sfor<0, T>([&](auto qc) {
      constexpr int cq = s * T + decltype(qc)::value;
      y = (T == 1 || t == decltype(qc)::value) ? bv[cq] : y;
    });
    yo[s] = y;
  });
  float* xp = xout + offb[sys];
  const bool store = act && fail < 0;
  sfor<0, NM>([&](auto jc) {
    constexpr int j = NM - 1 - decltype(jc)::value;
    if (j >= N) return;
    constexpr int sj = j / T, o = j % T;
    const float xj = gbcast<T, o>(yo[sj] / col[sj][j]);
    if (store && (T == 1 || t == o)) xp[j] = xj;
    // row j of the columns c < j (for c > j the row is above the diagonal
    // and held at zero; the owner's own y_j is spent)
    sfor<0, sj + 1>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      yo[s] = fmaf(-col[s][j], xj, yo[s]);
    });
  });
  if (act && t == 0) info[sys] = fail;
}

// The compiled (padded size, lanes per system) paths: sizes up to 16 on one
// lane (or two from 9), 17..24 on two (or four), 25..32 on four -- every
// path within 256 registers.
constexpr int kPaths = 10;
__host__ __device__ constexpr int path_nm(int p) {
  return p == 0 ? 4 : p == 1 ? 8 : p == 2 ? 12 : p == 3 ? 16 : p == 4 ? 12 : p == 5 ? 16
       : p == 6 ? 20 : p == 7 ? 24 : p == 8 ? 24 : 32;
}
__host__ __device__ constexpr int path_lanes(int p) {
  return p <= 3 ? 1 : p <= 7 ? 2 : 4;
}

// the persistent launch's schedule: buckets of one size and path, largest
// size first; bucket b covers tasks [first_task[b], first_task[b + 1])
constexpr int kMaxBuckets = kMaxBatchN;
struct LltSched {
  int nbk;
  int first_task[kMaxBuckets + 1];
  int start[kMaxBuckets];   // first position in `order`
  int count[kMaxBuckets];   // systems
  unsigned char n[kMaxBuckets], path[kMaxBuckets];
};

__global__ __launch_bounds__(256) void llt_batch_persistent(LltSched sc, const int* __restrict__ order,
                                                           const int64_t* __restrict__ offA,
                                                           const float* __restrict__ A,
                                                           const int64_t* __restrict__ offb,
                                                           const float* __restrict__ bin,
                                                           float* __restrict__ xout,
                                                           int32_t* __restrict__ info) {
  const int lane = threadIdx.x & 63;
  const int gw = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int nwaves = gridDim.x * 4;
  const int total = sc.first_task[sc.nbk];
  int b = 0;
  for (int task = gw; task < total; task += nwaves) {
    while (task >= sc.first_task[b + 1]) ++b;  // tasks ascend: the bucket only moves on
    const int N = sc.n[b], P = sc.path[b];
    const int spw = 64 / path_lanes(P);
    const int first = sc.start[b] + (task - sc.first_task[b]) * spw;
    const int cnt = min(spw, sc.start[b] + sc.count[b] - first);
    sfor<0, kPaths>([&](auto pc) {
      constexpr int PP = decltype(pc)::value;
      if (P == PP)
        llt_task<path_nm(PP), path_lanes(PP)>(N, lane, order, first, cnt, offA, A, offb, bin,
                                              xout, info);
    });
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


// The paths that serve size N: the tightest padding on the narrowest group,
// and where the sizes allow it one on a wider group.
std::vector<int> paths_for(int N) {
  if (N <= 4) return {0};
  if (N <= 8) return {1};
  if (N <= 12) return {2, 4};
  if (N <= 16) return {3, 5};
  if (N <= 20) return {6, 8};
  if (N <= 24) return {7, 8};
  return {9};
}

// Persistent grid: two workgroups of four waves per CU (every path fits two
// waves per SIMD), never more than there are tasks.
int persistent_grid(eamd_handle* h, int tasks) {
  const int cap = 2 * (h->num_cus > 0 ? h->num_cus : 256);
  return std::max(1, std::min(cap, ceil_div(tasks, 4)));
}

// Appends bucket (N, path P) over positions [start, start + count) of `order`.
void sched_add(LltSched& sc, int N, int P, int start, int count) {
  if (count <= 0) return;
  const int b = sc.nbk++;
  sc.n[b] = (unsigned char)N;
  sc.path[b] = (unsigned char)P;
  sc.start[b] = start;
  sc.count[b] = count;
  sc.first_task[b + 1] = sc.first_task[b] + ceil_div(count, 64 / path_lanes(P));
}

int launch_sched(eamd_handle* h, const LltSched& sc, const int* order, const int64_t* offA,
                 const float* A, const int64_t* offb, const float* bin, float* xout,
                 int32_t* info) {
  const int tasks = sc.first_task[sc.nbk];
  if (tasks == 0) return kOk;
  hipLaunchKernelGGL(llt_batch_persistent, dim3(persistent_grid(h, tasks)), dim3(256), 0,
                     h->stream, sc, order, offA, A, offb, bin, xout, info);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// process-wide autotune table: path + 1 for each size (0 = unknown)
std::mutex g_tune_mu;
int g_tune[kMaxBatchN + 1] = {0};

// Times each path that serves size N on 32768 synthetic systems of that
// size (one persistent launch each, best of two after a warm-up) and keeps
// the fastest.  EAMD_BATCH_AUTOTUNE=0 keeps the first (narrowest) path.
int autotune_size(eamd_handle* h, int N) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune[N]) return g_tune[N] - 1;
  }
  const std::vector<int> ws = paths_for(N);
  int best = ws.front();
  const char* env = getenv("EAMD_BATCH_AUTOTUNE");
  if (ws.size() > 1 && !(env && env[0] == '0')) {
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
      for (int P : ws) {
        LltSched sc{};
        sched_add(sc, N, P, 0, cnt);
        float ms_min = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
          (void)hipEventRecord(e0, h->stream);
          (void)launch_sched(h, sc, order, offA, A, offb, b, x, info);
          (void)hipEventRecord(e1, h->stream);
          (void)hipEventSynchronize(e1);
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) ms_min = std::min(ms_min, ms);
        }
        if (ms_min < best_ms) {
          best_ms = ms_min;
          best = P;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (hipGetLastError() != hipSuccess) best = ws.front();
    }
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[N] = best + 1;
  return best;
}

}  // namespace


}  // namespace eamd

namespace eamd {

// Size sort into caller-provided device buffers: one histogram of n[] into
// hist (2 * (kMaxPlanN + 2) ints: counts, then scatter cursors; one host
// synchronisation), descending-size buckets (bucket_start / bucket_count,
// host arrays of kMaxPlanN + 2), a scatter of the system indices into
// order (count ints).  Returns -3 when a size lies outside [1, max_n].
int size_sort(eamd_handle* h, int64_t count, const int32_t* n, int max_n, int* hist, int* order,
              int* bucket_start, int* bucket_count) {
  int hh[kMaxPlanN + 2] = {0};
  hipError_t e = hipSuccess;
  do {
    if ((e = hipMemsetAsync(hist, 0, sizeof(int) * 2 * (kMaxPlanN + 2), h->stream))) break;
    int grid = (int)std::min<int64_t>(1024, (count + 255) / 256);
    hipLaunchKernelGGL(size_histogram_kernel, dim3(std::max(1, grid)), dim3(256), 0, h->stream,
                       (int)count, max_n, n, hist);
    if ((e = hipMemcpyAsync(hh, hist, sizeof(hh), hipMemcpyDeviceToHost, h->stream))) break;
    if ((e = hipStreamSynchronize(h->stream))) break;
  } while (0);
  if (e != hipSuccess) return hip_fail(e, "batched plan", __FILE__, __LINE__);
  if (hh[kMaxPlanN + 1] > 0) {
    set_error("batched solve: %d systems have a size outside [1, %d]", hh[kMaxPlanN + 1], max_n);
    return -3;
  }
  int acc = 0;
  // descending size: the largest (slowest) systems are scheduled first
  int cur[kMaxPlanN + 2] = {0};
  for (int v = kMaxPlanN; v >= 1; --v) {
    bucket_start[v] = acc;
    bucket_count[v] = hh[v];
    cur[v] = acc;
    acc += hh[v];
  }
  bucket_start[0] = acc;
  bucket_count[0] = 0;
  int* cursor = hist + (kMaxPlanN + 2);
  e = hipMemcpyAsync(cursor, cur, sizeof(cur), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    int grid = (int)std::min<int64_t>(1024, (count + 255) / 256);
    hipLaunchKernelGGL(size_scatter_kernel, dim3(std::max(1, grid)), dim3(256), 0, h->stream,
                       (int)count, max_n, n, cursor, order);
    e = hipStreamSynchronize(h->stream);
  }
  if (e != hipSuccess) return hip_fail(e, "batched plan scatter", __FILE__, __LINE__);
  return kOk;
}

// The plan's size sort, into buffers the plan owns.
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
  const int st = size_sort(h, count, n, max_n, hist, p->order, p->bucket_start, p->bucket_count);
  (void)hipFree(hist);
  if (st != kOk) {
    (void)hipFree(p->order);
    delete p;
    return st;
  }
  *out = p;
  return kOk;
}

// ---- allocation-free batched LLT (eamd_spotrf_solve_batched_noalloc) ----
// The plan's device buffers (the size histogram with its scatter cursors,
// and the size-sorted order) live in one file-scope device array, one copy
// per GPU, whose address each device looks up once (batched_noalloc_init,
// called by eamd_create with the handle's device current).  Host code only
// ever passes that looked-up address, never the array's name.
constexpr int64_t kNoAllocMaxSystems = 1 << 20;
constexpr size_t kNoAllocInts = 2 * (kMaxPlanN + 2) + (size_t)kNoAllocMaxSystems;
static __device__ __align__(256) int scratch[kNoAllocInts];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kNoAllocMaxDevices] = {};

int batched_noalloc_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) {
    set_error("allocation-free batched solve: device %d beyond %d", device, kNoAllocMaxDevices);
    return kErrUnsupported;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return kOk;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
  g_scratch_ptr[device] = static_cast<int*>(ptr);
  return kOk;
}

int spotrf_solve_batched_noalloc(eamd_handle* h, int64_t count, const int64_t* offA,
                                 const int32_t* n, const float* A, const int64_t* offb, float* b,
                                 int32_t* info) {
  if (count > kNoAllocMaxSystems) {
    set_error("allocation-free batched solve: %lld systems exceed its maximum of %lld",
              (long long)count, (long long)kNoAllocMaxSystems);
    return kErrUnsupported;
  }
  int* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (h->device >= 0 && h->device < kNoAllocMaxDevices) base = g_scratch_ptr[h->device];
  }
  if (!base) {
    set_error("allocation-free batched solve: no scratch address for device %d", h->device);
    return kErrHip;
  }
  int* hist = base;                       // cleared by size_sort on every call
  int* order = base + 2 * (kMaxPlanN + 2);  // fully written for this call's count
  int bstart[kMaxPlanN + 2], bcount[kMaxPlanN + 2];
  EAMD_CHECK(size_sort(h, count, n, kMaxBatchN, hist, order, bstart, bcount));
  // the widths an earlier planned call tuned, else the narrowest (tuning
  // would allocate its synthetic systems)
  LltSched sc{};
  for (int v = kMaxBatchN; v >= 1; --v) {
    int path = paths_for(v).front();
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      if (g_tune[v]) path = g_tune[v] - 1;
    }
    sched_add(sc, v, path, bstart[v], bcount[v]);
  }
  return launch_sched(h, sc, order, offA, A, offb, b, b, info);
}

int batched_plan_create(eamd_handle* h, int64_t count, const int32_t* n, int max_n,
                        eamd_batched_plan** out) {
  EAMD_CHECK(batched_plan_sort(h, count, n, max_n, out));
  eamd_batched_plan* p = *out;
  for (int v = 1; v <= kMaxBatchN; ++v)
    p->llt_path[v] = (unsigned char)(p->bucket_count[v] ? autotune_size(h, v)
                                                      : paths_for(v).front());
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
  LltSched sc{};
  for (int v = kMaxBatchN; v >= 1; --v)
    sched_add(sc, v, p->llt_path[v], p->bucket_start[v], p->bucket_count[v]);
  return launch_sched(h, sc, p->order, offA, A, offb, b, b, info);
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
  if (n < 1 || n > p->max_n || n > kMaxPlanN) return 0;
  return p->kind == 0 ? path_lanes(p->llt_path[n]) : p->lanes[n];
}

}  // namespace eamd
