// trsv.hip -- blocked triangular solves with few right-hand sides.
//
// Replaces triangular_solve_vector (TriangularSolverVector.h:88-139) and the
// few-RHS use of triangular_solve_matrix (TriangularSolverMatrix.h:38-184)
// inside PartialPivLU::_solve_impl (PartialPivLU.h:226-244) and
// LLT::_solve_impl (LLT.h:475-505).
//
// One workgroup per BS-row block, dispatched in dependency order.  Workgroup
// b first inverts its diagonal block in LDS (independent of the others, so
// it overlaps the wait), accumulates B_b - sum_j T_bj X_j as each
// predecessor's flag appears, then X_b = inv(T_bb) * (...) and releases its
// own flag.  Only the last predecessor's product is on the critical path.
// Flags carry a per-call epoch so no clearing pass is needed.
#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kSpin = 1 << 26;

template <typename T>
__device__ __forceinline__ T tload(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// element (r, c) of the triangular matrix in the requested layout
template <typename T, bool ROWMAJOR>
__device__ __forceinline__ const T* tptr(const T* A, int64_t ld, int r, int c) {
  return ROWMAJOR ? A + (int64_t)r * ld + c : A + r + (int64_t)c * ld;
}

template <typename T, bool ROWMAJOR, bool LOWER, bool UNIT, int BS, int NRC>
__global__ __launch_bounds__(256) void trsv_chain_kernel(int n, const T* __restrict__ A,
                                                         int64_t lda, T* __restrict__ X,
                                                         int64_t ldx, int c0, int nrhs,
                                                         int* __restrict__ flags, int epoch,
                                                         int* __restrict__ err) {
  __shared__ T Dinv[BS][BS + 1];
  __shared__ T Tb[BS][BS + 1];
  __shared__ T Xj[BS][NRC];
  __shared__ T acc_s[BS][NRC];
  __shared__ int s_abort;
  const int tid = threadIdx.x;
  const int nblk = (n + BS - 1) / BS;
  const int w = blockIdx.x;
  const int b = LOWER ? w : nblk - 1 - w;
  const int rb0 = b * BS;
  const int bs = min(BS, n - rb0);
  const int nc = min(NRC, nrhs - c0);

  // ---- diagonal block into LDS
  for (int e = tid; e < BS * BS; e += 256) {
    int r = e / BS, c = e - r * BS;
    T v = T(0);
    if (r < bs && c < bs) v = *tptr<T, ROWMAJOR>(A, lda, rb0 + r, rb0 + c);
    Tb[r][c] = v;
  }
  for (int e = tid; e < BS * NRC; e += 256) {
    int r = e / NRC, c = e - r * NRC;
    acc_s[r][c] = (r < bs && c < nc) ? X[(int64_t)(c0 + c) * ldx + rb0 + r] : T(0);
  }
  // the padding of a partial last block must be exactly zero
  for (int e = tid; e < BS * BS; e += 256) Dinv[e / BS][e % BS] = T(0);
  if (tid == 0) s_abort = 0;
  __syncthreads();
  // ---- invert it: thread c owns column c of the inverse
  if (tid < BS) {
    const int c = tid;
    if (c < bs) {
      if (LOWER) {
        for (int r = 0; r < c; ++r) Dinv[r][c] = T(0);
        T d = UNIT ? T(1) : T(1) / Tb[c][c];
        Dinv[c][c] = d;
        for (int r = c + 1; r < bs; ++r) {
          T s = T(0);
          for (int q = c; q < r; ++q) s += Tb[r][q] * Dinv[q][c];
          Dinv[r][c] = UNIT ? -s : -s / Tb[r][r];
        }
      } else {
        for (int r = c + 1; r < bs; ++r) Dinv[r][c] = T(0);
        T d = UNIT ? T(1) : T(1) / Tb[c][c];
        Dinv[c][c] = d;
        for (int r = c - 1; r >= 0; --r) {
          T s = T(0);
          for (int q = r + 1; q <= c; ++q) s += Tb[r][q] * Dinv[q][c];
          Dinv[r][c] = UNIT ? -s : -s / Tb[r][r];
        }
      }
    }
  }
  // ---- accumulate the predecessors in dependency order
  const int ndep = LOWER ? b : nblk - 1 - b;
  for (int d = 0; d < ndep; ++d) {
    const int jb = LOWER ? d : nblk - 1 - d;
    const int cb0 = jb * BS;
    const int cbs = min(BS, n - cb0);
    __syncthreads();
    // off-diagonal block T[rb0:, cb0:] (independent of X: load before waiting)
    for (int e = tid; e < BS * BS; e += 256) {
      int r, c;
      if (ROWMAJOR) { r = e / BS; c = e - r * BS; } else { c = e / BS; r = e - c * BS; }
      Tb[r][c] = (r < bs && c < cbs) ? *tptr<T, ROWMAJOR>(A, lda, rb0 + r, cb0 + c) : T(0);
    }
    if (tid == 0) {
      int spins = 0;
      while (__hip_atomic_load(&flags[jb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++spins > kSpin ||
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          s_abort = 1;
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __syncthreads();
    if (s_abort) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int e = tid; e < BS * NRC; e += 256) {
      int r = e / NRC, c = e - r * NRC;
      Xj[r][c] = (r < cbs && c < nc) ? tload(&X[(int64_t)(c0 + c) * ldx + cb0 + r]) : T(0);
    }
    __syncthreads();
    for (int e = tid; e < BS * NRC; e += 256) {
      int r = e % BS, c = e / BS;
      T s = T(0);
#pragma unroll 8
      for (int q = 0; q < BS; ++q) s += Tb[r][q] * Xj[q][c];
      acc_s[r][c] -= s;
    }
  }
  __syncthreads();
  // ---- X_b = inv(T_bb) * acc
  for (int e = tid; e < BS * NRC; e += 256) {
    int r = e % BS, c = e / BS;
    T s = T(0);
#pragma unroll 8
    for (int q = 0; q < BS; ++q) s += Dinv[r][q] * acc_s[q][c];
    if (r < bs && c < nc)
      __hip_atomic_store(&X[(int64_t)(c0 + c) * ldx + rb0 + r], s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0)
    __hip_atomic_store(&flags[b], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

// Solves op(T) X = X in place.  T is n x n; `rowmajor` selects its layout;
// lower/unit select the triangle.  X is column-major n x nrhs.
template <typename T>
int trsv_device(eamd_handle* h, int n, const T* A, int64_t lda, bool rowmajor, bool lower,
                bool unit, T* X, int64_t ldx, int nrhs) {
  if (n <= 0 || nrhs <= 0) return kOk;
  constexpr int BS = 64, NRC = 4;
  const int nblk = ceil_div(n, BS);
  // flags[nblk] + err word; reuse the status slot
  size_t need = sizeof(int) * ((size_t)nblk + 64);
  void* base = nullptr;
  bool fresh = h->ws.bytes[kWsStatus] < need;
  EAMD_CHECK(h->ws.get(kWsStatus, need, &base));
  int* err = static_cast<int*>(base);
  int* flags = err + 64;
  if (fresh || h->trsv_epoch > (1 << 30)) {
    EAMD_HIP(hipMemsetAsync(base, 0, need, h->stream));
    h->trsv_epoch = 0;
  }
  PhaseScope ps(h, kPhTrsv, sizeof(T) * (0.5 * (double)n * n + 2.0 * (double)n * nrhs),
                (double)n * n * nrhs);
  for (int c0 = 0; c0 < nrhs; c0 += NRC) {
    int epoch = ++h->trsv_epoch;
#define EAMD_TRSV(RM, LO, UN)                                                                   \
  hipLaunchKernelGGL((trsv_chain_kernel<T, RM, LO, UN, BS, NRC>), dim3(nblk), dim3(256), 0,     \
                     h->stream, n, A, lda, X, ldx, c0, nrhs, flags, epoch, err)
    if (rowmajor) {
      if (lower) { if (unit) EAMD_TRSV(true, true, true); else EAMD_TRSV(true, true, false); }
      else { if (unit) EAMD_TRSV(true, false, true); else EAMD_TRSV(true, false, false); }
    } else {
      if (lower) { if (unit) EAMD_TRSV(false, true, true); else EAMD_TRSV(false, true, false); }
      else { if (unit) EAMD_TRSV(false, false, true); else EAMD_TRSV(false, false, false); }
    }
#undef EAMD_TRSV
    EAMD_HIP(hipGetLastError());
  }
  return kOk;
}

int trsv_check_error(eamd_handle* h) {
  if (!h->ws.ptr[kWsStatus]) return kOk;
  int e = 0;
  EAMD_HIP(hipMemcpyAsync(&e, h->ws.ptr[kWsStatus], sizeof(int), hipMemcpyDeviceToHost,
                          h->stream));
  EAMD_HIP(hipStreamSynchronize(h->stream));
  if (e) {
    EAMD_HIP(hipMemset(h->ws.ptr[kWsStatus], 0, sizeof(int)));
    set_error("triangular solve: dependency wait timed out");
    return kErrHip;
  }
  return kOk;
}

template int trsv_device<double>(eamd_handle*, int, const double*, int64_t, bool, bool, bool,
                                 double*, int64_t, int);
template int trsv_device<float>(eamd_handle*, int, const float*, int64_t, bool, bool, bool,
                                float*, int64_t, int);

// ---------------------------------------------------------------------------
// Row interchanges of a column-major right-hand side by a transposition
// sequence (Transpositions applied in order, PermutationMatrix.h:82-88 /
// ProductEvaluators.h:927-980).  The sequence is applied by one workgroup per
// right-hand side column with the column resident in LDS (n <= kMaxLdsRows),
// else directly in global memory.
namespace {
constexpr int kMaxLdsRows = 18432;

template <typename T>
__global__ __launch_bounds__(256) void apply_transp_kernel(int n, const int* __restrict__ t,
                                                           int forward, T* __restrict__ X,
                                                           int64_t ldx) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* col = reinterpret_cast<T*>(smem_raw);
  T* x = X + (int64_t)blockIdx.x * ldx;
  const bool in_lds = n <= kMaxLdsRows;
  if (in_lds) {
    for (int i = threadIdx.x; i < n; i += 256) col[i] = x[i];
    __syncthreads();
  }
  T* v = in_lds ? col : x;
  if (threadIdx.x == 0) {
    if (forward) {
      for (int k = 0; k < n; ++k) {
        int p = t[k];
        if (p != k) { T a = v[k]; v[k] = v[p]; v[p] = a; }
      }
    } else {
      for (int k = n - 1; k >= 0; --k) {
        int p = t[k];
        if (p != k) { T a = v[k]; v[k] = v[p]; v[p] = a; }
      }
    }
  }
  if (in_lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) x[i] = col[i];
  }
}
}  // namespace

template <typename T>
int apply_transpositions(eamd_handle* h, int n, const int* t, bool forward, T* X, int64_t ldx,
                         int nrhs) {
  if (n <= 0 || nrhs <= 0) return kOk;
  size_t lds = n <= kMaxLdsRows ? sizeof(T) * (size_t)n : 0;
  auto kern = apply_transp_kernel<T>;
  EAMD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               150 * 1024));
  hipLaunchKernelGGL(kern, dim3(nrhs), dim3(256), lds, h->stream, n, t, forward ? 1 : 0, X, ldx);
  EAMD_HIP(hipGetLastError());
  return kOk;
}
template int apply_transpositions<double>(eamd_handle*, int, const int*, bool, double*, int64_t,
                                          int);
template int apply_transpositions<float>(eamd_handle*, int, const int*, bool, float*, int64_t,
                                         int);

}  // namespace eamd
