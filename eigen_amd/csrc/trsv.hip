// trsv.hip -- blocked triangular solves with few right-hand sides.
//
// Replaces triangular_solve_vector (TriangularSolverVector.h:88-139) and the
// few-RHS use of triangular_solve_matrix (TriangularSolverMatrix.h:38-184)
// inside PartialPivLU::_solve_impl (PartialPivLU.h:226-244) and
// LLT::_solve_impl (LLT.h:475-505).
//
// One workgroup per BS-row block, dispatched in dependency order.  Workgroup
// b first inverts its diagonal block in LDS (independent of the others, so
// it overlaps the wait), accumulates B_b - sum_j T_bj X_j over its
// predecessors j in order, then X_b = inv(T_bb) * (...).  Solved values are
// published as 8-byte words carrying 32 bits of payload and the call's
// epoch (an aligned 8-byte store is single-copy atomic, so a word with the
// right epoch has the right payload): consumers poll the words themselves,
// with no flag and no cache write-back or invalidate on the critical path.
// Several right-hand sides run as up to four independent chains of four
// columns in one launch, their workgroups interleaved block by block.
#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

// polls a wait may take before it gives up (and reports a time-out) unless
// EAMD_TRSV_SPIN sets another limit
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

typedef unsigned long long u64;

template <typename T>
struct XWords {
  static constexpr int n = sizeof(T) / 4;  // 32-bit payload words per value
};

template <typename T, bool ROWMAJOR, bool LOWER, bool UNIT, int BS, int NRC>
__global__ __launch_bounds__(4 * BS) void trsv_chain_kernel(int n, const T* __restrict__ A,
                                                         int64_t lda, T* __restrict__ X,
                                                         int64_t ldx, int c0, int nrhs,
                                                         u64* __restrict__ xw, unsigned epoch,
                                                         int* __restrict__ err, int nch,
                                                         int spin_limit) {
  static_assert((BS == 64 || BS == 128) && (NRC == 16 || NRC == 4 || NRC == 1),
                "four threads per row of a 64- or 128-row block");
  constexpr int NT = 4 * BS;   // threads
  constexpr int SEG = BS / 4;  // columns of a predecessor block per thread
  constexpr int NW = XWords<T>::n;
  constexpr int CHUNK = BS * NRC * NW;          // words of one block's solution
  constexpr int WPT = (CHUNK + NT - 1) / NT;    // words staged per thread
  __shared__ T Dinv[BS][BS + 1];
  __shared__ T Tb[BS][BS + 1];
  __shared__ T accs[BS][NRC];
  __shared__ unsigned xsw[2][CHUNK];  // staged solution of the current predecessor
  __shared__ int s_abort;
  // FP32, one right-hand side: M = inv(T_bb) T_bl for the last predecessor l
  // is formed while the block waits, and inv(T_bb) (B_b - sum over the other
  // predecessors) before l's solution arrives, so the step after it is one
  // row of M times X_l (one barrier instead of two, no reduction through LDS)
  constexpr bool kM = NRC == 1 && sizeof(T) == 4;
  __shared__ T Tl[kM ? BS : 1][BS + 1];
  __shared__ unsigned xwv[kM ? NT / 64 : 1][BS];  // (kM) each wave's copy of X_l
  const int tid = threadIdx.x;
  const int nblk = (n + BS - 1) / BS;
  // nch independent chains of NRC columns each, interleaved block by block
  // (workgroup w * nch + k is block w of chain k: every chain's blocks are
  // dispatched in dependency order), each with its own exchange words
  const int chain = blockIdx.x % nch;
  const int w = blockIdx.x / nch;
  c0 += chain * NRC;
  xw += (size_t)chain * nblk * BS * NRC * NW;
  const int b = LOWER ? w : nblk - 1 - w;
  const int rb0 = b * BS;
  const int bs = min(BS, n - rb0);
  const int nc = min(NRC, nrhs - c0);
  const int r = tid >> 2, q = tid & 3;  // my row of the block, my quarter of each column block
  const int ndep = LOWER ? b : nblk - 1 - b;
  const int jl = LOWER ? b - 1 : b + 1;  // the last predecessor (ndep > 0)

  // ---- diagonal block into LDS and its inverse (independent of the
  // predecessors: overlaps their solves)
  for (int e = tid; e < BS * BS; e += NT) {
    const int rr = e / BS, cc = e - rr * BS;
    Tb[rr][cc] = (rr < bs && cc < bs) ? *tptr<T, ROWMAJOR>(A, lda, rb0 + rr, rb0 + cc) : T(0);
    Dinv[rr][cc] = T(0);  // the padding of a partial last block stays exactly zero
    if constexpr (kM)
      Tl[rr][cc] = (ndep > 0 && rr < bs && jl * BS + cc < n)
                       ? *tptr<T, ROWMAJOR>(A, lda, rb0 + rr, jl * BS + cc)
                       : T(0);
  }
  if (tid == 0) s_abort = 0;
  __syncthreads();
  if (tid < BS) {
    const int c = tid;
    if (c < bs) {
      if (LOWER) {
        const T d = UNIT ? T(1) : T(1) / Tb[c][c];
        Dinv[c][c] = d;
        for (int rr = c + 1; rr < bs; ++rr) {
          T sum = T(0);
          for (int qq = c; qq < rr; ++qq) sum += Tb[rr][qq] * Dinv[qq][c];
          Dinv[rr][c] = UNIT ? -sum : -sum / Tb[rr][rr];
        }
      } else {
        const T d = UNIT ? T(1) : T(1) / Tb[c][c];
        Dinv[c][c] = d;
        for (int rr = c - 1; rr >= 0; --rr) {
          T sum = T(0);
          for (int qq = rr + 1; qq <= c; ++qq) sum += Tb[rr][qq] * Dinv[qq][c];
          Dinv[rr][c] = UNIT ? -sum : -sum / Tb[rr][rr];
        }
      }
    }
  }

  // ---- off-diagonal part: thread (r, q) accumulates row r over columns
  // q*SEG .. q*SEG+SEG-1 of every predecessor block.  Per predecessor the
  // workgroup stages its solution from the exchange words into LDS (one
  // barrier); the matrix values, which do not depend on X, are prefetched
  // one predecessor ahead.
  // (kM) this thread's quarter of row r of M
  T mrow[kM ? SEG : 1];
  if constexpr (kM) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SEG; ++i) mrow[i] = T(0);
    if (ndep > 0) {
#pragma unroll 4
      for (int k = 0; k < BS; ++k) {
        const T dk = Dinv[r][k];
#pragma unroll
        for (int i = 0; i < SEG; ++i) mrow[i] += dk * Tl[k][q * SEG + i];
      }
    }
  }
  T acc[NRC];
#pragma unroll
  for (int c = 0; c < NRC; ++c) acc[c] = T(0);
  // predecessors taken in the loop (kM: all but the last)
  const int nloop = kM ? max(ndep - 1, 0) : ndep;
  const bool rvalid = r < bs;
  // the block's right-hand sides, requested now: read after the loop they
  // would put a memory round trip on every link of the chain
  T bval[NRC];
#pragma unroll
  for (int c = 0; c < NRC; ++c)
    bval[c] = (q == 0 && rvalid && c < nc) ? X[(int64_t)(c0 + c) * ldx + rb0 + r] : T(0);
  // a row-major segment is contiguous: 16-byte loads when aligned (a quarter
  // or half of the address work of element loads, which the CU's address
  // unit otherwise limits)
  constexpr int PW = 16 / sizeof(T);  // values per 16-byte word
  const bool vec = ROWMAJOR && ((uintptr_t)A & 15) == 0 && lda % PW == 0;
  auto load_seg = [&](int d, T (&tv)[SEG]) {
    const int jb = LOWER ? d : nblk - 1 - d;
    const int cb = jb * BS + q * SEG;
    if (vec && cb + SEG <= n) {
      const T* p = tptr<T, ROWMAJOR>(A, lda, rb0 + r, cb);
#pragma unroll
      for (int i = 0; i < SEG; i += PW) {
        if (sizeof(T) == 4) {
          const v4f w = rvalid ? *reinterpret_cast<const v4f*>(p + i) : v4f{0.f, 0.f, 0.f, 0.f};
          tv[i] = (T)w.x;
          tv[i + 1] = (T)w.y;
          tv[i + 2] = (T)w.z;
          tv[i + 3] = (T)w.w;
        } else {
          typedef double v2d __attribute__((ext_vector_type(2)));
          const v2d w = rvalid ? *reinterpret_cast<const v2d*>(p + i) : v2d{0.0, 0.0};
          tv[i] = (T)w.x;
          tv[i + 1] = (T)w.y;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < SEG; ++i)
        tv[i] = (rvalid && cb + i < n) ? *tptr<T, ROWMAJOR>(A, lda, rb0 + r, cb + i) : T(0);
    }
  };
  int bad = 0;
  // the solution of predecessor step d, from its exchange words into
  // xsw[d & 1] (my words all requested together, then only the words not
  // yet published read again), then one barrier
  auto stage = [&](int d) {
    const int jb = LOWER ? d : nblk - 1 - d;
    const int par = d & 1;
    // my words of the predecessor's solution (row jr, column c, half h): all
    // loads in flight together, then only the words not yet published are
    // read again
    const u64* wsrc[WPT];
    u64 wv[WPT];
    bool need[WPT];
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      const int widx = tid + k * NT;
      const int row = jb * BS + widx / (NRC * NW);
      const int c = (widx / NW) % NRC;
      need[k] = widx < CHUNK && row < n && c < nc;
      wsrc[k] = xw + (size_t)row * NRC * NW + (widx % (NRC * NW));
      wv[k] = need[k] ? __hip_atomic_load(const_cast<u64*>(wsrc[k]), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                      : 0ull;
      if (widx < CHUNK && !need[k]) xsw[par][widx] = 0u;
    }
    for (int spins = 0;;) {
      bool all = true;
#pragma unroll
      for (int k = 0; k < WPT; ++k)
        if (need[k]) {
          if ((unsigned)(wv[k] >> 32) == epoch) {
            xsw[par][tid + k * NT] = (unsigned)wv[k];
            need[k] = false;
          } else {
            all = false;
          }
        }
      if (all) break;
      // the limit every poll, another workgroup's give-up every 256th
      ++spins;
      if (spins > spin_limit ||
          ((spins & 255) == 0 &&
           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        bad = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int k = 0; k < WPT; ++k)
        if (need[k])
          wv[k] = __hip_atomic_load(const_cast<u64*>(wsrc[k]), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    }
    if (bad) s_abort = 1;
    __syncthreads();
    if (s_abort) bad = 1;
  };
  T tcur[SEG], tnext[SEG];
  if (nloop > 0) load_seg(0, tcur);
  for (int d = 0; d < nloop && !bad; ++d) {
    if (d + 1 < nloop) load_seg(d + 1, tnext);
    const int jb = LOWER ? d : nblk - 1 - d;
    const int par = d & 1;
    stage(d);
    if (bad) break;
    const int cb = jb * BS + q * SEG;
#pragma unroll
    for (int c = 0; c < NRC; ++c) {
      if (c < nc) {
        T sum = T(0);
#pragma unroll
        for (int i = 0; i < SEG; ++i) {
          const unsigned* wp = &xsw[par][((q * SEG + i) * NRC + c) * NW];
          T xv;
          if (NW == 2)
            xv = (T)__longlong_as_double((long long)(((u64)wp[1] << 32) | wp[0]));
          else
            xv = (T)__uint_as_float(wp[0]);
          if (cb + i < n) sum += tcur[i] * xv;
        }
        acc[c] += sum;
      }
    }
#pragma unroll
    for (int i = 0; i < SEG; ++i) tcur[i] = tnext[i];
  }
  if (bad) {
    s_abort = 1;
    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // the four quarters of a row sit in lanes 4r .. 4r+3 of one wave
#pragma unroll
  for (int c = 0; c < NRC; ++c) {
    acc[c] += __shfl_xor(acc[c], 1);
    acc[c] += __shfl_xor(acc[c], 2);
  }
  if (q == 0) {
#pragma unroll
    for (int c = 0; c < NRC; ++c)
      accs[r][c] = (rvalid && c < nc) ? bval[c] - acc[c] : T(0);
  }
  __syncthreads();
  if (s_abort) return;
  // ---- X_b = inv(T_bb) * (B_b - sum): thread (r, q) takes a quarter of the row of Dinv
  T xs[NRC];
#pragma unroll
  for (int c = 0; c < NRC; ++c) {
    T sum = T(0);
#pragma unroll
    for (int i = 0; i < SEG; ++i) sum += Dinv[r][q * SEG + i] * accs[q * SEG + i][c];
    xs[c] = sum;
  }
#pragma unroll
  for (int c = 0; c < NRC; ++c) {
    xs[c] += __shfl_xor(xs[c], 1);
    xs[c] += __shfl_xor(xs[c], 2);
  }
  if constexpr (kM) {
    // X_b -= M X_l once l's solution is here: every wave polls X_l's BS
    // words itself (lane l: words l, l + 64) and publishes its own rows, so
    // no workgroup barrier sits on the chain
    if (ndep > 0) {
      constexpr int WL = BS / 64;
      const int lane = tid & 63, wv = tid >> 6;
      const u64* src[WL];
      bool need[WL];
      u64 w[WL];
#pragma unroll
      for (int k = 0; k < WL; ++k) {
        const int row = jl * BS + lane + 64 * k;
        need[k] = row < n;
        src[k] = xw + (size_t)row;
        w[k] = need[k] ? __hip_atomic_load(const_cast<u64*>(src[k]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)
                       : (u64)epoch << 32;
      }
      auto missing = [&]() {
        bool m = false;
#pragma unroll
        for (int k = 0; k < WL; ++k) m = m || (unsigned)(w[k] >> 32) != epoch;
        return m;
      };
      for (int spins = 0; __ballot(missing()) != 0ull;) {
        ++spins;
        if (spins > spin_limit ||
            ((spins & 255) == 0 &&
             __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          return;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int k = 0; k < WL; ++k)
          if ((unsigned)(w[k] >> 32) != epoch)
            w[k] = __hip_atomic_load(const_cast<u64*>(src[k]), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      }
      // the wave's own LDS copy (a wave's LDS accesses execute in order)
#pragma unroll
      for (int k = 0; k < WL; ++k) xwv[wv][lane + 64 * k] = (unsigned)w[k];
      asm volatile("" ::: "memory");
      T m = T(0);
#pragma unroll
      for (int i = 0; i < SEG; ++i) m += mrow[i] * __uint_as_float(xwv[wv][q * SEG + i]);
      m += __shfl_xor(m, 1);
      m += __shfl_xor(m, 2);
      xs[0] -= m;
    }
  }
  // publish: the result, and its exchange words for the successors
  if (q == 0 && rvalid) {
#pragma unroll
    for (int c = 0; c < NRC; ++c)
      if (c < nc) {
        X[(int64_t)(c0 + c) * ldx + rb0 + r] = xs[c];
        u64* dst = xw + ((size_t)(rb0 + r) * NRC + c) * NW;
        if (NW == 2) {
          const u64 bits = (u64)__double_as_longlong((double)xs[c]);
          __hip_atomic_store(dst, ((u64)epoch << 32) | (unsigned)bits, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(dst + 1, ((u64)epoch << 32) | (unsigned)(bits >> 32),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          __hip_atomic_store(dst, ((u64)epoch << 32) | __float_as_uint((float)xs[c]),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
  }
}

}  // namespace

// Solves op(T) X = X in place.  T is n x n; `rowmajor` selects its layout;
// lower/unit select the triangle.  X is column-major n x nrhs.
template <typename T>
int trsv_device(eamd_handle* h, int n, const T* A, int64_t lda, bool rowmajor, bool lower,
                bool unit, T* X, int64_t ldx, int nrhs) {
  if (n <= 0 || nrhs <= 0) return kOk;
  static const int spin_limit = env_int("EAMD_TRSV_SPIN", kSpin);
  constexpr int BS = 64, NRC = 4, kChains = 4;
  const int nblk = ceil_div(n, BS);
  // error word | exchange words (kChains chains of NRC values per row, 32
  // payload bits each)
  const size_t nwords = (size_t)nblk * BS * NRC * kChains * (sizeof(T) / 4);
  size_t need = 256 + sizeof(unsigned long long) * nwords;
  void* base = nullptr;
  bool fresh = h->ws.bytes[kWsStatus] < need;
  EAMD_CHECK(h->ws.get(kWsStatus, need, &base));
  int* err = static_cast<int*>(base);
  unsigned long long* xw =
      reinterpret_cast<unsigned long long*>(static_cast<char*>(base) + 256);
  if (fresh || h->trsv_epoch > (1 << 30)) {
    EAMD_HIP(hipMemsetAsync(base, 0, h->ws.bytes[kWsStatus], h->stream));
    h->trsv_epoch = 0;
  }
  PhaseScope ps(h, kPhTrsv, sizeof(T) * (0.5 * (double)n * n + 2.0 * (double)n * nrhs),
                (double)n * n * nrhs);
  if (nrhs == 1) {
    // one right-hand side: one exchange word (per 32 payload bits) per row
    // (128-row blocks measured slower: 7.4 vs 3.7 ms of solves per step at
    // n = 16384)
    constexpr int BS1 = 64;
    const int nblk1 = ceil_div(n, BS1);
    const unsigned epoch = (unsigned)++h->trsv_epoch;
#define EAMD_TRSV1(RM, LO, UN)                                                                   \
  hipLaunchKernelGGL((trsv_chain_kernel<T, RM, LO, UN, BS1, 1>), dim3(nblk1), dim3(4 * BS1), 0, \
                     h->stream, n, A, lda, X, ldx, 0, 1, xw, epoch, err, 1, spin_limit)
    if (rowmajor) {
      if (lower) { if (unit) EAMD_TRSV1(true, true, true); else EAMD_TRSV1(true, true, false); }
      else { if (unit) EAMD_TRSV1(true, false, true); else EAMD_TRSV1(true, false, false); }
    } else {
      if (lower) { if (unit) EAMD_TRSV1(false, true, true); else EAMD_TRSV1(false, true, false); }
      else { if (unit) EAMD_TRSV1(false, false, true); else EAMD_TRSV1(false, false, false); }
    }
#undef EAMD_TRSV1
    EAMD_HIP(hipGetLastError());
    return kOk;
  }
  // several right-hand sides: chains of four columns, up to four chains in
  // one launch.  A chain's step costs about linearly in its columns (n =
  // 8192: 24 us per 64-row step with sixteen, 6.5 us with four), so chains
  // run side by side instead of one wide chain (LLT solve with 16
  // right-hand sides: 6.3 ms with one chain of sixteen, 6.7 ms with four
  // chains of four one after the other).
  for (int c0 = 0; c0 < nrhs;) {
    const unsigned epoch = (unsigned)++h->trsv_epoch;
    const int nch = std::min(kChains, ceil_div(nrhs - c0, NRC));
#define EAMD_TRSV(RM, LO, UN)                                                                  \
  hipLaunchKernelGGL((trsv_chain_kernel<T, RM, LO, UN, BS, NRC>), dim3(nblk * nch), dim3(256), \
                     0, h->stream, n, A, lda, X, ldx, c0, nrhs, xw, epoch, err, nch, spin_limit)
    if (rowmajor) {
      if (lower) { if (unit) EAMD_TRSV(true, true, true); else EAMD_TRSV(true, true, false); }
      else { if (unit) EAMD_TRSV(true, false, true); else EAMD_TRSV(true, false, false); }
    } else {
      if (lower) { if (unit) EAMD_TRSV(false, true, true); else EAMD_TRSV(false, true, false); }
      else { if (unit) EAMD_TRSV(false, false, true); else EAMD_TRSV(false, false, false); }
    }
#undef EAMD_TRSV
    EAMD_HIP(hipGetLastError());
    c0 += nch * NRC;
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
