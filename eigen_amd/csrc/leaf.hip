// leaf.hip -- the partial-pivoting column loop of the LU panel.
//
// Restates unblocked_lu (Eigen/src/LU/PartialPivLU.h:365-409) for a leaf of
// at most 32 columns of a tall row-major sub-panel: per column, the first
// maximum of |a| (strict '>' scan in row order, Visitor.h:179; a NaN wins only
// from the diagonal position), the row interchange, the division of the
// column by the pivot and the rank-1 update of the leaf.
//
// MI355X mapping.  Every thread keeps one row of the leaf (32 values) in
// registers, so the rank-1 update is register FMAs; a workgroup holds
// BLOCK rows (512 in FP64, 1024 in FP32) and at most 32 workgroups cover a
// 16384-row panel.  With several workgroups, each column costs ONE memory
// round trip: every workgroup publishes its candidate (value, row) together
// with the whole candidate row, and the owner of row j publishes row j, all
// as 8-byte words carrying 32 bits of payload and the column's epoch (an
// aligned 8-byte store is single-copy atomic, so a word with the right epoch
// has the right payload: no fences).  Every workgroup then reads all
// candidates and row j in one parallel poll and resolves the winner locally.
#include <type_traits>
#include <utility>

#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int LW = 32;       // leaf width
constexpr int kMaxG = 32;    // workgroups in the exchange
constexpr int kLeafSpin = 1 << 22;
typedef unsigned long long u64;

template <typename T>
struct Words {
  static constexpr int n = sizeof(T) / 4;  // 32-bit payload words per value
};

template <typename T>
struct LeafBlock {
  static constexpr int v = sizeof(T) == 8 ? 512 : 1024;  // rows per workgroup
};

__device__ __forceinline__ void put_word(u64* p, unsigned payload, unsigned epoch) {
  __hip_atomic_store(p, ((u64)epoch << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void put_value(u64* p, T v, unsigned epoch) {
  if (sizeof(T) == 8) {
    u64 bits = (u64)__double_as_longlong((double)v);
    put_word(p, (unsigned)bits, epoch);
    put_word(p + 1, (unsigned)(bits >> 32), epoch);
  } else {
    put_word(p, __float_as_uint((float)v), epoch);
  }
}

// Bit-mask selection (a ? x : y with a mask of all ones or zeros).  Written
// on the integer bits so the compiler does not fold a chain of them back
// into a runtime-indexed access of the register array (which it would place
// in scratch memory).
__device__ __forceinline__ float mask_sel(unsigned m, float x, float y) {
  return __uint_as_float((__float_as_uint(x) & m) | (__float_as_uint(y) & ~m));
}
__device__ __forceinline__ double mask_sel(unsigned m, double x, double y) {
  const unsigned long long mm = ((unsigned long long)m << 32) | m;
  return __longlong_as_double((long long)(((unsigned long long)__double_as_longlong(x) & mm) |
                                          ((unsigned long long)__double_as_longlong(y) & ~mm)));
}

template <typename T>
__device__ __forceinline__ T get_value(const unsigned* w) {
  if (sizeof(T) == 8) return (T)__longlong_as_double((long long)(((u64)w[1] << 32) | w[0]));
  return (T)__uint_as_float(w[0]);
}

__device__ __forceinline__ bool beats(double s, int i, double bs, int bi) {
  return s > bs || (s == bs && i < bi);
}

__device__ __forceinline__ double score_of(double v, bool first) {
  double s = fabs(v);
  if (s != s) s = first ? INFINITY : -1.0;
  return s;
}

// Rotation of a register row by 8 slots (compile-time moves): the column
// loops work on groups of 8 columns that always sit in slots 0..7, so every
// register access has a compile-time index; after 4 rotations the row is back
// in its original order.
template <typename T>
__device__ __forceinline__ void rotate8(T (&v)[32]) {
  T t[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) t[q] = v[q];
#pragma unroll
  for (int c = 0; c < 24; ++c) v[c] = v[c + 8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[24 + q] = t[q];
}

// Lane permutations of the DPP unit, applied to a double as two halves.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((u64)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double readlane_d(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((u64)(unsigned)hi << 32) | (unsigned)lo));
}

// Maximum over the 64 lanes of a wave (inputs never NaN): xor-1 and xor-2
// within quads, half-row and row mirrors within each 16-lane row, then the
// four row results read as uniform values.
__device__ __forceinline__ double wave_max(double s) {
  s = fmax(s, dpp_d<0xB1>(s));   // quad_perm [1,0,3,2]
  s = fmax(s, dpp_d<0x4E>(s));   // quad_perm [2,3,0,1]
  s = fmax(s, dpp_d<0x141>(s));  // row_half_mirror
  s = fmax(s, dpp_d<0x140>(s));  // row_mirror
  return fmax(fmax(readlane_d(s, 0), readlane_d(s, 16)), fmax(readlane_d(s, 32), readlane_d(s, 48)));
}

// Maximum of a 64-bit key over the 64 lanes of a wave (same lane pattern).
template <int CTRL>
__device__ __forceinline__ u64 dpp_u64(u64 x) {
  const int lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xf, 0xf, false);
  return ((u64)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ u64 readlane_u64(u64 x, int lane) {
  const int lo = __builtin_amdgcn_readlane((int)x, lane);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
  return ((u64)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ u64 umax64(u64 a, u64 b) { return a > b ? a : b; }
__device__ __forceinline__ u64 wave_max_u64(u64 k) {
  k = umax64(k, dpp_u64<0xB1>(k));
  k = umax64(k, dpp_u64<0x4E>(k));
  k = umax64(k, dpp_u64<0x141>(k));
  k = umax64(k, dpp_u64<0x140>(k));
  return umax64(umax64(readlane_u64(k, 0), readlane_u64(k, 16)),
                umax64(readlane_u64(k, 32), readlane_u64(k, 48)));
}

// Slot (words): [0..1] candidate value (FP64 bits), [2] row or ~0u (no
// candidate), [3] unused, [4, 4 + LW*Words) candidate row.  The slot after
// the last workgroup's holds row j (from word 4).
template <typename T>
constexpr int slot_words() {
  return 4 + LW * Words<T>::n;
}

__device__ __forceinline__ int readlane_i(int x, int lane) {
  return __builtin_amdgcn_readlane(x, lane);
}
__device__ __forceinline__ float readlane_f(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// The leaf-level permutation of a leaf's row interchanges, in registers (one
// wave): entry k (lane k < ln) is leaf row tpos, which receives the content
// of entry tsrc and so holds original row trow (rows 0..w-1 are entries
// 0..w-1, rows swapped in from below are appended).  FROM_SEL: lane c < w
// holds x = the original row selected for row c, and x returns the
// transposition of column c (the row swapped with row c); otherwise x is
// that transposition.
template <bool FROM_SEL>
__device__ __forceinline__ void leaf_perm(int& x, int w, int lane, int& tpos, int& trow,
                                          int& tsrc, int& ln) {
  tpos = lane < w ? lane : -1;
  trow = tpos;
  tsrc = lane;
  int nent = w, pivv = 0;
  for (int c = 0; c < w; ++c) {
    const int p = readlane_i(x, c);
    const u64 hit = __ballot((FROM_SEL ? trow : tpos) == p);
    int k;
    if (hit) {
      k = __ffsll((long long)hit) - 1;
    } else {  // an untouched row p (>= w) sits at position p
      k = nent++;
      if (lane == k) { tpos = p; trow = p; tsrc = k; }
    }
    if (FROM_SEL) {
      const int pk = readlane_i(tpos, k);
      if (lane == c) pivv = pk;
    }
    const int rc = readlane_i(trow, c), rk = readlane_i(trow, k);
    const int sc = readlane_i(tsrc, c), sk = readlane_i(tsrc, k);
    if (lane == c) { trow = rk; tsrc = sk; }
    if (lane == k) { trow = rc; tsrc = sc; }
  }
  if (FROM_SEL) x = pivv;
  ln = nent;
}

// Composes a leaf's list (lane k < ln: leaf row tpos receives the content of
// entry tsrc) into the panel-level list ppos/psrc/pperm_n (the leaf starts at
// panel row row0; panel rows 0..pnb-1 are slots 0..pnb-1, rows below them are
// appended slots; slot s: row ppos[s] receives the content of slot psrc[s]).
// All entries at once: a touched slot's new source is the old source of its
// entry's source slot, and a newly appended slot's old source is itself.
// The panel's first leaf (row0 == 0) starts the list as the identity.  One
// wave; s_pp is LDS scratch of at least pnb + 4 ints.
__device__ void panel_compose(int tpos, int tsrc, int ln, int row0, int pnb, int lane,
                              int* pperm_n, int* ppos, int* psrc, int* s_pp, const int* s_ps,
                              int pn0_pre) {
  const bool first = row0 == 0;
  if (first) {
    for (int e = lane; e < pnb; e += 64) {
      ppos[e] = e;
      psrc[e] = e;
    }
    // complete before the touched slots are stored over them
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  }
  // pn0_pre >= 0: the incoming list was prefetched (count pn0_pre, appended
  // rows in s_pp padded with -1 to a multiple of 4, sources in s_ps)
  const bool pre = pn0_pre >= 0;
  const int pn0 = first ? pnb : (pre ? pn0_pre : *pperm_n), na = pn0 - pnb, na4 = (na + 3) & ~3;
  if (!pre)
    for (int e = lane; e < na4; e += 64) s_pp[e] = e < na ? ppos[pnb + e] : -1;
  asm volatile("" ::: "memory");  // LDS accesses of one wave execute in order
  const bool mine = lane < ln;
  const int prow = row0 + tpos;
  int slot = prow;
  bool need = false;
  if (mine && prow >= pnb) {
    int f = -1;
    for (int e = 0; e < na4; e += 4) {
      const int4 q = *reinterpret_cast<const int4*>(&s_pp[e]);
      f = q.x == prow ? pnb + e : f;
      f = q.y == prow ? pnb + e + 1 : f;
      f = q.z == prow ? pnb + e + 2 : f;
      f = q.w == prow ? pnb + e + 3 : f;
    }
    slot = f;
    need = f < 0;
  }
  const u64 nb = __ballot(need);
  if (need) slot = pn0 + __popcll(nb & ((1ull << lane) - 1ull));
  int old = slot;
  if (mine && !need && !first) old = pre ? s_ps[slot] : psrc[slot];
  const int src_old = __shfl(old, tsrc);
  if (mine) {
    psrc[slot] = src_old;
    if (need) ppos[slot] = prow;
  }
  if (lane == 0) *pperm_n = pn0 + __popcll(nb);
}

// All lists of a leaf from its interchanges (one wave): piv[] (plus
// piv_offset), the leaf-level list lperm_n/lpos/lsrc and the panel-level one.
// FROM_SEL as for leaf_perm; tpos/trow/tsrc return the leaf list's entries.
template <bool FROM_SEL>
__device__ void leaf_lists(int x, int w, int row0, int pnb, int lane, int* piv, int piv_offset,
                           int* lperm_n, int* lpos, int* lsrc, int* pperm_n, int* ppos,
                           int* psrc, int* s_pp, int& tpos, int& trow, int& tsrc, int& ln,
                           const int* s_ps = nullptr, int pn0_pre = -1) {
  leaf_perm<FROM_SEL>(x, w, lane, tpos, trow, tsrc, ln);
  if (lane < w) piv[lane] = x + piv_offset;
  if (lane < ln) {
    lpos[lane] = tpos;
    lsrc[lane] = tsrc;
  }
  if (lane == 0) *lperm_n = ln;
  panel_compose(tpos, tsrc, ln, row0, pnb, lane, pperm_n, ppos, psrc, s_pp, s_ps, pn0_pre);
}

}  // namespace

// Outputs: piv[c] leaf-relative transposition of leaf column c; leaf
// permutation lists (rows touched, leaf-relative) lperm_n/lpos/lsrc; the
// panel-level lists pperm_n/ppos/psrc (panel-relative rows, composed across
// the leaves of a panel: on entry they hold the permutation of the previous
// leaves; `row0` is the leaf's first row inside the panel, `pnb` the panel
// width); zero_col the first exactly-zero pivot column of the leaf or -1.
// piv[] is stored plus piv_offset, and a zero pivot is also recorded in
// *fz_global (nullable) as first_zero + piv_offset when that is still < 0.
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void leaf_lu_kernel(
    T* __restrict__ P, int64_t ldp, int m, int w, int G, int* __restrict__ piv,
    int* __restrict__ lperm_n, int* __restrict__ lpos, int* __restrict__ lsrc,
    int* __restrict__ pperm_n, int* __restrict__ ppos, int* __restrict__ psrc, int row0,
    int pnb, int* __restrict__ zero_col, u64* __restrict__ xbuf, int* __restrict__ err,
    unsigned epoch_base, int piv_offset, int* __restrict__ fz_global, int spin_limit) {
  __builtin_amdgcn_s_setprio(3);  // on the critical path (see tsl_kernel)
  constexpr int NW = Words<T>::n;
  constexpr int SW = slot_words<T>();
  constexpr int NWAVE = BLOCK / 64;
  constexpr int STAGE = (kMaxG + 1) * SW;
  constexpr int KMAX = (STAGE + BLOCK - 1) / BLOCK;  // words polled per thread
  // per column, double-buffered by parity (one barrier per column): each
  // wave's best score, row index, value and row, and row j when owned here
  __shared__ double red_s[2][NWAVE];
  __shared__ int red_i[2][NWAVE];
  __shared__ double red_v[2][NWAVE];
  __shared__ T candw[2][NWAVE][LW];
  __shared__ T urow[2][LW];   // pivot row received from the exchange
  __shared__ T rowj[2][LW];   // old row j received from the exchange
  __shared__ T rowjL[2][LW];  // row j, when this workgroup owns it
  __shared__ unsigned stage[STAGE];
  __shared__ int s_abort;
  __shared__ int s_win[2];  // the exchange's winning row, by column parity
  // bookkeeping (workgroup 0): the leaf's pivots and the panel list's rows
  __shared__ int s_piv[LW];
  __shared__ __align__(16) int s_pp[256 + 4];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x;
  const int i = g * BLOCK + tid;  // my leaf-relative row
  const bool valid = i < m;

  T v[LW];
  {
    const T* src = P + (int64_t)i * ldp;
#pragma unroll
    for (int c = 0; c < LW; ++c) v[c] = (valid && c < w) ? src[c] : T(0);
  }
  if (tid == 0) s_abort = 0;
  __syncthreads();

  int first_zero = -1;
  // A runtime column loop over a register row: every access to v[] uses a
  // compile-time index, with the column selected by branch-free selects and
  // masked FMAs (a runtime index into a register array would go to scratch
  // memory; a fully unrolled loop overflows the instruction cache).
  // Columns in groups of 8 that sit in register slots 0..7 (see rotate8):
  // column j = 8*g8 + jj is slot jj, a compile-time index; slots [live, 32)
  // hold columns finished in earlier groups and take no update.
  auto column = [&](auto jjc, int j, int live) __attribute__((always_inline)) -> bool {
    constexpr int jj = decltype(jjc)::value;
    const int par = j & 1;
    const unsigned epoch = epoch_base + (unsigned)j + 1u;
    const T vj = v[jj];
    // ---- 1. my candidate, the wave's (maximum score, then the lowest lane
    // holding it: lanes are in row order), then the workgroup's
    double sc = -2.0;
    if (valid && i >= j) sc = score_of((double)vj, i == j);
    const double wmax = wave_max(sc);
    const int wl = __ffsll((long long)__ballot(sc == wmax)) - 1;
    // the wave's candidate lane also stores its row, and the owner of row j
    // stores that row, before the column's single barrier
    if (lane == wl) {
      red_s[par][wave] = wmax;
      red_i[par][wave] = wmax > -2.0 ? g * BLOCK + wave * 64 + wl : 0x7fffffff;
      red_v[par][wave] = (double)vj;
#pragma unroll
      for (int c = 0; c < LW; ++c) candw[par][wave][c] = v[c];
    }
    if (i == j) {
#pragma unroll
      for (int c = 0; c < LW; ++c) rowjL[par][c] = v[c];
    }
    __syncthreads();
    double best = red_s[par][0];
    int bq = 0;
#pragma unroll
    for (int q = 1; q < NWAVE; ++q) {
      const double s2 = red_s[par][q];
      if (s2 > best) { best = s2; bq = q; }
    }
    const int bi = red_i[par][bq];
    const double bval = red_v[par][bq];
    // ---- 2. the workgroup's candidate row is the winning wave's
    const bool none = bi == 0x7fffffff;
    const T* up = candw[par][bq];  // pivot row
    const T* rp = rowjL[par];      // old row j
    int p = bi;
    if (G > 1) {
      // ---- 3. publish: header, candidate row and (owner) row j, one word per lane
      u64* myslot = xbuf + ((size_t)par * (G + 1) + g) * SW;
      u64* rj = xbuf + ((size_t)par * (G + 1) + G) * SW;
      if (tid == 0) {
        u64 bits = (u64)__double_as_longlong(bval);
        put_word(myslot, (unsigned)bits, epoch);
        put_word(myslot + 1, (unsigned)(bits >> 32), epoch);
        put_word(myslot + 2, none ? 0xffffffffu : (unsigned)bi, epoch);
        put_word(myslot + 3, 0u, epoch);
      } else if (tid >= 64 && tid < 64 + LW * NW) {
        const int q = tid - 64, c = q / NW, half = q - c * NW;
        const T val = none ? T(0) : up[c];
        unsigned pl;
        if (sizeof(T) == 8) {
          u64 bits = (u64)__double_as_longlong((double)val);
          pl = half ? (unsigned)(bits >> 32) : (unsigned)bits;
        } else {
          pl = __float_as_uint((float)val);
        }
        put_word(myslot + 4 + q, pl, epoch);
      } else if (j >= g * BLOCK && j < (g + 1) * BLOCK && tid >= 128 && tid < 128 + 4 + LW * NW) {
        const int q = tid - 128;
        unsigned pl = 0;
        if (q >= 4) {
          const int c = (q - 4) / NW, half = (q - 4) - c * NW;
          const T val = rowjL[par][c];
          if (sizeof(T) == 8) {
            u64 bits = (u64)__double_as_longlong((double)val);
            pl = half ? (unsigned)(bits >> 32) : (unsigned)bits;
          } else {
            pl = __float_as_uint((float)val);
          }
        }
        put_word(rj + q, pl, epoch);
      }
      // ---- 4. one parallel poll of every slot and of row j: each thread
      // issues the loads of its words together, then re-polls the missing ones
      const int total = (G + 1) * SW;
      const u64* base = xbuf + (size_t)par * (G + 1) * SW;
      bool done[KMAX];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) done[k] = tid + k * BLOCK >= total;
      int bad = 0;
      for (int spins = 0;; ++spins) {
        u64 wv[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
          if (!done[k])
            wv[k] = __hip_atomic_load(const_cast<u64*>(base + tid + k * BLOCK), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        bool all = true;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
          if (!done[k]) {
            if ((unsigned)(wv[k] >> 32) == epoch) {
              stage[tid + k * BLOCK] = (unsigned)wv[k];
              done[k] = true;
            } else {
              all = false;
            }
          }
        if (all) break;
        // the limit every poll, another workgroup's give-up every 256th
        if (spins >= spin_limit ||
            ((spins & 255) == 255 &&
             __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          bad = 1;
          break;
        }
      }
      if (bad) s_abort = 1;
      __syncthreads();
      if (s_abort) {
        if (tid == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
      // ---- resolve the winner: wave 0, lane s holding slot s's candidate;
      // the highest score, then the lowest row among those (the `beats`
      // order), by two wave reductions
      if (wave == 0) {
        double sc = -2.0;
        int idx = 0x7fffffff;
        if (lane < G) {
          const unsigned* hh = &stage[lane * SW];
          const unsigned id = hh[2];
          if (id != 0xffffffffu) {
            sc = score_of(get_value<double>(hh), (int)id == j);
            idx = (int)id;
          }
        }
        const double mx = wave_max(sc);
        const int wi = (int)-wave_max(sc == mx ? -(double)idx : -4294967296.0);
        const int wslot = __ffsll((long long)__ballot(sc == mx && idx == wi)) - 1;
        if (lane < LW) {
          urow[par][lane] = get_value<T>(&stage[wslot * SW + 4 + lane * NW]);
          rowj[par][lane] = get_value<T>(&stage[G * SW + 4 + lane * NW]);
        }
        if (lane == 0) s_win[par] = wi;
      }
      __syncthreads();
      p = s_win[par];
      up = urow[par];
      rp = rowj[par];
    }
    // ---- 5. interchange in registers
    if (p != j) {
      if (i == j) {
#pragma unroll
        for (int c = 0; c < LW; ++c) v[c] = up[c];
      } else if (i == p) {
#pragma unroll
        for (int c = 0; c < LW; ++c) v[c] = rp[c];
      }
    }
    if (g == 0 && tid == 0) s_piv[j] = p;
    // ---- 6. scale and rank-1 update of the rows below j
    // branch-free (a branch around the register row makes the compiler copy
    // the whole row at the join); l = 0 leaves a row unchanged
    const T pv = up[jj];
    if (pv == T(0) && first_zero < 0) first_zero = j;
    const bool act = pv != T(0) && valid && i > j;
    const T qv = (i == p ? rp[jj] : vj) / pv;
    T l = act ? qv : T(0);
    asm volatile("" : "+v"(l));  // keep the update unconditional
    v[jj] = act ? l : v[jj];
    const unsigned lm = 0u - (unsigned)(live > 8), lm16 = 0u - (unsigned)(live > 16),
                   lm24 = 0u - (unsigned)(live > 24);
#pragma unroll
    for (int c = jj + 1; c < LW; ++c) {
      // slot c is live iff c < live; an integer mask on the loaded value keeps
      // the LDS reads unconditional
      const unsigned m = c < 8 ? ~0u : (c < 16 ? lm : (c < 24 ? lm16 : lm24));
      const T u = mask_sel(m, up[c], T(0));
      v[c] -= l * u;
    }
    return true;
  };
  const int ngroups = (w + 7) / 8;
  bool ok = true;
#pragma unroll 1
  for (int g8 = 0; g8 < ngroups && ok; ++g8) {
    const int live = LW - 8 * g8, j0 = 8 * g8;
#define EAMD_COL(JJ)                                                                   \
    if (ok && j0 + JJ < w) ok = column(std::integral_constant<int, JJ>{}, j0 + JJ, live);
    EAMD_COL(0) EAMD_COL(1) EAMD_COL(2) EAMD_COL(3) EAMD_COL(4) EAMD_COL(5) EAMD_COL(6)
    EAMD_COL(7)
#undef EAMD_COL
    rotate8(v);
  }
#pragma unroll 1
  for (int r = ngroups; r < 4; ++r) rotate8(v);
  if (s_abort) {
    // gave up (another workgroup's words never came; the call reports a
    // time-out): empty permutation lists and no zero pivot, so the kernels
    // that apply the lists after this leaf do nothing instead of indexing
    // rows with stale entries.  Workgroup 0 always takes this path when any
    // workgroup does (it needs every workgroup's words for each column).
    if (g == 0 && tid == 0) {
      *lperm_n = 0;
      *pperm_n = 0;
      *zero_col = -1;
    }
    return;
  }
  __syncthreads();
  // ---- bookkeeping (workgroup 0, wave 0): the leaf-level and panel-level
  // permutation lists, replaying the interchanges in column order
  if (g == 0 && wave == 0) {
    int tpos, trow, tsrc, ln;
    leaf_lists<false>(lane < w ? s_piv[lane] : 0, w, row0, pnb, lane, piv, piv_offset, lperm_n,
                      lpos, lsrc, pperm_n, ppos, psrc, s_pp, tpos, trow, tsrc, ln);
  }
  if (valid) {
    T* dst = P + (int64_t)i * ldp;
#pragma unroll
    for (int c = 0; c < LW; ++c)
      if (c < w) dst[c] = v[c];
  }
  if (g == 0 && tid == 0) {
    *zero_col = first_zero;
    if (fz_global && first_zero >= 0 && *fz_global < 0) *fz_global = first_zero + piv_offset;
  }
}


// Launches one leaf over rows [0, m) of the sub-panel at P (row-major, ldp),
// columns [0, w), w <= 32.
template <typename T>
int leaf_lu(eamd_handle* h, T* P, int64_t ldp, int m, int w, int* piv, int* lperm_n, int* lpos,
            int* lsrc, int* pperm_n, int* ppos, int* psrc, int row0, int pnb, int* zero_col,
            int piv_offset, int* fz_global) {
  constexpr int BLOCK = LeafBlock<T>::v;
  if (w > LW || pnb > 256) {
    set_error("leaf width %d / panel width %d unsupported", w, pnb);
    return kErrUnsupported;
  }
  const int G = ceil_div(m, BLOCK);
  if (G > kMaxG) {
    set_error("panel of %d rows exceeds the leaf exchange (%d workgroups of %d rows)", m, kMaxG,
              BLOCK);
    return kErrUnsupported;
  }
  const int SW = slot_words<T>();
  size_t need = 256 + sizeof(u64) * 2 * (size_t)(G + 1) * SW;
  void* base = nullptr;
  bool fresh = h->ws.bytes[kWsPanelSync] < need;
  EAMD_CHECK(h->ws.get(kWsPanelSync, need, &base));
  if (fresh || h->panel_epoch > (1u << 31)) {
    EAMD_HIP(hipMemsetAsync(base, 0, h->ws.bytes[kWsPanelSync], h->stream));
    h->panel_epoch = 0;
  }
  unsigned epoch_base = h->panel_epoch;
  h->panel_epoch += (unsigned)w + 1u;
  char* b = static_cast<char*>(base);
  // polls an exchange may take before the leaf gives up (the call then
  // reports a time-out); EAMD_LEAF_SPIN sets another limit
  static const int spin_limit = env_int("EAMD_LEAF_SPIN", kLeafSpin);
  hipLaunchKernelGGL((leaf_lu_kernel<T, BLOCK>), dim3(G), dim3(BLOCK), 0, h->stream, P, ldp, m, w,
                     G, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc, row0, pnb, zero_col,
                     reinterpret_cast<u64*>(b + 256), reinterpret_cast<int*>(b), epoch_base,
                     piv_offset, fz_global, spin_limit);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// ---------------------------------------------------------------------------
// Tournament pivoting (CALU: Grigori, Demmel, Xiang, "CALU: a communication
// optimal LU factorization algorithm", SIAM J. Matrix Anal. Appl. 32, 2011)
// for the FP32 panel of the mixed-precision solver, whose result is judged
// by the backward error after refinement rather than by Eigen's pivot
// sequence.  A leaf of m rows costs a few short launches instead of one
// exchange round trip per column:
//   1. one wave per 256 rows (two waves per 512 rows when more than 32
//      groups would follow) runs partial pivoting on them, in registers, and
//      records the rows it picked (up to 32 candidates);
//   2. while more than 32 candidate groups remain, one wave per 8 groups runs
//      partial pivoting on the original values of their candidates and keeps
//      32 of them;
//   3. one workgroup does the same over the last groups (one wave for at
//      most 8, four waves for at most 32, each column's pivot the largest of
//      the waves' maxima): the rows it picks, in order, are the leaf's pivots
//      and its first 32 rows after elimination are the factored top block
//      L11\U11, written in place; its first wave also saves the original rows
//      0..w-1, derives the transpositions, replays them into the permutation
//      lists and lists where the displaced rows go;
//   4. every other row is solved against U11 (l = a U11^-1) in parallel.
namespace {

constexpr int kTslR = 4;                 // rows per lane
constexpr int kTslBlock = 64 * kTslR;    // rows (candidates) per wave
constexpr int kTslFan = kTslBlock / LW;  // candidate groups merged per wave
constexpr int kMoveCols = 96;            // panel columns moved per pass in the final round

// Maximum of a 32-bit key over the 64 lanes of a wave: xor-1 and xor-2
// within quads, half-row and row mirrors within each 16-lane row, then the
// row maxima are carried into lane 63 (row_bcast:15 into rows 1 and 3,
// row_bcast:31 into rows 2 and 3).
__device__ __forceinline__ unsigned wave_max_u32(unsigned k) {
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0xB1, 0xf, 0xf, false));
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0x4E, 0xf, 0xf, false));
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0x141, 0xf, 0xf, false));
  k = max(k, (unsigned)__builtin_amdgcn_mov_dpp((int)k, 0x140, 0xf, 0xf, false));
  k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x142, 0xa, 0xf, false));
  k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x143, 0xc, 0xf, false));
  return (unsigned)readlane_i((int)k, 63);
}

// Lane `me` stores its register row v[rs] (rs wave-uniform) to LDS as eight
// 16-byte words: a scalar branch over the slot, then the one lane's stores.
__device__ __forceinline__ void put_row(float* dst, const float (&v)[kTslR][LW], int rs,
                                        bool me) {
#define EAMD_PUT(R_)                                                                   \
  if (rs == R_) {                                                                      \
    if (me) {                                                                          \
      _Pragma("unroll") for (int q = 0; q < LW / 4; ++q)                               \
        reinterpret_cast<float4*>(dst)[q] =                                            \
            make_float4(v[R_][4 * q], v[R_][4 * q + 1], v[R_][4 * q + 2], v[R_][4 * q + 3]); \
    }                                                                                  \
    asm volatile("; slot " #R_ ::: "memory");                                          \
  }
  EAMD_PUT(0) else EAMD_PUT(1) else EAMD_PUT(2) else EAMD_PUT(3)
#undef EAMD_PUT
}

// One wave per block of up to 256 candidate rows, 4 rows per lane, all in
// registers.  MODE 0: local selection over rows blockIdx.x*256 + [0, 256),
// candidates to cand.  MODE 2: merge of candidate groups [8 b, 8 b + 8) of
// cand_in into group b of cand.  MODE 1: the final round over all G_in
// (<= 8) groups.  With NW > 1 waves a workgroup covers NW such blocks (MODE
// 0: NW*256 rows; MODE 1: up to 8*NW groups, wave q holding groups
// [8q, 8q + 8)): every column's pivot is the maximum over the waves' own
// maxima, exchanged through LDS under one barrier.
template <int MODE, int NW>
__global__ __launch_bounds__(64 * NW) void tsl_kernel(
    float* __restrict__ P, int64_t ldp, int m, int w, const int* __restrict__ cand_in,
    const int* __restrict__ cand_n_in, int G_in, int* __restrict__ cand,
    int* __restrict__ cand_n, float* __restrict__ saved, int* __restrict__ moved,
    int* __restrict__ piv, int* __restrict__ lperm_n, int* __restrict__ lpos,
    int* __restrict__ lsrc, int* __restrict__ pperm_n, int* __restrict__ ppos,
    int* __restrict__ psrc, int row0, int pnb, int* __restrict__ zero_col, int piv_offset,
    int* __restrict__ fz_global, int fuse, float* __restrict__ xtop) {
  // the panel is the factorization's critical path while the trailing
  // update's waves share its SIMDs: its waves take issue priority
  __builtin_amdgcn_s_setprio(3);
  constexpr int R = kTslR;
  static_assert(MODE != 2 || NW == 1, "merge rounds run one wave per workgroup");
  // slot ids in the key's low KB bits: 2^KB = 2 * NW * 256 keeps every
  // id's code (2^KB - 1 - slot) nonzero
  constexpr int KB = NW == 1 ? 9 : (NW == 2 ? 10 : 11);
  constexpr unsigned kSlotMask = (1u << KB) - 1u;
  static_assert((1 << KB) == 2 * NW * kTslBlock, "slot code width");
  __shared__ int s_sel[LW];
  __shared__ __align__(16) int s_pp[256 + 4];
  __shared__ __align__(16) float s_u[LW];  // the pivot row of the current column
  __shared__ __align__(16) float s_rows[NW > 1 ? 2 : 1][NW][LW];  // each wave's best row
  __shared__ unsigned s_key[NW > 1 ? 2 : 1][NW];                  // and its key
  const int lane = threadIdx.x & 63;
  const int wave = NW > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
  // MODE 1: the incoming panel list's count (the lists themselves are
  // prefetched below, while the columns are eliminated)
  int pn0_pre = -1;
  if (MODE == 1) pn0_pre = row0 == 0 ? pnb : *pperm_n;
  // MODE 1: the original rows 0..w-1 (saved before the top block overwrites
  // them), loaded first and stored last
  float sv[LW * LW / 64];
  if (MODE == 1 && wave == 0) {
#pragma unroll
    for (int k = 0; k < LW * LW / 64; ++k) {
      const int e = lane + 64 * k, r = e / LW, c = e % LW;
      sv[k] = P[(int64_t)r * ldp + min(c, w - 1)];
    }
  }
  int orig[R];
  bool valid[R];
  int nrows;  // candidates of the workgroup
  if (MODE == 0) {
    const int wg0 = blockIdx.x * (NW * kTslBlock), base = wg0 + wave * kTslBlock;
    nrows = min(NW * kTslBlock, m - wg0);
    const int nw = min(kTslBlock, m - base);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = r * 64 + lane;
      valid[r] = q < nw;
      orig[r] = valid[r] ? base + q : 0;
    }
  } else {
    // the workgroup's group counts, one per lane, and their inclusive
    // prefix sums; wave q's groups are [8q, 8q + 8) of the workgroup's
    constexpr int NG = NW * kTslFan;
    const int gb = MODE == 2 ? blockIdx.x * kTslFan : 0;
    const int nga = max(0, min(NG, G_in - gb));
    int inc = lane < nga ? cand_n_in[gb + lane] : 0;
#pragma unroll
    for (int d = 1; d < NG; d <<= 1) {
      const int t = __shfl_up(inc, d);
      if (lane >= d) inc += t;
    }
    nrows = readlane_i(inc, NG - 1);
    const int wg = wave * kTslFan;
    const int pre = wave > 0 ? readlane_i(inc, wg - 1) : 0;
    int off[kTslFan + 1];  // wave-uniform offsets of the wave's groups
    off[0] = 0;
#pragma unroll
    for (int g = 0; g < kTslFan; ++g) off[g + 1] = readlane_i(inc, wg + g) - pre;
    const int nw = off[kTslFan];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = r * 64 + lane;
      valid[r] = q < nw;
      // group of slot q: the last g whose offset is <= q (empty groups
      // share their successor's offset)
      int gs = 0, o = 0;
#pragma unroll
      for (int g = 1; g < kTslFan; ++g) {
        const bool ge = q >= off[g];
        gs = ge ? g : gs;
        o = ge ? off[g] : o;
      }
      orig[r] = valid[r] ? cand_in[(gb + wg + gs) * LW + q - o] : 0;
    }
  }
  // An empty slot loads row 0 and a leaf narrower than 32 repeats its last
  // column: such values are never keyed, selected or stored.
  float v[R][LW];
  if (w == LW && ((uintptr_t)P & 15) == 0 && (ldp & 3) == 0) {
    // 16-byte loads: a quarter of the address work of word loads, which
    // otherwise saturates the CU's address unit when several waves load
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4f* src = reinterpret_cast<const v4f*>(P + (int64_t)orig[r] * ldp);
#pragma unroll
      for (int q = 0; q < LW / 4; ++q) {
        const v4f t = src[q];
        v[r][4 * q] = t.x;
        v[r][4 * q + 1] = t.y;
        v[r][4 * q + 2] = t.z;
        v[r][4 * q + 3] = t.w;
      }
    }
  } else if (w == LW) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float* src = P + (int64_t)orig[r] * ldp;
#pragma unroll
      for (int c = 0; c < LW; ++c) v[r][c] = src[c];
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float* src = P + (int64_t)orig[r] * ldp;
#pragma unroll
      for (int c = 0; c < LW; ++c) v[r][c] = src[min(c, w - 1)];
    }
  }
  int ppv[4], psv[8];
  if (MODE == 1) {
    const bool has = row0 != 0 && wave == 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 64 * q + lane;
      ppv[q] = (has && e < pn0_pre - pnb) ? ppos[pnb + e] : -1;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = 64 * q + lane;
      psv[q] = (has && e < pn0_pre) ? psrc[e] : 0;
    }
  }
  const int steps = min(w, nrows);
  int first_zero = -1;
  // Rows never move between lanes.  Slot q = 256 wave + 64 r + lane holds a
  // candidate row while em[r] is all ones (valid and not yet selected);
  // st[r] is the step that selected it.  The pivot key of slot q is |a| with
  // its KB low mantissa bits replaced by 2^KB - 1 - q (km = em & 0x7fffffff
  // with those bits clear, kq = em & (2^KB - 1 - q); zero for a slot that is
  // out), so the maximum names its slot, exact ties go to the lowest slot
  // (the first row in order for the local round) and scores within
  // 2^(KB-23) relative count as ties, which keeps every multiplier within
  // 1 + 2^(KB-23) (2^-14 for one wave, 2^-12 for four).  Column j of group g8 is register slot jj
  // (see rotate8); slots [live, 32) hold columns of earlier groups and are
  // skipped by uniform branches on 8-slot chunks, so the update does exactly
  // the triangle.  The multipliers use the reciprocal of the pivot.
  unsigned em[R], km[R], kq[R];
  int st[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    em[r] = valid[r] ? ~0u : 0u;
    km[r] = em[r] & (0x7fffffffu & ~kSlotMask);
    kq[r] = em[r] & (kSlotMask - (unsigned)(wave * kTslBlock + r * 64 + lane));
    st[r] = -1;
  }
  auto column = [&](auto jjc, int j, int live) __attribute__((always_inline)) {
    constexpr int jj = decltype(jjc)::value;
    unsigned kr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) kr[r] = (__float_as_uint(v[r][jj]) & km[r]) | kq[r];
    const unsigned key = max(max(kr[0], kr[1]), max(kr[2], kr[3]));
    // a valid unselected slot exists (j < steps), so the winner does
    const unsigned wkey = wave_max_u32(key);
    const int ql = (int)(kSlotMask - (wkey & kSlotMask)) - wave * kTslBlock;
    int rs = ql >> 6, wl = ql & 63;
    bool me = lane == wl;
    const float* ub = s_u;
    if constexpr (NW == 1) {
      // the pivot row: stored by its lane, read back by all (a wave's LDS
      // accesses execute in order; the barriers keep the compiler from
      // moving the reads above the store)
      put_row(s_u, v, rs, me);
      asm volatile("" ::: "memory");
    } else {
      // every wave stores its best row and key (a wave without candidates
      // has key 0 and decodes a slot outside its 256, so stores nothing),
      // then all take the largest key: its wave's row is the pivot row.
      // Column parity double-buffers the exchange: the stores of column
      // j + 2 follow the barrier of column j + 1, after every read of j.
      const int par = j & 1;
      put_row(s_rows[par][wave], v, rs, me);
      if (lane == 0) s_key[par][wave] = wkey;
      __syncthreads();
      unsigned gkey = s_key[par][0];
#pragma unroll
      for (int q = 1; q < NW; ++q) gkey = max(gkey, s_key[par][q]);
      const int gw = (int)(kSlotMask - (gkey & kSlotMask)) >> 8;
      me = me && gw == wave;
      ub = s_rows[par][gw];
    }
    float u[LW];
#pragma unroll
    for (int q = 0; q < LW / 4; ++q) {
      const float4 t = reinterpret_cast<const float4*>(ub)[q];
      u[4 * q] = t.x;
      u[4 * q + 1] = t.y;
      u[4 * q + 2] = t.z;
      u[4 * q + 3] = t.w;
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (rs == r && me) {
        em[r] = 0u;
        km[r] = 0u;
        kq[r] = 0u;
        st[r] = j;
      }
    const float pv = u[jj];
    if (pv == 0.f && first_zero < 0) first_zero = j;
    const float rinv = pv != 0.f ? __builtin_amdgcn_rcpf(pv) : 0.f;
    float l[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const unsigned lb = __float_as_uint(v[r][jj] * rinv);
      l[r] = __uint_as_float(lb & em[r]);
      v[r][jj] = __uint_as_float((lb & em[r]) | (__float_as_uint(v[r][jj]) & ~em[r]));
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = jj + 1; c < 8; ++c) v[r][c] = fmaf(-l[r], u[c], v[r][c]);
    if (live > 8) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 8; c < 16; ++c) v[r][c] = fmaf(-l[r], u[c], v[r][c]);
    }
    if (live > 16) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 16; c < 24; ++c) v[r][c] = fmaf(-l[r], u[c], v[r][c]);
    }
    if (live > 24) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 24; c < 32; ++c) v[r][c] = fmaf(-l[r], u[c], v[r][c]);
    }
  };
  const int ngroups = (steps + 7) / 8;
#pragma unroll 1
  for (int g8 = 0; g8 < ngroups; ++g8) {
    const int live = LW - 8 * g8, j0 = 8 * g8;
#define EAMD_COL(JJ) \
    if (j0 + JJ < steps) column(std::integral_constant<int, JJ>{}, j0 + JJ, live);
    EAMD_COL(0) EAMD_COL(1) EAMD_COL(2) EAMD_COL(3) EAMD_COL(4) EAMD_COL(5) EAMD_COL(6)
    EAMD_COL(7)
#undef EAMD_COL
#pragma unroll
    for (int r = 0; r < R; ++r) rotate8(v[r]);
  }
  if (MODE != 1) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (st[r] >= 0) cand[blockIdx.x * LW + st[r]] = orig[r];
    if (threadIdx.x == 0) cand_n[blockIdx.x] = steps;
    return;
  }
  if constexpr (MODE == 1) {
#pragma unroll 1
  for (int q = ngroups; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < R; ++r) rotate8(v[r]);
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k < LW * LW / 64; ++k) {
      const int e = lane + 64 * k;
      saved[e] = e % LW < w ? sv[k] : 0.f;
    }
  }
  // ---- final round: the top block in place (rows 0..w-1) and the selection
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (st[r] >= 0) {
      float* dst = P + (int64_t)st[r] * ldp;
      if (w == LW && ((uintptr_t)P & 15) == 0 && (ldp & 3) == 0) {
#pragma unroll
        for (int q = 0; q < LW / 4; ++q)
          reinterpret_cast<v4f*>(dst)[q] =
              v4f{v[r][4 * q], v[r][4 * q + 1], v[r][4 * q + 2], v[r][4 * q + 3]};
      } else {
#pragma unroll
        for (int c = 0; c < LW; ++c)
          if (c < w) dst[c] = v[r][c];
      }
      s_sel[st[r]] = orig[r];
    }
  __syncthreads();
  // the rest is one wave's (no barrier follows)
  if (NW > 1 && wave != 0) return;
  // ---- transpositions and permutation lists from the selection order
  __shared__ int s_ps[512];
#pragma unroll
  for (int q = 0; q < 4; ++q) s_pp[64 * q + lane] = ppv[q];
#pragma unroll
  for (int q = 0; q < 8; ++q) s_ps[64 * q + lane] = psv[q];
  asm volatile("" ::: "memory");  // a wave's LDS accesses execute in order
  int tpos, trow, tsrc, ln;
  leaf_lists<true>(lane < w ? s_sel[lane] : 0, w, row0, pnb, lane, piv, piv_offset, lperm_n,
                   lpos, lsrc, pperm_n, ppos, psrc, s_pp, tpos, trow, tsrc, ln, s_ps, pn0_pre);
  // positions >= w now hold displaced original rows < w: list them
  const bool mv = tpos >= w && lane < ln;
  const u64 mb = __ballot(mv);
  const int idx = __popcll(mb & ((1ull << lane) - 1ull));
  if (mv) {
    moved[1 + 2 * idx] = tpos;
    moved[2 + 2 * idx] = trow;
  }
  if (lane == 0) moved[0] = __popcll(mb);
  if (fuse) {
    // (w == 32) The leaf's interchanges on the other columns of the panel
    // (row0 = j0 is the leaf's first panel column, pnb the panel width): the
    // lane of each touched row moves its source row's values, 96 columns at
    // a time (all loads of a chunk before its stores); the top rows' values
    // right of the leaf go to xtop (LW x (pnb - j0 - LW)), from which the
    // apply kernel solves U12.
    const int srow = __shfl(tpos, tsrc);
    const int j0 = row0, nc = pnb - LW, rest = pnb - j0 - LW;
    float* base = P - j0;  // panel column 0 of leaf row 0
    // 16-byte words: the host fuses only panels whose rows start 16-byte
    // aligned with pnb and ldp multiples of 4 (a word never straddles the
    // leaf, whose columns start at a multiple of 32)
    if (lane < ln) {
      const v4f* sp = reinterpret_cast<const v4f*>(base + (int64_t)srow * ldp);
      v4f* dp = reinterpret_cast<v4f*>(base + (int64_t)tpos * ldp);
      v4f* xp = tpos < LW ? reinterpret_cast<v4f*>(xtop + tpos * rest) : nullptr;
      const int nq = nc / 4;
      for (int q0 = 0; q0 < nq; q0 += kMoveCols / 4) {
        v4f x[kMoveCols / 4];
#pragma unroll
        for (int t = 0; t < kMoveCols / 4; ++t) {
          const int c = 4 * min(q0 + t, nq - 1);
          x[t] = sp[(c < j0 ? c : c + LW) / 4];
        }
#pragma unroll
        for (int t = 0; t < kMoveCols / 4; ++t) {
          const int c = 4 * (q0 + t);
          v4f* d = xp && c >= j0 ? xp + (c - j0) / 4 : dp + (c < j0 ? c : c + LW) / 4;
          if (q0 + t < nq) *d = x[t];
        }
      }
    }
  }
  if (lane == 0) {
    *zero_col = first_zero;
    if (fz_global && first_zero >= 0 && *fz_global < 0) *fz_global = first_zero + piv_offset;
  }
  }
}

// Rows w..m-1 of the leaf: l = a U11^-1, where a is the row's original
// value (its own row, or a saved row 0..w-1 for a position listed in moved).
// FULL: w == 32, without the column guards (which would multiply the code).
template <bool FULL>
__global__ __launch_bounds__(256) void tsl_apply_kernel(float* __restrict__ P, int64_t ldp, int m,
                                                        int w, const float* __restrict__ saved,
                                                        const int* __restrict__ moved) {
  __builtin_amdgcn_s_setprio(3);  // on the critical path (see tsl_kernel)
  if (FULL) w = LW;
  __shared__ float U[LW][LW + 1];
  __shared__ int mpos[LW], msrc[LW];
  __shared__ int s_nm;
  const int tid = threadIdx.x;
  for (int e = tid; e < LW * LW; e += 256) {
    const int r = e / LW, c = e - r * LW;
    U[r][c] = (r < w && c < w) ? P[(int64_t)r * ldp + c] : 0.f;
  }
  if (tid == 0) s_nm = moved[0];
  if (tid < LW) {
    mpos[tid] = moved[1 + 2 * tid];
    msrc[tid] = moved[2 + 2 * tid];
  }
  __syncthreads();
  const int i = w + blockIdx.x * 256 + tid;
  if (i >= m) return;
  int src = -1;
  for (int k = 0; k < s_nm; ++k)
    if (mpos[k] == i) src = msrc[k];
  float a[LW];
  // 16-byte words when the rows allow (saved rows always do)
  const bool vec = FULL && ((uintptr_t)P & 15) == 0 && (ldp & 3) == 0;
  if (FULL && (src >= 0 || vec)) {
    const v4f* ar = reinterpret_cast<const v4f*>(src >= 0 ? saved + src * LW
                                                          : P + (int64_t)i * ldp);
#pragma unroll
    for (int q = 0; q < LW / 4; ++q) {
      const v4f t = ar[q];
      a[4 * q] = t.x;
      a[4 * q + 1] = t.y;
      a[4 * q + 2] = t.z;
      a[4 * q + 3] = t.w;
    }
  } else {
    // one load path (a saved row or the row itself), word by word, every
    // load unconditional (columns past w re-read column w - 1 and are
    // masked): conditional loads of a narrow leaf's row keep `a` in scratch
    const float* row = src >= 0 ? saved + src * LW : P + (int64_t)i * ldp;
#pragma unroll
    for (int c = 0; c < LW; ++c) {
      const float v = row[FULL ? c : min(c, w - 1)];
      a[c] = (FULL || c < w) ? v : 0.f;
    }
  }
#pragma unroll
  for (int k = 0; k < LW; ++k) {
    const float d = U[k][k];
    const float x = d != 0.f ? a[k] / d : a[k];  // a zero pivot leaves its column unscaled
    a[k] = x;
#pragma unroll
    for (int c = k + 1; c < LW; ++c) a[c] -= x * U[k][c];
  }
  float* dst = P + (int64_t)i * ldp;
  if (vec) {
#pragma unroll
    for (int q = 0; q < LW / 4; ++q)
      reinterpret_cast<v4f*>(dst)[q] = v4f{a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]};
  } else {
    // unconditional stores, last column first: columns past w land on
    // column w - 1, which its own value overwrites last (conditional stores
    // of a narrow leaf keep `a` in scratch memory)
#pragma unroll
    for (int c = LW - 1; c >= 0; --c) dst[FULL ? c : min(c, w - 1)] = a[c];
  }
}


// As tsl_apply_kernel (w == 32), and the row's `rest` panel columns right of
// the leaf take the update x -= l U12 (the rows' values already permuted).
// Every workgroup solves U12 = L11^-1 xtop (the permuted top rows right of
// the leaf) in LDS, one column per thread; workgroup 0 also stores it as the
// top rows.  Four threads per row: each computes l and updates every fourth
// column.
constexpr int kApplyTpr = 4;
constexpr int kFuseRestMax = 96;  // panel columns right of a fused leaf (panels <= 128)
__global__ __launch_bounds__(256) void tsl_apply_rest_kernel(float* __restrict__ P, int64_t ldp,
                                                             int m, const float* __restrict__ saved,
                                                             const int* __restrict__ moved,
                                                             const float* __restrict__ xtop,
                                                             int rest) {
  __builtin_amdgcn_s_setprio(3);  // on the critical path (see tsl_kernel)
  extern __shared__ float s_u12[];  // LW x rest
  __shared__ float U[LW][LW + 1];
  __shared__ float dinv[LW];
  __shared__ int mpos[LW], msrc[LW];
  __shared__ int s_nm;
  const int tid = threadIdx.x;
  for (int e = tid; e < LW * LW; e += 256) {
    const int r = e / LW, c = e - r * LW;
    U[r][c] = P[(int64_t)r * ldp + c];
  }
  if (tid == 0) s_nm = moved[0];
  if (tid < LW) {
    mpos[tid] = moved[1 + 2 * tid];
    msrc[tid] = moved[2 + 2 * tid];
  }
  __syncthreads();
  if (tid < LW) {
    const float d = U[tid][tid];
    dinv[tid] = d != 0.f ? 1.0f / d : 1.0f;  // a zero pivot leaves its column unscaled
  }
  for (int c = tid; c < rest; c += 256) {
    float u[LW];
#pragma unroll
    for (int i = 0; i < LW; ++i) {
      float sum = xtop[i * rest + c];
#pragma unroll
      for (int k = 0; k < i; ++k) sum -= U[i][k] * u[k];
      u[i] = sum;
      s_u12[i * rest + c] = sum;
    }
    if (blockIdx.x == 0)
#pragma unroll
      for (int i = 0; i < LW; ++i) P[(int64_t)i * ldp + LW + c] = u[i];
  }
  __syncthreads();
  const int q = tid % kApplyTpr;
  const int i = LW + blockIdx.x * (256 / kApplyTpr) + tid / kApplyTpr;
  if (i >= m) return;
  int src = -1;
  for (int k = 0; k < s_nm; ++k)
    if (mpos[k] == i) src = msrc[k];
  // 16-byte words throughout: the host fuses only panels whose rows start
  // 16-byte aligned, with ldp and rest multiples of 4
  float a[LW];
  const v4f* arow = reinterpret_cast<const v4f*>(src >= 0 ? saved + src * LW
                                                          : P + (int64_t)i * ldp);
#pragma unroll
  for (int c = 0; c < LW / 4; ++c) {
    const v4f t = arow[c];
    a[4 * c] = t.x;
    a[4 * c + 1] = t.y;
    a[4 * c + 2] = t.z;
    a[4 * c + 3] = t.w;
  }
#pragma unroll
  for (int k = 0; k < LW; ++k) {
    const float x = a[k] * dinv[k];
    a[k] = x;
#pragma unroll
    for (int c = k + 1; c < LW; ++c) a[c] -= x * U[k][c];
  }
  float* row = P + (int64_t)i * ldp;
  if (q == 0) {
#pragma unroll
    for (int c = 0; c < LW / 4; ++c)
      reinterpret_cast<v4f*>(row)[c] = v4f{a[4 * c], a[4 * c + 1], a[4 * c + 2], a[4 * c + 3]};
  }
  // the row's rest columns as 16-byte words q, q + 4, ... (rest is a
  // multiple of 4; LDS reads of U12 stay 4-byte, broadcast across the 16
  // rows of a wave)
  v4f* xr = reinterpret_cast<v4f*>(row + LW);
  const int nq = rest / 4;
  for (int w0 = q; w0 < nq; w0 += kApplyTpr * 2) {
    v4f x[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int w = w0 + kApplyTpr * t;
      x[t] = w < nq ? xr[w] : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll 4
    for (int k = 0; k < LW; ++k) {
      const float* uk = s_u12 + k * rest;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = 4 * min(w0 + kApplyTpr * t, nq - 1);
        x[t][0] -= a[k] * uk[c];
        x[t][1] -= a[k] * uk[c + 1];
        x[t][2] -= a[k] * uk[c + 2];
        x[t][3] -= a[k] * uk[c + 3];
      }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int w = w0 + kApplyTpr * t;
      if (w < nq) xr[w] = x[t];
    }
  }
}

}  // namespace

// Same outputs as leaf_lu; FP32 only.  Panels of up to 256 rows take the
// partial-pivoting leaf.  Candidate groups are merged 8 at a time until at
// most 32 remain for the final round (at most 8 for a one-wave final round
// when the wide rounds are switched off).
int leaf_lu_tournament(eamd_handle* h, float* P, int64_t ldp, int m, int w, int* piv,
                       int* lperm_n, int* lpos, int* lsrc, int* pperm_n, int* ppos, int* psrc,
                       int row0, int pnb, int* zero_col, int piv_offset, int* fz_global,
                       bool* fused) {
  const int G1 = ceil_div(m, kTslBlock);
  // the leaf's interchanges on the other panel columns, the U rows right of
  // it and their update are done here too when w == 32 (the caller skips
  // its own)
  const bool fuse = fused && w == LW && G1 > 1 && pnb - row0 - w <= kFuseRestMax &&
                    ((uintptr_t)(P - row0) & 15) == 0 && ldp % 4 == 0 && pnb % 4 == 0;
  if (fused) *fused = fuse;
  if (G1 <= 1)
    return leaf_lu<float>(h, P, ldp, m, w, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc, row0,
                          pnb, zero_col, piv_offset, fz_global);
  if (w > LW || pnb > 256) {
    set_error("tournament leaf: %d rows x %d columns unsupported", m, w);
    return kErrUnsupported;
  }
  void* base = nullptr;
  // two candidate buffers (ping-pong between merge levels), counts, moved
  // list and the saved rows
  const size_t need = sizeof(int) * (2 * (size_t)G1 * LW + 2 * (size_t)G1 + 2 * LW + 8) +
                      sizeof(float) * LW * LW + sizeof(float) * LW * 256;
  EAMD_CHECK(h->ws.get(kWsTsl, need, &base));
  int* candA = static_cast<int*>(base);
  int* candB = candA + (size_t)G1 * LW;
  int* nA = candB + (size_t)G1 * LW;
  int* nB = nA + G1;
  int* moved = nB + G1;
  float* saved = reinterpret_cast<float*>(moved + 2 * LW + 8);
  float* xtop = saved + LW * LW;  // permuted top rows right of the leaf (fused)
  hipStream_t s = h->stream;
  // Rounds: one wave per 256 rows, or per 512 rows in pairs when more than
  // 32 groups would follow; merges of 8 groups while more than 32 remain;
  // the final round in one workgroup (one wave for up to 8 groups, four
  // waves for up to 32).
  static const bool wide = env_int("EAMD_TSL_WIDE", 1) != 0;
  const int fan1 = wide ? 4 * kTslFan : kTslFan;
  int G;
  if (wide && G1 > fan1) {
    G = ceil_div(m, 2 * kTslBlock);
    hipLaunchKernelGGL((tsl_kernel<0, 2>), dim3(G), dim3(128), 0, s, P, ldp, m, w, nullptr,
                       nullptr, 0, candA, nA, saved, moved, piv, lperm_n, lpos, lsrc, pperm_n,
                       ppos, psrc, row0, pnb, zero_col, piv_offset, fz_global, 0, xtop);
  } else {
    G = G1;
    hipLaunchKernelGGL((tsl_kernel<0, 1>), dim3(G), dim3(64), 0, s, P, ldp, m, w, nullptr,
                       nullptr, 0, candA, nA, saved, moved, piv, lperm_n, lpos, lsrc, pperm_n,
                       ppos, psrc, row0, pnb, zero_col, piv_offset, fz_global, 0, xtop);
  }
  int *cin = candA, *nin = nA, *cout = candB, *nout = nB;
  while (G > fan1) {
    const int Gn = ceil_div(G, kTslFan);
    hipLaunchKernelGGL((tsl_kernel<2, 1>), dim3(Gn), dim3(64), 0, s, P, ldp, m, w, cin, nin, G,
                       cout, nout, saved, moved, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc,
                       row0, pnb, zero_col, piv_offset, fz_global, 0, xtop);
    G = Gn;
    std::swap(cin, cout);
    std::swap(nin, nout);
  }
  if (G > kTslFan)
    hipLaunchKernelGGL((tsl_kernel<1, 4>), dim3(1), dim3(256), 0, s, P, ldp, m, w, cin, nin, G,
                       cout, nout, saved, moved, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc,
                       row0, pnb, zero_col, piv_offset, fz_global, fuse ? 1 : 0, xtop);
  else
    hipLaunchKernelGGL((tsl_kernel<1, 1>), dim3(1), dim3(64), 0, s, P, ldp, m, w, cin, nin, G,
                       cout, nout, saved, moved, piv, lperm_n, lpos, lsrc, pperm_n, ppos, psrc,
                       row0, pnb, zero_col, piv_offset, fz_global, fuse ? 1 : 0, xtop);
  const int rest = pnb - row0 - w;
  if (fuse && rest > 0) {
    static bool attr_set = false;
    if (!attr_set) {
      EAMD_HIP(hipFuncSetAttribute((const void*)tsl_apply_rest_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
      attr_set = true;
    }
    hipLaunchKernelGGL(tsl_apply_rest_kernel, dim3(ceil_div(m - w, 256 / kApplyTpr)), dim3(256),
                       sizeof(float) * LW * rest, s, P, ldp, m, saved, moved, xtop, rest);
  } else if (m > w) {
    if (w == LW)
      hipLaunchKernelGGL(tsl_apply_kernel<true>, dim3(ceil_div(m - w, 256)), dim3(256), 0, s, P,
                         ldp, m, w, saved, moved);
    else
      hipLaunchKernelGGL(tsl_apply_kernel<false>, dim3(ceil_div(m - w, 256)), dim3(256), 0, s, P,
                         ldp, m, w, saved, moved);
  }
  EAMD_HIP(hipGetLastError());
  return kOk;
}

template int leaf_lu<double>(eamd_handle*, double*, int64_t, int, int, int*, int*, int*, int*,
                             int*, int*, int*, int, int, int*, int, int*);
template int leaf_lu<float>(eamd_handle*, float*, int64_t, int, int, int*, int*, int*, int*,
                            int*, int*, int*, int, int, int*, int, int*);

}  // namespace eamd
