// batched_lu.hip -- many small general systems A_i X_i = B_i, each solved as
// PartialPivLU<MatrixXd>(A_i).solve(B_i) (Eigen/src/LU/PartialPivLU.h:128,
// 226-244, 426-496), computed as an FP32 LU on matrix cores with FP64
// iterative refinement; systems the FP32 factor cannot refine to the
// tolerance are refactored in FP64 by a second launch.
//
// MI355X mapping.  One workgroup per system (n <= 256), systems taken in
// descending size from the plan's order so the largest start first.  The
// workgroup converts its FP64 column-major A_i into a row-major working copy
// W_i (row interchanges are then contiguous), and runs a right-looking
// blocked LU with 16-column panels.  A panel is factored by wave 0 in its
// registers (pivot search by DPP wave reductions, interchanges as swapped
// row positions, the pivot row broadcast by readlane); the interchanges are
// applied to the other columns, U12 is solved against the unit lower L11,
// and the trailing update A22 -= L21 U12 runs as 16x16x4 MFMA tiles (FP32,
// transposed so a lane holds four adjacent columns of a row; FP64 for the
// refactor pass).  With more than one wave the next panel's columns are
// updated first and wave 0 factors that panel while the other waves update
// the rest.  The solves (right-looking in 64-row blocks, 16 lanes per row)
// and the refinement steps (residual B - A X in FP64 from the original A)
// stay in the same workgroup.  The workgroup size (64, 128 or 256 threads)
// is chosen per 16-wide size bucket by timing on synthetic systems, once per
// process.
#include <algorithm>
#include <type_traits>
#include <mutex>

#include "common.h"
#include "kernels.h"

namespace eamd {

namespace {

constexpr int kMaxLuN = 256;
constexpr int PB = 16;       // panel width = the MFMA k-depth used per trailing update
constexpr int kMaxRhs = 8;
constexpr int DB = 64;       // diagonal block of the triangular solves (one wave)
constexpr int kMaxIter = 10;
constexpr double kTol = 1e-14;

__device__ __forceinline__ v4f mfma16x4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma16x4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <typename T>
struct Acc4;
template <>
struct Acc4<float> {
  typedef v4f type;
};
template <>
struct Acc4<double> {
  typedef v4d type;
};
// row of accumulator register r in lane l (column is l % 16)
template <typename T>
__device__ __forceinline__ int acc_row(int l, int r) {
  return sizeof(T) == 4 ? 4 * (l >> 4) + r : (l >> 4) + 4 * r;
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double readlane_d(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// wave reductions (inputs must not be NaN): quad xor-1, xor-2, half-row and row
// mirrors, then the four 16-lane rows read as uniform values
__device__ __forceinline__ double wave_max(double s) {
  s = fmax(s, dpp_d<0xB1>(s));
  s = fmax(s, dpp_d<0x4E>(s));
  s = fmax(s, dpp_d<0x141>(s));
  s = fmax(s, dpp_d<0x140>(s));
  return fmax(fmax(readlane_d(s, 0), readlane_d(s, 16)), fmax(readlane_d(s, 32), readlane_d(s, 48)));
}
__device__ __forceinline__ double wave_sum(double s) {
  s += dpp_d<0xB1>(s);
  s += dpp_d<0x4E>(s);
  s += dpp_d<0x141>(s);
  s += dpp_d<0x140>(s);
  return (readlane_d(s, 0) + readlane_d(s, 16)) + (readlane_d(s, 32) + readlane_d(s, 48));
}
__device__ __forceinline__ int wave_min_i(int x) {
  x = min(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_mov_dpp(x, 0x141, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_mov_dpp(x, 0x140, 0xf, 0xf, false));
  return min(min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
             min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}

// Eigen's pivot score: |a|; the strict '>' scan makes a NaN win only from
// the first position (UnaryFunctors.h:63-66, Visitor.h:179)
template <typename T>
__device__ __forceinline__ T score_t(T v, bool first) {
  T s = fabs(v);
  if (s != s) s = first ? T(INFINITY) : T(-1);
  return s;
}

template <typename TF, int NMAX>
struct LuSmem {
  // factorization: the panel and U12
  static constexpr int kPanel = NMAX * (PB + 1);
  static constexpr int kU = PB * (NMAX + 1);
  // solves: the diagonal triangle, the right-hand sides and the residual
  static constexpr int kDiag = DB * (DB + 1);
  static constexpr int kVec = NMAX * kMaxRhs;
  static constexpr int kWords = (kPanel + kU) > (kDiag + 2 * kVec) ? (kPanel + kU)
                                                                   : (kDiag + 2 * kVec);
};

// |x| with a NaN mapped to +infinity, so maxima keep it
__device__ __forceinline__ double absinf(double x) { return x == x ? fabs(x) : INFINITY; }

// maximum over the workgroup; a NaN counts as +infinity (fmax alone would
// drop it and let a NaN residual pass the convergence test)
template <int NW>
__device__ __forceinline__ double block_max(double v, double* red, int lane, int wave) {
  const double w = wave_max(v == v ? v : INFINITY);
  if (lane == 0) red[wave] = w;
  __syncthreads();
  double m = red[0];
#pragma unroll
  for (int q = 1; q < NW; ++q) m = fmax(m, red[q]);
  __syncthreads();
  return m;
}

// x made opaque to the optimiser: index arithmetic derived from it is
// recomputed where it is used instead of hoisted out of the loops around and
// kept live (register pressure; without it the 256-thread instances spill)
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
// the thread index rebuilt from the wave index (a scalar, which is also the
// `wave` used for branches: uniform, so a wave's branch holds no registers
// through the other waves' branch) and the lane
// (mbcnt over a full mask, so exec does not enter; in a volatile asm block so
// it is recomputed where it is used): nothing vector-valued stays live for it
// between uses
__device__ __forceinline__ int thread_index(int swave) {
  int lane;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
  return swave * 64 + lane;
}
// four consecutive values from a 4-byte (float) or 8-byte (double) aligned
// address: 16-byte global loads (global memory takes unaligned vector loads)
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ void load4(const float* p, float* v) {
  const f4u x = *reinterpret_cast<const f4u*>(p);
  v[0] = x[0];
  v[1] = x[1];
  v[2] = x[2];
  v[3] = x[3];
}
__device__ __forceinline__ void load4(const double* p, double* v) {
  const d2u x = *reinterpret_cast<const d2u*>(p);
  const d2u z = *reinterpret_cast<const d2u*>(p + 2);
  v[0] = x[0];
  v[1] = x[1];
  v[2] = z[0];
  v[3] = z[1];
}
__device__ __forceinline__ void store4(float* p, const float* v) {
  *reinterpret_cast<f4u*>(p) = f4u{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void store4(double* p, const double* v) {
  *reinterpret_cast<d2u*>(p) = d2u{v[0], v[1]};
  *reinterpret_cast<d2u*>(p + 2) = d2u{v[2], v[3]};
}
// sum over the 16 lanes of a DPP row (quad xor-1, xor-2, half-row and row
// mirrors); every lane of the row receives it
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float sum16(float v) {
  v += dpp_f<0xB1>(v);
  v += dpp_f<0x4E>(v);
  v += dpp_f<0x141>(v);
  return v + dpp_f<0x140>(v);
}
__device__ __forceinline__ double sum16(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  return v + dpp_d<0x140>(v);
}
__device__ __forceinline__ float bcast(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double bcast(double x, int l) { return readlane_d(x, l); }
__device__ __forceinline__ float wave_max(float s) {
  s = fmaxf(s, dpp_f<0xB1>(s));
  s = fmaxf(s, dpp_f<0x4E>(s));
  s = fmaxf(s, dpp_f<0x141>(s));
  s = fmaxf(s, dpp_f<0x140>(s));
  return fmaxf(fmaxf(bcast(s, 0), bcast(s, 16)), fmaxf(bcast(s, 32), bcast(s, 48)));
}

// RC rows of the block's column segment per 16-lane group: row rb + g + k NG
// (k < RC), lane columns b0 + cb .. b0 + cb + 3, every load requested before
// the first use.  Branch-free: rows past the matrix read its last row (their
// sums are not stored), columns past the block read its last column and are
// zeroed.
template <typename TF, int RC, int NG>
__device__ __forceinline__ void load_rows(const TF* __restrict__ W, int n, int b0, int bn, int cb,
                                          int g, int rb, TF (&w)[RC][4]) {
#pragma unroll
  for (int k = 0; k < RC; ++k) {
    const int r = min(rb + g + k * NG, n - 1);
    if (bn == DB) {
      load4(W + r * n + b0 + cb, w[k]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const TF x = W[r * n + b0 + min(cb + e, bn - 1)];
        w[k][e] = cb + e < bn ? x : TF(0);
      }
    }
  }
}

// Forward (unit lower, FORWARD=true) or backward (upper) substitution of the
// NR (>= nrhs) columns y[r * kMaxRhs + q] against the factors in the
// row-major W, right-looking in blocks of DB rows.  The block's strict
// triangle is staged in LDS (zero-padded to DB, with U's reciprocal diagonal
// in column DB) and solved by wave 0, lanes over rows, the DB steps unrolled
// (broadcasts by readlane).  Every row beyond the block (after it going
// forward, before it going backward) then takes y[r] -= W[r][block] y[block]:
// 16 lanes per row, each lane four consecutive columns by one 16-byte load,
// the products summed over the 16 lanes; the first rows' loads are requested
// before the wait for the diagonal solve.
template <typename TF, int NT, bool FORWARD, int NR>
__device__ __forceinline__ void tri_solve(const TF* __restrict__ W, int n, TF* y, TF* diag,
                                          int tid0) {
  constexpr int LD = DB + 1, NG = NT / 16, RC = sizeof(TF) == 4 ? 8 : 4;
  const int nblk = (n + DB - 1) / DB;
  for (int bb = 0; bb < nblk; ++bb) {
    const int tid = opaque(tid0), lane = tid & 63;
    // the wave index as a scalar: branches on it are uniform
    const int wave = __builtin_amdgcn_readfirstlane(tid0 >> 6);
    const int g = tid >> 4, c16 = tid & 15, cb = 4 * c16;
    const int blk = FORWARD ? bb : nblk - 1 - bb;
    const int b0 = blk * DB, bn = min(DB, n - b0);
    {
      // chunks of at most 16 loads per thread, each chunk's loads issued
      // together (register pressure)
      constexpr int PER = DB * DB / NT, CH = PER < 16 ? PER : 16;
#pragma unroll
      for (int c0 = 0; c0 < PER; c0 += CH) {
        TF v[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int e = tid + (c0 + q) * NT, r = e / DB, t = e - r * DB;
          const bool keep = r < bn && t < bn && (FORWARD ? t < r : t >= r);
          v[q] = keep ? W[(b0 + r) * n + b0 + t] : TF(0);
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int e = tid + (c0 + q) * NT, r = e / DB, t = e - r * DB;
          if (!FORWARD && r == t) {
            diag[r * LD + DB] = r < bn ? TF(1) / v[q] : TF(1);
            diag[r * LD + t] = TF(0);
          } else {
            diag[r * LD + t] = v[q];
          }
        }
        asm volatile("" ::: "memory");
      }
    }
    const int r_lo = FORWARD ? b0 + bn : 0, r_hi = FORWARD ? n : b0;
    TF w[RC][4];
    load_rows<TF, RC, NG>(W, n, b0, bn, cb, g, r_lo, w);
    __syncthreads();
    if (wave == 0) {
      TF acc[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = lane < bn ? y[(b0 + lane) * kMaxRhs + q] : TF(0);
      if (FORWARD) {
        // l = 0 on and above the diagonal and in the padding
#pragma unroll
        for (int t = 0; t < DB - 1; ++t) {
          const TF l = diag[lane * LD + t];
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] -= l * bcast(acc[q], t);
          // bounds the LDS reads hoisted ahead (register pressure)
          if ((t & 7) == 7) asm volatile("" ::: "memory");
        }
      } else {
        // u = 0 on and below the diagonal; row t is final once scaled by
        // its reciprocal pivot (1 in the padding)
        const TF rd = diag[lane * LD + DB];
#pragma unroll
        for (int t = DB - 1; t >= 0; --t) {
          const TF u = diag[lane * LD + t];
          const TF rt = bcast(rd, t);
#pragma unroll
          for (int q = 0; q < NR; ++q) {
            const TF yt = bcast(acc[q], t) * rt;
            acc[q] = lane == t ? yt : acc[q] - u * yt;
          }
          if ((t & 7) == 0) asm volatile("" ::: "memory");
        }
      }
      if (lane < bn) {
#pragma unroll
        for (int q = 0; q < NR; ++q) y[(b0 + lane) * kMaxRhs + q] = acc[q];
      }
    }
    __syncthreads();
    if (r_lo < r_hi) {
      TF yv[4][NR];
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int q = 0; q < NR; ++q)
          yv[e][q] = cb + e < bn ? y[(b0 + cb + e) * kMaxRhs + q] : TF(0);
      for (int rb = r_lo;;) {
#pragma unroll
        for (int k = 0; k < RC; ++k) {
          const int r = rb + g + k * NG;
          TF s[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q)
            s[q] = sum16(w[k][0] * yv[0][q] + w[k][1] * yv[1][q] +
                         (w[k][2] * yv[2][q] + w[k][3] * yv[3][q]));
          if (r < r_hi) {
#pragma unroll
            for (int q = 0; q < NR; ++q)
              if (c16 == q) y[r * kMaxRhs + q] -= s[q];
          }
        }
        rb += NG * RC;
        if (rb >= r_hi) break;
        load_rows<TF, RC, NG>(W, n, b0, bn, cb, g, rb, w);
      }
    }
    __syncthreads();
  }
}

// u[j - C] = P[ks][j] of lane pl for j = C..PB-1, ks uniform: one readlane
// per column in the branch of slot ks
template <typename TF, int NS, int C, int K>
struct PivotRow {
  __device__ __forceinline__ static void get(TF (&P)[NS][PB], int ks, int pl, TF* u) {
    if (ks == K) {
#pragma unroll
      for (int j = C; j < PB; ++j) u[j - C] = bcast(P[K][j], pl);
    } else {
      PivotRow<TF, NS, C, K + 1>::get(P, ks, pl, u);
    }
  }
};
template <typename TF, int NS, int C>
struct PivotRow<TF, NS, C, NS> {
  __device__ __forceinline__ static void get(TF (&)[NS][PB], int, int, TF*) {}
};

// One column step C of the panel factorization in wave 0's registers.
// Lane l holds the physical panel rows l + 64 k in P[k] (NS = rows per
// lane); lg[k] is that row's current position in the interchanged order, so
// an interchange swaps two positions and moves no data.  The pivot is
// Eigen's first maximum of |a| over positions C..R-1 (the smallest position
// among equal maxima), the pivot row's columns C.. are broadcast by readlane,
// and the rows below take l = a * (1 / pivot) (one division per column; the
// reference divides each entry, which differs by at most an ulp in l) and the
// rank-1 update in registers (Eigen/src/LU/PartialPivLU.h:365-407).
template <typename TF, int NS, int C>
__device__ __forceinline__ void panel_col(TF (&P)[NS][PB], int (&lg)[NS], int lane, int R, int k0,
                                          int* piv) {
  TF sc = TF(-1);
  int ri = 0x7fffffff;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int g = lg[k];
    const TF v = score_t(P[k][C], g == C);
    if (g >= C && g < R && (v > sc || (v == sc && g < ri))) {
      sc = v;
      ri = g;
    }
  }
  const TF M = wave_max(sc);
  const int p = wave_min_i(sc == M ? ri : 0x7fffffff);
  // the pivot row: lane pl, slot ks (uniform), its columns C.. broadcast
  int kown = -1;
#pragma unroll
  for (int k = 0; k < NS; ++k)
    if (lg[k] == p) kown = k;
  const int pl = (int)__builtin_ctzll(__ballot(kown >= 0));
  const int ks = __builtin_amdgcn_readlane(kown, pl);
  TF u[PB - C];
  PivotRow<TF, NS, C, 0>::get(P, ks, pl, u);
  if (lane == 0) piv[k0 + C] = k0 + p;
#pragma unroll
  for (int k = 0; k < NS; ++k) lg[k] = lg[k] == C ? p : (lg[k] == p ? C : lg[k]);
  const TF pv = u[0];
  if (pv != TF(0)) {
    const TF rp = TF(1) / pv;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const bool below = lg[k] > C && lg[k] < R;
      const TF l = below ? P[k][C] * rp : TF(0);
      if (below) P[k][C] = l;
#pragma unroll
      for (int j = C + 1; j < PB; ++j) P[k][j] -= l * u[j - C];
    }
  }
}
// the column steps C.. while C < pb, unrolled by template recursion (every
// register index a constant)
template <typename TF, int NS, int C>
struct PanelCols {
  __device__ __forceinline__ static void run(TF (&P)[NS][PB], int (&lg)[NS], int lane, int R,
                                             int pb, int k0, int* piv) {
    if (C < pb) {
      panel_col<TF, NS, C>(P, lg, lane, R, k0, piv);
      PanelCols<TF, NS, C + 1>::run(P, lg, lane, R, pb, k0, piv);
    }
  }
};
template <typename TF, int NS>
struct PanelCols<TF, NS, PB> {
  __device__ __forceinline__ static void run(TF (&)[NS][PB], int (&)[NS], int, int, int, int,
                                             int*) {}
};

// R = B - A X in FP64 from the original A and the convergence test's norms,
// NR (>= nrhs) columns: one row per thread, 8 columns of A per step, the
// loads of the next RD steps in flight (the steps past n read column n - 1
// again); the first pass also sums |A| per row for ||A||_inf.  The maxima of
// all columns are reduced together (one LDS exchange); the normwise backward
// error of column q is ||r_q|| / (||A|| ||x_q|| + ||b_q||) in the infinity
// norm.
template <typename TF, int NT, int NR>
__device__ __forceinline__ void residual(const double* __restrict__ As,
                                         const double* __restrict__ Bs, int n, int nrhs,
                                         const double* xs, TF* rv, int tid, int lane, int wave,
                                         double* red, const double* bnorm, bool first,
                                         double& anorm, double& worst, bool& bad) {
  constexpr int NW = NT / 64, RD = NR > 4 ? 2 : 4, NV = 2 * NR + 1;
  double mx[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) mx[q] = 0.0;
#pragma unroll 1
  for (int r = tid; r < n; r += NT) {
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double b = Bs[r + min(q, nrhs - 1) * n];
      acc[q] = q < nrhs ? b : 0.0;
    }
    double asum = 0.0;
    double av[RD][8];
#pragma unroll
    for (int d = 0; d < RD; ++d)
#pragma unroll
      for (int j = 0; j < 8; ++j) av[d][j] = As[r + min(8 * d + j, n - 1) * n];
    for (int t0 = 0; t0 < n; t0 += 8 * RD) {
#pragma unroll
      for (int d = 0; d < RD; ++d) {
        const int c0 = t0 + 8 * d;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (c0 + j < n) {
            asum += fabs(av[d][j]);
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] -= av[d][j] * xs[(c0 + j) * kMaxRhs + q];
          }
          // bounds the X values read ahead (register pressure)
          constexpr int XM = NR > 4 ? 1 : 3;
          if ((j & XM) == XM) asm volatile("" ::: "memory");
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) av[d][j] = As[r + min(c0 + 8 * RD + j, n - 1) * n];
      }
    }
    mx[2 * NR] = fmax(mx[2 * NR], absinf(asum));
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      rv[r * kMaxRhs + q] = (TF)acc[q];
      mx[q] = fmax(mx[q], absinf(acc[q]));
      mx[NR + q] = fmax(mx[NR + q], absinf(xs[r * kMaxRhs + q]));
    }
  }
  // workgroup maxima (a NaN counts as +infinity): the waves' maxima through
  // LDS, one value at a time (register pressure)
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double w = wave_max(mx[q] == mx[q] ? mx[q] : INFINITY);
    if (lane == 0) red[wave * NV + q] = w;
    asm volatile("" ::: "memory");
  }
  __syncthreads();
  auto gmax = [&](int q) {
    double m = red[q];
#pragma unroll
    for (int w = 1; w < NW; ++w) m = fmax(m, red[w * NV + q]);
    return m;
  };
  if (first) anorm = gmax(2 * NR);
  worst = 0.0;
  bad = false;
#pragma unroll 1
  for (int q = 0; q < nrhs; ++q) {
    const double rn = gmax(q), den = anorm * gmax(NR + q) + bnorm[q];
    const double be = den > 0.0 ? rn / den : (rn > 0.0 ? INFINITY : 0.0);
    if (!(be <= kTol)) bad = true;
    worst = (be == be) ? fmax(worst, be) : INFINITY;
  }
  __syncthreads();
}

// One workgroup per system.  Pass 1 (TF = float): systems list[0..count);
// systems that do not reach the tolerance are appended to the refactor
// queue.  Pass 2 (TF = double): the systems of the queue (its length read
// on the device).  info: 0 refined from the FP32 factor, 1 after the FP64
// refactor, 2 not converged.  iters: refinement steps of the final pass.
//
// Systems are addressed by their position in the plan's size-sorted order:
// launch positions pos0 .. pos0+count-1, or the positions held in `list`
// (length *list_count).  The working copy of the system at position pos
// starts at wbase[n] + (pos - bstart[n]) n^2 (the plan lays the copies out
// by size).
template <typename TF, int NT, int NMAX>
// Occupancy: the kernel is latency-bound per system, so throughput follows
// the systems in flight.  The 256-thread FP32 instances are held to 128
// VGPRs (four workgroups per CU; 84 ms for 65,536 systems of sizes 16..256
// against 110 ms at 160 VGPRs and 132 ms at 256), which they fit without
// scratch once the loops rebuild the thread index (thread_index()).
__global__ __launch_bounds__(NT, sizeof(TF) == 8 ? 1 : (NT == 256 ? 4 : (NT == 128 ? 1 : 2))) void lu_ir_batch_kernel(
    int count, int pos0, const int* __restrict__ list, const int* __restrict__ list_count,
    const int* __restrict__ order, const int64_t* __restrict__ wbase,
    const int* __restrict__ bstart, const int64_t* __restrict__ offA,
    const int32_t* __restrict__ nsz, const double* __restrict__ A,
    const int64_t* __restrict__ offB, int nrhs, const double* __restrict__ B,
    double* __restrict__ X, void* __restrict__ Wbase, int32_t* __restrict__ info,
    int32_t* __restrict__ iters, int* __restrict__ queue, int* __restrict__ queue_n) {
  constexpr int NW = NT / 64;
  typedef typename Acc4<TF>::type AccT;
  __shared__ TF smem[LuSmem<TF, NMAX>::kWords];
  __shared__ int piv[NMAX];
  __shared__ int perm[NMAX + 1];
  __shared__ int s_tr[2 * PB], s_src[2 * PB], s_ntr;
  __shared__ double red_d[NW * (2 * kMaxRhs + 1)];
  __shared__ double s_bnorm[kMaxRhs];
  __shared__ int red_i[NW];
  const int total = list_count ? *list_count : count;
  constexpr bool kRefactor = sizeof(TF) == 8;
  const int swave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  for (int qi = blockIdx.x; qi < total; qi += gridDim.x) {
    // the thread index rebuilt (thread_index) here and in the inner loops
    const int tid = thread_index(swave), lane = tid & 63, wave = swave;
    const int pos = list ? list[qi] : pos0 + qi;
    const int s = order[pos];
    const int n = nsz[s];
    const double* __restrict__ As = A + offA[s];
    const double* __restrict__ Bs = B + offB[s];
    double* __restrict__ Xs = X + offB[s];
    TF* __restrict__ W =
        static_cast<TF*>(Wbase) + wbase[n] + (int64_t)(pos - bstart[n]) * n * n;

    // ---- 1. W = A in row-major, TF (64 x 64 tiles through LDS, a tile's
    // loads all requested before its first store)
    {
      constexpr int TS = NT >= 256 ? 64 : 32, PER = TS * TS / NT;
      TF* tile = smem;  // TS x (TS + 1)
      const int nt = (n + TS - 1) / TS;
      double v[PER];
      // tile tt's loads (branch-free: entries past the matrix read its last
      // row or column and are never stored); the next tile's are requested
      // before this tile's stores, so they are in flight meanwhile
      auto load_tile = [&](int tt) {
        const int tr = tt / nt, tc = tt - tr * nt;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT, i = e & (TS - 1), j = e / TS;
          const int r = min(tr * TS + i, n - 1), c = min(tc * TS + j, n - 1);
          v[q] = As[r + c * n];
        }
      };
      load_tile(0);
      for (int tt = 0; tt < nt * nt; ++tt) {
        const int tr = tt / nt, tc = tt - tr * nt;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT;
          tile[(e / TS) * (TS + 1) + (e & (TS - 1))] = (TF)v[q];
        }
        __syncthreads();
        if (tt + 1 < nt * nt) load_tile(tt + 1);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT, j = e & (TS - 1), i = e / TS;
          const int r = tr * TS + i, c = tc * TS + j;
          if (r < n && c < n) W[r * n + c] = tile[j * (TS + 1) + i];
        }
        __syncthreads();
      }
    }
    for (int r = tid; r <= n; r += NT) perm[r] = r;

    // ---- 2. blocked LU with partial pivoting
    TF* Pn = smem;                          // panel: [row][PB + 1]
    TF* Us = smem + LuSmem<TF, NMAX>::kPanel;     // U12: [PB][NMAX + 1]
    // Look-ahead: the next panel's columns (tile column 0 of the trailing
    // matrix) are updated first; then wave 0 loads and factors the next
    // panel in its registers while the other waves update the rest of the
    // trailing matrix, and writes it out after they are done.
    constexpr int NS = NMAX / 64;
    // wave 0: panel rows kp.. (columns kp .. kp + 15) from W into registers
    // (lane l holds rows kp + l + 64 k in P[k]), factored (PanelCols)
    auto factor_panel = [&](int kp, TF (&P)[NS][PB], int (&lg)[NS], int lane) {
      const int pbp = min(PB, n - kp), Rp = n - kp;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        lg[k] = lane + 64 * k;
        // rows past the matrix read its last row (never stored)
        const TF* p = W + (kp + min(lane + 64 * k, Rp - 1)) * n + kp;
        if (pbp == PB) {
#pragma unroll
          for (int j = 0; j < PB; j += 4) load4(p + j, &P[k][j]);
        } else {
#pragma unroll
          for (int j = 0; j < PB; ++j) {
            const TF x = p[min(j, pbp - 1)];
            P[k][j] = j < pbp ? x : TF(0);
          }
        }
      }
      PanelCols<TF, NS, 0>::run(P, lg, lane, Rp, pbp, kp, piv);
    };
    // wave 0: the factored panel to Pn and to W's panel columns in its final
    // row order, and the panel's interchanges as one permutation of the rows
    // they touch (s_tr / s_src / s_ntr): slot L < pb holds row kp + L, slot
    // pb + c the pivot row of column c unless an earlier slot holds it; src
    // is the row whose old content the slot's row receives
    auto write_panel = [&](int kp, TF (&P)[NS][PB], int (&lg)[NS], int lane) {
      const int pbp = min(PB, n - kp), Rp = n - kp;
#pragma unroll
      for (int k = 0; k < NS; ++k)
        if (lg[k] < Rp) {
#pragma unroll
          for (int j = 0; j < PB; ++j) Pn[lg[k] * (PB + 1) + j] = P[k][j];
          TF* w = W + (kp + lg[k]) * n + kp;
#pragma unroll
          for (int j = 0; j < PB; ++j)
            if (j < pbp) w[j] = P[k][j];
        }
      int row = lane < pbp ? kp + lane : (lane < 2 * pbp ? piv[kp + lane - pbp] : -1);
      int first = lane;
      for (int j = 0; j < 2 * PB; ++j) {
        const int rj = __builtin_amdgcn_readlane(row, j);
        if (j < lane && rj == row && first == lane) first = j;
      }
      const bool live = row >= 0 && first == lane;
      int src = row;
      for (int c = 0; c < pbp; ++c) {
        const int b = piv[kp + c];
        const int sb = (int)__builtin_ctzll(__ballot(live && row == b));
        const int va = __builtin_amdgcn_readlane(src, c);
        const int vb = __builtin_amdgcn_readlane(src, sb);
        if (lane == c) src = vb;
        if (lane == sb) src = va;
      }
      const bool moved = live && src != row;
      const unsigned long long mm = __ballot(moved);
      const int idx = __builtin_amdgcn_mbcnt_hi((unsigned)(mm >> 32),
                                                __builtin_amdgcn_mbcnt_lo((unsigned)mm, 0u));
      if (moved) {
        s_tr[idx] = row;
        s_src[idx] = src;
      }
      if (lane == 0) s_ntr = __popcll(mm);
    };
    // A22 += L21 (-U12) on the 16x16 tiles (ti, tj_begin .. tj_end - 1), in
    // groups of TB with all a group's loads issued first.  The MFMA A operand
    // is L21 of the tile row (panel at k0, trailing rows from T0), read from
    // the panel area or (A_FROM_W, while wave 0 rewrites the panel area)
    // from W's panel columns; rows past the matrix only reach products that
    // are never stored.  B is -U12 from the U12 area.
    constexpr int TB = sizeof(TF) == 8 ? 4 : (NMAX >= 128 ? 8 : 4);
    auto update_row = [&](auto a_from_w, int ti, int tj_begin, int tj_end, int k0, int T0,
                          int lane0) {
      const int lane = opaque(lane0), l16 = lane & 15, lk = lane >> 4;
      const int r0 = T0 + ti * 16, arow = r0 + l16;
      TF a[PB / 4];
      if constexpr (decltype(a_from_w)::value) {
        const TF* ap = W + min(arow, n - 1) * n + k0 + lk;
#pragma unroll
        for (int kk = 0; kk < PB / 4; ++kk) a[kk] = ap[kk * 4];
      } else {
#pragma unroll
        for (int kk = 0; kk < PB / 4; ++kk)
          a[kk] = arow < n ? Pn[(arow - k0) * (PB + 1) + kk * 4 + lk] : TF(0);
      }
      if constexpr (sizeof(TF) == 4) {
        // FP32: the transposed tile C^T += (-U12)^T L21^T (the operands
        // swapped), whose accumulator gives lane l row r0 + l % 16 and the
        // four consecutive columns 4 (l / 16) .. of the tile: one 16-byte
        // load and store per tile (rows past the matrix read its last row
        // and are not stored)
        const int row = r0 + l16;
        TF* wr = W + min(row, n - 1) * n + T0 + 4 * lk;
        for (int tj0 = tj_begin; tj0 < tj_end; tj0 += TB) {
          AccT acc[TB];
#pragma unroll
          for (int j = 0; j < TB; ++j)
            if (tj0 + j < tj_end) {
              const int c0 = (tj0 + j) * 16;
              if (T0 + c0 + 16 <= n) {
                acc[j] = *reinterpret_cast<const f4u*>(wr + c0);
              } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  acc[j][r] = T0 + c0 + 4 * lk + r < n ? wr[c0 + r] : TF(0);
              }
            }
#pragma unroll
          for (int j = 0; j < TB; ++j)
            if (tj0 + j < tj_end) {
              const int c0 = (tj0 + j) * 16;
#pragma unroll
              for (int kk = 0; kk < PB / 4; ++kk)
                acc[j] = mfma16x4(Us[(kk * 4 + lk) * (NMAX + 1) + c0 + l16], a[kk], acc[j]);
            }
#pragma unroll
          for (int j = 0; j < TB; ++j)
            if (tj0 + j < tj_end && row < n) {
              const int c0 = (tj0 + j) * 16;
              if (T0 + c0 + 16 <= n) {
                *reinterpret_cast<f4u*>(wr + c0) = acc[j];
              } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  if (T0 + c0 + 4 * lk + r < n) wr[c0 + r] = acc[j][r];
              }
            }
        }
        return;
      }
      for (int tj0 = tj_begin; tj0 < tj_end; tj0 += TB) {
        AccT acc[TB];
#pragma unroll
        for (int j = 0; j < TB; ++j) {
          const int col = T0 + (tj0 + j) * 16 + l16;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + acc_row<TF>(lane, r);
            acc[j][r] = (tj0 + j < tj_end && row < n && col < n) ? W[row * n + col] : TF(0);
          }
        }
#pragma unroll
        for (int j = 0; j < TB; ++j) {
          const int col = T0 + (tj0 + j) * 16 + l16;
#pragma unroll
          for (int kk = 0; kk < PB / 4; ++kk) {
            const TF bv = (tj0 + j < tj_end && col < n)
                              ? Us[(kk * 4 + lk) * (NMAX + 1) + col - T0]
                              : TF(0);
            acc[j] = mfma16x4(a[kk], bv, acc[j]);
          }
        }
#pragma unroll
        for (int j = 0; j < TB; ++j) {
          const int col = T0 + (tj0 + j) * 16 + l16;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + acc_row<TF>(lane, r);
            if (tj0 + j < tj_end && row < n && col < n) W[row * n + col] = acc[j][r];
          }
        }
      }
    };

    {
      const int tid = thread_index(swave), lane = tid & 63, wave = swave;
      if (wave == 0) {
        TF P[NS][PB];
        int lg[NS];
        factor_panel(0, P, lg, lane);
        write_panel(0, P, lg, lane);
      }
      __syncthreads();
    }
    for (int k0 = 0; k0 < n; k0 += PB) {
      const int tid = thread_index(swave), lane = tid & 63, wave = swave;
      const int pb = min(PB, n - k0);
      // the permutation on the other columns (every touched row read
      // before any is written, one column per thread) and on perm
      {
        const int ntr = s_ntr;
        // the first PB values in registers, the rest in this thread's slots
        // of the U12 area (written only after this step): fewer registers
        // (loads without branches: entries past ntr re-read the last source)
        TF* lds = Us + tid;
        for (int col = tid; ntr > 0 && col < n; col += NT) {
          if (col >= k0 && col < k0 + pb) continue;
          TF v[PB];
#pragma unroll
          for (int q = 0; q < PB; ++q) v[q] = W[s_src[min(q, ntr - 1)] * n + col];
#pragma unroll
          for (int q = 0; q < PB; ++q) lds[q * NT] = W[s_src[min(PB + q, ntr - 1)] * n + col];
#pragma unroll
          for (int q = 0; q < PB; ++q)
            if (q < ntr) W[s_tr[q] * n + col] = v[q];
#pragma unroll
          for (int q = 0; q < PB; ++q)
            if (PB + q < ntr) W[s_tr[PB + q] * n + col] = lds[q * NT];
        }
        const int pv = tid < ntr ? perm[s_src[tid]] : 0;
        __syncthreads();
        if (tid < ntr) perm[s_tr[tid]] = pv;
      }
      __syncthreads();
      const int T0 = k0 + pb, nr = n - T0;
      if (nr <= 0) break;
      // U12 = L11^-1 A12: four adjacent columns per thread (16-byte loads
      // and stores of W) up to a multiple of four, the last columns one per
      // thread
      const int nq = T0 + ((n - T0) & ~3);
      for (int c4 = T0 + 4 * tid; c4 < nq; c4 += 4 * NT) {
        TF u[PB][4];
#pragma unroll
        for (int t = 0; t < PB; ++t) load4(W + (k0 + t) * n + c4, u[t]);
#pragma unroll
        for (int t = 1; t < PB; ++t) {
#pragma unroll
          for (int q = 0; q < t; ++q) {
            const TF l = Pn[t * (PB + 1) + q];
#pragma unroll
            for (int e = 0; e < 4; ++e) u[t][e] -= l * u[q][e];
          }
          // bounds the L11 reads hoisted ahead (register pressure)
          asm volatile("" ::: "memory");
        }
#pragma unroll
        for (int t = 0; t < PB; ++t) {
          store4(W + (k0 + t) * n + c4, u[t]);
#pragma unroll
          for (int e = 0; e < 4; ++e) Us[t * (NMAX + 1) + c4 - T0 + e] = -u[t][e];
        }
      }
      for (int col = nq + tid; col < n; col += NT) {
        TF u[PB];
#pragma unroll
        for (int t = 0; t < PB; ++t) u[t] = W[(k0 + t) * n + col];
#pragma unroll
        for (int t = 1; t < PB; ++t)
#pragma unroll
          for (int q = 0; q < t; ++q) u[t] -= Pn[t * (PB + 1) + q] * u[q];
#pragma unroll
        for (int t = 0; t < PB; ++t) {
          W[(k0 + t) * n + col] = u[t];
          Us[t * (NMAX + 1) + col - T0] = -u[t];
        }
      }
      __syncthreads();
      const int tn = (nr + 15) / 16;
      const std::integral_constant<bool, false> from_lds;
      const std::integral_constant<bool, true> from_w;
      if constexpr (NW == 1) {
        // one wave: the whole update, then the next panel
        for (int ti = 0; ti < tn; ++ti) update_row(from_lds, ti, 0, tn, k0, T0, lane);
        TF P[NS][PB];
        int lg[NS];
        factor_panel(T0, P, lg, lane);
        write_panel(T0, P, lg, lane);
      } else {
        // (a) tile column 0 (the next panel's columns) by every wave
        for (int ti = wave; ti < tn; ti += NW) update_row(from_lds, ti, 0, 1, k0, T0, lane);
        __syncthreads();
        // (b) wave 0 factors the next panel and writes it out (panel area,
        // W's columns T0 .., interchange lists); the others update tile
        // columns 1.. with L21 from W
        if (wave == 0) {
          TF P[NS][PB];
          int lg[NS];
          factor_panel(T0, P, lg, lane);
          write_panel(T0, P, lg, lane);
        } else {
          for (int ti = wave - 1; ti < tn; ti += NW - 1)
            update_row(from_w, ti, 1, tn, k0, T0, lane);
        }
      }
      __syncthreads();
    }

    // ---- 3. solve and refine
    TF* diag = smem;
    TF* y = smem + LuSmem<TF, NMAX>::kDiag;
    TF* rv = y + LuSmem<TF, NMAX>::kVec;
    // ||B_q||_inf; ||A||_inf comes with the first residual
    // (held in LDS: registers are the scarce resource here)
    double anorm = 0.0;
#pragma unroll 1
    for (int q = 0; q < nrhs; ++q) {
      const int tid = thread_index(swave), lane = tid & 63, wave = swave;
      double v = 0.0;
      for (int r = tid; r < n; r += NT) v = fmax(v, absinf(Bs[r + q * n]));
      v = block_max<NW>(v, red_d, lane, wave);
      if (tid == 0) s_bnorm[q] = v;
    }

    int it = 0, converged = 0;
    double prev = INFINITY;
    for (;; ++it) {
      const int tid = thread_index(swave), lane = tid & 63, wave = swave;
      // the first pass solves P B, the later ones the correction P R:
      // X = U^-1 L^-1 P B, then X += U^-1 L^-1 P R (the columns past nrhs
      // zero; one call site of each triangular solve per column count)
#pragma unroll 1
      for (int e0 = tid; e0 < n * kMaxRhs; e0 += 8 * NT) {
        // eight entries per thread, their loads requested together
        // (branch-free: entries past the matrix or nrhs read a valid one and
        // are zeroed)
        double v[8];
        if (it == 0) {
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int e = min(e0 + k * NT, n * kMaxRhs - 1), r = e / kMaxRhs,
                      q = min(e - r * kMaxRhs, nrhs - 1);
            v[k] = Bs[perm[r] + q * n];
          }
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int e = min(e0 + k * NT, n * kMaxRhs - 1), r = e / kMaxRhs,
                      q = min(e - r * kMaxRhs, nrhs - 1);
            v[k] = (double)rv[perm[r] * kMaxRhs + q];
          }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int e = e0 + k * NT;
          if ((e & (kMaxRhs - 1)) >= nrhs) v[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (e0 + k * NT < n * kMaxRhs) y[e0 + k * NT] = (TF)v[k];
      }
      __syncthreads();
      if (nrhs <= 4) {
        tri_solve<TF, NT, true, 4>(W, n, y, diag, tid);
        tri_solve<TF, NT, false, 4>(W, n, y, diag, tid);
      } else {
        tri_solve<TF, NT, true, kMaxRhs>(W, n, y, diag, tid);
        tri_solve<TF, NT, false, kMaxRhs>(W, n, y, diag, tid);
      }
      // X = the solution (first pass) or X + the correction, written to X and
      // staged in LDS as FP64 for the residual (in the diagonal-block area,
      // which the solves no longer use); eight entries per thread, the old
      // values' loads requested together
      double* xs = reinterpret_cast<double*>(diag);
#pragma unroll 1
      for (int e0 = tid; e0 < n * nrhs; e0 += 8 * NT) {
        double xo[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int e = e0 + k * NT;
          const double x = Xs[min(e, n * nrhs - 1)];
          xo[k] = it > 0 ? x : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int e = e0 + k * NT, q = e / n, r = e - q * n;
          if (e < n * nrhs) {
            const double xn = xo[k] + (double)y[r * kMaxRhs + q];
            Xs[e] = xn;
            xs[r * kMaxRhs + q] = xn;
          }
        }
      }
      __syncthreads();
      // R = B - A X and the norms of the convergence test
      double worst = 0.0;
      bool bad = false;
      if (nrhs <= 4)
        residual<TF, NT, 4>(As, Bs, n, nrhs, xs, rv, tid, lane, wave, red_d, s_bnorm, it == 0,
                            anorm, worst, bad);
      else
        residual<TF, NT, kMaxRhs>(As, Bs, n, nrhs, xs, rv, tid, lane, wave, red_d, s_bnorm,
                                  it == 0, anorm, worst, bad);
      if (!bad) { converged = 1; break; }
      // stop on the iteration limit, or when a step gains less than 2x
      if (it >= kMaxIter || !(worst < 0.5 * prev)) break;
      prev = worst;
    }
    if (thread_index(swave) == 0) {
      if (converged) {
        info[s] = kRefactor ? 1 : 0;
        if (iters) iters[s] = it;
      } else if (kRefactor || !queue) {
        info[s] = 2;
        if (iters) iters[s] = it;
      } else {
        queue[atomicAdd(queue_n, 1)] = pos;
      }
    }
    __syncthreads();
  }
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

typedef void (*LuKernel)(int, int, const int*, const int*, const int*, const int64_t*,
                         const int*, const int64_t*, const int32_t*, const double*,
                         const int64_t*, int, const double*, double*, void*, int32_t*, int32_t*,
                         int*, int*);

// LDS is sized for the largest system of a launch: 64, 128 or 256
int nmax_for(int n) { return n <= 64 ? 64 : (n <= 128 ? 128 : 256); }

template <int NMAX>
LuKernel lu_kernel_nt(int nt) {
  if (nt == 64) return lu_ir_batch_kernel<float, 64, NMAX>;
  if (nt == 128) return lu_ir_batch_kernel<float, 128, NMAX>;
  return lu_ir_batch_kernel<float, 256, NMAX>;
}

LuKernel lu_kernel_for(int nt, int nmax) {
  if (nmax == 64) return lu_kernel_nt<64>(nt);
  if (nmax == 128) return lu_kernel_nt<128>(nt);
  return lu_kernel_nt<256>(nt);
}

int lu_default_threads(int n) { return n <= 48 ? 64 : (n <= 128 ? 128 : 256); }

int lu_grid(eamd_handle* h, int cnt) {
  return std::max(1, std::min(cnt, h->num_cus * 8));
}

// synthetic systems for the autotuner: A = n I + a small non-symmetric
// pattern (well conditioned: refinement converges from the FP32 factor)
__global__ void lu_tune_fill_kernel(int N, int count, int nrhs, double* A, int64_t* offA,
                                    double* B, int64_t* offB, int* order, int32_t* nn) {
  const long i = (long)blockIdx.x;
  if (i >= count) return;
  if (threadIdx.x == 0) {
    offA[i] = i * N * N;
    offB[i] = i * N * nrhs;
    order[i] = (int)i;
    nn[i] = N;
  }
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int r = e % N, c = e / N;
    A[i * N * N + e] = (r == c) ? (double)N : 0.5 / (1.0 + (double)((3 * r + c) % 11));
  }
  for (int e = threadIdx.x; e < N * nrhs; e += blockDim.x) B[i * N * nrhs + e] = 1.0 + 0.01 * e;
}

std::mutex g_lu_tune_mu;
int g_lu_tune[kMaxRhs + 1][kMaxLuN / 16 + 1] = {{0}};

int lu_autotune(eamd_handle* h, int n, int nrhs) {
  const int bkt = (n - 1) / 16;
  {
    std::lock_guard<std::mutex> lk(g_lu_tune_mu);
    if (g_lu_tune[nrhs][bkt]) return g_lu_tune[nrhs][bkt];
  }
  int best = lu_default_threads(n);
  const char* env = getenv("EAMD_BATCH_AUTOTUNE");
  if (!(env && env[0] == '0')) {
    const int N = std::min(kMaxLuN, bkt * 16 + 8);
    const int cnt = std::max(256, std::min(4096, (1 << 24) / (N * N)));
    const size_t nA = (size_t)cnt * N * N, nB = (size_t)cnt * N * nrhs;
    size_t need = sizeof(double) * (2 * nA + 2 * nB) + sizeof(int64_t) * (2 * cnt + 512) +
                  sizeof(int) * (4 * (size_t)cnt + 512) + 1024;
    void* p = nullptr;
    if (h->ws.get(kWsAux, need, &p) == kOk) {
      char* c = static_cast<char*>(p);
      double* Ad = reinterpret_cast<double*>(c);
      double* Wd = Ad + nA;
      double* Bd = Wd + nA;
      double* Xd = Bd + nB;
      int64_t* oA = reinterpret_cast<int64_t*>(Xd + nB);
      int64_t* oB = oA + cnt;
      // size tables of zeros: every system has size N, so the working copies
      // follow the position
      int64_t* wb = oB + cnt;
      int* bs = reinterpret_cast<int*>(wb + 512);
      int* order = bs + 512;
      int32_t* nn = order + cnt;
      int32_t* inf = nn + cnt;
      int32_t* its = inf + cnt;
      (void)hipMemsetAsync(wb, 0, sizeof(int64_t) * 512, h->stream);
      (void)hipMemsetAsync(bs, 0, sizeof(int) * 512, h->stream);
      hipLaunchKernelGGL(lu_tune_fill_kernel, dim3(cnt), dim3(256), 0, h->stream, N, cnt, nrhs,
                         Ad, oA, Bd, oB, order, nn);
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      float best_ms = 1e30f;
      for (int nt : {64, 128, 256}) {
        LuKernel k = lu_kernel_for(nt, nmax_for(N));
        float ms_min = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          (void)hipEventRecord(e0, h->stream);
          hipLaunchKernelGGL(k, dim3(lu_grid(h, cnt)), dim3(nt), 0, h->stream, cnt, 0, nullptr,
                             nullptr, order, wb, bs, oA, nn, Ad, oB, nrhs, Bd, Xd, Wd, inf, its,
                             nullptr, nullptr);
          (void)hipEventRecord(e1, h->stream);
          (void)hipEventSynchronize(e1);
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) ms_min = std::min(ms_min, ms);
        }
        if (ms_min < best_ms) {
          best_ms = ms_min;
          best = nt;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (hipGetLastError() != hipSuccess) best = lu_default_threads(n);
    }
  }
  std::lock_guard<std::mutex> lk(g_lu_tune_mu);
  g_lu_tune[nrhs][bkt] = best;
  return best;
}

}  // namespace

int batched_lu_plan_create(eamd_handle* h, int64_t count, const int32_t* n, int nrhs,
                           eamd_batched_plan** out) {
  *out = nullptr;
  if (nrhs < 1 || nrhs > kMaxRhs) {
    set_error("batched LU: nrhs %d outside [1, %d]", nrhs, kMaxRhs);
    return -6;
  }
  eamd_batched_plan* p = nullptr;
  EAMD_CHECK(batched_plan_sort(h, count, n, kMaxLuN, &p));
  p->kind = 1;
  p->nrhs = nrhs;
  // working copies laid out by size, largest first (the order of positions)
  int64_t wb[kMaxLuN + 1] = {0};
  int bs[kMaxLuN + 1] = {0};
  int64_t acc = 0;
  double flops = 0.0, bytes = 0.0;
  for (int v = kMaxLuN; v >= 1; --v) {
    wb[v] = acc;
    bs[v] = p->bucket_start[v];
    const double c = p->bucket_count[v], dv = v;
    acc += (int64_t)p->bucket_count[v] * v * v;
    // LU (2/3 n^3) plus the two triangular solves per right-hand side
    flops += c * (2.0 / 3.0 * dv * dv * dv + 2.0 * dv * dv * nrhs);
    // A read, B read, X written, offsets, size and info
    bytes += c * (8.0 * dv * dv + 16.0 * dv * nrhs + 16.0 + 8.0);
    p->lanes[v] = p->bucket_count[v] ? lu_autotune(h, v, nrhs) : 0;
  }
  p->total_elems = acc;
  p->lu_flops = flops;
  p->lu_bytes = bytes;
  hipError_t e = hipMalloc(&p->wbase, sizeof(wb));
  if (e == hipSuccess) e = hipMalloc(&p->bstart, sizeof(bs));
  if (e == hipSuccess) e = hipMemcpy(p->wbase, wb, sizeof(wb), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->bstart, bs, sizeof(bs), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    batched_plan_destroy(p);
    return hip_fail(e, "batched LU plan", __FILE__, __LINE__);
  }
  *out = p;
  return kOk;
}

int dgesv_ir_batched_exec(eamd_handle* h, const eamd_batched_plan* p, const int64_t* offA,
                          const int32_t* n, const double* A, const int64_t* offB,
                          const double* B, double* X, int32_t* info, int32_t* iters) {
  if (p->kind != 1) {
    set_error("batched LU: the plan was made for another solver");
    return kErrUnsupported;
  }
  const int64_t count = p->count;
  if (count == 0) return kOk;
  // workspace: the working copies (FP64-sized, so the refactor pass fits
  // at the same offsets), the refactor queue and its length
  const int64_t elems = std::max<int64_t>(1, p->total_elems);
  size_t need = sizeof(double) * (size_t)elems + sizeof(int) * (size_t)(count + 64);
  void* base = nullptr;
  EAMD_CHECK(h->ws.get(kWsBatch, need, &base));
  double* W = static_cast<double*>(base);
  int* queue_n = reinterpret_cast<int*>(W + elems);
  int* queue = queue_n + 64;
  const double flops = p->lu_flops;
  PhaseScope ps(h, kPhBatched, p->lu_bytes, flops);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, h->stream, queue_n);
  // consecutive sizes (descending) that share a workgroup size form one launch
  int v = kMaxLuN;
  while (v >= 1) {
    if (!p->bucket_count[v]) { --v; continue; }
    const int nt = p->lanes[v], nmax = nmax_for(v);
    const int start = p->bucket_start[v];
    int cnt = 0, u = v;
    while (u >= 1 && (!p->bucket_count[u] || (p->lanes[u] == nt && nmax_for(u) == nmax))) {
      cnt += p->bucket_count[u];
      --u;
    }
    hipLaunchKernelGGL(lu_kernel_for(nt, nmax), dim3(lu_grid(h, cnt)), dim3(nt), 0, h->stream, cnt,
                       start, nullptr, nullptr, p->order, p->wbase, p->bstart, offA, n, A, offB,
                       p->nrhs, B, X, W, info, iters, queue, queue_n);
    v = u;
  }
  // FP64 refactor of the systems the FP32 factor could not refine
  hipLaunchKernelGGL((lu_ir_batch_kernel<double, 256, kMaxLuN>), dim3(h->num_cus), dim3(256), 0,
                     h->stream, 0, 0, queue, queue_n, p->order, p->wbase, p->bstart, offA, n, A,
                     offB, p->nrhs, B, X, W, info, iters, nullptr, nullptr);
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int dgesv_ir_batched_impl(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                          const double* A, const int64_t* offB, int64_t nrhs, const double* B,
                          double* X, int32_t* info, int32_t* iters) {
  eamd_batched_plan* p = nullptr;
  EAMD_CHECK(batched_lu_plan_create(h, count, n, (int)nrhs, &p));
  int st = dgesv_ir_batched_exec(h, p, offA, n, A, offB, B, X, info, iters);
  (void)hipStreamSynchronize(h->stream);
  batched_plan_destroy(p);
  return st;
}

}  // namespace eamd
