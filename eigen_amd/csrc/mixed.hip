// mixed.hip -- low-precision trailing update of the FP32 LU used by the
// mixed-precision solver.
//
// The LU of the refinement solver keeps its working matrix in FP32 (row-major)
// and factors panels in FP32 (panel_lu_kernel<float>).  The trailing update
// A22 -= A21 * A12 runs on FP16 / BF16 matrix cores with FP32 accumulation:
//
//   * both operands are packed once per block step into K-contiguous FP16
//     (or BF16) copies, scaled per 128-row block by a power of two chosen so
//     that the block maximum lands just below 2^14 (exact scaling, no range
//     loss in FP16);
//   * FP16X3 splits each scaled value x into hi = fp16(x) and
//     lo = fp16((x - hi) * 2^11) and accumulates hi*hi + 2^-11 (hi*lo + lo*hi)
//     in FP32 (two accumulator sets): 22 significant bits per operand, an
//     update accurate to about FP32 unit roundoff on three MFMA products;
//   * the MFMA tile computes the transposed product so each lane's four
//     results are four columns of one row run of C (64-byte coalesced FP32
//     read-modify-write).
// The update is HBM-bound (C is read and written once per block step); see
// the roofline accounting in the benchmark.
#include "../../include/eigen_amd.h"
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace eamd {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

constexpr int MT = 128;   // C tile (M and N), also the scale block
constexpr int KS = 32;    // K per MFMA slab (v_mfma_f32_16x16x32_{f16,bf16})
constexpr int LROW = 32;  // LDS row length in halves (64 B, four 16-byte chunks)

// 16-byte chunk c of LDS row r is stored at chunk c ^ lds_swz(r).  Without
// padding, the 16 rows that one MFMA operand read touches then fall on
// distinct 16-byte bank groups, for 8 lanes per LDS cycle on 32 banks and 16
// lanes on 64 banks alike.  Unpadded rows keep the two buffers of the four
// arrays at 64 KB per workgroup: two workgroups leave 32 KB of a CU's LDS to
// the panel kernels that run beside the update.
__device__ __forceinline__ int lds_swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

enum { kModeF16 = 0, kModeF16x3 = 1, kModeBF16 = 2 };

__device__ __forceinline__ float pow2_scale(float amax) {
  if (!(amax > 0.0f) || !isfinite(amax)) return 1.0f;
  int ex;
  frexpf(amax, &ex);  // amax < 2^ex
  return ldexpf(1.0f, 14 - ex);
}

template <int MODE>
__device__ __forceinline__ void split_store(float x, float s, unsigned short* hi,
                                            unsigned short* lo) {
  float y = x * s;
  if (MODE == kModeBF16) {
    __bf16 h = (__bf16)y;
    *hi = *reinterpret_cast<unsigned short*>(&h);
  } else {
    _Float16 h = (_Float16)y;
    *hi = *reinterpret_cast<unsigned short*>(&h);
    if (MODE == kModeF16x3) {
      // the low part unscaled: below 2^-14 it is an FP16 subnormal, an
      // absolute error of at most 2^-25 against a block maximum near 2^14
      _Float16 l = (_Float16)(y - (float)h);
      *lo = *reinterpret_cast<unsigned short*>(&l);
    }
  }
}

// Block maximum of |x| over a workgroup of 256 threads.
__device__ __forceinline__ float block_absmax(float m, float* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Eight consecutive k of one packed row: converted and stored as one 16-byte
// write per array.
template <int MODE>
__device__ __forceinline__ void store8(const float (&x)[8], float s, unsigned short* ph,
                                       unsigned short* pl) {
  unsigned short hb[8], lb[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) split_store<MODE>(x[q], s, &hb[q], &lb[q]);
  uint4 h, l;
  __builtin_memcpy(&h, hb, 16);
  *reinterpret_cast<uint4*>(ph) = h;
  if (MODE == kModeF16x3) {
    __builtin_memcpy(&l, lb, 16);
    *reinterpret_cast<uint4*>(pl) = l;
  }
}

// Pack rows of a row-major source (N rows x K, element (j,k) at S[j*ld+k])
// into P[j*KP + k] (+ lo copy), scale per 128-row block.  One workgroup per
// 128-row block; a thread takes 8 consecutive k of a row per step (one
// 16-byte store per array); first pass the block maximum, second the
// conversion (its reads hit the cache).
template <int MODE>
__global__ __launch_bounds__(256) void pack_rows_kernel(const float* __restrict__ S, int64_t ld,
                                                        int N, int K, int KP,
                                                        unsigned short* __restrict__ Ph,
                                                        unsigned short* __restrict__ Pl,
                                                        float* __restrict__ scale) {
  __shared__ float red[4];
  const int j0 = blockIdx.x * MT;
  const int rows = min(MT, N - j0);
  const int kc = KP / 8, chunks = MT * kc;
  float m = 0.0f;
  for (int e = threadIdx.x; e < chunks; e += 256) {
    const int jj = e / kc, k = (e - jj * kc) * 8;
    if (jj < rows) {
      const float* src = S + (int64_t)(j0 + jj) * ld + k;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (k + q < K) m = fmaxf(m, fabsf(src[q]));
    }
  }
  const float s = pow2_scale(block_absmax(m, red));
  if (threadIdx.x == 0) scale[blockIdx.x] = s;
  for (int e = threadIdx.x; e < chunks; e += 256) {
    const int jj = e / kc, k = (e - jj * kc) * 8;
    const float* src = S + (int64_t)(j0 + jj) * ld + k;
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = (jj < rows && k + q < K) ? src[q] : 0.0f;
    const size_t o = (size_t)(j0 + jj) * KP + k;
    store8<MODE>(x, s, Ph + o, Pl + o);
  }
}

// Pack columns of a row-major source (K rows x M, element (k,i) at
// S[k*ld+i]) into P[i*KP + k]: 128 columns per workgroup; a thread takes 8
// consecutive k of one column per step (each of the 8 loads is coalesced
// across the threads of a wave, the store is one 16-byte write per array).
template <int MODE>
__global__ __launch_bounds__(256) void pack_cols_kernel(const float* __restrict__ S, int64_t ld,
                                                        int M, int K, int KP,
                                                        unsigned short* __restrict__ Ph,
                                                        unsigned short* __restrict__ Pl,
                                                        float* __restrict__ scale) {
  __shared__ float red[4];
  const int i0 = blockIdx.x * MT;
  const int cols = min(MT, M - i0);
  const int ii = threadIdx.x & (MT - 1), kg = threadIdx.x >> 7;  // 2 groups of k chunks
  const int kc = KP / 8;
  const bool cv = ii < cols;
  float m = 0.0f;
  for (int c = kg; c < kc; c += 2) {
    const int k = c * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (cv && k + q < K) m = fmaxf(m, fabsf(S[(int64_t)(k + q) * ld + i0 + ii]));
  }
  const float s = pow2_scale(block_absmax(m, red));
  if (threadIdx.x == 0) scale[blockIdx.x] = s;
  for (int c = kg; c < kc; c += 2) {
    const int k = c * 8;
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      x[q] = (cv && k + q < K) ? S[(int64_t)(k + q) * ld + i0 + ii] : 0.0f;
    const size_t o = (size_t)(i0 + ii) * KP + k;
    store8<MODE>(x, s, Ph + o, Pl + o);
  }
}

__device__ __forceinline__ void tile_coords(int bid, int nblocks, int tiles_m, int tiles_n,
                                            int& tm, int& tn) {
  int xcd = bid & 7, loc = bid >> 3;
  int per = nblocks >> 3, rem = nblocks & 7;
  int logical = (xcd < rem) ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
  const int G = 8;
  int group = G * tiles_n;
  int g = logical / group;
  int first_m = g * G;
  int gm = min(tiles_m - first_m, G);
  int r = logical - g * group;
  tm = first_m + r % gm;
  tn = r / gm;
}

template <int MODE>
__device__ __forceinline__ v4f mfma16(const uint4& a, const uint4& b, v4f c) {
  if (MODE == kModeBF16) {
    v8bf x, y;
    __builtin_memcpy(&x, &a, 16);
    __builtin_memcpy(&y, &b, 16);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, c, 0, 0, 0);
  } else {
    v8h x, y;
    __builtin_memcpy(&x, &a, 16);
    __builtin_memcpy(&y, &b, 16);
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(x, y, c, 0, 0, 0);
  }
}

// C (M x N column-major, ldc) -= (Pa . Pb) / (sa * sb).  Pa: M x KP
// K-contiguous, Pb: N x KP K-contiguous.
template <int MODE>
__global__ __launch_bounds__(256, 2) void mixed_update_kernel(int M, int N, int KP,
                                                           const unsigned short* __restrict__ Pah,
                                                           const unsigned short* __restrict__ Pal,
                                                           const unsigned short* __restrict__ Pbh,
                                                           const unsigned short* __restrict__ Pbl,
                                                           const float* __restrict__ sa,
                                                           const float* __restrict__ sb,
                                                           float* __restrict__ C, int64_t ldc) {
  constexpr bool X3 = MODE == kModeF16x3;
  constexpr int NARR = X3 ? 4 : 2;
  __shared__ __align__(16) unsigned short lds[2][NARR][MT * LROW];

  const int tiles_m = (M + MT - 1) / MT, tiles_n = (N + MT - 1) / MT;
  const int ntiles = tiles_m * tiles_n;
  // a grid smaller than the tile count loops (a multiple of 8 workgroups
  // keeps each tile on the XCD the mapping assigns it)
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
  int tm, tn;
  tile_coords(t, ntiles, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * MT, n0 = tn * MT;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int l16 = lane & 15, lk = lane >> 4;

  // one accumulator per 16x16 tile: hi*hi and both cross terms (the low
  // parts are unscaled) accumulate together in FP32
  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // staging: each thread moves 32 bytes of each array per slab:
  // row = tid / 2, byte half = (tid & 1) * 32 of the 64-byte K slab
  const int srow = tid >> 1, shalf = (tid & 1) * 16;  // in halves
  const int swz = lds_swz(srow);
  // the operand rows below are 16-aligned bases + l16: their swizzle is l16's
  const int rch = (lk ^ lds_swz(l16)) * 8;
  const unsigned short* src[4] = {Pah, Pbh, Pal, Pbl};
  uint4 stg[4][2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int a = 0; a < NARR; ++a) {
      const bool isB = (a & 1);
      const int base = isB ? n0 : m0;
      const int lim = isB ? N : M;
      const int r = base + srow;
      if (r < lim) {
        const uint4* p = reinterpret_cast<const uint4*>(src[a] + (size_t)r * KP + k0 + shalf);
        stg[a][0] = p[0];
        stg[a][1] = p[1];
      } else {
        stg[a][0] = make_uint4(0, 0, 0, 0);
        stg[a][1] = make_uint4(0, 0, 0, 0);
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int a = 0; a < NARR; ++a) {
      unsigned short* d = &lds[buf][a][srow * LROW];
      *reinterpret_cast<uint4*>(d + (((shalf >> 3) ^ swz) << 3)) = stg[a][0];
      *reinterpret_cast<uint4*>(d + ((((shalf >> 3) + 1) ^ swz) << 3)) = stg[a][1];
    }
  };

  const int kslabs = KP / KS;
  gload(0);
  // The C tile is read in four parts of 16 rows per wave, part p beside
  // slab p + 1's operand loads (the operand stores of a slab then wait only
  // for their own loads), and is used only in the epilogue: a 64 KB burst of
  // C per workgroup ahead of the first products made every workgroup of a
  // launch wait for it at once.  The products are formed transposed (A-side
  // operand first), so a lane's four results are four consecutive M indices
  // 4*lk + r of one N index l16: one 16-byte word of the row-major C when
  // ldc and C allow it.
  const bool cvec = (ldc & 3) == 0 && ((uintptr_t)C & 15) == 0;
  v4f cv[4][4];
  auto cpart = [&](auto I) {
    constexpr int i = decltype(I)::value;
    const int gi0 = m0 + wm + i * 16 + lk * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gj = n0 + wn + j * 16 + l16;
      const float* pc = C + (int64_t)gj * ldc + gi0;
      if (gj < N && cvec && gi0 + 3 < M) {
        cv[i][j] = *reinterpret_cast<const v4f*>(pc);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][j][r] = (gj < N && gi0 + r < M) ? pc[r] : 0.f;
      }
    }
  };
  // the products of slab ks (staged in LDS buffer ks & 1)
  auto slab = [&](int ks) {
    const int buf = ks & 1;
    uint4 af[4], bf[4], afl[4], bfl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      af[i] = *reinterpret_cast<const uint4*>(&lds[buf][0][(wm + i * 16 + l16) * LROW + rch]);
      if (X3)
        afl[i] = *reinterpret_cast<const uint4*>(&lds[buf][2][(wm + i * 16 + l16) * LROW + rch]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bf[j] = *reinterpret_cast<const uint4*>(&lds[buf][1][(wn + j * 16 + l16) * LROW + rch]);
      if (X3)
        bfl[j] = *reinterpret_cast<const uint4*>(&lds[buf][3][(wn + j * 16 + l16) * LROW + rch]);
    }
    // small terms first, so each tile's sum ends with its largest part
    if (X3) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16<MODE>(af[i], bfl[j], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16<MODE>(afl[i], bf[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma16<MODE>(af[i], bf[j], acc[i][j]);
  };
  lstore(0);
  __syncthreads();
  // slabs 0..3 with C part p (a slab that does not exist only reads C)
  auto step = [&](auto I) {
    constexpr int p = decltype(I)::value;
    if (p < kslabs) {
      if (p + 1 < kslabs) gload((p + 1) * KS);
      cpart(I);
      slab(p);
      if (p + 1 < kslabs) lstore((p + 1) & 1);
      __syncthreads();
    } else {
      cpart(I);
    }
  };
  step(std::integral_constant<int, 0>{});
  step(std::integral_constant<int, 1>{});
  step(std::integral_constant<int, 2>{});
  step(std::integral_constant<int, 3>{});
  for (int ks = 4; ks < kslabs; ++ks) {
    if (ks + 1 < kslabs) gload((ks + 1) * KS);
    slab(ks);
    if (ks + 1 < kslabs) lstore((ks + 1) & 1);
    __syncthreads();
  }

  const float inv = 1.0f / (sa[tm] * sb[tn]);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi0 = m0 + wm + i * 16 + lk * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gj = n0 + wn + j * 16 + l16;
      if (gj >= N) continue;
      float* pc = C + (int64_t)gj * ldc + gi0;
      const v4f res = cv[i][j] - acc[i][j] * inv;
      if (cvec && gi0 + 3 < M) {
        *reinterpret_cast<v4f*>(pc) = res;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (gi0 + r < M) pc[r] = res[r];
      }
    }
  }
  }
}

// FP32 update on the FP32 matrix cores (v_mfma_f32_16x16x4f32), operands
// read straight from the row-major working matrix through LDS.
__global__ __launch_bounds__(256) void f32_update_kernel(int M, int N, int K,
                                                         const float* __restrict__ A12,
                                                         const float* __restrict__ A21,
                                                         int64_t ld, float* __restrict__ C) {
  // Xa(i,k) = A12[k*ld + i]; Xb(k,j) = A21[j*ld + k]; C(i,j) = C[j*ld + i]
  constexpr int BK = 16;
  __shared__ float As[2][BK][MT + 4];
  __shared__ float Bs[2][BK][MT + 4];
  const int tiles_m = (M + MT - 1) / MT, tiles_n = (N + MT - 1) / MT;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * MT, n0 = tn * MT;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int l16 = lane & 15, lk = lane >> 4;
  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  float ra[8], rb[8];
  auto gload = [&](int k0) {
    {  // A12 rows k (16) x 128 i: thread -> k = tid/16, i = (tid%16)*8
      int k = k0 + (tid >> 4), i = m0 + (tid & 15) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) ra[q] = (k < K && i + q < M) ? A12[(int64_t)k * ld + i + q] : 0.f;
    }
    {  // A21 rows j (128) x 16 k: thread -> j = tid/2, k = (tid%2)*8
      int j = n0 + (tid >> 1), k = k0 + (tid & 1) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) rb[q] = (j < N && k + q < K) ? A21[(int64_t)j * ld + k + q] : 0.f;
    }
  };
  auto lstore = [&](int buf) {
    int k = tid >> 4, i = (tid & 15) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) As[buf][k][i + q] = ra[q];
    int j = tid >> 1, kk = (tid & 1) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) Bs[buf][kk + q][j] = rb[q];
  };
  const int kt = (K + BK - 1) / BK;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int t = 0; t < kt; ++t) {
    const int buf = t & 1;
    if (t + 1 < kt) gload((t + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      float af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[buf][kk * 4 + lk][wm + i * 16 + l16];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][kk * 4 + lk][wn + j * 16 + l16];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (t + 1 < kt) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = m0 + wm + i * 16 + l16;
    if (gi >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn + j * 16 + lk * 4 + r;
        if (gj >= N) continue;
        C[(int64_t)gj * ld + gi] -= acc[i][j][r];
      }
  }
}

// FP32 update of at most 128 columns with an inner dimension of at most 128
// (the next panel's columns in the look-ahead LU, a leaf's panel columns):
// C (N rows x M columns) -= A21 (N x K) A12 (K x M), row-major, ld.  One
// workgroup per 32 rows, each wave 32 of the columns.  Every operand word is
// loaded up front (the whole A12, the workgroup's A21 rows and its C tile),
// so a single memory round trip precedes the FP32 MFMA work; K passes
// through LDS in four quarters of 32 (21 KB of LDS: the workgroup fits on a
// CU beside two trailing-update workgroups).  M, K and ld are multiples of 4
// and the operands 16-byte aligned (checked by the caller).
__global__ __launch_bounds__(256) void f32_update_narrow_kernel(int M, int N, int K,
                                                                const float* __restrict__ A12,
                                                                const float* __restrict__ A21,
                                                                int64_t ld, float* __restrict__ C) {
  // the next panel's columns: on the critical path beside the trailing
  // update's waves (issue priority)
  __builtin_amdgcn_s_setprio(3);
  constexpr int NT = 32, KH = 32;
  __shared__ float As[KH][128 + 4];  // A12: k x column
  __shared__ float Bs[KH][NT + 4];   // A21 transposed: k x row
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n0 = blockIdx.x * NT;
  const int l16 = lane & 15, lk = lane >> 4, wm = wave * 32;
  // A12 rows k = tid/32 + 8t, words (tid%32); A21 row tid/8, words (tid%8) + 8t
  v4f ra[16], rb[4];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int k = (tid >> 5) + 8 * t, i = (tid & 31) * 4;
    ra[t] = (k < K && i < M) ? *reinterpret_cast<const v4f*>(A12 + (int64_t)k * ld + i)
                             : v4f{0.f, 0.f, 0.f, 0.f};
  }
  {
    const int j = n0 + (tid >> 3);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = (tid & 7) * 4 + 32 * t;
      rb[t] = (j < N && k < K) ? *reinterpret_cast<const v4f*>(A21 + (int64_t)j * ld + k)
                               : v4f{0.f, 0.f, 0.f, 0.f};
    }
  }
  // the C tile in the MFMA result layout: row n0 + 16 j + 4 lk + r, column
  // wm + 16 i + l16
  float cv[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + j * 16 + lk * 4 + r, gi = wm + i * 16 + l16;
        cv[i][j][r] = (gj < N && gi < M) ? C[(int64_t)gj * ld + gi] : 0.f;
      }
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    if (h) __syncthreads();  // every read of the previous quarter is done
#pragma unroll
    for (int t = 0; t < 4; ++t)
      *reinterpret_cast<v4f*>(&As[(tid >> 5) + 8 * t][(tid & 31) * 4]) = ra[4 * h + t];
    {
      const int kk = (tid & 7) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) Bs[kk + e][tid >> 3] = rb[h][e];
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < KH / 4; ++k4) {
      float af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = As[k4 * 4 + lk][wm + i * 16 + l16];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Bs[k4 * 4 + lk][j * 16 + l16];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + j * 16 + lk * 4 + r, gi = wm + i * 16 + l16;
        if (gj < N && gi < M) C[(int64_t)gj * ld + gi] = cv[i][j][r] - acc[i][j][r];
      }
}

}  // namespace

// row-major A22 (mrows x ncols) -= A21 (mrows x k) * A12 (k x ncols), ld
namespace {
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
int mixed_mode(int precision) {
  return precision == EAMD_PREC_FP16X3 ? kModeF16x3
         : precision == EAMD_PREC_BF16 ? kModeBF16 : kModeF16;
}
}  // namespace

size_t mixed_operand_bytes(int len, int k, int precision) {
  const size_t KP = (size_t)ceil_div(k, KS) * KS, LB = (size_t)ceil_div(len, MT);
  const size_t narr = mixed_mode(precision) == kModeF16x3 ? 2 : 1;
  return narr * align256(sizeof(unsigned short) * LB * MT * KP) + align256(sizeof(float) * LB);
}

// Packs one operand of the mixed update into buf (mixed_operand_bytes):
// rows = true takes `len` rows of K values (A21), rows = false takes `len`
// columns of K rows (A12), both row-major with leading dimension ld.
int mixed_pack(eamd_handle* h, bool rows, const float* S, int64_t ld, int len, int k,
               int precision, void* buf, MixedOperand* out) {
  if (len <= 0 || k <= 0) return kOk;
  const int mode = mixed_mode(precision);
  const int KP = ceil_div(k, KS) * KS, LB = ceil_div(len, MT);
  const size_t arr = align256(sizeof(unsigned short) * (size_t)LB * MT * KP);
  char* c = static_cast<char*>(buf);
  out->hi = reinterpret_cast<unsigned short*>(c);
  out->lo = mode == kModeF16x3 ? reinterpret_cast<unsigned short*>(c + arr) : out->hi;
  out->scale = reinterpret_cast<float*>(c + (mode == kModeF16x3 ? 2 : 1) * arr);
  const int narr = mode == kModeF16x3 ? 2 : 1;
  PhaseScope ps(h, kPhPack, (4.0 + 2.0 * narr) * (double)k * len, 0.0);
#define EAMD_PACK(MODE)                                                                    \
  if (rows)                                                                                \
    hipLaunchKernelGGL(pack_rows_kernel<MODE>, dim3(LB), dim3(256), 0, h->stream, S, ld, len, \
                       k, KP, out->hi, out->lo, out->scale);                               \
  else                                                                                     \
    hipLaunchKernelGGL(pack_cols_kernel<MODE>, dim3(LB), dim3(256), 0, h->stream, S, ld, len, \
                       k, KP, out->hi, out->lo, out->scale);
  if (mode == kModeF16x3) {
    EAMD_PACK(kModeF16x3)
  } else if (mode == kModeBF16) {
    EAMD_PACK(kModeBF16)
  } else {
    EAMD_PACK(kModeF16)
  }
#undef EAMD_PACK
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// A22 (N rows x M columns, row-major, ld) -= A21 A12 from packed operands:
// a = A12 (M columns), b = A21 (N rows), inner dimension k.
int mixed_update_packed(eamd_handle* h, const MixedOperand& a, int M, const MixedOperand& b, int N,
                        int k, int precision, float* A22, int64_t ld, int max_wgs) {
  if (M <= 0 || N <= 0 || k <= 0) return kOk;
  const int mode = mixed_mode(precision);
  const int KP = ceil_div(k, KS) * KS;
  int tiles = ceil_div(M, MT) * ceil_div(N, MT);
  if (max_wgs > 0 && tiles > max_wgs) tiles = max_wgs & ~7;
  const int narr = mode == kModeF16x3 ? 2 : 1;
  // algorithmic traffic: FP32 C read + write, packed operands read once
  PhaseScope ps(h, kPhUpdate, 8.0 * M * N + 2.0 * narr * (double)KP * (M + N),
                2.0 * M * N * (double)k);
#define EAMD_UPD(MODE)                                                                         \
  hipLaunchKernelGGL(mixed_update_kernel<MODE>, dim3(tiles), dim3(256), 0, h->stream, M, N, KP, \
                     a.hi, a.lo, b.hi, b.lo, a.scale, b.scale, A22, ld);
  if (mode == kModeF16x3) {
    EAMD_UPD(kModeF16x3)
  } else if (mode == kModeBF16) {
    EAMD_UPD(kModeBF16)
  } else {
    EAMD_UPD(kModeF16)
  }
#undef EAMD_UPD
  EAMD_HIP(hipGetLastError());
  return kOk;
}

// FP32 update A22 -= A21 A12 (row-major, ld) on the FP32 matrix cores,
// timed under `phase` (the panel's own updates add no bytes or flops).
int f32_update_rowmajor(eamd_handle* h, const float* A12, const float* A21, float* A22,
                        int64_t ld, int mrows, int ncols, int k, int phase) {
  if (mrows <= 0 || ncols <= 0 || k <= 0) return kOk;
  const int M = ncols, N = mrows;
  const bool in_panel = phase == kPhPanel;
  PhaseScope ps(h, phase, in_panel ? 0.0 : 4.0 * (2.0 * M * N + (double)k * (M + N)),
                in_panel ? 0.0 : 2.0 * M * N * k);
  const bool narrow = M <= 128 && k <= 128 && M % 4 == 0 && k % 4 == 0 && ld % 4 == 0 &&
                      ((uintptr_t)A12 & 15) == 0 && ((uintptr_t)A21 & 15) == 0;
  if (narrow) {
    hipLaunchKernelGGL(f32_update_narrow_kernel, dim3(ceil_div(N, 32)), dim3(256), 0, h->stream,
                       M, N, k, A12, A21, ld, A22);
  } else {
    const int tiles = ceil_div(M, MT) * ceil_div(N, MT);
    hipLaunchKernelGGL(f32_update_kernel, dim3(tiles), dim3(256), 0, h->stream, M, N, k, A12,
                       A21, ld, A22);
  }
  EAMD_HIP(hipGetLastError());
  return kOk;
}

int sgemm_update_rowmajor(eamd_handle* h, const float* A12, const float* A21, float* A22,
                          int64_t ld, int mrows, int ncols, int k, int precision) {
  if (mrows <= 0 || ncols <= 0 || k <= 0) return kOk;
  const int M = ncols, N = mrows;
  if (precision == EAMD_PREC_FP32 || precision == EAMD_PREC_FP64 || precision < 0)
    // precision < 0: an update inside the panel (FP32, accounted to the panel)
    return f32_update_rowmajor(h, A12, A21, A22, ld, mrows, ncols, k,
                               precision < 0 ? kPhPanel : kPhUpdate);
  const size_t oa = mixed_operand_bytes(M, k, precision), ob = mixed_operand_bytes(N, k, precision);
  void* p = nullptr;
  EAMD_CHECK(h->ws.get(kWsSplit, oa + ob, &p));
  MixedOperand a, b;
  EAMD_CHECK(mixed_pack(h, false, A12, ld, M, k, precision, p, &a));
  EAMD_CHECK(mixed_pack(h, true, A21, ld, N, k, precision, static_cast<char*>(p) + oa, &b));
  return mixed_update_packed(h, a, M, b, N, k, precision, A22, ld);
}

}  // namespace eamd
