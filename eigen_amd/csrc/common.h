// common.h -- shared device/host definitions for the eigen_amd HIP library
// (MI355X / gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace eamd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------- errors --
// Status codes follow the reference's LAPACK convention (lapack/lu.cpp:14-41):
// 0 ok, <0 = -(index of the bad argument), >0 numerical.  Library failures
// (HIP errors, allocation) use the reserved range below.
enum : int {
  kOk = 0,
  kErrHip = -1000,
  kErrAlloc = -1001,
  kErrUnsupported = -1002,
  kErrNotConverged = -1003,
};

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define EAMD_HIP(...)                                                        \
  do {                                                                       \
    hipError_t _e = (__VA_ARGS__);                                           \
    if (_e != hipSuccess) return ::eamd::hip_fail(_e, #__VA_ARGS__, __FILE__, __LINE__); \
  } while (0)

#define EAMD_CHECK(...)           \
  do {                            \
    int _s = (__VA_ARGS__);       \
    if (_s != 0) return _s;       \
  } while (0)

// ------------------------------------------------------------- workspace --
// A grow-only device arena owned by a handle.  Slots are named so that
// concurrent users inside one call do not alias.
struct Workspace {
  static constexpr int kSlots = 24;
  void* ptr[kSlots] = {};
  size_t bytes[kSlots] = {};
  int get(int slot, size_t need, void** out);
  void release();
};

}  // namespace eamd

namespace eamd {
// Device phase timers: hipEvent pairs recorded around the library's own
// launches on the handle's stream (enabled per handle; off by default).
enum Phase : int {
  kPhPanel = 0, kPhLaswp, kPhTrsm, kPhUpdate, kPhPack, kPhTrsv, kPhResidual, kPhGemm,
  kPhPotf2, kPhBatched, kPhNext, kPhCount
};
struct PhaseTimer {
  bool enabled = false;
  struct Rec {
    int phase;
    hipEvent_t a, b;
    double bytes, flops;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Rec> open;
  double ms[kPhCount] = {}, bytes[kPhCount] = {}, flops[kPhCount] = {};
  long count[kPhCount] = {};
  hipEvent_t take();
  void begin(hipStream_t s, int phase, hipEvent_t* a);
  void end(hipStream_t s, int phase, hipEvent_t a, double bytes, double flops);
  int resolve();  // waits for the recorded pairs and accumulates
  void reset();
  ~PhaseTimer();
};
}  // namespace eamd

// The opaque handle of the C-ABI.
struct eamd_handle {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  int num_cus = 256;
  unsigned panel_epoch = 0;  // monotonic epoch of the panel grid exchange
  int trsv_epoch = 0;   // monotonic epoch of the triangular-solve chain
  eamd::Workspace ws;
  eamd::PhaseTimer prof;
  // look-ahead LU: a stream for the panels (on a reserved set of CUs) and
  // one for the trailing update (the other CUs), both at the default
  // priority, and the events that order them; pstream is null when no CUs
  // are reserved (the panels then run on the handle's stream)
  hipStream_t side = nullptr;
  hipStream_t pstream = nullptr;
  hipEvent_t la_ev[6] = {};
};

// Wraps one launch (or a short sequence) in a phase timer when enabled.
struct PhaseScope {
  eamd_handle* h;
  int phase;
  double bytes, flops;
  hipEvent_t a = nullptr;
  PhaseScope(eamd_handle* h_, int ph, double by = 0, double fl = 0)
      : h(h_), phase(ph), bytes(by), flops(fl) {
    if (h->prof.enabled) h->prof.begin(h->stream, phase, &a);
  }
  ~PhaseScope() {
    if (h->prof.enabled && a) h->prof.end(h->stream, phase, a, bytes, flops);
  }
};

namespace eamd {
// Workspace slot ids
enum WsSlot : int {
  kWsWork = 0,     // working matrix copy (row-major LU, FP32 copy, ...)
  kWsPanelSync,    // grid-barrier flags and candidate rows of the panel kernel
  kWsTmp,          // strip temporaries (U12 / W21)
  kWsInv,          // inverted diagonal blocks
  kWsPiv,          // device pivot vectors
  kWsVec,          // residual / correction vectors
  kWsGmres,        // Krylov basis
  kWsSplit,        // FP16 hi/lo split panels
  kWsStatus,       // device status words
  kWsBatch,        // batched scheduling arrays
  kWsAux,
  kWsAux2,
  kWsFactor32,     // FP32 factors kept for refinement
  kWsTsl,          // tournament-pivoting candidates and saved rows
  kWsPackL,        // packed panel columns (A21) of two consecutive panels
  kWsPackU,        // packed U rows (A12) of the next panel and of the rest
  kWsConv,         // FP64 copies of the FP32 single-matrix entry points' operands
};

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace eamd
