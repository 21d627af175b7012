// batched_lu.hip -- placeholder until the batched LU kernels land.
#include "common.h"
#include "kernels.h"
namespace eamd {
int dgesv_ir_batched_impl(eamd_handle* h, int64_t count, const int64_t* offA, const int32_t* n,
                          const double* A, const int64_t* offB, int64_t nrhs, const double* B,
                          double* X, int32_t* info, int32_t* iters) {
  set_error("batched LU not built yet");
  return kErrUnsupported;
}
}  // namespace eamd
