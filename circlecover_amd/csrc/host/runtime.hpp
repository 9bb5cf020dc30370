// Host-side context of the gfx950 prover: device, stream, the round-down
// kernel module and its functions, and error bookkeeping.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "circlecover_amd.h"

// Device-constant block written by ccrd_init_konst (must cover sizeof(Konst)
// of the device code, checked by the init kernel's static_assert twin).
static const size_t kKonstBytes = 256;

struct cc_ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  hipModule_t mod = nullptr;
  hipFunction_t f_init_konst = nullptr;
  hipFunction_t f_c2_eval = nullptr;
  hipFunction_t f_prim = nullptr;
  hipFunction_t f_c2_eval_f32 = nullptr;
  hipFunction_t f_prim_rn = nullptr;
  hipFunction_t f_trace_strip = nullptr;
  hipFunction_t f_rsb_verdict = nullptr;
  hipFunction_t f_rsb_masks = nullptr;
  void* d_konst = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  volatile int interrupt = 0;
  std::string err;
};

namespace cchost {

int fail(cc_ctx* ctx, int code, const std::string& msg);
int hip_fail(cc_ctx* ctx, hipError_t e, const char* what);
#define CC_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ::cchost::hip_fail((ctx), e_, #call); \
  } while (0)

// Launch helper over hipModuleLaunchKernel with a packed argument list.
template <typename... A>
hipError_t launch(hipFunction_t f, dim3 grid, dim3 block, size_t shmem, hipStream_t s, A... args) {
  void* argv[] = {(void*)&args...};
  return hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y, block.z, (unsigned)shmem, s, argv,
                               nullptr);
}

}  // namespace cchost
