// The batch entry point that allocates nothing (cc_eval_batch_noalloc_async;
// no counterpart in the reference).  Its scratch is one file-scope device
// array per GPU, held until the process ends.  This unit is compiled by
// hipcc as ordinary HIP code, so the runtime registers the array with every
// device; the kernels themselves live in the round-down code object and
// take the array's device address as an argument, like any other buffer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "runtime.hpp"

namespace {

constexpr size_t kScratchBytes = cchost::c2_scratch_bytes(CC_NOALLOC_MAX_BOXES);
static_assert(kScratchBytes == CC_NOALLOC_SCRATCH_BYTES, "header and layout disagree on the scratch size");

// Device memory of the batch scratch (C2Scratch layout for up to
// CC_NOALLOC_MAX_BOXES boxes).  In host code `scratch` names a host-side
// stand-in, never the device memory: only the address from
// hipGetSymbolAddress, kept per device below, is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];  // per device: the address of its copy of `scratch`

// The current device's address of the array.  Defined here, where `scratch`
// names the array: inside namespace cchost the same name is the function
// cchost::scratch(cc_ctx*).
hipError_t scratch_address(void** p) { return hipGetSymbolAddress(p, HIP_SYMBOL(scratch)); }

}  // namespace

namespace cchost {

int noalloc_setup(cc_ctx* ctx) {
  ctx->noalloc = nullptr;
  if (ctx->device < 0 || ctx->device >= kMaxDevices) {
    ctx->noalloc_err = "device index out of range";
    return CC_E_NODEVICE;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[ctx->device]) {
    // cc_init made ctx->device current: the address is that device's copy.
    void* p = nullptr;
    const hipError_t e = scratch_address(&p);
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      ctx->noalloc_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return CC_E_HIP;
    }
    g_ptr[ctx->device] = p;
  }
  ctx->noalloc = g_ptr[ctx->device];
  return CC_OK;
}

}  // namespace cchost

extern "C" int cc_eval_batch_noalloc_async(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n,
                                           size_t ld, uint32_t* d_verdict_mask, uint8_t* d_first_ok) {
  using namespace cchost;
  if (int rc = c2_batch_check(ctx, prover, precision, d_soa, n, ld, d_verdict_mask, d_first_ok)) return rc;
  if (n > CC_NOALLOC_MAX_BOXES)
    return fail(ctx, CC_E_ARG, "cc_eval_batch_noalloc_async: n exceeds CC_NOALLOC_MAX_BOXES (262144 boxes)");
  if (n == 0) return CC_OK;
  if (!ctx->noalloc) return fail(ctx, CC_E_HIP, "no static batch scratch on this device: " + ctx->noalloc_err);
  return c2_batch_launch(ctx, prover, precision, d_soa, n, ld, d_verdict_mask, d_first_ok,
                         c2_scratch_at(ctx->noalloc, n));
}
