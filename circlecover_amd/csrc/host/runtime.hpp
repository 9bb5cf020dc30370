// Host-side context of the gfx950 prover: device, stream, the round-down
// kernel module and its functions, and error bookkeeping.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>

#include "circlecover_amd.h"

// Device-constant block written by ccrd_init_konst (must cover sizeof(Konst)
// of the device code, checked by the init kernel's static_assert twin).
static const size_t kKonstBytes = 256;

namespace cchost {
struct Scratch;
}

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
  std::unordered_map<std::string, hipFunction_t> fns;  // kernels looked up by name on first use
  void* d_konst = nullptr;
  unsigned long long* h_counts = nullptr;  // pinned host mirror of level counters
  void* d_counts = nullptr;                // device level counters (64 x u64)
  cchost::Scratch* scratch = nullptr;       // level buffers of the enumeration drivers
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

// Frees the context's level buffers (scratch.cpp).
void release_scratch(cc_ctx* ctx);

// Kernel `name` of the context's module (cached after the first lookup).
hipError_t kernel(cc_ctx* ctx, const char* name, hipFunction_t* f);

// A device allocation that only grows (contents are not preserved).
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void swap(DevBuf& o) {
    void* t = p;
    p = o.p;
    o.p = t;
    size_t c = cap;
    cap = o.cap;
    o.cap = c;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    size_t want = bytes < 4096 ? 4096 : bytes + bytes / 4;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Launch helper over hipModuleLaunchKernel with a packed argument list.
template <typename... A>
hipError_t launch(hipFunction_t f, dim3 grid, dim3 block, size_t shmem, hipStream_t s, A... args) {
  // The dispatch packet holds the grid in work-items as 32-bit numbers per
  // dimension: a larger grid would wrap and silently run fewer work-items.
  if ((uint64_t)grid.x * block.x > 0xffffffffull || (uint64_t)grid.y * block.y > 0xffffffffull ||
      (uint64_t)grid.z * block.z > 0xffffffffull)
    return hipErrorInvalidConfiguration;
  void* argv[] = {(void*)&args...};
  return hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y, block.z, (unsigned)shmem, s, argv,
                               nullptr);
}

// Level buffers shared by the enumeration and branch-and-bound drivers.
struct Scratch {
  DevBuf cells, aux, s1, s2, s3, leaves, next, verdict;
  DevBuf tr_piece, tr_buckets;          // TR leaves' r1 piece (4 B each); per-(cell, piece) hulls
  DevBuf c2_defer;                      // batch kernel: deferred-box counter (first 256 B), data, list
  DevBuf enum_list, enum_b;             // RSB enumeration: non-empty leaves, leaves for C14 (8 B each)
  DevBuf c2_r1;                         // batch kernel: per-wave staging of the boxes criterion 0's
                                        // first strategy leaves open
  void* fast = nullptr;                 // batched branch-and-bound state (rsb_fast.cpp)
  void (*fast_free)(void*) = nullptr;
  void* ml = nullptr;                   // ranker buffers (ml.cpp)
  void (*ml_free)(void*) = nullptr;
  ~Scratch() {
    if (fast_free) fast_free(fast);
    if (ml_free) ml_free(ml);
  }
};
Scratch& scratch(cc_ctx* ctx);

}  // namespace cchost
