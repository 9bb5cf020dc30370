// Host-side context of the gfx950 prover: device, stream, the round-down
// kernel module and its functions, and error bookkeeping.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
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
  hipFunction_t f_rsb_verdict = nullptr;
  hipFunction_t f_rsb_masks = nullptr;
  hipFunction_t f_c2_eval_f32pre = nullptr;  // the batch with criterion 0's FP32 pre-filter
  hipFunction_t f_zero_u64 = nullptr;
  hipFunction_t f_c2_deferred = nullptr;     // the batch's second passes
  hipFunction_t f_c2_r1 = nullptr;
  void* noalloc = nullptr;                   // this device's static batch scratch (noalloc.hip)
  std::string noalloc_err;                   // why noalloc is null, if it is
  std::unordered_map<std::string, hipFunction_t> fns;  // kernels looked up by name on first use
  void* d_konst = nullptr;
  unsigned long long* h_counts = nullptr;  // pinned host mirror of level counters
  void* d_counts = nullptr;                // device level counters (64 x u64)
  cchost::Scratch* scratch = nullptr;       // level buffers of the enumeration drivers
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::atomic<int> interrupt{0};           // cc_interrupt's flag (any thread)
  // Piece counts and generation grid of the two proofs (cc_tr_prove_shape,
  // cc_rsb_prove_shape); the reference's constants unless a caller changes
  // them (tight_rectangle.cu:44-62, rectangle_size_bound.cu:36-62).
  int tr_init_counts[4] = {128, 128, 128, 128}, tr_gen_counts[4] = {128, 128, 128, 64};
  int rsb_init_counts[4] = {64, 64, 64, 64}, rsb_gen_counts[4] = {16, 16, 16, 32};
  int rsb_gen_grid[3] = {64, 32, 32};
  int enum_form = 0;  // CC_ENUM_* of both proofs' enumerations (cc_set_enumeration_form)
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

// cc_interrupt's flag.  An entry point that polls it clears it when it
// starts, as the reference's start_interruptible_computation clears its
// SIGINT flag (P/interrupt.cpp:38-41): an interrupt stops the call that is
// running when it arrives, never a later one.
inline void clear_interrupt(cc_ctx* ctx) { ctx->interrupt.store(0, std::memory_order_relaxed); }
inline bool interrupted(const cc_ctx* ctx) { return ctx->interrupt.load(std::memory_order_relaxed) != 0; }

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

// Scratch of one batch launch of cc_eval_batch_async (n boxes): the
// deferred-box counter (first 256 B), the deferred boxes' rows (8 x n
// doubles) and indices (n), criterion 0's per-wave staging rows (5 x ns
// doubles, ns = n rounded up to the 256-lane block) with their indices (ns)
// and one count per wave (ns / 64).  Nothing in it needs to start at zero
// except the counter, which the launch clears with a kernel.
struct C2Scratch {
  unsigned long long* cnt;
  double* dstage;
  uint32_t* list;
  double* r1_stage;
  uint32_t* r1_id;
  uint32_t* r1_cnt;
};
constexpr size_t c2_round8(size_t b) { return (b + 7) / 8 * 8; }
constexpr size_t c2_lanes(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t c2_scratch_bytes(size_t n) {
  return 256 + 8 * 8 * n + c2_round8(4 * n) + 5 * 8 * c2_lanes(n) + 4 * c2_lanes(n) + 4 * (c2_lanes(n) / 64);
}
inline C2Scratch c2_scratch_at(void* base, size_t n) {
  char* p = static_cast<char*>(base);
  C2Scratch s;
  s.cnt = reinterpret_cast<unsigned long long*>(p);
  s.dstage = reinterpret_cast<double*>(p + 256);
  s.list = reinterpret_cast<uint32_t*>(p + 256 + 8 * 8 * n);
  s.r1_stage = reinterpret_cast<double*>(p + 256 + 8 * 8 * n + c2_round8(4 * n));
  s.r1_id = reinterpret_cast<uint32_t*>(s.r1_stage + 5 * c2_lanes(n));
  s.r1_cnt = s.r1_id + c2_lanes(n);
  return s;
}
// Validates the arguments of a batch and enqueues its launches with scratch
// `s` (laid out by c2_scratch_at for n); allocates nothing.
int c2_batch_launch(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                    uint32_t* d_mask, uint8_t* d_first, const C2Scratch& s);
// Argument checks shared by the batch entry points (CC_OK when the batch may
// be launched; n == 0 is valid and launches nothing).
int c2_batch_check(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                   const uint32_t* d_mask, const uint8_t* d_first);
// Uploads the inference weights of a ranker (ml.cpp) and returns the device
// blocks ccml_rank takes: feature normalisation, BF16 weights, FP32 biases.
int ml_upload_inference(cc_ctx* ctx, const cc_ranker* model, const void** norm, const void** wb, const void** bias);

// Looks up this device's static batch scratch (once per device, the device
// being current) and stores it in ctx->noalloc; on failure ctx->noalloc
// stays null, the reason goes to ctx->noalloc_err, and only the entry point
// that needs the array reports it.
int noalloc_setup(cc_ctx* ctx);

// Level buffers shared by the enumeration and branch-and-bound drivers.
struct Scratch {
  DevBuf cells, aux, s1, s2, s3, leaves, next, verdict;
  DevBuf tr_piece, tr_buckets;          // TR leaves' r1 piece (4 B each); per-(cell, piece) hulls
  DevBuf c2_defer;                      // batch kernel scratch (C2Scratch layout)
  DevBuf enum_list, enum_b;             // RSB enumeration: non-empty leaves, leaves for C14 (8 B each)
  DevBuf log;                           // records of cc_rsb_bnb_log
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
