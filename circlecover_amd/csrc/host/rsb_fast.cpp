// Batched branch-and-bound driver: persistent stage kernels (rsb_fast.hip)
// with device-resident counts, the whole depth loop captured once as a
// hipGraph, and a small autotuner for the launch shape.
#include <chrono>
#include <cstring>
#include <vector>

#include "runtime.hpp"

using namespace cchost;

namespace {

constexpr size_t kRec = 112;
constexpr int kMaxDepth = 64;

// Host mirror of the device fast_ctrl block (rsb_fast.hip).
struct FastCtrl {
  unsigned long long n0;
  unsigned long long pairs[kMaxDepth + 1];
  unsigned long long nb[kMaxDepth + 1];
  unsigned long long head_a[kMaxDepth + 1];
  unsigned long long head_b[kMaxDepth + 1];
  unsigned long long unproven;
  unsigned long long trap_cells;
  unsigned long long overflow;
  unsigned long long hist[21];
  unsigned long long nb2[kMaxDepth + 1];
  unsigned long long head_b2[kMaxDepth + 1];
  unsigned long long head_b3[kMaxDepth + 1];
};
constexpr size_t kB2Rec = 48;  // struct b2rec of rsb_fast.hip
static_assert(sizeof(FastCtrl) == 8 * (1 + 7 * (kMaxDepth + 1) + 3 + 21), "mirror of fast_ctrl (rsb_fast.hip)");

struct FastState {
  DevBuf lvl[2], listb, listb2, unproven, ctrl;
  // learned order: class per cell, cells grouped by class, the grouping's
  // counts and offsets, per-depth cell counts, class -> step table
  DevBuf top1, order, bcounts, boffsets, bstart, n32, class_step;
  bool g_learned = false;
  FastCtrl* h_ctrl = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  // what the captured graph was built for
  size_t g_ncells = 0, g_cap = 0;
  int g_depth = -1;
  cc_fast_params g_params{};
  const void* g_src = nullptr;
  cc_fast_params tuned{};
  bool have_tuned = false;
  ~FastState() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
  }
};

FastState* state(cc_ctx* ctx) {
  Scratch& S = scratch(ctx);
  if (!S.fast) {
    S.fast = new FastState();
    S.fast_free = [](void* p) { delete static_cast<FastState*>(p); };
  }
  return static_cast<FastState*>(S.fast);
}

// Fastest shape measured on the 256x64x64 grid (one MI355X): 0.54 s.
const cc_fast_params kDefault = {/*cpw_a=*/64, /*cpw_b=*/64, /*blocks_per_cu=*/4, /*use_graph=*/1,
                                  /*waves_per_simd=*/4};

bool same(const cc_fast_params& a, const cc_fast_params& b) { return std::memcmp(&a, &b, sizeof a) == 0; }

int cu_count(cc_ctx* ctx, int* n) {
  CC_HIP(ctx, hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, ctx->device));
  return CC_OK;
}

struct Fns {
  hipFunction_t reset, a, b1, b2, b3;
  hipFunction_t ar, ncount, rank, bcount, bprefix, bscatter;  // learned order
  int cus;
};

// Grouping of a level's cells by predicted class: blocks of the counting
// sort (at most 1024, one prefix workgroup) and the ranker's persistent
// grid (four 32-cell tiles per workgroup round).
constexpr unsigned kGroupBlocks = 256, kRankBlocks = 512;

// Device blocks of a learned run's ranker.
struct RankerDev {
  const void *norm = nullptr, *wb = nullptr, *bias = nullptr;
};

int resolve(cc_ctx* ctx, const cc_fast_params& p, bool learned, Fns* f) {
  const int w = p.waves_per_simd;
  const char* a = w == 1 ? "ccrd_fast_stage_a_w1" : w == 2 ? "ccrd_fast_stage_a_w2" : "ccrd_fast_stage_a_w4";
  const char* b1 = w == 1 ? "ccrd_fast_stage_b1_w1" : w == 2 ? "ccrd_fast_stage_b1_w2" : "ccrd_fast_stage_b1_w4";
  const char* b2 = w == 1 ? "ccrd_fast_stage_b2_w1" : w == 2 ? "ccrd_fast_stage_b2_w2" : "ccrd_fast_stage_b2_w4";
  CC_HIP(ctx, kernel(ctx, "ccrd_fast_reset", &f->reset));
  CC_HIP(ctx, kernel(ctx, a, &f->a));
  CC_HIP(ctx, kernel(ctx, b1, &f->b1));
  CC_HIP(ctx, kernel(ctx, b2, &f->b2));
  CC_HIP(ctx, kernel(ctx, "ccrd_fast_stage_b3", &f->b3));
  if (learned) {
    const char* ar = w == 1 ? "ccrd_fast_stage_ar_w1" : w == 2 ? "ccrd_fast_stage_ar_w2" : "ccrd_fast_stage_ar_w4";
    CC_HIP(ctx, kernel(ctx, ar, &f->ar));
    CC_HIP(ctx, kernel(ctx, "cc_fast_ncount", &f->ncount));
    CC_HIP(ctx, kernel(ctx, "ccml_rank", &f->rank));
    CC_HIP(ctx, kernel(ctx, "ccml_bucket_count", &f->bcount));
    CC_HIP(ctx, kernel(ctx, "ccml_bucket_prefix", &f->bprefix));
    CC_HIP(ctx, kernel(ctx, "ccml_bucket_scatter", &f->bscatter));
  }
  return cu_count(ctx, &f->cus);
}

// Enqueues the whole run (reset, then stages A, B1, B2, B3 per depth) on
// the stream.
// With a ranker, each depth first classifies its cells (ccml_rank on the
// cell records), groups them by class (ccml_bucket_*) and runs the ranked
// stage A over the grouped order.
int enqueue(cc_ctx* ctx, const Fns& f, FastState& F, size_t ncells, size_t cap, int max_depth,
            const cc_fast_params& p, const RankerDev* rk) {
  const hipFunction_t fr = f.reset, fa = f.a;
  const unsigned grid = (unsigned)(f.cus * p.blocks_per_cu);
  const unsigned long long n0 = ncells;
  CC_HIP(ctx, launch(fr, dim3(1), dim3(256), 0, ctx->stream, F.ctrl.p, n0));
  const uint64_t cap64 = cap;
  for (int d = 0; d <= max_depth; ++d) {
    void* in = F.lvl[d & 1].p;
    void* next = F.lvl[(d + 1) & 1].p;
    const int at_limit = d == max_depth;
    const unsigned cpw_a = (unsigned)p.cpw_a, cpw_b = (unsigned)p.cpw_b;
    if (rk) {
      uint32_t* n32 = static_cast<uint32_t*>(F.n32.p) + d;
      const uint64_t one = 1, fourteen = 14, zero = 0;
      const void* none = nullptr;
      CC_HIP(ctx, launch(f.ncount, dim3(1), dim3(64), 0, ctx->stream, F.ctrl.p, d, cap64, F.n32.p));
      CC_HIP(ctx, launch(f.rank, dim3(kRankBlocks), dim3(256), 0, ctx->stream, in, cap64, one, fourteen, none,
                         (const uint32_t*)n32, rk->norm, rk->wb, rk->bias, F.top1.p, none));
      CC_HIP(ctx, launch(f.bcount, dim3(kGroupBlocks), dim3(256), 0, ctx->stream, F.top1.p, cap64,
                         (const uint32_t*)n32, F.bcounts.p));
      CC_HIP(ctx, launch(f.bprefix, dim3(1), dim3(1024), 0, ctx->stream, F.bcounts.p, kGroupBlocks, F.boffsets.p,
                         F.bstart.p));
      CC_HIP(ctx, launch(f.bscatter, dim3(kGroupBlocks), dim3(256), 0, ctx->stream, F.top1.p, cap64,
                         (const uint32_t*)n32, F.boffsets.p, none, F.order.p, none, zero, none));
      CC_HIP(ctx, launch(f.ar, dim3(grid), dim3(256), 0, ctx->stream, in, cap64, d, cpw_a, ctx->d_konst, F.listb.p,
                         cap64, F.ctrl.p, F.order.p, F.top1.p, F.class_step.p));
    } else {
      CC_HIP(ctx, launch(fa, dim3(grid), dim3(256), 0, ctx->stream, in, cap64, d, cpw_a, ctx->d_konst, F.listb.p,
                         cap64, F.ctrl.p));
    }
    CC_HIP(ctx, launch(f.b1, dim3(grid), dim3(256), 0, ctx->stream, in, F.listb.p, cap64, d, at_limit, cpw_b,
                       ctx->d_konst, next, cap64, F.unproven.p, cap64, F.listb2.p, cap64, F.ctrl.p));
    CC_HIP(ctx, launch(f.b2, dim3(grid), dim3(256), 0, ctx->stream, in, F.listb2.p, cap64, d, cpw_b, ctx->d_konst,
                       F.ctrl.p));
    CC_HIP(ctx, launch(f.b3, dim3(grid), dim3(256), 0, ctx->stream, in, F.listb2.p, cap64, d, at_limit, cpw_b, next,
                       cap64, F.unproven.p, cap64, F.ctrl.p));
  }
  return CC_OK;
}


// The batched branch-and-bound; ranker != NULL gives the learned order.
int bnb_batched(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, const cc_fast_params* params,
                const cc_rsb_ranker* ranker, cc_rsb_cell* out, size_t cap_out, size_t* n_out, cc_bnb_stats* stats) {
  if (!ctx || !n_out || (ncells && !cells) || (cap_out && !out) || max_depth < 0 || max_depth >= CC_BNB_MAX_DEPTH)
    return CC_E_ARG;
  FastState& F = *state(ctx);
  cc_fast_params p = params ? *params : (F.have_tuned ? F.tuned : kDefault);
  if (p.cpw_a < 64 || p.cpw_b < 64 || p.cpw_a % 64 || p.cpw_b % 64 || p.blocks_per_cu < 1 || p.blocks_per_cu > 16)
    return fail(ctx, CC_E_ARG, "cells per wave must be positive multiples of 64, 1..16 blocks per CU");
  if (p.waves_per_simd != 1 && p.waves_per_simd != 2 && p.waves_per_simd != 4)
    return fail(ctx, CC_E_ARG, "waves per SIMD must be 1, 2 or 4");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  // Queue capacity per level: twice the initial cells, at least 2^20 (the
  // widest level of the 256x64x64 grid holds 0.6x its initial cells).
  size_t cap = 2 * ncells > ((size_t)1 << 20) ? 2 * ncells : ((size_t)1 << 20);
  if (cap > 0x7fffffffull) cap = 0x7fffffffull;  // list entries carry the index in 31 bits
  if (ncells > cap) return fail(ctx, CC_E_ARG, "too many initial cells");
  CC_HIP(ctx, F.lvl[0].ensure(cap * kRec));
  CC_HIP(ctx, F.lvl[1].ensure(cap * kRec));
  CC_HIP(ctx, F.listb.ensure(cap * sizeof(uint32_t)));
  CC_HIP(ctx, F.listb2.ensure(cap * kB2Rec));
  CC_HIP(ctx, F.unproven.ensure(cap * kRec));
  CC_HIP(ctx, F.ctrl.ensure(sizeof(FastCtrl)));
  if (!F.h_ctrl) CC_HIP(ctx, hipHostMalloc((void**)&F.h_ctrl, sizeof(FastCtrl), hipHostMallocDefault));
  if (ncells) CC_HIP(ctx, hipMemcpyAsync(F.lvl[0].p, cells, ncells * kRec, hipMemcpyHostToDevice, ctx->stream));
  const bool learned = ranker != nullptr;
  RankerDev rk;
  if (learned) {
    for (int c = 0; c < 11; ++c)
      if (ranker->class_step[c] < -1 || ranker->class_step[c] > 18)
        return fail(ctx, CC_E_ARG, "ranker class_step entries must be -1..18 (steps before C14)");
    int steps[16];
    for (int c = 0; c < 16; ++c) steps[c] = c < 11 ? ranker->class_step[c] : -1;
    CC_HIP(ctx, F.top1.ensure(cap + 64));
    CC_HIP(ctx, F.order.ensure(cap * sizeof(uint32_t)));
    CC_HIP(ctx, F.bcounts.ensure(kGroupBlocks * 11 * sizeof(unsigned long long)));
    CC_HIP(ctx, F.boffsets.ensure(kGroupBlocks * 11 * sizeof(unsigned long long)));
    CC_HIP(ctx, F.bstart.ensure(16 * sizeof(uint64_t)));
    CC_HIP(ctx, F.n32.ensure((kMaxDepth + 1) * sizeof(uint32_t)));
    CC_HIP(ctx, F.class_step.ensure(sizeof steps));
    CC_HIP(ctx, hipMemcpyAsync(F.class_step.p, steps, sizeof steps, hipMemcpyHostToDevice, ctx->stream));
    if (int rc = ml_upload_inference(ctx, &ranker->net, &rk.norm, &rk.wb, &rk.bias)) return rc;
  }

  Fns fns;
  int rc = resolve(ctx, p, learned, &fns);
  if (rc) return rc;
  CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  bool graphed = false;
  if (p.use_graph) {
    const bool reuse = F.exec && F.g_ncells == ncells && F.g_cap == cap && F.g_depth == max_depth &&
                       same(F.g_params, p) && F.g_src == F.lvl[0].p && F.g_learned == learned;
    if (!reuse) {
      if (F.exec) (void)hipGraphExecDestroy(F.exec);
      if (F.graph) (void)hipGraphDestroy(F.graph);
      F.exec = nullptr;
      F.graph = nullptr;
      hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        int erc = enqueue(ctx, fns, F, ncells, cap, max_depth, p, learned ? &rk : nullptr);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (erc == CC_OK && e2 == hipSuccess && g) {
          if (hipGraphInstantiate(&F.exec, g, nullptr, nullptr, 0) == hipSuccess) {
            F.graph = g;
            F.g_ncells = ncells;
            F.g_cap = cap;
            F.g_depth = max_depth;
            F.g_params = p;
            F.g_src = F.lvl[0].p;
            F.g_learned = learned;
          } else {
            (void)hipGraphDestroy(g);
            F.exec = nullptr;
          }
        } else if (g) {
          (void)hipGraphDestroy(g);
        }
      }
      (void)hipGetLastError();
    }
    if (F.exec) {
      CC_HIP(ctx, hipGraphLaunch(F.exec, ctx->stream));
      graphed = true;
    }
  }
  if (!graphed && (rc = enqueue(ctx, fns, F, ncells, cap, max_depth, p, learned ? &rk : nullptr))) return rc;
  CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(F.h_ctrl, F.ctrl.p, sizeof(FastCtrl), hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const FastCtrl& C = *F.h_ctrl;
  if (C.overflow) {
    // A level outgrew the queues: redo the run with the level-synchronous
    // driver, whose buffers grow on demand (same result).
    return cc_rsb_bnb(ctx, cells, ncells, max_depth, out, cap_out, n_out, stats);
  }
  float ms = 0.f;
  CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  cc_bnb_stats st;
  std::memset(&st, 0, sizeof st);
  int last = -1;
  for (int d = 0; d <= max_depth; ++d) {
    st.cells_per_depth[d] = d == 0 ? C.n0 : 2 * C.pairs[d];
    if (st.cells_per_depth[d]) last = d;
  }
  st.depth_reached = last;
  for (int v = 0; v < 21; ++v) {
    st.verdict_hist[v] = C.hist[v];
    st.cell_evals += C.hist[v] * (uint64_t)(v == 0 ? 20 : v);
  }
  st.trap_cells = C.trap_cells;
  st.device_ms = ms;
  *n_out = C.unproven;
  const size_t m = C.unproven < cap_out ? C.unproven : cap_out;
  if (m) {
    CC_HIP(ctx, hipMemcpyAsync(out, F.unproven.p, m * kRec, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  if (C.unproven > cap_out) return fail(ctx, CC_E_OVERFLOW, "output capacity too small");
  return CC_OK;
}

}  // namespace

extern "C" {

int cc_rsb_bnb_batched(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth,
                       const cc_fast_params* params, cc_rsb_cell* out, size_t cap_out, size_t* n_out,
                       cc_bnb_stats* stats) {
  return bnb_batched(ctx, cells, ncells, max_depth, params, nullptr, out, cap_out, n_out, stats);
}

int cc_rsb_bnb_learned(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth,
                       const cc_rsb_ranker* ranker, const cc_fast_params* params, cc_rsb_cell* out, size_t cap,
                       size_t* n_out, cc_bnb_stats* stats) {
  if (!ranker) return CC_E_ARG;
  return bnb_batched(ctx, cells, ncells, max_depth, params, ranker, out, cap, n_out, stats);
}

int cc_rsb_autotune(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_fast_params* best,
                    double* best_ms) {
  if (!ctx || (ncells && !cells)) return CC_E_ARG;
  FastState& F = *state(ctx);
  // blocks per CU follows the occupancy target (4 waves per block, 4 SIMDs)
  static const int kCpw[] = {64, 128};
  static const int kWaves[] = {2, 4};
  std::vector<cc_rsb_cell> sink(1 << 12);
  double best_t = 1e300;
  cc_fast_params bp = kDefault;
  for (int a : kCpw)
    for (int w : kWaves)
      for (int extra = 0; extra < 2; ++extra) {
        cc_fast_params p = {a, 64, w * (1 + extra), 1, w};
        double t = 0;
        for (int rep = 0; rep < 2; ++rep) {  // the first run captures the graph
          size_t n = 0;
          cc_bnb_stats st;
          int rc = cc_rsb_bnb_batched(ctx, cells, ncells, max_depth, &p, sink.data(), sink.size(), &n, &st);
          if (rc != CC_OK && rc != CC_E_OVERFLOW) return rc;
          t = st.device_ms;
        }
        if (t < best_t) {
          best_t = t;
          bp = p;
        }
      }
  F.tuned = bp;
  F.have_tuned = true;
  if (best) *best = bp;
  if (best_ms) *best_ms = best_t;
  return CC_OK;
}

}  // extern "C"
