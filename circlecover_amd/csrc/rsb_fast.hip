// Batched rectangle_size_bound branch-and-bound (gfx950, round-down mode).
//
// Same verdicts and the same cells as the level-synchronous kernel of
// rsb_bnb.hip, organised for the hardware:
//   * persistent grids: every kernel is launched with a fixed number of
//     workgroups that pull chunks of `cpw` cells per wave from a device
//     work counter until the level is drained, so a few expensive cells no
//     longer hold a whole launch's tail;
//   * two stages per depth: stage A runs every step before C14; the cells it
//     leaves undecided are compacted into a list that stage B runs C14 on,
//     so the waves of stage B hold only cells that need the expensive step;
//   * every count lives in device memory (struct fast_ctrl), so the host
//     loop over depths needs no synchronisation and is captured once as a
//     hipGraph;
//   * exact early exits of bisections whose test already fails at their
//     smallest possible argument (CC_EXACT_SHORTCUTS, see three_height).
#define CC_EXACT_SHORTCUTS 1
#include "dev/queue.hpp"
#include "dev/rsb_verdict.hpp"

using namespace ccd;

namespace {

constexpr int kMaxDepth = 64;

struct rec112 {
  double v[14];
};

}  // namespace

// Device-resident bookkeeping of one run.  Zeroed by ccrd_fast_reset.
struct fast_ctrl {
  unsigned long long n0;                    // cells at depth 0
  unsigned long long pairs[kMaxDepth + 1];  // split cells of depth d-1 (2 cells each at depth d)
  unsigned long long nb[kMaxDepth + 1];     // cells of depth d that reach C14
  unsigned long long head_a[kMaxDepth + 1]; // work counters
  unsigned long long head_b[kMaxDepth + 1];
  unsigned long long unproven;              // cells unproven at the depth limit
  unsigned long long trap_cells;
  unsigned long long overflow;              // a queue append exceeded its capacity
  unsigned long long hist[21];              // verdict v -> [v+1]
};

CCD cell7 load7(const rec112* p) { return load_cell7(p->v); }
CCD void store7(rec112* out, const cell7& c) {
  rec112 r;
  r.v[0] = c.la.lo;
  r.v[1] = c.la.hi;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    r.v[2 + 2 * k] = c.r[k].lo;
    r.v[3 + 2 * k] = c.r[k].hi;
  }
  *out = r;
}
CCD cell7 half7(cell7 c, int h) {
  int best = 0;
  double bw = rn_span_div(c.la.hi, c.la.lo, 1.0);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double w = rn_span_div(c.r[k].hi, c.r[k].lo, 1.0);
    if (w > bw) {
      bw = w;
      best = k + 1;
    }
  }
  if (best == 0) {
    c.la = subinterval(c.la, h, 2);
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (best == k + 1) c.r[k] = subinterval(c.r[k], h, 2);
  }
  return c;
}

// Next chunk of `cpw` items for this wave: one atomic by the first active
// lane, broadcast to the wave.
CCD uint64_t next_chunk(unsigned long long* head, unsigned cpw) {
  const uint64_t active = __ballot(1);
  const int leader = __ffsll((unsigned long long)active) - 1;
  unsigned long long base = 0;
  if (__lane_id() == leader) base = atomicAdd(head, (unsigned long long)cpw);
  return __shfl(base, leader);
}

extern "C" __global__ void ccrd_fast_reset(fast_ctrl* __restrict__ ctrl, unsigned long long n0) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(ctrl);
  const size_t words = sizeof(fast_ctrl) / sizeof(unsigned long long);
  for (size_t i = threadIdx.x; i < words; i += blockDim.x) p[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) ctrl->n0 = n0;
}

// Stage A of depth d: every step before C14 on each cell of `in`.  Proven
// cells are counted; cells that only C14 can still prove go to `listb` as
// (index | trap << 31).
__device__ __forceinline__ void stage_a(const rec112* __restrict__ in, uint64_t cap_in, int d, unsigned cpw,
                                        const Konst* __restrict__ kp, uint32_t* __restrict__ listb, uint64_t cap,
                                        fast_ctrl* __restrict__ ctrl) {
  __shared__ unsigned int bins[21];
  __shared__ unsigned int traps;
  if (threadIdx.x < 21) bins[threadIdx.x] = 0;
  if (threadIdx.x == 0) traps = 0;
  __syncthreads();
  uint64_t n = d == 0 ? ctrl->n0 : 2 * ctrl->pairs[d];
  if (n > cap_in) n = cap_in;  // records past the capacity were never written (overflow is flagged)
  const Konst K = *kp;
  for (;;) {
    const uint64_t base = next_chunk(&ctrl->head_a[d], cpw);
    if (base >= n) break;
    const uint64_t end = base + cpw < n ? base + cpw : n;
    for (uint64_t j0 = base; j0 < end; j0 += 64) {
      const uint64_t j = j0 + __lane_id();
      const bool live = j < end;
      int v = 0;
      Trap tp{0};
      if (live) {
        v = rsb_verdict_before_c14(K, load7(&in[j]), tp);
        if (v >= 0) {
          atomicAdd(&bins[v + 1], 1u);
          if (tp.flag) atomicAdd(&traps, 1u);
        }
      }
      const bool to_b = live && v == -2;
      const uint64_t s = wave_append(to_b, &ctrl->nb[d]);
      if (to_b) {
        if (s < cap) listb[s] = (uint32_t)j | ((uint32_t)tp.flag << 31);
        else atomicOr(&ctrl->overflow, 1ull);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 21 && bins[threadIdx.x]) atomicAdd(&ctrl->hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
  if (threadIdx.x == 0 && traps) atomicAdd(&ctrl->trap_cells, (unsigned long long)traps);
}

// Stage B of depth d: C14 on the listed cells.  A proven cell is counted;
// an unproven cell is split into `next` (depth d+1) or, at the depth limit,
// appended to `unproven`.
__device__ __forceinline__ void stage_b(const rec112* __restrict__ in, const uint32_t* __restrict__ listb,
                                        uint64_t cap_list, int d, int at_limit, unsigned cpw,
                                        const Konst* __restrict__ kp, rec112* __restrict__ next, uint64_t cap_next,
                                        rec112* __restrict__ unproven, uint64_t cap_u,
                                        fast_ctrl* __restrict__ ctrl) {
  __shared__ unsigned int bins[2];
  __shared__ unsigned int traps;
  if (threadIdx.x < 2) bins[threadIdx.x] = 0;
  if (threadIdx.x == 0) traps = 0;
  __syncthreads();
  uint64_t n = ctrl->nb[d];
  if (n > cap_list) n = cap_list;  // entries past the list capacity were never written (overflow is flagged)
  const Konst K = *kp;
  for (;;) {
    const uint64_t base = next_chunk(&ctrl->head_b[d], cpw);
    if (base >= n) break;
    const uint64_t end = base + cpw < n ? base + cpw : n;
    for (uint64_t j0 = base; j0 < end; j0 += 64) {
      const uint64_t j = j0 + __lane_id();
      const bool live = j < end;
      int v = 0;
      cell7 c;
      if (live) {
        const uint32_t e = listb[j];
        Trap tp{(int)(e >> 31)};
        c = load7(&in[e & 0x7fffffffu]);
        v = rsb_verdict_c14(K, c, tp);
        atomicAdd(&bins[v < 0 ? 0 : 1], 1u);
        if (tp.flag) atomicAdd(&traps, 1u);
      }
      const bool split = live && v < 0 && !at_limit;
      const bool keep = live && v < 0 && at_limit;
      const uint64_t ps = wave_append(split, &ctrl->pairs[d + 1]);
      if (split) {
        if (2 * ps + 1 < cap_next) {
          store7(&next[2 * ps], half7(c, 0));
          store7(&next[2 * ps + 1], half7(c, 1));
        } else {
          atomicOr(&ctrl->overflow, 1ull);
        }
      }
      const uint64_t us = wave_append(keep, &ctrl->unproven);
      if (keep) {
        if (us < cap_u) store7(&unproven[us], c);
        else atomicOr(&ctrl->overflow, 1ull);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (bins[0]) atomicAdd(&ctrl->hist[0], (unsigned long long)bins[0]);
    if (bins[1]) atomicAdd(&ctrl->hist[20], (unsigned long long)bins[1]);
    if (traps) atomicAdd(&ctrl->trap_cells, (unsigned long long)traps);
  }
}

// Kernel variants per occupancy target: waves per SIMD W caps the register
// budget at 512/W VGPRs (spills go to scratch); the autotuner picks W.
#define CC_STAGE_KERNELS(W)                                                                                    \
  extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))                 \
  ccrd_fast_stage_a_w##W(const rec112* __restrict__ in, uint64_t cap_in, int d, unsigned cpw,                  \
                         const Konst* __restrict__ kp, uint32_t* __restrict__ listb, uint64_t cap,            \
                         fast_ctrl* __restrict__ ctrl) {                                                      \
    stage_a(in, cap_in, d, cpw, kp, listb, cap, ctrl);                                                         \
  }                                                                                                           \
  extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))                 \
  ccrd_fast_stage_b_w##W(const rec112* __restrict__ in, const uint32_t* __restrict__ listb, uint64_t cap_list, \
                         int d, int at_limit, unsigned cpw, const Konst* __restrict__ kp,                      \
                         rec112* __restrict__ next, uint64_t cap_next, rec112* __restrict__ unproven,         \
                         uint64_t cap_u, fast_ctrl* __restrict__ ctrl) {                                      \
    stage_b(in, listb, cap_list, d, at_limit, cpw, kp, next, cap_next, unproven, cap_u, ctrl);                \
  }

CC_STAGE_KERNELS(1)
CC_STAGE_KERNELS(2)
CC_STAGE_KERNELS(4)
