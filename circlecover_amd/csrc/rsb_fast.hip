// Batched rectangle_size_bound branch-and-bound (gfx950, round-down mode).
//
// Same verdicts and the same cells as the level-synchronous kernel of
// rsb_bnb.hip, organised for the hardware:
//   * persistent grids: every kernel is launched with a fixed number of
//     workgroups that pull chunks of `cpw` cells per wave from a device
//     work counter until the level is drained, so a few expensive cells no
//     longer hold a whole launch's tail;
//   * stages per depth: stage A runs every step before C14; the cells it
//     leaves undecided are compacted into a list that stage B runs C14 on,
//     so the waves of stage B hold only cells that need the expensive step.
//     C14 itself is split: B1 runs its prefix and the first of its 16 r7
//     pieces on one lane per cell (most cells fail there), B2 runs every
//     further piece of the remaining cells on a lane of its own, B3 combines
//     the pieces' outcomes per cell.  A cell that C14 proves costs 16
//     pieces on one lane otherwise, about 15 times what a failing cell
//     costs, and its wave waits for it;
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
  unsigned long long nb2[kMaxDepth + 1];    // cells of depth d whose C14 pieces 1..15 decide
  unsigned long long head_b2[kMaxDepth + 1];
  unsigned long long head_b3[kMaxDepth + 1];
};

// A cell whose C14 pieces 1..15 decide: its index in the level buffer, the
// trap sites reached so far (stage A, the prefix, piece 0), the C14 state of
// the prefix, and the outcomes of the pieces (bit k), filled in by stage B2.
struct b2rec {
  uint32_t j;
  uint32_t trap;
  uint32_t fail;   // bit k: piece k is not covered
  uint32_t traps;  // bit k: piece k reached a trap site
  double range_lo, range_hi, w_rem_hi, h34;
};

static_assert(sizeof(b2rec) == 48, "b2rec layout (host: kB2Rec)");
static_assert(sizeof(fast_ctrl) == 8 * (1 + 7 * (kMaxDepth + 1) + 3 + 21), "fast_ctrl layout (host: FastCtrl)");

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

// Stage A with the learned order: positions [0, n) of `order` (the level's
// cell indices grouped by predicted class, so a wave mostly holds one
// class); cell j first tries step class_step[cls[j]] (-1: none), then the
// others in the reference's order.  Same outputs as stage_a otherwise.
__device__ __forceinline__ void stage_a_ranked(const rec112* __restrict__ in, uint64_t cap_in, int d, unsigned cpw,
                                               const Konst* __restrict__ kp, uint32_t* __restrict__ listb,
                                               uint64_t cap, fast_ctrl* __restrict__ ctrl,
                                               const uint32_t* __restrict__ order, const uint8_t* __restrict__ cls,
                                               const int* __restrict__ class_step) {
  __shared__ unsigned int bins[21];
  __shared__ unsigned int traps;
  __shared__ int steps[16];
  if (threadIdx.x < 21) bins[threadIdx.x] = 0;
  if (threadIdx.x < 16) steps[threadIdx.x] = threadIdx.x < 11 ? class_step[threadIdx.x] : -1;
  if (threadIdx.x == 0) traps = 0;
  __syncthreads();
  uint64_t n = d == 0 ? ctrl->n0 : 2 * ctrl->pairs[d];
  if (n > cap_in) n = cap_in;
  const Konst K = *kp;
  for (;;) {
    const uint64_t base = next_chunk(&ctrl->head_a[d], cpw);
    if (base >= n) break;
    const uint64_t end = base + cpw < n ? base + cpw : n;
    for (uint64_t p0 = base; p0 < end; p0 += 64) {
      const uint64_t p = p0 + __lane_id();
      const bool live = p < end;
      int v = 0;
      uint32_t j = 0;
      Trap tp{0};
      if (live) {
        j = order[p];
        const uint32_t c = cls[j];
        v = rsb_verdict_before_c14_ranked(K, load7(&in[j]), c < 11 ? steps[c] : -1, tp);
        if (v >= 0) {
          atomicAdd(&bins[v + 1], 1u);
          if (tp.flag) atomicAdd(&traps, 1u);
        }
      }
      const bool to_b = live && v == -2;
      const uint64_t s = wave_append(to_b, &ctrl->nb[d]);
      if (to_b) {
        if (s < cap) listb[s] = j | ((uint32_t)tp.flag << 31);
        else atomicOr(&ctrl->overflow, 1ull);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 21 && bins[threadIdx.x]) atomicAdd(&ctrl->hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
  if (threadIdx.x == 0 && traps) atomicAdd(&ctrl->trap_cells, (unsigned long long)traps);
}

// The number of cells at depth d (capped at the level capacity) as the
// 32-bit count the ranking and grouping kernels read.
extern "C" __global__ void cc_fast_ncount(const fast_ctrl* __restrict__ ctrl, int d, uint64_t cap,
                                          uint32_t* __restrict__ n32) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const uint64_t n = d == 0 ? ctrl->n0 : 2 * ctrl->pairs[d];
    n32[d] = (uint32_t)(n < cap ? n : cap);
  }
}

// An unproven cell of depth d (wave-uniform call): split into `next`
// (depth d+1), or, at the depth limit, appended to `unproven`.
CCD void put_unproven(const cell7& c, bool split, bool keep, int d, rec112* __restrict__ next, uint64_t cap_next,
                      rec112* __restrict__ unproven, uint64_t cap_u, fast_ctrl* __restrict__ ctrl) {
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

// Per-workgroup tallies of stage B, added to the control block at the end.
struct b_tally {
  unsigned int unproven, proven, traps;
};
CCD void tally_flush(b_tally* t, fast_ctrl* __restrict__ ctrl) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (t->unproven) atomicAdd(&ctrl->hist[0], (unsigned long long)t->unproven);
    if (t->proven) atomicAdd(&ctrl->hist[20], (unsigned long long)t->proven);
    if (t->traps) atomicAdd(&ctrl->trap_cells, (unsigned long long)t->traps);
  }
}

// Stage B1 of depth d: C14's prefix and piece 0 on the listed cells.  A
// cell the prefix proves is counted; a cell whose piece 0 is not covered is
// unproven (split or kept); the others go to `listb2` for their pieces
// 1..15 (stage B2).
__device__ __forceinline__ void stage_b1(const rec112* __restrict__ in, const uint32_t* __restrict__ listb,
                                         uint64_t cap_list, int d, int at_limit, unsigned cpw,
                                         const Konst* __restrict__ kp, rec112* __restrict__ next, uint64_t cap_next,
                                         rec112* __restrict__ unproven, uint64_t cap_u, b2rec* __restrict__ listb2,
                                         uint64_t cap_b2, fast_ctrl* __restrict__ ctrl) {
  __shared__ b_tally t;
  if (threadIdx.x == 0) t = b_tally{0, 0, 0};
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
      rsbc::c14_state s;
      uint32_t idx = 0;
      Trap tp{0};
      if (live) {
        const uint32_t e = listb[j];
        idx = e & 0x7fffffffu;
        tp.flag = (int)(e >> 31);
        c = load7(&in[idx]);
        v = rsb_c14_head(K, c, tp, s);
        if (v != -3) {
          atomicAdd(v < 0 ? &t.unproven : &t.proven, 1u);
          if (tp.flag) atomicAdd(&t.traps, 1u);
        }
      }
      const bool to_b2 = live && v == -3;
      const uint64_t s2 = wave_append(to_b2, &ctrl->nb2[d]);
      if (to_b2) {
        if (s2 < cap_b2) {
          b2rec r;
          r.j = idx;
          r.trap = (uint32_t)tp.flag;
          r.fail = 0u;
          r.traps = 0u;
          r.range_lo = s.range.lo;
          r.range_hi = s.range.hi;
          r.w_rem_hi = s.w_rem_hi;
          r.h34 = s.h34;
          listb2[s2] = r;
        } else {
          atomicOr(&ctrl->overflow, 1ull);
        }
      }
      put_unproven(c, live && v == -1 && !at_limit, live && v == -1 && at_limit, d, next, cap_next, unproven,
                   cap_u, ctrl);
    }
  }
  tally_flush(&t, ctrl);
}

// Stage B2 of depth d: one lane per (listed cell, C14 piece 1..15).  The
// outcome goes to the cell's b2rec (bit k of fail / traps); every piece is
// evaluated, and stage B3 keeps the trap sites of the pieces up to the first
// uncovered one, as the reference's loop reaches them.
__device__ __forceinline__ void stage_b2(const rec112* __restrict__ in, b2rec* __restrict__ listb2, uint64_t cap_b2,
                                         int d, unsigned cpw, const Konst* __restrict__ kp,
                                         fast_ctrl* __restrict__ ctrl) {
  uint64_t n2 = ctrl->nb2[d];
  if (n2 > cap_b2) n2 = cap_b2;
  const uint64_t items = 15 * n2;
  const Konst K = *kp;
  for (;;) {
    const uint64_t base = next_chunk(&ctrl->head_b2[d], cpw);
    if (base >= items) break;
    const uint64_t end = base + cpw < items ? base + cpw : items;
    for (uint64_t j0 = base; j0 < end; j0 += 64) {
      const uint64_t i = j0 + __lane_id();
      if (i >= end) continue;
      const uint64_t r = i / 15;
      const int k = 1 + (int)(i - 15 * r);
      const b2rec rec = listb2[r];
      rsbc::c14_state s;
      s.range = iv{rec.range_lo, rec.range_hi};
      s.w_rem_hi = rec.w_rem_hi;
      s.h34 = rec.h34;
      Trap tp{0};
      const bool ok = rsb_c14_piece(K, load7(&in[rec.j]), s, k, tp);
      if (!ok) atomicOr(&listb2[r].fail, 1u << k);
      if (tp.flag) atomicOr(&listb2[r].traps, 1u << k);
    }
  }
}

// Stage B3 of depth d: the verdict of each cell of `listb2` from its
// pieces: proven if every piece is covered, else unproven (split or kept);
// its trap sites are those up to its first uncovered piece.
__device__ __forceinline__ void stage_b3(const rec112* __restrict__ in, const b2rec* __restrict__ listb2,
                                         uint64_t cap_b2, int d, int at_limit, unsigned cpw,
                                         rec112* __restrict__ next, uint64_t cap_next, rec112* __restrict__ unproven,
                                         uint64_t cap_u, fast_ctrl* __restrict__ ctrl) {
  __shared__ b_tally t;
  if (threadIdx.x == 0) t = b_tally{0, 0, 0};
  __syncthreads();
  uint64_t n2 = ctrl->nb2[d];
  if (n2 > cap_b2) n2 = cap_b2;
  for (;;) {
    const uint64_t base = next_chunk(&ctrl->head_b3[d], cpw);
    if (base >= n2) break;
    const uint64_t end = base + cpw < n2 ? base + cpw : n2;
    for (uint64_t j0 = base; j0 < end; j0 += 64) {
      const uint64_t j = j0 + __lane_id();
      const bool live = j < end;
      bool failed = false;
      cell7 c;
      if (live) {
        const b2rec rec = listb2[j];
        failed = rec.fail != 0u;
        // trap sites of pieces 1..f, f the first uncovered piece (all pieces if none)
        const uint32_t upto = failed ? (2u << __builtin_ctz(rec.fail)) - 1u : 0xffffffffu;
        const bool trap = rec.trap != 0u || (rec.traps & upto) != 0u;
        atomicAdd(failed ? &t.unproven : &t.proven, 1u);
        if (trap) atomicAdd(&t.traps, 1u);
        if (failed) c = load7(&in[rec.j]);
      }
      put_unproven(c, live && failed && !at_limit, live && failed && at_limit, d, next, cap_next, unproven, cap_u,
                   ctrl);
    }
  }
  tally_flush(&t, ctrl);
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
  ccrd_fast_stage_ar_w##W(const rec112* __restrict__ in, uint64_t cap_in, int d, unsigned cpw,                 \
                          const Konst* __restrict__ kp, uint32_t* __restrict__ listb, uint64_t cap,           \
                          fast_ctrl* __restrict__ ctrl, const uint32_t* __restrict__ order,                   \
                          const uint8_t* __restrict__ cls, const int* __restrict__ class_step) {              \
    stage_a_ranked(in, cap_in, d, cpw, kp, listb, cap, ctrl, order, cls, class_step);                          \
  }                                                                                                           \
  extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))                 \
  ccrd_fast_stage_b1_w##W(const rec112* __restrict__ in, const uint32_t* __restrict__ listb, uint64_t cap_list,\
                          int d, int at_limit, unsigned cpw, const Konst* __restrict__ kp,                     \
                          rec112* __restrict__ next, uint64_t cap_next, rec112* __restrict__ unproven,        \
                          uint64_t cap_u, b2rec* __restrict__ listb2, uint64_t cap_b2,                        \
                          fast_ctrl* __restrict__ ctrl) {                                                     \
    stage_b1(in, listb, cap_list, d, at_limit, cpw, kp, next, cap_next, unproven, cap_u, listb2, cap_b2, ctrl); \
  }                                                                                                           \
  extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))                 \
  ccrd_fast_stage_b2_w##W(const rec112* __restrict__ in, b2rec* __restrict__ listb2, uint64_t cap_b2, int d,   \
                          unsigned cpw, const Konst* __restrict__ kp, fast_ctrl* __restrict__ ctrl) {          \
    stage_b2(in, listb2, cap_b2, d, cpw, kp, ctrl);                                                            \
  }

CC_STAGE_KERNELS(1)
CC_STAGE_KERNELS(2)
CC_STAGE_KERNELS(4)

// Stage B3 is a few loads per cell: one variant.
extern "C" __global__ void __launch_bounds__(256) ccrd_fast_stage_b3(const rec112* __restrict__ in,
                                                                      const b2rec* __restrict__ listb2,
                                                                      uint64_t cap_b2, int d, int at_limit,
                                                                      unsigned cpw, rec112* __restrict__ next,
                                                                      uint64_t cap_next,
                                                                      rec112* __restrict__ unproven, uint64_t cap_u,
                                                                      fast_ctrl* __restrict__ ctrl) {
  stage_b3(in, listb2, cap_b2, d, at_limit, cpw, next, cap_next, unproven, cap_u, ctrl);
}
