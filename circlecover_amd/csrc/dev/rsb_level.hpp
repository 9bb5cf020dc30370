// rectangle_size_bound level step of the level-synchronous
// branch-and-bound: the widest-dimension bisection and the per-cell body of
// one depth's launch.  Included by rsb_bnb.hip (the reference's operation
// sequence, CC_EXACT_SHORTCUTS 0) and rsb_bnb_x.hip (with the exact
// shortcuts of the batched form, CC_EXACT_SHORTCUTS 1).
#pragma once
#include "queue.hpp"
#include "rsb_rec.hpp"

namespace ccd {

// Widest of (lambda, r1..r6) by absolute width (round-to-nearest ub - lb,
// ties to the lowest index); half h = get_subinterval(x, h, 2).
CCD cell7 bisect7(cell7 c, int h) {
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

// One level of the branch-and-bound: one lane per cell of `in`.  A proven
// cell is dropped; an unproven cell below the depth limit appends its two
// halves to `children` (pair slot from counts[0]); at the limit it is
// appended to `unproven` (counts[1], records past cap_unproven dropped).
// counts[2] counts cells that reached a trap site; hist[v+1] += cells with
// verdict v.  log (may be null): every cell's record (the cell, its verdict,
// `depth`) is appended (slot from counts[3], records past log_cap dropped).
struct bnb_log_rec {
  rsb_rec cell;
  int32_t verdict;
  int32_t depth;
};
__device__ __forceinline__ void bnb_level_body(const rsb_rec* __restrict__ in, uint64_t n, int at_limit,
                                               const Konst* __restrict__ kp, rsb_rec* __restrict__ children,
                                               rsb_rec* __restrict__ unproven, uint64_t cap_unproven,
                                               unsigned long long* __restrict__ counts,
                                               unsigned long long* __restrict__ hist, bnb_log_rec* __restrict__ log,
                                               uint64_t log_cap, int depth) {
  __shared__ unsigned int bins[21];
  if (threadIdx.x < 21) bins[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int verdict = 0;
  bool live = i < n;
  cell7 c;
  Trap tp{0};
  if (live) {
    const Konst K = *kp;
    c = load_cell7(in[i].v);
    verdict = rsb_cell_verdict(K, c, tp);
    atomicAdd(&bins[verdict + 1], 1u);
  }
  const bool split = live && verdict < 0 && !at_limit;
  const bool keep = live && verdict < 0 && at_limit;
  const uint64_t ps = wave_append(split, &counts[0]);
  if (split) {
    store_cell7(&children[2 * ps], bisect7(c, 0));
    store_cell7(&children[2 * ps + 1], bisect7(c, 1));
  }
  const uint64_t us = wave_append(keep, &counts[1]);
  if (keep && us < cap_unproven) store_cell7(&unproven[us], c);
  if (live && tp.flag) atomicAdd(&counts[2], 1ull);
  if (log) {
    const uint64_t ls = wave_append(live, &counts[3]);
    if (live && ls < log_cap) {
      store_cell7(&log[ls].cell, c);
      log[ls].verdict = verdict;
      log[ls].depth = depth;
    }
  }
  __syncthreads();
  if (threadIdx.x < 21 && bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
}

}  // namespace ccd
