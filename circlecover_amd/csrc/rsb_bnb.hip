// rectangle_size_bound level kernels (gfx950, round-down mode): the
// widest-dimension branch-and-bound over 7-D cells, and the enumeration of
// the reference's per-cell kernel (kernel_find_critical_intervals_in,
// Prover/rectangle_size_bound/rectangle_size_bound.cu:161-236) flattened
// into two levels over a batch of cells.
#include "dev/queue.hpp"
#include "dev/rsb_verdict.hpp"

using namespace ccd;

// 7-D cell record of the device queues: the 112-byte reference layout.
struct rsb_rec {
  double v[14];
};

CCD void store_cell7(rsb_rec* __restrict__ out, const cell7& c) {
  rsb_rec r;
  r.v[0] = c.la.lo;
  r.v[1] = c.la.hi;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    r.v[2 + 2 * k] = c.r[k].lo;
    r.v[3 + 2 * k] = c.r[k].hi;
  }
  *out = r;
}

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
// verdict v.
extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_bnb_level(const rsb_rec* __restrict__ in, uint64_t n,
                                                                      int at_limit, const Konst* __restrict__ kp,
                                                                      rsb_rec* __restrict__ children,
                                                                      rsb_rec* __restrict__ unproven,
                                                                      uint64_t cap_unproven,
                                                                      unsigned long long* __restrict__ counts,
                                                                      unsigned long long* __restrict__ hist) {
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
  __syncthreads();
  if (threadIdx.x < 21 && bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// The reference kernel's enumeration.  Level A, one lane per (cell, r3
// piece): the cell shortcuts S1/S2 (rectangle_size_bound.cu:162) and the r3
// shortcuts S3..S5 (:172-178); surviving (cell, r3) pairs are compacted.
struct rsb_s3 {
  uint32_t cell, pad;
  iv r3;
};

extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_enum_r3(const rsb_rec* __restrict__ cells, uint64_t ncells,
                                                                    int n3, const Konst* __restrict__ kp,
                                                                    rsb_s3* __restrict__ out,
                                                                    unsigned long long* __restrict__ count,
                                                                    int* __restrict__ trap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells * (uint64_t)n3) return;
  const Konst K = *kp;
  const uint32_t ci = (uint32_t)(i / (uint64_t)n3);
  const int i3 = (int)(i % (uint64_t)n3);
  const cell7 c = load_cell7(cells[ci].v);
  rsb_s3 rec;
  rec.cell = ci;
  rec.pad = 0;
  rec.r3 = subinterval(c.r[2], i3, n3);
  const bool keep = !rsbc::s1_even_split(c.la, c.r[0], c.r[1]) && !rsbc::s23_uneven_split(c.la, c.r[0], c.r[1]) &&
                    !rsbc::s23_uneven_split(c.la, c.r[0], rec.r3) &&
                    !rsbc::s4_opposite_corners(c.la, c.r[0], c.r[1], rec.r3) &&
                    !rsbc::s5_vertical_strip(K, c.la, c.r[0], c.r[1], rec.r3);
  const uint64_t slot = wave_append(keep, count);
  if (keep) out[slot] = rec;
  (void)trap;
}

// Level B, one lane per (level-A record, r4, r5, r6 pieces): the leaf
// (rectangle_size_bound.cu:180-233).  Unproven leaves, tightened as the
// reference pushes them, go to `leaves` (counter counts all; records past
// cap are dropped).
extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_enum_leaf(const rsb_rec* __restrict__ cells,
                                                                      const rsb_s3* __restrict__ in, uint64_t nin,
                                                                      int n4, int n5, int n6,
                                                                      const Konst* __restrict__ kp,
                                                                      rsb_rec* __restrict__ leaves, uint64_t cap,
                                                                      unsigned long long* __restrict__ count,
                                                                      int* __restrict__ trap) {
  const uint64_t per = (uint64_t)n4 * n5 * n6;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin * per) return;
  const Konst K = *kp;
  const rsb_s3 s = in[i / per];
  uint64_t j = i % per;
  const int i6 = (int)(j % (uint64_t)n6);
  j /= (uint64_t)n6;
  const int i5 = (int)(j % (uint64_t)n5);
  const int i4 = (int)(j / (uint64_t)n5);
  const cell7 c = load_cell7(cells[s.cell].v);
  cell7 v;
  v.la = c.la;
  v.r[0] = c.r[0];
  v.r[1] = c.r[1];
  v.r[2] = s.r3;
  v.r[3] = subinterval(c.r[3], i4, n4);
  v.r[4] = subinterval(c.r[4], i5, n5);
  v.r[5] = subinterval(c.r[5], i6, n6);
  v = rsbc::tighten(v);
  Trap tp{0};
  const bool keep = rsb_chain_verdict(K, v, tp) < 0;
  if (tp.flag) atomicOr(trap, 1);
  const uint64_t slot = wave_append(keep, count);
  if (keep && slot < cap) store_cell7(&leaves[slot], v);
}
