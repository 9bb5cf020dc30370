// The reference's per-cell enumeration kernels in their own shape, for the
// plain form of the drop-in proofs: one block per (outer piece pair), the
// inner two pieces in a 16x16 block's threads, every criterion re-evaluated
// by each thread that reaches it, no compaction between levels, one atomic
// per unproven leaf.  A launch covers a batch of cells (blockIdx.z), each
// as the reference's launch for one cell covers it.  Included by the
// translation units that build it with and without the exact shortcuts.
#pragma once
#include "rsb_rec.hpp"
#include "rsb_verdict.hpp"
#include "tr.hpp"

namespace ccd {

// tight_rectangle kernel_find_critical_intervals_in
// (Prover/tight_rectangle/tight_rectangle.cu:169-262); the leaves' cell and
// r1 piece go with them for the per-launch merge.
CCD void tr_plain_cell(const double* __restrict__ cells, int n1, int n2, int n3, int n4, const Konst& K,
                       trc::cell5* __restrict__ leaves, uint64_t cap, uint32_t* __restrict__ leaf_cell,
                       uint32_t* __restrict__ leaf_piece, unsigned long long* __restrict__ count,
                       int* __restrict__ trap) {
  using namespace trc;
  const uint32_t ci = blockIdx.z;
  const double* p = cells + 10 * (uint64_t)ci;
  cell5 c;
  c.la = iv{p[0], p[1]};
#pragma unroll
  for (int k = 0; k < 4; ++k) c.r[k] = iv{p[2 + 2 * k], p[3 + 2 * k]};
  const iv W = weight_for(K, c.la);
  const iv r1r = intersect(r1_range(K, c.la), c.r[0]);
  if (empty(r1r)) return;
  Trap tp{0};
  for (int i1 = blockIdx.x; i1 < n1; i1 += gridDim.x) {
    const iv r1 = subinterval(r1r, i1, n1);
    const iv r2r = intersect(r2_range(W, r1), c.r[1]);
    if (empty(r2r)) continue;
    if (r1_strategies(K, c.la, W, r1, r2r.hi)) continue;
    for (int i2 = blockIdx.y; i2 < n2; i2 += gridDim.y) {
      const iv r2 = subinterval(r2r, i2, n2);
      if (r1_strategies(K, c.la, W, r1, r2.hi)) continue;
      const iv r3r = intersect(r3_range(W, r1, r2), c.r[2]);
      if (empty(r3r)) continue;
      if (r1_r2_strategies(K, c.la, W, r1, r2, r3r.hi, tp)) continue;
      for (int i3 = threadIdx.x; i3 < n3; i3 += blockDim.x) {
        const iv r3 = subinterval(r3r, i3, n3);
        if (r1_r2_strategies(K, c.la, W, r1, r2, r3.hi, tp)) continue;
        const iv r4r = intersect(r4_range(W, r1, r2, r3), c.r[3]);
        if (empty(r4r)) continue;
        if (r123_strategies(K, c.la, W, r1, r2, r3, r4r.hi)) continue;
        iv R4 = W - r1 - r2 - r3;
        tighten_lo(R4, 0.0);
        if (empty(R4)) continue;
        for (int i4 = threadIdx.y; i4 < n4; i4 += blockDim.y) {
          const iv r4 = subinterval(r4r, i4, n4);
          if (r123_strategies(K, c.la, W, r1, r2, r3, r4.hi)) continue;
          cell5 v;
          v.la = c.la;
          v.r[0] = r1;
          v.r[1] = r2;
          v.r[2] = r3;
          v.r[3] = r4;
          tighten_lo(v.r[2], v.r[3].lo);
          tighten_lo(v.r[1], v.r[2].lo);
          tighten_lo(v.r[0], v.r[1].lo);
          iv R = W - v.r[0] - v.r[1] - v.r[2] - v.r[3];
          tighten_lo(R, 0.0);
          if (empty(R)) continue;
          if (two_by_two_leaf(K, v, R) || three_disk_leaf(K, v, R4, tp) || gaps_leaf(K, v, R, tp)) continue;
          const uint64_t slot = atomicAdd(count, 1ull);
          if (slot < cap) {
            leaves[slot] = v;
            leaf_cell[slot] = ci;
            leaf_piece[slot] = (uint32_t)i1;
          }
        }
      }
    }
  }
  if (tp.flag) atomicOr(trap, 1);
}

// rectangle_size_bound kernel_find_critical_intervals_in
// (Prover/rectangle_size_bound/rectangle_size_bound.cu:161-236); the
// unproven leaves are written tightened, as the reference pushes them.
CCD void rsb_plain_cell(const rsb_rec* __restrict__ cells, int n3, int n4, int n5, int n6, const Konst& K,
                        rsb_rec* __restrict__ leaves, uint64_t cap, unsigned long long* __restrict__ count,
                        int* __restrict__ trap) {
  const uint32_t ci = blockIdx.z;
  const cell7 c = load_cell7(cells[ci].v);
  if (rsbc::s1_even_split(c.la, c.r[0], c.r[1]) || rsbc::s23_uneven_split(c.la, c.r[0], c.r[1])) return;
  Trap tp{0};
  for (int i3 = blockIdx.x; i3 < n3; i3 += gridDim.x) {
    const iv r3 = subinterval(c.r[2], i3, n3);
    if (rsbc::s23_uneven_split(c.la, c.r[0], r3) || rsbc::s4_opposite_corners(c.la, c.r[0], c.r[1], r3) ||
        rsbc::s5_vertical_strip(K, c.la, c.r[0], c.r[1], r3))
      continue;
    for (int i4 = blockIdx.y; i4 < n4; i4 += gridDim.y) {
      const iv r4 = subinterval(c.r[3], i4, n4);
      for (int i5 = threadIdx.x; i5 < n5; i5 += blockDim.x) {
        const iv r5 = subinterval(c.r[4], i5, n5);
        for (int i6 = threadIdx.y; i6 < n6; i6 += blockDim.y) {
          cell7 v;
          v.la = c.la;
          v.r[0] = c.r[0];
          v.r[1] = c.r[1];
          v.r[2] = r3;
          v.r[3] = r4;
          v.r[4] = r5;
          v.r[5] = subinterval(c.r[5], i6, n6);
          if (rsb_chain_verdict(K, v, tp) != -1) continue;
          const uint64_t slot = atomicAdd(count, 1ull);
          if (slot < cap) store_cell7(&leaves[slot], rsbc::tighten(v));
        }
      }
    }
  }
  if (tp.flag) atomicOr(trap, 1);
}

}  // namespace ccd
