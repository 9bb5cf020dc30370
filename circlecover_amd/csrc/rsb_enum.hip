// The enumeration of the reference's per-cell kernel
// (kernel_find_critical_intervals_in,
// Prover/rectangle_size_bound/rectangle_size_bound.cu:161-236) flattened
// into two levels over a batch of cells (gfx950, round-down mode), for the
// drop-in proof.  Built with the exact shortcuts: every criterion returns
// what the reference's operation sequence returns, sooner.  The leaf
// chain's steps and can_recurse are inlined: stage A runs the chain on about
// 10^12 leaves of a full proof, and the out-of-line form paid a call with
// its register spills to scratch wherever a wave's lanes reached a step
// (9.3 -> 7.9 s on 26 lambda pieces of the proof grid).
#define CC_EXACT_SHORTCUTS 1
#define CC_INLINE_RECURSE 1
#define CC_CHAIN_FN static __device__ __forceinline__
// constants as VGPR operands (see round.hpp)
#define CC_CONST_OPERANDS 0
#include "dev/queue.hpp"
#include "dev/rsb_rec.hpp"

using namespace ccd;

// Level A, one lane per (cell, r3 piece): the cell shortcuts S1/S2
// (rectangle_size_bound.cu:162) and the r3 shortcuts S3..S5 (:172-178);
// surviving (cell, r3) pairs are compacted.
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

// Level B: the leaves (level-A record, r4, r5, r6 pieces) and their chain
// (rectangle_size_bound.cu:180-233), in three steps, each dense on the
// leaves the step before leaves open:
//   rows      one lane per (record, r4, r5): the r6 pieces whose leaf is
//             not empty after tightening (the chain's first test; about 94%
//             of the leaves of a proof are empty) are listed, 8 bytes each;
//   stage A   the chain up to C13 on the listed leaves; those only C14 can
//             still prove are listed again;
//   stage B   C14 on that list.
// Run per leaf, every wave would wait for its slowest leaf: nearly every
// wave of 64 leaves holds one that goes deep into the chain.  Unproven
// leaves, tightened as the reference pushes them, go to `leaves` (counter
// counts all; records past cap are dropped).
struct leaf_ref {
  uint32_t rec;
  uint8_t i4, i5, i6, pad;
};

CCD cell7 leaf_of(const rsb_rec* __restrict__ cells, const rsb_s3& s, int i4, int n4, int i5, int n5, int i6,
                  int n6) {
  const cell7 c = load_cell7(cells[s.cell].v);
  cell7 v;
  v.la = c.la;
  v.r[0] = c.r[0];
  v.r[1] = c.r[1];
  v.r[2] = s.r3;
  v.r[3] = subinterval(c.r[3], i4, n4);
  v.r[4] = subinterval(c.r[4], i5, n5);
  v.r[5] = subinterval(c.r[5], i6, n6);
  return rsbc::tighten(v);
}

// Estimate (any rounding) of the number of open leaves in a row: 0 when the
// row's own part is empty, else the first r6 piece index whose lower bound
// reaches the least of the thresholds on l6 at which the tightened leaf
// becomes empty (l6 >= min(hi of r2..r5)) or its remaining weight's upper
// bound la.hi * 0.61 - r1.lo - sum of the tightened lower bounds of r2..r6
// falls below 0.  v holds the row's la, r1..r5 (r5 and r4 as their pieces).
// A CPU check against the oracle's emptiness verdicts matched all of 400
// sampled rows of the proof grid.
CCD int estimate_open_prefix(const cell7& v, double lo6, double d6, int n6) {
  double H = v.r[1].hi;
  for (int k = 2; k < 5; ++k) H = v.r[k].hi < H ? v.r[k].hi : H;
  // m[k]: the tightened lower bound of r(k+2) for l6 below it, non-increasing in k
  double m[4];
  double run = 0.0;
  for (int k = 3; k >= 0; --k) {
    run = v.r[k + 1].lo > run ? v.r[k + 1].lo : run;
    m[k] = run;
  }
  // a row whose own part is empty (r1, or r2..r5 with the row's tightened
  // lower bounds) has no open leaf
  if (v.r[0].lo >= v.r[0].hi) return 0;
  for (int k = 0; k < 4; ++k)
    if (m[k] >= v.r[k + 1].hi) return 0;
  double rest = v.la.hi * kRsbRatio - v.r[0].lo;
  double T = H;
  // l6 in [m[k], m[k-1]]: r2..r(k+1) keep m, the other 4 - k and r6 take l6
  double fixed = 0.0;
  for (int k = 0; k <= 4; ++k) {
    const double root = (rest - fixed) / (double)(5 - k);
    const double lo_seg = k < 4 ? m[k] : -1e300;
    if (root >= lo_seg) {
      T = root < T ? root : T;
      break;
    }
    if (k < 4) fixed += m[k];
  }
  if (!(T > lo6)) return 0;
  const double x = (T - lo6) / d6;
  if (!(x < (double)n6)) return n6;
  const int g = (int)ceil(x);
  return g < 0 ? 0 : g > n6 ? n6 : g;
}

// Rows of the level-A records [rec0, rec0 + nrec) (n4, n5, n6 <= 256), in
// workgroups of 1024 lanes.  The list has room for every leaf of those
// records.  One atomic per workgroup claims its part of the list: nearly
// every wave lists some leaves, and one atomic per wave on the one counter
// cost more than the emptiness tests themselves.
static constexpr int kRowsBlock = 1024;

extern "C" __global__ void __launch_bounds__(kRowsBlock) ccrd_rsb_enum_rows(const rsb_rec* __restrict__ cells,
                                                                      const rsb_s3* __restrict__ in, uint64_t rec0,
                                                                      uint64_t nrec, int n4, int n5, int n6,
                                                                      leaf_ref* __restrict__ list,
                                                                      unsigned long long* __restrict__ count_list) {
  const uint64_t per_rec = (uint64_t)n4 * n5;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < nrec * per_rec;
  const int lane = __lane_id();
  uint32_t rec = 0;
  int i4 = 0, i5 = 0;
  cell7 v{};
  iv r6{0.0, 0.0};
  if (live) {
    const uint64_t r = t / per_rec;
    const uint32_t q = (uint32_t)(t - r * per_rec);
    i4 = (int)(q / (uint32_t)n5);
    i5 = (int)(q % (uint32_t)n5);
    rec = (uint32_t)(rec0 + r);
    const rsb_s3 s = in[rec];
    const cell7 c = load_cell7(cells[s.cell].v);
    v.la = c.la;
    v.r[0] = c.r[0];
    v.r[1] = c.r[1];
    v.r[2] = s.r3;
    v.r[3] = subinterval(c.r[3], i4, n4);
    v.r[4] = subinterval(c.r[4], i5, n5);
    r6 = c.r[5];
  }
  // the r6 pieces as subinterval(r6, i, n6) computes them: one quotient for
  // the row, and each piece's upper bound is the next piece's lower bound
  const double d6 = live ? rn_span_div(r6.hi, r6.lo, (double)n6) : 0.0;
  // The open (not empty) leaves of a row are a prefix of its r6 pieces
  // when the pieces' bounds strictly increase, which a quotient above two
  // ulps of r6.hi ensures: with l6, u6 the piece's bounds, the tightened
  // leaf is empty exactly when the row's part is, or l6 >= u6, or l6 >= the
  // least upper bound of r2..r5 (tightening raises their lower bounds to
  // l6); and its remaining weight R (lower bound <= upper bound for a
  // non-empty leaf) is empty exactly when R.hi < 0, where R.hi is the
  // upward-rounded la * 0.61 minus the tightened lower bounds, none of which
  // falls as l6 grows.  So a binary search with the full test finds the
  // prefix; rows with narrower pieces test every piece.
  int open_prefix = -1;
  if (live && d6 > 2.0 * (next_up(r6.hi) - r6.hi)) {
    auto open_at = [&](int i) {
      double l6, u6;
      rn_fma2(d6, (double)i, (double)(i + 1), r6.lo, l6, u6);
      v.r[5] = iv{l6, i + 1 == n6 ? r6.hi : u6};
      const cell7 c = rsbc::tighten(v);
      return !(rsbc::is_empty(c) || empty(rsbc::remaining_weight(c)));
    };
    int a = 0, b = n6;  // open on [0, a), closed on [b, n6)
    // The prefix's end estimated from the two thresholds on l6 (the least
    // upper bound H of r2..r5, and where R.hi, a piecewise linear function
    // of l6, reaches 0), then confirmed by the full test on the pieces
    // either side of it; an unconfirmed estimate falls back to the search.
    const int g = estimate_open_prefix(v, r6.lo, d6, n6);
    if ((g == 0 || open_at(g - 1)) && (g == n6 || !open_at(g))) a = b = g;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (open_at(mid)) a = mid + 1;
      else b = mid;
    }
    open_prefix = a;
  }
  for (int b6 = 0; b6 < n6; b6 += 64) {  // the same trip count on every lane
    uint64_t mask = 0;
    if (live && open_prefix >= 0) {
      const int m = open_prefix - b6 < 64 ? open_prefix - b6 : 64;
      if (m > 0) mask = m == 64 ? ~0ull : (1ull << m) - 1ull;
    } else if (live) {
      const int m = n6 - b6 < 64 ? n6 - b6 : 64;
      double l = rn_fma(d6, (double)b6, r6.lo);
      for (int k = 0; k < m; ++k) {
        const int i = b6 + k;
        const double u = i + 1 == n6 ? r6.hi : rn_fma(d6, (double)(i + 1), r6.lo);
        v.r[5] = iv{l, u};
        l = u;
        const cell7 c = rsbc::tighten(v);
        if (!(rsbc::is_empty(c) || empty(rsbc::remaining_weight(c)))) mask |= 1ull << k;
      }
    }
    // an exclusive scan of the lanes' counts in the wave, then of the
    // waves' totals in the workgroup, and one atomic for the workgroup
    __shared__ uint32_t wave_total[kRowsBlock / 64];
    __shared__ unsigned long long wave_base[kRowsBlock / 64];
    const int wv = threadIdx.x >> 6;
    const uint32_t mine = (uint32_t)__popcll(mask);
    uint32_t x = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wave_total[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long sum = 0;
      for (int w = 0; w < kRowsBlock / 64; ++w) sum += wave_total[w];
      unsigned long long base = sum ? atomicAdd(count_list, sum) : 0ull;
      for (int w = 0; w < kRowsBlock / 64; ++w) {
        wave_base[w] = base;
        base += wave_total[w];
      }
    }
    __syncthreads();
    unsigned long long at = wave_base[wv] + (x - mine);
    while (mask) {
      const int k = __builtin_ctzll(mask);
      mask &= mask - 1;
      leaf_ref ref;
      ref.rec = rec;
      ref.i4 = (uint8_t)i4;
      ref.i5 = (uint8_t)i5;
      ref.i6 = (uint8_t)(b6 + k);
      ref.pad = 0;
      list[at++] = ref;
    }
  }
}

// Stage A over the *count_list listed leaves, grid-stride.  listb has room
// for all of them.  Four waves per SIMD (128 VGPRs, the rest of the chain's
// state in scratch): the chain's long FP64 dependency chains need the
// latency hiding more than the registers (3 or 2 waves were slower).
extern "C" __global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 4))) ccrd_rsb_enum_leaf_a(const rsb_rec* __restrict__ cells,
                                                                        const rsb_s3* __restrict__ in,
                                                                        const leaf_ref* __restrict__ list,
                                                                        const unsigned long long* __restrict__ count_list,
                                                                        int n4, int n5, int n6,
                                                                        const Konst* __restrict__ kp,
                                                                        leaf_ref* __restrict__ listb,
                                                                        unsigned long long* __restrict__ count_b,
                                                                        int* __restrict__ trap) {
  const uint64_t n = *count_list;
  const Konst K = *kp;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    bool keep = false;
    leaf_ref ref{};
    if (i < n) {
      ref = list[i];
      const cell7 v = leaf_of(cells, in[ref.rec], ref.i4, n4, ref.i5, n5, ref.i6, n6);
      Trap tp{0};
      keep = rsb_chain_before_c14(K, v, tp) == -2;
      if (tp.flag) atomicOr(trap, 1);
    }
    const uint64_t slot = wave_append(keep, count_b);
    if (keep) listb[slot] = ref;
  }
}

// Stage B: C14 on the *count_b listed leaves, grid-stride.
extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_enum_leaf_b(const rsb_rec* __restrict__ cells,
                                                                        const rsb_s3* __restrict__ in,
                                                                        const leaf_ref* __restrict__ listb,
                                                                        const unsigned long long* __restrict__ count_b,
                                                                        int n4, int n5, int n6,
                                                                        const Konst* __restrict__ kp,
                                                                        rsb_rec* __restrict__ leaves, uint64_t cap,
                                                                        unsigned long long* __restrict__ count,
                                                                        int* __restrict__ trap) {
  const uint64_t nb = *count_b;
  const Konst K = *kp;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nb; base += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    bool keep = false;
    cell7 v{};
    if (i < nb) {
      const leaf_ref ref = listb[i];
      v = leaf_of(cells, in[ref.rec], ref.i4, n4, ref.i5, n5, ref.i6, n6);
      Trap tp{0};
      keep = rsb_verdict_c14(K, v, tp) < 0;
      if (tp.flag) atomicOr(trap, 1);
    }
    const uint64_t slot = wave_append(keep, count);
    if (keep && slot < cap) store_cell7(&leaves[slot], v);
  }
}
