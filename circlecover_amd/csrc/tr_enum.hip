// tight_rectangle enumeration kernels (gfx950, round-down mode).
//
// The reference evaluates one lambda cell per launch as a four-deep loop
// nest over r1..r4 pieces with early exits (kernel_find_critical_intervals_in,
// Prover/tight_rectangle/tight_rectangle.cu:169-262).  Here the loop nest is
// flattened into four level kernels over a batch of cells; each level keeps
// the (cell, piece) pairs that no criterion of that level proves and compacts
// them with a wave-aggregated append into the next level's input.  Every
// pair is visited once, as in the reference, so the multiset of unproven
// leaves is the same; only their order differs (the reference's is that of
// its atomicAdd too).
#define CC_EXACT_SHORTCUTS 1
#include "dev/queue.hpp"
#include "dev/tr.hpp"

using namespace ccd;
using namespace ccd::trc;

struct tr_aux {  // per input cell
  iv W;          // required_weight_for(la)
  iv r1r;        // r1_range(la) intersected with the cell's r1
};
struct tr_s1 {
  uint32_t cell, pad;
  iv r1, r2r;
};
struct tr_s2 {
  uint32_t cell, pad;
  iv r1, r2, r3r;
};
struct tr_s3 {
  uint32_t cell, pad;
  iv r1, r2, r3, r4r, R4;
};

CCD cell5 load_cell(const double* __restrict__ p) {
  cell5 c;
  c.la = iv{p[0], p[1]};
#pragma unroll
  for (int k = 0; k < 4; ++k) c.r[k] = iv{p[2 + 2 * k], p[3 + 2 * k]};
  return c;
}

// Level 0, one lane per cell: weight and r1 range (tight_rectangle.cu:175-180).
extern "C" __global__ void __launch_bounds__(256) ccrd_tr_level0(const double* __restrict__ cells, uint64_t ncells,
                                                                  const Konst* __restrict__ kp,
                                                                  tr_aux* __restrict__ aux) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells) return;
  const Konst K = *kp;
  const cell5 c = load_cell(cells + 10 * i);
  tr_aux a;
  a.W = weight_for(K, c.la);
  a.r1r = intersect(r1_range(K, c.la), c.r[0]);
  aux[i] = a;
}

// Level 1, one lane per (cell, r1 piece): tight_rectangle.cu:182-191.
extern "C" __global__ void __launch_bounds__(256) ccrd_tr_level1(const double* __restrict__ cells, uint64_t ncells,
                                                                  int n1, const Konst* __restrict__ kp,
                                                                  const tr_aux* __restrict__ aux,
                                                                  tr_s1* __restrict__ out,
                                                                  unsigned long long* __restrict__ count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells * (uint64_t)n1) return;
  const Konst K = *kp;
  const uint32_t ci = (uint32_t)(i / (uint64_t)n1);
  const int i1 = (int)(i % (uint64_t)n1);
  const tr_aux a = aux[ci];
  bool keep = false;
  tr_s1 rec;
  if (!empty(a.r1r)) {
    const cell5 c = load_cell(cells + 10 * (uint64_t)ci);
    rec.cell = ci;
    rec.pad = (uint32_t)i1;  // the r1 piece, carried down to the leaves
    rec.r1 = subinterval(a.r1r, i1, n1);
    rec.r2r = intersect(r2_range(a.W, rec.r1), c.r[1]);
    keep = !empty(rec.r2r) && !r1_strategies(K, c.la, a.W, rec.r1, rec.r2r.hi);
  }
  const uint64_t slot = wave_append(keep, count);
  if (keep) out[slot] = rec;
}

// Level 2, one lane per (level-1 record, r2 piece): tight_rectangle.cu:193-206.
extern "C" __global__ void __launch_bounds__(256) ccrd_tr_level2(const double* __restrict__ cells,
                                                                  const tr_s1* __restrict__ in, uint64_t nin, int n2,
                                                                  const Konst* __restrict__ kp,
                                                                  const tr_aux* __restrict__ aux,
                                                                  tr_s2* __restrict__ out,
                                                                  unsigned long long* __restrict__ count,
                                                                  int* __restrict__ trap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin * (uint64_t)n2) return;
  const Konst K = *kp;
  const tr_s1 s = in[i / (uint64_t)n2];
  const int i2 = (int)(i % (uint64_t)n2);
  const cell5 c = load_cell(cells + 10 * (uint64_t)s.cell);
  const iv W = aux[s.cell].W;
  Trap tp{0};
  tr_s2 rec;
  rec.cell = s.cell;
  rec.pad = s.pad;
  rec.r1 = s.r1;
  rec.r2 = subinterval(s.r2r, i2, n2);
  bool keep = !r1_strategies(K, c.la, W, s.r1, rec.r2.hi);
  if (keep) {
    rec.r3r = intersect(r3_range(W, s.r1, rec.r2), c.r[2]);
    keep = !empty(rec.r3r) && !r1_r2_strategies(K, c.la, W, s.r1, rec.r2, rec.r3r.hi, tp);
  }
  if (tp.flag) atomicOr(trap, 1);
  const uint64_t slot = wave_append(keep, count);
  if (keep) out[slot] = rec;
}

// Level 3, one lane per (level-2 record, r3 piece): tight_rectangle.cu:208-227.
extern "C" __global__ void __launch_bounds__(256) ccrd_tr_level3(const double* __restrict__ cells,
                                                                  const tr_s2* __restrict__ in, uint64_t nin, int n3,
                                                                  const Konst* __restrict__ kp,
                                                                  const tr_aux* __restrict__ aux,
                                                                  tr_s3* __restrict__ out,
                                                                  unsigned long long* __restrict__ count,
                                                                  int* __restrict__ trap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin * (uint64_t)n3) return;
  const Konst K = *kp;
  const tr_s2 s = in[i / (uint64_t)n3];
  const int i3 = (int)(i % (uint64_t)n3);
  const cell5 c = load_cell(cells + 10 * (uint64_t)s.cell);
  const iv W = aux[s.cell].W;
  Trap tp{0};
  tr_s3 rec;
  rec.cell = s.cell;
  rec.pad = s.pad;
  rec.r1 = s.r1;
  rec.r2 = s.r2;
  rec.r3 = subinterval(s.r3r, i3, n3);
  bool keep = !r1_r2_strategies(K, c.la, W, s.r1, s.r2, rec.r3.hi, tp);
  if (keep) {
    rec.r4r = intersect(r4_range(W, s.r1, s.r2, rec.r3), c.r[3]);
    keep = !empty(rec.r4r) && !r123_strategies(K, c.la, W, s.r1, s.r2, rec.r3, rec.r4r.hi);
    if (keep) {
      rec.R4 = W - s.r1 - s.r2 - rec.r3;
      tighten_lo(rec.R4, 0.0);
      keep = !empty(rec.R4);
    }
  }
  if (tp.flag) atomicOr(trap, 1);
  const uint64_t slot = wave_append(keep, count);
  if (keep) out[slot] = rec;
}

// Level 4, one lane per (level-3 record, r4 piece): tight_rectangle.cu:229-258.
// Unproven leaves are appended to `leaves`; the counter counts all of them,
// records past `cap` are dropped (the caller re-runs with a larger buffer).
extern "C" __global__ void __launch_bounds__(128) ccrd_tr_level4(const double* __restrict__ cells,
                                                                  const tr_s3* __restrict__ in, uint64_t nin, int n4,
                                                                  const Konst* __restrict__ kp,
                                                                  const tr_aux* __restrict__ aux,
                                                                  cell5* __restrict__ leaves, uint64_t cap,
                                                                  uint32_t* __restrict__ leaf_cell,
                                                                  uint32_t* __restrict__ leaf_piece,
                                                                  unsigned long long* __restrict__ count,
                                                                  int* __restrict__ trap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nin * (uint64_t)n4) return;
  const Konst K = *kp;
  const tr_s3 s = in[i / (uint64_t)n4];
  const int i4 = (int)(i % (uint64_t)n4);
  const iv la = iv{cells[10 * (uint64_t)s.cell], cells[10 * (uint64_t)s.cell + 1]};
  const iv W = aux[s.cell].W;
  Trap tp{0};
  const iv r4 = subinterval(s.r4r, i4, n4);
  bool keep = !r123_strategies(K, la, W, s.r1, s.r2, s.r3, r4.hi);
  cell5 v;
  if (keep) {
    v.la = la;
    v.r[0] = s.r1;
    v.r[1] = s.r2;
    v.r[2] = s.r3;
    v.r[3] = r4;
    tighten_lo(v.r[2], v.r[3].lo);
    tighten_lo(v.r[1], v.r[2].lo);
    tighten_lo(v.r[0], v.r[1].lo);
    iv R = W - v.r[0] - v.r[1] - v.r[2] - v.r[3];
    tighten_lo(R, 0.0);
    keep = !empty(R) && !two_by_two_leaf(K, v, R) && !three_disk_leaf(K, v, s.R4, tp) && !gaps_leaf(K, v, R, tp);
  }
  if (tp.flag) atomicOr(trap, 1);
  const uint64_t slot = wave_append(keep, count);
  if (keep && slot < cap) {
    leaves[slot] = v;
    leaf_cell[slot] = s.cell;
    leaf_piece[slot] = s.pad;
  }
}

// The per-launch merge of leaves (merge_adjacent_r1,
// tight_rectangle.cu:129-159) joins the leaves whose r1 intervals form one
// connected component of their union, hull over all radii.  Every leaf of
// one (cell, r1 piece) bucket has the piece's r1 upper bound (tightening
// only raises r1's lower bound, to at most r2's upper bound, which r2's
// range keeps below r1's), so a bucket lies in one component and the merge
// only needs each bucket's hull.  This kernel forms them: per bucket, the
// least lower bound and greatest upper bound of r1..r4, the least r1 upper
// bound (equal to the greatest one when the premise holds), and the leaf
// count, as order-preserving 64-bit keys (10 words per bucket, initialised
// by the caller to the identities).  A leaf with r1's lower bound above its
// upper bound sets *bad.  Within a wave, an inclusive segmented scan over
// runs of equal bucket keys (the leaves of a run come from the same few
// level-3 records) leaves each run's hull in its last lane, which alone
// issues the run's atomics: one set per run instead of one per leaf, so
// the few hot buckets of a launch do not serialise the kernel.
CCD unsigned long long ord_key(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : b | (1ull << 63);
}

extern "C" __global__ void __launch_bounds__(256) cc_tr_bucket_hull(const cell5* __restrict__ leaves,
                                                                    const uint32_t* __restrict__ leaf_cell,
                                                                    const uint32_t* __restrict__ leaf_piece,
                                                                    uint64_t n, int n1,
                                                                    unsigned long long* __restrict__ buckets,
                                                                    int* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const uint64_t lives = __ballot(live);
  if (!lives) return;  // the whole wave
  uint64_t key = 0;
  unsigned long long lo[4], hi[4], hmin;
  if (live) {
    const cell5 c = leaves[i];
    key = (uint64_t)leaf_cell[i] * (uint64_t)n1 + leaf_piece[i];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lo[q] = ord_key(c.r[q].lo);
      hi[q] = ord_key(c.r[q].hi);
    }
    hmin = hi[0];
    if (!(c.r[0].lo <= c.r[0].hi)) atomicOr(bad, 1);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lo[q] = ~0ull;
      hi[q] = 0ull;
    }
    hmin = ~0ull;
  }
  const int lane = __lane_id();
  // (head, hull) pairs combined left to right; a head starts a new run
  const uint64_t prev_key = __shfl_up(key, 1);
  bool head = lane == 0 || prev_key != key || !live;
  unsigned cnt = live ? 1u : 0u;
  for (int off = 1; off < 64; off <<= 1) {
    const bool ok = lane >= off;
    const bool h = __shfl_up((int)head, off) != 0;
    const unsigned c = __shfl_up(cnt, off);
    unsigned long long l[4], u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      l[q] = __shfl_up(lo[q], off);
      u[q] = __shfl_up(hi[q], off);
    }
    const unsigned long long m = __shfl_up(hmin, off);
    if (ok && !head) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        lo[q] = l[q] < lo[q] ? l[q] : lo[q];
        hi[q] = u[q] > hi[q] ? u[q] : hi[q];
      }
      hmin = m < hmin ? m : hmin;
      cnt += c;
    }
    if (ok) head = head || h;
  }
  // the last lane of a run: the next lane starts another run or is dead
  const uint64_t next_key = __shfl_down(key, 1);
  const bool next_live = __shfl_down((int)live, 1) != 0;
  const bool last = live && (lane == 63 || !next_live || next_key != key);
  if (last) {
    unsigned long long* bk = buckets + 10 * key;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      atomicMin(bk + q, lo[q]);
      atomicMax(bk + 4 + q, hi[q]);
    }
    atomicMin(bk + 8, hmin);
    atomicAdd(bk + 9, (unsigned long long)cnt);
  }
}

// Bisection of unproven leaves for the next refinement level: the widest of
// (lambda, r1..r4) by absolute width (round-to-nearest ub - lb, ties to the
// lowest index) is split into get_subinterval(x, 0, 2) and (x, 1, 2).
extern "C" __global__ void __launch_bounds__(256) ccrd_tr_bisect(const cell5* __restrict__ in, uint64_t n,
                                                                  cell5* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  cell5 c = in[i >> 1];
  iv* d[5] = {&c.la, &c.r[0], &c.r[1], &c.r[2], &c.r[3]};
  int best = 0;
  double bw = rn_span_div(c.la.hi, c.la.lo, 1.0);
#pragma unroll
  for (int k = 1; k < 5; ++k) {
    const double w = rn_span_div(d[k]->hi, d[k]->lo, 1.0);
    if (w > bw) {
      bw = w;
      best = k;
    }
  }
  *d[best] = subinterval(*d[best], (int)(i & 1), 2);
  out[i] = c;
}

// Distances of the final critical cells from the theoretical worst cases
// (kernel_compute_distance_from_theoretical_worst_cases,
// tight_rectangle.cu:343-456): one lane walks the cells in order, as the
// reference's <<<1,1>>> launch does, so the "dominating critical" records
// are the same.  Output layout: struct tr_maxdist below (host mirror in
// host/prover_api.cpp).
struct tr_maxdist {
  int have_2d, have_3d;
  iv la_2d, la_3d;
  iv d_r1_2d;
  cell5 d_r1_2d_crit[2];
  iv d_r2_2d;
  cell5 d_r2_2d_crit[2];
  double lb_h1_2d, ub_la_by_S1, ub_R3_2d;
  iv d_r1_3d;
  cell5 d_r1_3d_crit[2];
  iv d_r2_3d;
  cell5 d_r2_3d_crit[2];
  iv d_r3_3d;
  cell5 d_r3_3d_crit[2];
  double ub_R4_3d;
  iv S1, h2, S3;
  double pocket_3d;
  int trap;
};

extern "C" __global__ void ccrd_tr_max_distances(const cell5* __restrict__ crit, uint64_t n,
                                                 const Konst* __restrict__ kp, tr_maxdist* __restrict__ o) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Konst K = *kp;
  Trap tp{0};
  tr_maxdist r;
  r.have_2d = r.have_3d = 0;
  r.la_3d = pt(1.0);
  r.la_2d = pt(2.0);
  r.d_r1_2d = r.d_r2_2d = pt(0.0);
  r.ub_R3_2d = r.ub_R4_3d = 0.0;
  r.d_r1_3d = r.d_r2_3d = r.d_r3_3d = pt(0.0);
  r.pocket_3d = -1.0;
  r.lb_h1_2d = 1.0;
  r.ub_la_by_S1 = 1.0;
  const cell5 zero{pt(0.0), {pt(0.0), pt(0.0), pt(0.0), pt(0.0)}};
  for (int k = 0; k < 2; ++k)
    r.d_r1_2d_crit[k] = r.d_r2_2d_crit[k] = r.d_r1_3d_crit[k] = r.d_r2_3d_crit[k] = r.d_r3_3d_crit[k] = zero;
  r.S1 = r.h2 = r.S3 = pt(0.0);
  for (uint64_t i = 0; i < n; ++i) {
    const cell5 v = crit[i];
    if (v.r[2].hi < 0.1) {  // two-disk critical: r1 the circumcircle, r2^2 = 1/4
      r.have_2d = 1;
      join_into(r.la_2d, v.la);
      const iv r1t = 0.25 * (square(v.la) + 1.0);
      const iv d1 = v.r[0] - r1t;
      if (d1.lo < r.d_r1_2d.lo) r.d_r1_2d_crit[0] = v;
      if (d1.hi > r.d_r1_2d.hi) r.d_r1_2d_crit[1] = v;
      join_into(r.d_r1_2d, d1);
      const iv d2 = v.r[1] - pt(0.25);
      if (d2.lo < r.d_r2_2d.lo) r.d_r2_2d_crit[0] = v;
      if (d2.hi > r.d_r2_2d.hi) r.d_r2_2d_crit[1] = v;
      join_into(r.d_r2_2d, d2);
      const iv R3 = weight_for(K, v.la) - v.r[0] - v.r[1];
      if (R3.hi > r.ub_R3_2d) r.ub_R3_2d = R3.hi;
      const double S1 = mul_rd(2.0, sqrt_rd(add_rd(v.r[0].lo, -0.25)));
      const double q = div_ru(v.la.hi, S1);
      if (q > r.ub_la_by_S1) r.ub_la_by_S1 = q;
      const double half = add_ru(v.la.hi, -mul_rd(0.5, S1));
      const double h1 = mul_rd(2.0, sqrt_rd(add_rd(v.r[0].lo, -mul_ru(half, half))));
      if (h1 < r.lb_h1_2d) r.lb_h1_2d = h1;
    } else {  // three-disk critical
      r.have_3d = 1;
      join_into(r.la_3d, v.la);
      const iv la2 = square(v.la);
      const iv la4 = square(la2);
      const iv th = (16.0 * la4 + 40.0 * la2 + 9.0) / (256.0 * la2);
      const iv d1 = v.r[0] - th;
      if (d1.lo < r.d_r1_3d.lo) r.d_r1_3d_crit[0] = v;
      if (d1.hi > r.d_r1_3d.hi) r.d_r1_3d_crit[1] = v;
      const iv d2 = v.r[1] - th;
      if (d2.lo < r.d_r2_3d.lo) r.d_r2_3d_crit[0] = v;
      if (d2.hi > r.d_r2_3d.hi) r.d_r2_3d_crit[1] = v;
      const iv d3 = v.r[2] - th;
      if (d3.lo < r.d_r3_3d.lo) r.d_r3_3d_crit[0] = v;
      if (d3.hi > r.d_r3_3d.hi) r.d_r3_3d_crit[1] = v;
      join_into(r.d_r1_3d, d1);
      join_into(r.d_r2_3d, d2);
      join_into(r.d_r3_3d, d3);
      const iv R4 = weight_for(K, v.la) - v.r[0] - v.r[1] - v.r[2];
      if (R4.hi > r.ub_R4_3d) r.ub_R4_3d = R4.hi;
      const iv S1 = sqrt(4.0 * v.r[0] - 1.0, tp);
      const iv h2 = sqrt(4.0 * v.r[1] - square(v.la - S1), tp);
      const iv S3 = sqrt(4.0 * v.r[2] - square(1.0 - h2), tp);
      const double pw = (v.la - S1 - S3).hi;
      if (pw > r.pocket_3d) {
        if (r.pocket_3d < 0.0) {
          r.S1 = S1;
          r.h2 = h2;
          r.S3 = S3;
        }
        r.pocket_3d = pw;
      }
      join_into(r.S1, S1);
      join_into(r.h2, h2);
      join_into(r.S3, S3);
    }
  }
  r.trap = tp.flag;
  *o = r;
}
