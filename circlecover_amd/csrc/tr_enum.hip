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
// tight_rectangle.cu:343-456), as a reduction over one workgroup.  The
// reference walks the cells in order on one thread.  Each of its running
// extremes starts from a fixed value and takes a cell's value only if it is
// strictly beyond (so NaN never enters), and its "dominating" cells are the
// ones that set the final extremes: every extreme is therefore the pair
// (value, first cell attaining it), position 0 standing for the start value,
// and pairs combine by value and then by the lower position, which gives the
// same value bits (also for zeros of either sign) and the same cell in any
// order of combination.  S1, h2 and S3 restart from a cell's own values
// whenever its pocket width exceeds a running maximum that is still negative;
// the last restart is at the first cell whose width is >= 0, or, without
// one, at the first cell with the largest width above -1.  A first pass finds
// that cell, a second forms the three hulls from it on.  Output layout:
// struct tr_maxdist below (host mirror in host/prover_api.cpp).
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

// Lanes of the workgroup (one launch of one workgroup).
constexpr int kMdLanes = 256;
constexpr uint64_t kMdNone = ~0ull;  // no value yet

// (value, position): position 0 is the start value, cell i is i + 1.
struct md_best {
  double v;
  uint64_t i;
};
// A lane meets its cells in increasing order: a strictly better value wins.
CCD void md_lane_min(md_best& a, double v, uint64_t i) {
  if (a.i == kMdNone || v < a.v) a = md_best{v, i};
}
CCD void md_lane_max(md_best& a, double v, uint64_t i) {
  if (a.i == kMdNone || v > a.v) a = md_best{v, i};
}
// Two partial results: the better value, then the earlier position.
template <bool Min>
CCD void md_merge(md_best& a, md_best b) {
  if (b.i == kMdNone) return;
  const bool better = Min ? b.v < a.v : b.v > a.v;
  if (a.i == kMdNone || better || (b.v == a.v && b.i < a.i)) a = b;
}
template <bool Min>
CCD md_best md_reduce(md_best a, md_best* lds) {
  for (int off = 32; off > 0; off >>= 1) {
    md_best b;
    b.v = __shfl_xor(a.v, off);
    b.i = (uint64_t)__shfl_xor((unsigned long long)a.i, off);
    md_merge<Min>(a, b);
  }
  __syncthreads();  // lds is reused by the next reduction
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
  __syncthreads();
  a = lds[0];
  for (int w = 1; w < kMdLanes / 64; ++w) md_merge<Min>(a, lds[w]);
  return a;
}

// The reference's per-cell terms (tight_rectangle.cu:372-387 and :389-414).
struct md_two {
  iv d1, d2;
  double R3_hi, la_by_S1, h1;
};
CCD md_two md_two_disk(const Konst& K, const cell5& v) {
  md_two t;
  const iv r1t = 0.25 * (square(v.la) + 1.0);
  t.d1 = v.r[0] - r1t;
  t.d2 = v.r[1] - pt(0.25);
  t.R3_hi = (weight_for(K, v.la) - v.r[0] - v.r[1]).hi;
  const double S1 = mul_rd(2.0, sqrt_rd(add_rd(v.r[0].lo, -0.25)));
  t.la_by_S1 = div_ru(v.la.hi, S1);
  const double half = add_ru(v.la.hi, -mul_rd(0.5, S1));
  t.h1 = mul_rd(2.0, sqrt_rd(add_rd(v.r[0].lo, -mul_ru(half, half))));
  return t;
}
struct md_three {
  iv d1, d2, d3;
  double R4_hi;
  iv S1, h2, S3;
  double pw;
};
CCD md_three md_three_disk(const Konst& K, const cell5& v, Trap& tp) {
  md_three t;
  const iv la2 = square(v.la);
  const iv la4 = square(la2);
  const iv th = (16.0 * la4 + 40.0 * la2 + 9.0) / (256.0 * la2);
  t.d1 = v.r[0] - th;
  t.d2 = v.r[1] - th;
  t.d3 = v.r[2] - th;
  t.R4_hi = (weight_for(K, v.la) - v.r[0] - v.r[1] - v.r[2]).hi;
  t.S1 = sqrt(4.0 * v.r[0] - 1.0, tp);
  t.h2 = sqrt(4.0 * v.r[1] - square(v.la - t.S1), tp);
  t.S3 = sqrt(4.0 * v.r[2] - square(1.0 - t.h2), tp);
  t.pw = (v.la - t.S1 - t.S3).hi;
  return t;
}
CCD bool md_is_two_disk(const cell5& v) { return v.r[2].hi < 0.1; }  // r1 the circumcircle, r2^2 = 1/4

extern "C" __global__ void __launch_bounds__(kMdLanes)
    ccrd_tr_max_distances(const cell5* __restrict__ crit, uint64_t n, const Konst* __restrict__ kp,
                          tr_maxdist* __restrict__ o) {
  __shared__ md_best lds[kMdLanes / 64];
  const Konst K = *kp;
  const uint64_t lane = threadIdx.x;
  Trap tp{0};
  // start values of the reference's running extremes, at position 0
  md_best have2{0.0, 0}, la2_lo{2.0, 0}, la2_hi{2.0, 0}, d1_2_lo{0.0, 0}, d1_2_hi{0.0, 0}, d2_2_lo{0.0, 0},
      d2_2_hi{0.0, 0}, R3{0.0, 0}, laS1{1.0, 0}, h1{1.0, 0};
  md_best have3{0.0, 0}, la3_lo{1.0, 0}, la3_hi{1.0, 0}, d1_3_lo{0.0, 0}, d1_3_hi{0.0, 0}, d2_3_lo{0.0, 0},
      d2_3_hi{0.0, 0}, d3_3_lo{0.0, 0}, d3_3_hi{0.0, 0}, R4{0.0, 0}, pocket{-1.0, 0};
  md_best first_nonneg{0.0, kMdNone};  // first cell with pocket width >= 0 (value unused)
  for (uint64_t i = lane; i < n; i += kMdLanes) {
    const cell5 v = crit[i];
    const uint64_t p = i + 1;
    if (md_is_two_disk(v)) {
      const md_two t = md_two_disk(K, v);
      have2.v = 1.0;
      md_lane_min(la2_lo, v.la.lo, p);
      md_lane_max(la2_hi, v.la.hi, p);
      md_lane_min(d1_2_lo, t.d1.lo, p);
      md_lane_max(d1_2_hi, t.d1.hi, p);
      md_lane_min(d2_2_lo, t.d2.lo, p);
      md_lane_max(d2_2_hi, t.d2.hi, p);
      md_lane_max(R3, t.R3_hi, p);
      md_lane_max(laS1, t.la_by_S1, p);
      md_lane_min(h1, t.h1, p);
    } else {
      const md_three t = md_three_disk(K, v, tp);
      have3.v = 1.0;
      md_lane_min(la3_lo, v.la.lo, p);
      md_lane_max(la3_hi, v.la.hi, p);
      md_lane_min(d1_3_lo, t.d1.lo, p);
      md_lane_max(d1_3_hi, t.d1.hi, p);
      md_lane_min(d2_3_lo, t.d2.lo, p);
      md_lane_max(d2_3_hi, t.d2.hi, p);
      md_lane_min(d3_3_lo, t.d3.lo, p);
      md_lane_max(d3_3_hi, t.d3.hi, p);
      md_lane_max(R4, t.R4_hi, p);
      md_lane_max(pocket, t.pw, p);
      if (t.pw >= 0.0 && first_nonneg.i == kMdNone) first_nonneg.i = p;
    }
  }
  md_best trap{(double)tp.flag, 0};
  trap = md_reduce<false>(trap, lds);
  have2 = md_reduce<false>(have2, lds);
  have3 = md_reduce<false>(have3, lds);
  la2_lo = md_reduce<true>(la2_lo, lds);
  la2_hi = md_reduce<false>(la2_hi, lds);
  d1_2_lo = md_reduce<true>(d1_2_lo, lds);
  d1_2_hi = md_reduce<false>(d1_2_hi, lds);
  d2_2_lo = md_reduce<true>(d2_2_lo, lds);
  d2_2_hi = md_reduce<false>(d2_2_hi, lds);
  R3 = md_reduce<false>(R3, lds);
  laS1 = md_reduce<false>(laS1, lds);
  h1 = md_reduce<true>(h1, lds);
  la3_lo = md_reduce<true>(la3_lo, lds);
  la3_hi = md_reduce<false>(la3_hi, lds);
  d1_3_lo = md_reduce<true>(d1_3_lo, lds);
  d1_3_hi = md_reduce<false>(d1_3_hi, lds);
  d2_3_lo = md_reduce<true>(d2_3_lo, lds);
  d2_3_hi = md_reduce<false>(d2_3_hi, lds);
  d3_3_lo = md_reduce<true>(d3_3_lo, lds);
  d3_3_hi = md_reduce<false>(d3_3_hi, lds);
  R4 = md_reduce<false>(R4, lds);
  pocket = md_reduce<false>(pocket, lds);
  first_nonneg.v = first_nonneg.i == kMdNone ? 0.0 : -(double)first_nonneg.i;
  first_nonneg = md_reduce<false>(first_nonneg, lds);  // the lowest position
  // The last restart of S1, h2, S3 (position 0: none, the hulls start at 0).
  const uint64_t k = first_nonneg.i != kMdNone ? first_nonneg.i : (pocket.v > -1.0 ? pocket.i : 0);
  md_best S_lo[3], S_hi[3];
  for (int q = 0; q < 3; ++q) S_lo[q] = S_hi[q] = k ? md_best{0.0, kMdNone} : md_best{0.0, 0};
  for (uint64_t i = lane; i < n; i += kMdLanes) {
    if (i + 1 < k) continue;
    const cell5 v = crit[i];
    if (md_is_two_disk(v)) continue;
    Trap unused{0};
    const md_three t = md_three_disk(K, v, unused);
    const iv S[3] = {t.S1, t.h2, t.S3};
    for (int q = 0; q < 3; ++q) {
      md_lane_min(S_lo[q], S[q].lo, i + 1);
      md_lane_max(S_hi[q], S[q].hi, i + 1);
    }
  }
  for (int q = 0; q < 3; ++q) {
    S_lo[q] = md_reduce<true>(S_lo[q], lds);
    S_hi[q] = md_reduce<false>(S_hi[q], lds);
  }
  if (lane != 0) return;
  const cell5 zero{pt(0.0), {pt(0.0), pt(0.0), pt(0.0), pt(0.0)}};
  auto cell_at = [&](uint64_t p) { return p ? crit[p - 1] : zero; };
  tr_maxdist r;
  r.have_2d = have2.v != 0.0;
  r.have_3d = have3.v != 0.0;
  r.la_2d = iv{la2_lo.v, la2_hi.v};
  r.la_3d = iv{la3_lo.v, la3_hi.v};
  r.d_r1_2d = iv{d1_2_lo.v, d1_2_hi.v};
  r.d_r1_2d_crit[0] = cell_at(d1_2_lo.i);
  r.d_r1_2d_crit[1] = cell_at(d1_2_hi.i);
  r.d_r2_2d = iv{d2_2_lo.v, d2_2_hi.v};
  r.d_r2_2d_crit[0] = cell_at(d2_2_lo.i);
  r.d_r2_2d_crit[1] = cell_at(d2_2_hi.i);
  r.lb_h1_2d = h1.v;
  r.ub_la_by_S1 = laS1.v;
  r.ub_R3_2d = R3.v;
  r.d_r1_3d = iv{d1_3_lo.v, d1_3_hi.v};
  r.d_r1_3d_crit[0] = cell_at(d1_3_lo.i);
  r.d_r1_3d_crit[1] = cell_at(d1_3_hi.i);
  r.d_r2_3d = iv{d2_3_lo.v, d2_3_hi.v};
  r.d_r2_3d_crit[0] = cell_at(d2_3_lo.i);
  r.d_r2_3d_crit[1] = cell_at(d2_3_hi.i);
  r.d_r3_3d = iv{d3_3_lo.v, d3_3_hi.v};
  r.d_r3_3d_crit[0] = cell_at(d3_3_lo.i);
  r.d_r3_3d_crit[1] = cell_at(d3_3_hi.i);
  r.ub_R4_3d = R4.v;
  r.S1 = iv{S_lo[0].v, S_hi[0].v};
  r.h2 = iv{S_lo[1].v, S_hi[1].v};
  r.S3 = iv{S_lo[2].v, S_hi[2].v};
  r.pocket_3d = pocket.v;
  r.trap = trap.v != 0.0;
  *o = r;
}
