// rectangle_size_bound leaf-criterion chain C1..C14 for the gfx950 kernels.
//
// Same rounded-operation sequence, branches and comparisons as the cited
// reference lines; GPU-side structure:
//   * the cell-independent Intermediate_values constants come from the
//     uniform Konst block (SGPRs), only R is per cell (vals7);
//   * of the 2x2 cover only what C8 reads later (width, circle of disk 4)
//     is kept live;
//   * trap sites of the reference set the per-lane Trap flag and return
//     false / 0 / "r7 unchanged" at that site;
//   * the chain is split into separately launched stages (see
//     prover_rd.hip) so each kernel's register budget covers only its stage.
#pragma once
// CC_EXACT_SHORTCUTS (0 by default) enables early exits that provably return
// what the full loops return (see three_height); the branch-and-bound's
// batched kernels are built with it.
#ifndef CC_EXACT_SHORTCUTS
#define CC_EXACT_SHORTCUTS 0
#endif
#include "rsb.hpp"
// The leaf chain's larger steps are out-of-line functions (smaller code for
// the kernels that run few leaves); CC_CHAIN_FN inlines them where a kernel
// runs the chain on very many leaves.
#ifndef CC_CHAIN_FN
#define CC_CHAIN_FN static __device__ __noinline__
#endif

namespace ccd {

struct cell7 {
  iv la;
  iv r[6];
};
struct vals7 {
  iv R;
  double cover_width;  // Two_by_two_cover::width
  circle cover_c4;     // Two_by_two_cover::circles[3]
};
struct strip2 {
  circle c1, c2;
  double h;
};

namespace rsbc {

CCD double allowed_weight(double h_lb) { return mul_rd(kDiskWeightMax, mul_rd(h_lb, h_lb)); }  // :34-37

// compute_efficient_rectangle_cover_weight_range.cu:25-35
CCD iv efficient_cover_range(double lb_w, iv h) {
  double wmax = mul_rd(kRsbRatio, mul_rd(h.lo, lb_w));
  double lo_esq = mul_ru(mul_ru(kRsbRatio, kRsbRatio), mul_ru(h.hi, h.hi));
  double hi_esq = mul_rd(mul_rd(kRsbRatio, kRsbRatio), mul_rd(h.lo, h.lo));
  double root = sqrt_rd(add_rd(4.0, -rcp_ru(mul_rd(kRsbRatio, kRsbRatio))));
  double lo = mul_ru(lo_esq, add_ru(2.0, -root));
  double hi = mul_rd(hi_esq, add_rd(2.0, root));
  return iv{lo, wmax < hi ? wmax : hi};
}

// bound_critical_ratio: can_recurse.cu:90-161
static __device__ __noinline__ double bound_crit_ratio(const Konst& K, double lb_rem, double ub_largest, double ub_w,
                                                double ub_h) {
  const double inf = __builtin_inf();
  if (lb_rem < 0.0) return inf;
  if (ub_w <= 0.0 || ub_h <= 0.0) return 0.0;
  double smin = ub_w < ub_h ? ub_w : ub_h;
  double smax = ub_w < ub_h ? ub_h : ub_w;
  double best = inf;
  double area = mul_ru(ub_w, ub_h);
  double req_side = sqrt_ru(div_ru(ub_largest, kDiskWeightMax));
  if (smin >= req_side) return mul_ru(kRsbRatio, area) <= lb_rem ? kRsbRatio : inf;
  if (req_side > smax) {
    double sq_side = sqrt_rd(div_rd(lb_rem, kRsbRatio));
    if (sq_side >= req_side) best = div_ru(mul_ru(mul_ru(req_side, req_side), kRsbRatio), mul_rd(ub_w, ub_h));
  } else {
    double req = mul_ru(kRsbRatio, mul_ru(req_side, smax));
    if (req <= lb_rem) best = div_ru(req, mul_rd(ub_w, ub_h));
  }
  double sig = div_ru(ub_largest, mul_rd(smin, smin));
  sig = mul_ru(sig, sig);
  double eff = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(sig, 1.0)), 1.0)));
  if (eff < kWorstRatio) eff = kWorstRatio;
  if (eff < best && mul_ru(eff, area) <= lb_rem) best = eff;
  double cr = trc::crit_ratio(K, iv{div_rd(smax, smin), div_ru(smax, smin)}).hi;
  if (cr < best && mul_ru(cr, area) <= lb_rem) best = cr;
  return best;
}

// recursion_bound_size: can_recurse.cu:163-218
static __device__ __noinline__ double recursion_bound(const Konst& K, double lb_rem, double ub_w, double ub_h) {
  const double inf = __builtin_inf();
  if (lb_rem <= 0.0) {
    if (lb_rem < 0.0) return 0.0;
    return (ub_w <= 0.0 || ub_h <= 0.0) ? inf : 0.0;
  }
  if (ub_w <= 0.0 || ub_h <= 0.0) return inf;
  double max_cr = div_rd(lb_rem, mul_ru(ub_w, ub_h));
  double smin = ub_w < ub_h ? ub_w : ub_h;
  double smax = ub_w < ub_h ? ub_h : ub_w;
  double cr_wc = trc::crit_ratio(K, iv{div_rd(smax, smin), div_ru(smax, smin)}).hi;
  if (cr_wc <= max_cr) return inf;
  double sb_area = div_rd(lb_rem, kRsbRatio);
  double side = div_rd(sb_area, smax);
  if (side < smin) return 0.0;
  if (side > smax) side = sqrt_rd(sb_area);
  double best = mul_rd(kDiskWeightMax, mul_rd(side, side));
  if (max_cr >= kWorstRatio) {
    double lam = mul_rd(2.0, max_cr);
    lam = add_rd(lam, sqrt_rd(add_rd(mul_rd(4.0, mul_rd(max_cr, max_cr)), -2.0)));
    double sigma = mul_rd(max_cr, add_rd(lam, -div_ru(2.0, lam)));
    double scaled = mul_rd(sigma, mul_rd(smin, smin));
    if (scaled >= best) best = scaled;
  }
  return best;
}

// can_recurse_worst_case: can_recurse.cu:237-261
CCD bool can_recurse_wc(double lb_rem, double ub_w, double ub_h) {
  double lam, l, s;
  if (ub_w < ub_h) {
    lam = div_ru(ub_h, ub_w);
    l = ub_h;
    s = ub_w;
  } else {
    lam = div_ru(ub_w, ub_h);
    l = ub_w;
    s = ub_h;
  }
  if (lam <= kRsbToleratedLambda) return mul_ru(mul_ru(ub_w, ub_h), kWorstRatio) <= lb_rem;
  l = mul_ru(l, l);
  s = mul_ru(s, s);
  return add_ru(mul_ru(0.5, s), mul_ru(0.25, l)) <= lb_rem;
}

// Width of an h-tall strip covered by a disk of weight lb (4*lb - h^2), the
// building block of maximize_covered.cu:28-51, :171-237.
CCD double chord(double q4, double hsq) {
  double w = add_rd(q4, -hsq);
  return w <= 0 ? 0.0 : sqrt_rd(w);
}
// nd_can_cover_no_small: maximize_covered.cu:28-51 (d = lower bounds)
CCD bool nd_can_cover_no_small(double ub_w, double ub_h, const double* d, int n) {
  double htot = 0.0, wtot = 0.0;
  double wsq = mul_ru(ub_w, ub_w), hsq = mul_ru(ub_h, ub_h);
  for (int i = 0; i < n; ++i) {
    double q = mul_rd(4.0, d[i]);
    double wc = chord(q, hsq);
    double hc = chord(q, wsq);
    wtot = add_rd(wtot, wc);
    htot = add_rd(htot, hc);
  }
  return htot >= ub_h || wtot >= ub_w;
}
#if CC_EXACT_SHORTCUTS
// The reference's bisection loop `mid = RN((lo + hi) / 2); stop once mid is
// not strictly inside (lo, hi); test(mid) ? lo = mid : hi = mid` keeps lo at
// a passing midpoint (or its start) and hi at a failing one (or its start)
// until no double lies between them.  For a test that passes, then fails,
// on (lo, hi) it therefore ends at the largest passing double of (lo, hi),
// or at lo if none passes.  That value is found here with one test at the
// double above lo and a bisection of the bit patterns, at most 64 tests,
// where the loop needs up to about 1075 when that value is tiny or lo itself.
// Requires 0 <= lo < hi < inf.
template <class Test>
CCD double monotone_last_pass(double lo, double hi, Test test) {
  uint64_t L = lo == 0.0 ? 0u : bits(lo), H = bits(hi);  // -0.0 orders as +0.0
  if (H - L <= 1u || !test(from_bits(L + 1u))) return lo;
  ++L;
  while (H - L > 1u) {
    const uint64_t M = L + (H - L) / 2u;
    if (test(from_bits(M))) L = M;
    else H = M;
  }
  return from_bits(L);
}
#endif

// three_disks_can_cover: maximize_covered.cu:171-196
CCD bool three_can_cover(double r1, double r2, double r3, double w, double h) {
  double hsq = mul_ru(h, h);
  double w3 = chord(mul_rd(4.0, r3), hsq);
  double w2 = chord(mul_rd(4.0, r2), hsq);
  double w1 = chord(mul_rd(4.0, r1), hsq);
  return add_rd(w1, add_rd(w2, w3)) >= w;
}
#if CC_EXACT_SHORTCUTS
// Real-arithmetic estimate of the height at which three chords cover ub_w.
// In the third chord's half-width s = sqrt(q3 - h^2) the covered width is
// G(s) = s + sum over i = 1, 2 of sqrt(max(q_i - q3 + s^2, 0)), increasing
// with G' >= 1 and convex when q3 <= q1, q2 (the callers' order), so
// Newton's method from s = sqrt(q3) (h = 0) descends to the root without
// overshoot; a bracket keeps the other case in range.  Only the starting
// point of three_height's search: its accuracy changes the number of tests
// there, never the result.
CCD double three_height_estimate(double q1, double q2, double q3, double ub_w) {
  const double c1 = q1 - q3, c2 = q2 - q3;
  double s_lo = 0.0, s_hi = __builtin_sqrt(q3), s = s_hi;
  for (int it = 0; it < 10; ++it) {
    const double ss = s * s;
    double g = s - ub_w, dg = 1.0;
    if (c1 + ss > 0.0) {
      const double t = __builtin_sqrt(c1 + ss);
      g += t;
      dg += s / t;
    }
    if (c2 + ss > 0.0) {
      const double t = __builtin_sqrt(c2 + ss);
      g += t;
      dg += s / t;
    }
    if (g >= 0.0) s_hi = s;  // still covers: the root lies at a smaller s (a larger h)
    else s_lo = s;
    double next = s - g / dg;
    const bool done = __builtin_fabs(next - s) <= 0x1p-52 * s;
    if (!done && !(next > s_lo && next < s_hi)) next = 0.5 * (s_lo + s_hi);
    s = next;
    if (done) break;
  }
  const double u = q3 - s * s;
  return u > 0.0 ? __builtin_sqrt(u) : 0.0;
}
// The largest double h in (0, hi) at which the chords cover ub_w (their
// covered width, non-increasing in h, is >= ub_w), or 0 if there is none:
// the value three_height's bisection loop returns (see monotone_last_pass).
// Searched over bit patterns from the estimate: a test there, a gallop of
// doubling steps to a test of the other outcome, then a bisection of the
// bit patterns between the two, about 2 log2(d) + 2 tests for an estimate d
// doubles off (d is a few doubles for most operands), where the loop needs
// about 50.  Requires 0 < hi < inf.
CCD double three_height_search(double q1, double q2, double q3, double ub_w, double hi) {
  const uint64_t H = bits(hi);
  auto covers = [&](uint64_t B) {
    const double h = from_bits(B);
    const double hsq = mul_ru(h, h);
    return !(add_rd(chord(q1, hsq), add_rd(chord(q2, hsq), chord(q3, hsq))) < ub_w);
  };
  if (H <= 1u) return 0.0;            // no double lies in (0, hi)
  if (covers(H - 1u)) return from_bits(H - 1u);  // every double of (0, hi) passes
  uint64_t E = bits(three_height_estimate(q1, q2, q3, ub_w));
  if (E == 0u || E >= H) E = H / 2u;  // estimate outside (0, hi), or NaN: start mid-range
  uint64_t lo, up;  // covers(lo) or lo == 0; !covers(up) or up == H
  if (covers(E)) {
    lo = E;
    up = H;
    for (uint64_t step = 1u;; step <<= 1) {
      if (H - lo <= step) break;
      if (!covers(lo + step)) {
        up = lo + step;
        break;
      }
      lo += step;
    }
  } else {
    up = E;
    lo = 0u;
    for (uint64_t step = 1u;; step <<= 1) {
      if (up <= step) break;
      if (covers(up - step)) {
        lo = up - step;
        break;
      }
      up -= step;
    }
  }
  while (up - lo > 1u) {
    const uint64_t M = lo + (up - lo) / 2u;
    if (covers(M)) lo = M;
    else up = M;
  }
  return from_bits(lo);
}
#endif

// three_disks_maximize_height: maximize_covered.cu:198-237
static __device__ __noinline__ double three_height(double r1, double r2, double r3, double ub_w) {
  const double q1 = mul_rd(4.0, r1), q2 = mul_rd(4.0, r2), q3 = mul_rd(4.0, r3);
  double lo = 0.0, hi = mul_rd(2.0, sqrt_rd(r3));
#if CC_EXACT_SHORTCUTS
  // The covered width sum(chord(q_i, RU(h*h))) is non-increasing in h >= 0
  // (every operation is monotone), every midpoint the loop below visits is a
  // positive double, and lo stays 0 as long as the test fails.  So when the
  // test already fails at the smallest positive double, the loop would halve
  // hi down to zero (up to ~1075 iterations) and return 0: return it now.
  // hi > 0 excludes NaN bounds, for which the loop behaves differently.
  // Otherwise the loop ends at the largest double of (0, hi) that passes
  // (monotone_last_pass), which three_height_search finds with fewer tests
  // (finite operands: the test is then monotone and defined throughout).
  if (hi > 0.0) {
    const double tiny = 0x1p-1074;
    const double hsq = mul_ru(tiny, tiny);
    if (add_rd(chord(q1, hsq), add_rd(chord(q2, hsq), chord(q3, hsq))) < ub_w) return 0.0;
    if (hi < __builtin_inf() && __builtin_isfinite(q1) && __builtin_isfinite(q2) && __builtin_isfinite(q3) &&
        !__builtin_isnan(ub_w))
      return three_height_search(q1, q2, q3, ub_w, hi);
  }
#endif
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) return lo;
    double hsq = mul_ru(mid, mid);
    double w3 = chord(q3, hsq);
    double w2 = chord(q2, hsq);
    double w1 = chord(q1, hsq);
    if (add_rd(w1, add_rd(w2, w3)) < ub_w) hi = mid;
    else lo = mid;
  }
  return lo;
}
// nd_can_cover: maximize_covered.cu:53-71
CCD bool nd_can_cover(double ub_w, double ub_h, const double* d, int n) {
  if (n == 1) return mul_rd(4.0, d[0]) >= add_ru(mul_ru(ub_w, ub_w), mul_ru(ub_h, ub_h));
  if (n == 2) {
    double h = trc::two_disk_height(d[0], d[1], ub_w);
    double w = trc::two_disk_height(d[0], d[1], ub_h);
    return h >= ub_h || w >= ub_w;
  }
  if (n == 3) return three_can_cover(d[0], d[1], d[2], ub_w, ub_h);
  return nd_can_cover_no_small(ub_w, ub_h, d, n);
}
// nd_maximize_covered: maximize_covered.cu:73-104
static __device__ __noinline__ double nd_maximize(double ub_w, const double* d, int n) {
  if (n == 1) {
    double h = add_rd(mul_rd(4.0, d[0]), -mul_ru(ub_w, ub_w));
    return h <= 0 ? 0.0 : sqrt_rd(h);
  }
  if (n == 2) return trc::two_disk_height(d[0], d[1], ub_w);
  if (n == 3) return three_height(d[0], d[1], d[2], ub_w);
  double hi = mul_rd(2.0, sqrt_rd(d[n - 1]));
  double lo = 0.0;
#if CC_EXACT_SHORTCUTS
  // nd_can_cover_no_small(ub_w, h) = (sum of chords at ub_w >= h) or
  // (sum of chords at h >= ub_w): both parts are non-increasing in h, so a
  // test that fails at the smallest positive double fails at every midpoint
  // and the loop returns 0 (as in three_height).
  if (hi > 0.0 && !nd_can_cover_no_small(ub_w, 0x1p-1074, d, n)) return 0.0;
#endif
  for (int it = 0; it < kBisectCap; ++it) {
    double h = rn_mid(lo, hi);
    if (h <= lo || h >= hi) return lo;
    if (nd_can_cover_no_small(ub_w, h, d, n)) lo = h;
    else hi = h;
  }
  return lo;
}

// two_disks_maximal_height_strip: maximize_covered.cu:132-150
CCD strip2 max_height_strip(double la, double r1, double r2, Trap& tp) {
  iv rdiff{add_rd(r1, -r2), add_ru(r1, -r2)};
  iv wdiff = (2.0 * rdiff) / la;
  iv w1 = pt(0.5) * la + wdiff;
  iv w2 = pt(0.5) * la - wdiff;
  iv h1 = r1 - 0.25 * square(w1);
  iv h2 = r2 - 0.25 * square(w2);
  const iv z = pt(0.0);
  if (possibly(h1 <= 0) || possibly(h2 <= 0)) return strip2{circle{{z, z}, pt(r1)}, circle{{z, z}, pt(r2)}, 0.0};
  iv h = sqrt(h1, tp);
  return strip2{circle{{0.5 * w1, h}, pt(r1)}, circle{{la - 0.5 * w2, h}, pt(r2)}, mul_rd(2.0, h.lo)};
}
// two_disks_maximal_height_strip_wc_recursion: maximize_covered.cu:239-278
static __device__ __noinline__ strip2 max_height_strip_wc(double lr1, double lr2, double lb_R, double ub_w) {
  double lo = 0.0, hi = mul_rd(2.0, lr2);
  iv c1x = pt(0.0), c2x = pt(0.0);
  const double q1 = mul_rd(4.0, lr1), q2 = mul_rd(4.0, lr2);
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) break;
    double hsq = mul_ru(mid, mid);
    double w1 = add_rd(q1, -hsq);
    double w2 = add_rd(q2, -hsq);
    if (w1 < 0 || w2 < 0) {
      hi = mid;
      continue;
    }
    w1 = sqrt_rd(w1);
    w2 = sqrt_rd(w2);
    double wrem = add_ru(ub_w, -add_rd(w1, w2));
    if (wrem <= 0.0 || can_recurse_wc(lb_R, wrem, mid)) {
      lo = mid;
      c1x = iv{mul_rd(0.5, w1), mul_ru(0.5, w1)};
      c2x = iv{add_rd(ub_w, -mul_ru(0.5, w2)), add_ru(ub_w, -mul_rd(0.5, w2))};
    } else {
      hi = mid;
    }
  }
  const iv z = pt(0.0);
  if (lo <= 0.0) return strip2{circle{{z, z}, z}, circle{{z, z}, z}, 0.0};
  iv y{mul_rd(0.5, lo), mul_ru(0.5, lo)};
  return strip2{circle{{c1x, y}, pt(lr1)}, circle{{c2x, y}, pt(lr2)}, lo};
}

// ------------------------------------------------------------------ cell prep
// variables.hpp:53-66 (radii[0].lb is never raised)
CCD cell7 tighten(const cell7& c) {
  cell7 r = c;
  double lb = 0.0;
#pragma unroll
  for (int i = 5; i > 0; --i) {
    if (lb < c.r[i].lo) lb = c.r[i].lo;
    else r.r[i].lo = lb;
  }
  return r;
}
// variables.hpp:68-80
CCD bool is_empty(const cell7& c) {
  if (empty(c.la)) return true;
#pragma unroll
  for (int i = 0; i < 6; ++i)
    if (c.r[i].lo >= c.r[i].hi) return true;
  return false;
}
// variables.cuh:32-40: the per-cell part of Intermediate_values.
CCD iv remaining_weight(const cell7& c) {
  iv R = c.la * kRsbRatio;
#pragma unroll
  for (int i = 0; i < 6; ++i) R = R - c.r[i];
  if (R.lo < 0.0) R.lo = 0.0;
  return R;
}

// ------------------------------------------------------------------ C1
// compute_even_split_distance_bound.cu:25-76 + even_split_recursion.cu:25-40
CCD bool c1_even_split(const cell7& c, const vals7& v) {
  iv g1 = c.r[0];
  iv g2 = c.r[1] + c.r[2];
#pragma unroll
  for (int i = 3; i < 6; ++i) {
    if (g1.hi <= g2.hi) g1 = g1 + c.r[i];
    else g2 = g2 + c.r[i];
  }
  double d = add_ru(g1.hi, -g2.lo), d2 = add_ru(g2.hi, -g1.lo);
  if (d < d2) d = d2;
  double dist;
  if (add_ru(d, -c.r[5].lo) <= v.R.lo) {
    dist = c.r[5].hi;
  } else {
    if (g1.hi <= g2.hi) g1 = g1 + v.R;
    else g2 = g2 + v.R;
    d = add_ru(g1.hi, -g2.lo);
    d2 = add_ru(g2.hi, -g1.lo);
    if (d < d2) d = d2;
    dist = c.r[0].hi;
    if (d <= dist) dist = d;
  }
  double frac = add_rd(0.5, -div_ru(dist, mul_rd(2.0, mul_rd(c.la.lo, kRsbRatio))));
  double w = mul_rd(frac, c.la.lo);
  if (w >= 1.0) return true;
  return disk_fits(c.r[0].hi, w);
}

// ------------------------------------------------------------------ C2
// uneven_split_recursion.cu:25-76
CCD bool c2_uneven_split(const cell7& c) {
  if (c.r[0].lo <= 0.0) return false;
  double need_w = required_height(c.r[0].hi);
  if (need_w >= c.la.lo) return false;
  double need = mul_ru(need_w, kRsbRatio);
  int possibly_at = -1, definitely_at = -1;
  iv wl = pt(0.0);
  for (int i = 0; i < 6; ++i) {
    wl = wl + c.r[i];
    if (possibly_at < 0 && wl.hi >= need) possibly_at = i;
    if (wl.lo >= need) {
      definitely_at = i;
      break;
    }
  }
  double largest = (possibly_at < 0 || possibly_at == 5) ? c.r[5].hi : c.r[possibly_at + 1].hi;
  double D1 = definitely_at < 0 ? add_ru(need, c.r[5].hi) : wl.hi;
  if (D1 >= mul_rd(c.la.lo, kRsbRatio)) return false;
  double wrem = add_rd(c.la.lo, -div_ru(D1, kRsbRatio));
  return wrem >= 1.0 || disk_fits(largest, wrem);
}

// ------------------------------------------------------------------ C3
// multi_disk_strip.cu:31-59
CCD bool strip_nd(const cell7& c, int k) {
  double wgt = 0.0;
  for (int j = 0; j < k; ++j) wgt = add_ru(wgt, c.r[j].hi);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h = 0.0;
  for (int j = 0; j < k; ++j) {
    double hc = add_rd(mul_rd(4.0, c.r[j].lo), -wsq);
    if (hc <= 0.0) break;
    h = add_rd(h, sqrt_rd(hc));
  }
  if (h < 1.0) return false;
  double largest = k == 6 ? c.r[5].hi : c.r[k].hi;
  double rem = add_rd(c.la.lo, -w);
  return rem >= 1.0 || disk_fits(largest, rem);
}
// multi_disk_strip.cu:61-93
CCD bool stacked_34_horizontal(const cell7& c) {
  double wgt = 0.0;
  for (int j = 0; j < 4; ++j) wgt = add_ru(wgt, c.r[j].hi);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h1 = add_rd(mul_rd(4.0, c.r[0].lo), -wsq);
  double h2 = add_rd(mul_rd(4.0, c.r[1].lo), -wsq);
  if (h1 <= 0.0 || h2 <= 0.0) return false;
  double hrem = add_ru(1.0, -add_rd(sqrt_rd(h1), sqrt_rd(h2)));
  double hrem_sq = mul_ru(hrem, hrem);
  double w4 = add_rd(mul_rd(4.0, c.r[3].lo), -hrem_sq);
  double w3 = add_rd(mul_rd(4.0, c.r[2].lo), -hrem_sq);
  if (w4 <= 0.0 || w3 <= 0.0) return false;
  if (add_rd(sqrt_rd(w3), sqrt_rd(w4)) < w) return false;
  double rem = add_rd(c.la.lo, -w);
  return rem >= 1.0 || disk_fits(c.r[4].hi, rem);
}
// multi_disk_strip.cu:95-145
CC_CHAIN_FN bool stacked_34_56_horizontal(const cell7& c) {
  double wgt = 0.0;
  for (int j = 0; j < 6; ++j) wgt = add_ru(wgt, c.r[j].hi);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h1 = add_rd(mul_rd(4.0, c.r[0].lo), -wsq);
  double h2 = add_rd(mul_rd(4.0, c.r[1].lo), -wsq);
  if (h1 <= 0.0 || h2 <= 0.0) return false;
  iv d34 = c.r[2] - c.r[3];
  tighten_lo(d34, 0.0);
  iv w3 = (2.0 * d34) / w + 0.5 * pt(w);
  iv w3sq = w3 * w3;
  if (!definitely(w3sq < 4.0 * c.r[2]) || !definitely((w - w3) * (w - w3) < 4.0 * c.r[3])) return false;
  double h34 = sqrt_rd(add_rd(mul_rd(4.0, c.r[2].lo), -w3sq.hi));
  iv d56 = c.r[4] - c.r[5];
  tighten_lo(d56, 0.0);
  iv w5 = (2.0 * d56) / w + 0.5 * pt(w);
  iv w5sq = w5 * w5;
  if (!definitely(w5sq < 4.0 * c.r[4]) || !definitely((w - w5) * (w - w5) < 4.0 * c.r[5])) return false;
  double h56 = sqrt_rd(add_rd(mul_rd(4.0, c.r[4].lo), -w5sq.hi));
  double h12 = add_rd(sqrt_rd(h1), sqrt_rd(h2));
  if (add_rd(h12, add_rd(h34, h56)) < 1.0) return false;
  double rem = add_rd(c.la.lo, -w);
  return rem >= 1.0 || disk_fits(c.r[5].hi, rem);
}
// multi_disk_strip.cu:147-240 (lb_h14 / lb_h23 take sqrt twice, as the reference)
CC_CHAIN_FN bool six_strip_three_rows(const cell7& c, const vals7& v, Trap& tp) {
  double wgt = 0.0;
  for (int i = 0; i < 6; ++i) wgt = add_ru(wgt, c.r[i].hi);
  double wn = div_ru(wgt, kRsbRatio);
  double ub_r7 = c.r[5].hi;
  if (v.R.hi < ub_r7) ub_r7 = v.R.hi;
  double wrem = add_rd(c.la.lo, -wn);
  if (wrem < 1.0 && !disk_fits(ub_r7, wrem)) return false;
  iv w = pt(wn);
  iv wsq = w * wn;
  iv w1 = 2.0 / w * (c.r[0] - c.r[3]) + 0.5 * w;
  iv w2 = 2.0 / w * (c.r[1] - c.r[2]) + 0.5 * w;
  double d14 = add_ru(c.r[0].hi, -c.r[3].lo);
  double h14 = sqrt_rd(add_rd(add_rd(mul_rd(2.0, add_rd(c.r[0].lo, c.r[3].lo)), -mul_ru(0.25, wsq.hi)),
                              -div_ru(mul_ru(4.0, mul_ru(d14, d14)), wsq.lo)));
  if (h14 < 0) return false;
  h14 = sqrt_rd(h14);
  double d23 = add_ru(c.r[1].hi, -c.r[2].lo);
  double h23 = sqrt_rd(add_rd(add_rd(mul_rd(2.0, add_rd(c.r[1].lo, c.r[2].lo)), -mul_ru(0.25, wsq.hi)),
                              -div_ru(mul_ru(4.0, mul_ru(d23, d23)), wsq.lo)));
  if (h23 < 0) return false;
  h23 = sqrt_rd(h23);
  iv hrem = 1.0 - pt(h23) - h14;
  if (hrem.hi <= 0.0) return true;
  double q = mul_ru(0.25, mul_ru(hrem.hi, hrem.hi));
  double x5 = add_rd(c.r[4].lo, -q);
  double w6 = add_rd(c.r[5].lo, -q);
  if (x5 < 0 || w6 < 0) return false;
  x5 = sqrt_rd(x5);
  w6 = sqrt_rd(w6);
  double x6 = add_ru(wn, -w6);
  iv y5 = 0.5 * hrem + h23;
  double y1 = add_ru(1.0, -mul_rd(0.5, h14));
  double y2 = mul_rd(0.5, h23);
  point p4{0.5 * (w - w1), pt(y1)};
  point p3{0.5 * (w - w2), pt(y2)};
  point p6{pt(x6), y5};
  circle c1{{w - 0.5 * w1, pt(y1)}, pt(c.r[0].lo)};
  circle c2{{w - 0.5 * w2, pt(y2)}, pt(c.r[1].lo)};
  circle c5{{pt(x5), y5}, pt(c.r[4].lo)};
  crossing x15 = intersect_circles(c5, c1, tp);
  crossing x25 = intersect_circles(c5, c2, tp);
  if (!x15.sure || !x25.sure) return false;
  if (dist2(p4, x15.p[0]).hi > c.r[3].lo) return false;
  if (dist2(p3, x25.p[0]).hi > c.r[2].lo) return false;
  return dist2(p6, x15.p[1]).hi <= c.r[5].lo && dist2(p6, x25.p[1]).hi <= c.r[5].lo;
}
// multi_disk_strip.cu:242-262
CCD bool c3_multi_disk_strip(const cell7& c, const vals7& v, Trap& tp) {
  for (int i = 2; i <= 6; ++i)
    if (strip_nd(c, i)) return true;
  if (stacked_34_horizontal(c)) return true;
  if (stacked_34_56_horizontal(c)) return true;
  return six_strip_three_rows(c, v, tp);
}

// ------------------------------------------------------------------ C4
// vertical_wall_building_recursion.cu:25-78 ("remaining width" adds, as the reference)
CCD bool c4_wall_building(const Konst& K, const cell7& c, const vals7& v) {
  double rem_w = add_rd(c.la.lo, sqrt_ru(mul_ru(c.r[5].hi, 2.0)));
  if (c.r[5].lo <= 0 || v.R.lo <= 0) return false;
  if (rem_w < 1.0 && !disk_fits(c.r[0].hi, rem_w)) return false;
  if (c.r[5].hi > K.weight_bound_wall_building) return false;
  const double phase = K.phase_factor_wall_building.hi;
  double R6 = add_rd(c.r[5].lo, v.R.lo);
  double below = mul_ru(c.r[5].hi, phase);
  double in_phases = mul_ru(kRsbRatio, sqrt_ru(mul_ru(2.0, c.r[5].hi)));
  for (int k = 0; k < 32; ++k) {
    double total_below = add_rd(R6, -in_phases);
    if (total_below <= 0) break;
    double req = mul_ru(kRsbRatio, required_height(below));
    if (req <= total_below) {
      double rem = add_rd(c.la.lo, div_ru(add_ru(req, below), kRsbRatio));
      if (rem >= 1.0 || disk_fits(c.r[0].hi, rem)) return true;
    }
    in_phases = add_ru(in_phases, mul_ru(kRsbRatio, sqrt_ru(mul_ru(2.0, below))));
    below = mul_ru(below, phase);
  }
  return false;
}

// ------------------------------------------------------------------ C5
// r1_in_corner_explicit_recursion.cu:36-187: r1 in the corner, the other
// disks stacked above it (right = false) or beside it (right = true) in
// backtracked groups.
CC_CHAIN_FN bool corner_explicit(const cell7& c, const vals7& v, int nd, bool right) {
  double lbs[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) lbs[i] = c.r[i].lo;
  double wgt = 0.0;
  for (int j = nd; j >= 0; --j) wgt = add_ru(wgt, c.r[j].hi);
  double max_rem = (nd == 5) ? (c.r[5].hi > v.R.hi ? v.R.hi : c.r[5].hi) : c.r[nd + 1].hi;
  double req_rem = required_height(max_rem);
  double span;
  if (!right) {
    span = div_ru(wgt, kRsbRatio);
    if (span > add_rd(c.la.lo, -req_rem)) return false;
  } else {
    span = div_ru(div_ru(wgt, kRsbRatio), c.la.lo);
    if (span > add_rd(1.0, -req_rem)) return false;
  }
  double start = add_rd(mul_rd(4.0, lbs[0]), -mul_ru(span, span));
  if (start <= 0) return false;
  start = sqrt_rd(start);
  const double goal = right ? c.la.hi : 1.0;
  double reached[6];
  int used[6], choice[6];
  int sh = 1;
  reached[0] = start;
  used[0] = 0;
  choice[0] = 0;
  while (sh > 0) {
    const int s = sh - 1;
    const int remaining = nd - used[s];
    const int next = ++choice[s];
    if (next > remaining) {
      --sh;
      continue;
    }
    const int first = 1 + used[s];
    if (next == remaining) {
      if (nd_can_cover(span, add_ru(goal, -reached[s]), lbs + first, next)) return true;
      --sh;
      continue;
    }
    double got = nd_maximize(span, lbs + first, next);
    if (got == 0) continue;
    double reach = add_rd(reached[s], got);
    if (reach >= goal) return true;
    reached[sh] = reach;
    used[sh] = used[s] + next;
    choice[sh] = 0;
    ++sh;
  }
  return false;
}
// r1_in_corner_explicit_recursion.cu:189-251
CCD bool corner_right_34_2(const cell7& c, Trap& tp) {
  iv w1 = sqrt(2.0 * c.r[0], tp);
  iv wrem = c.la - w1;
  if (!disk_fits(c.r[4].hi, add_rd(1.0, -w1.hi))) return false;
  double h34 = trc::two_disk_height(c.r[2].lo, c.r[3].lo, wrem.hi);
  if (h34 <= 0.0 || h34 >= w1.lo) return false;
  double hrem = add_ru(w1.hi, -h34);
  double w2 = add_rd(mul_rd(4.0, c.r[1].lo), -mul_ru(hrem, hrem));
  if (w2 <= 0.0) return false;
  w2 = sqrt_rd(w2);
  bool no_B = false;
  if (w2 >= wrem.hi) {
    no_B = true;
    w2 = wrem.lo;
  }
  if (w2 > wrem.lo) w2 = wrem.lo;
  double a234 = add_rd(mul_rd(w2, w1.lo), mul_rd(h34, add_rd(wrem.lo, -w2)));
  double wt234 = add_ru(c.r[1].hi, add_ru(c.r[2].hi, c.r[3].hi));
  double extra = add_rd(mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), c.r[0].lo),
                        add_rd(mul_rd(kRsbRatio, a234), -wt234));
  if (extra <= 0.0) return false;
  if (no_B) return true;
  iv hB = w1 - h34;
  iv wB = wrem - w2;
  double minB = hB.lo > wB.lo ? wB.lo : hB.lo;
  double cost = c.r[5].hi;
  if (!disk_fits(c.r[5].hi, minB))
    cost = add_ru(cost, mul_ru(add_ru(worst_case_ratio(wB, hB), -kRsbRatio), mul_ru(hB.hi, wB.hi)));
  return cost <= extra;
}
// r1_in_corner_explicit_recursion.cu:253-261
CCD bool c5_corner_explicit(const cell7& c, const vals7& v, Trap& tp) {
  for (int i = 2; i <= 5; ++i)
    if (corner_explicit(c, v, i, false) || corner_explicit(c, v, i, true)) return true;
  return corner_right_34_2(c, tp);
}

// ------------------------------------------------------------------ C6
struct short_long {
  iv shorter, longer;
};
struct size_check {
  double cost_bigger, cost_smaller;
  bool works;
};
// r1_r2_opposite_corners_recursion.cu:146-179
CCD short_long shorter_longer(iv w, iv h) {
  UB cmp = w <= h;
  if (is_certain(cmp)) return cmp.lo ? short_long{w, h} : short_long{h, w};
  short_long r;
  if (w.lo < h.lo) {
    r.shorter.lo = w.lo;
    r.longer.lo = h.lo;
  } else {
    r.shorter.lo = h.lo;
    r.longer.lo = w.lo;
  }
  if (w.hi < h.hi) {
    r.shorter.hi = w.hi;
    r.longer.hi = h.hi;
  } else {
    r.shorter.hi = h.hi;
    r.longer.hi = w.hi;
  }
  return r;
}
// r1_r2_opposite_corners_recursion.cu:188-252
CCD size_check opposite_sizes(const cell7& c, const vals7& v, iv wA, iv hA, iv wB, iv hB, int k) {
  double minA = hA.lo < wA.lo ? hA.lo : wA.lo;
  double minB = hB.lo < wB.lo ? hB.lo : wB.lo;
  double areaA = mul_rd(wA.lo, hA.lo);
  double areaB = mul_rd(wB.lo, hB.lo);
  double big = minA < minB ? minB : minA;
  double small = minA < minB ? minA : minB;
  if (k < 6) {
    if (!disk_fits(c.r[k].hi, big)) return size_check{0.0, 0.0, false};
    double big_weight = mul_rd(kRsbRatio, minA < minB ? areaB : areaA);
    double in_big = 0.0;
    for (; k < 6; ++k) {
      in_big = add_ru(in_big, c.r[k].hi);
      if (in_big >= big_weight) break;
    }
  }
  double r7 = c.r[5].hi;
  if (v.R.hi < r7) r7 = v.R.hi;
  if (k < 6) {
    if (!disk_fits(c.r[k].hi, small)) return size_check{0.0, 0.0, false};
    double cost_big = c.r[k].hi;
    double small_weight = mul_rd(kRsbRatio, minA < minB ? areaA : areaB);
    double in_small = 0.0;
    for (; k < 6; ++k) {
      in_small = add_ru(in_small, c.r[k].hi);
      if (in_small >= small_weight) break;
    }
    return size_check{cost_big, k < 6 ? c.r[k].hi : r7, true};
  }
  if (!disk_fits(r7, small)) return size_check{0.0, 0.0, false};
  return size_check{r7, r7, true};
}
// r1_r2_opposite_corners_recursion.cu:254-340
CC_CHAIN_FN bool c6_opposite_corners(const cell7& c, const vals7& v, Trap& tp) {
  iv w1{sqrt_rd(mul_rd(2.0, c.r[0].lo)), sqrt_ru(mul_ru(2.0, c.r[0].hi))};
  iv w2{sqrt_rd(mul_rd(2.0, c.r[1].lo)), sqrt_ru(mul_ru(2.0, c.r[1].hi))};
  iv both = w1 + w2;
  double extra12 = mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), add_rd(c.r[0].lo, c.r[1].lo));
  if (!(both.hi < 1.0)) return false;
  iv wA = c.la - w2, wB = c.la - w1, hA = 1.0 - w1, hB = w1, wC = w2, hC = 1.0 - both;
  short_long sC = shorter_longer(wC, hC);
  iv range = efficient_cover_range(sC.longer.lo, sC.shorter);
  int k;
  for (k = 2; k < 6; ++k) {
    if (range.lo > c.r[k].lo || c.r[k].hi > range.hi) break;
    iv len = 4.0 * c.r[k] - square(sC.shorter);
    if (possibly(len <= 0.0)) break;
    len = sqrt(len, tp);
    sC.longer = sC.longer - len;
    tighten_lo(sC.longer, 0.0);
    if (empty(sC.longer)) {
      size_check sb = opposite_sizes(c, v, wA, hA, wB, hB, k + 1);
      if (!sb.works) return false;
      return sb.cost_bigger <= extra12;
    }
    sC = shorter_longer(sC.shorter, sC.longer);
    range = efficient_cover_range(sC.longer.lo, sC.shorter);
  }
  size_check first = opposite_sizes(c, v, wA, hA, wB, hB, k);
  if (first.works && disk_fits(first.cost_smaller, sC.shorter.lo))
    return add_ru(first.cost_smaller, first.cost_bigger) <= extra12;
  double crC = worst_case_ratio(wC, hC);
  double minweightC = mul_rd(crC, mul_rd(wC.lo, hC.lo));
  double big_in_C = 0.0;
  for (; k < 6; ++k) {
    big_in_C = add_ru(big_in_C, c.r[k].hi);
    if (big_in_C >= minweightC) break;
  }
  size_check second = opposite_sizes(c, v, wA, hA, wB, hB, k);
  if (!second.works) return false;
  double split = add_ru(second.cost_smaller, second.cost_bigger);
  double extraC = mul_ru(add_ru(crC, -kRsbRatio), mul_ru(wC.hi, hC.hi));
  return add_ru(split, extraC) <= extra12;
}

// ------------------------------------------------------------------ C7
// r1_r2_vertical_strip.cu:43-95
CCD bool c7_vertical_strip(const Konst& K, const cell7& c, const vals7& v) {
  double w = trc::two_disk_height(c.r[0].lo, c.r[1].lo, 1.0);
  if (w <= 0.0) return false;
  double R = add_rd(mul_rd(c.la.lo, kRsbRatio), -add_ru(c.r[0].hi, c.r[1].hi));
  double ub_w = add_ru(c.la.hi, -w), ub_h = 1.0;
  for (int k = 2; k < 6; ++k) {
    if (can_recurse(K, R, c.r[k].hi, ub_w, ub_h)) return true;
    double m = ub_w < ub_h ? ub_w : ub_h;
    double cut = add_rd(mul_rd(4.0, c.r[k].lo), -mul_ru(m, m));
    if (cut <= 0) return false;
    cut = sqrt_rd(cut);
    if (ub_w < ub_h) {
      ub_h = add_ru(ub_h, -cut);
      if (ub_h <= 0.0) return true;
    } else {
      ub_w = add_ru(ub_w, -cut);
      if (ub_w <= 0.0) return true;
    }
    R = add_rd(R, -c.r[k].hi);
  }
  double r7 = c.r[5].hi < v.R.hi ? c.r[5].hi : v.R.hi;
  return can_recurse(K, R, r7, ub_w, ub_h);
}

}  // namespace rsbc
}  // namespace ccd

#include "rsb_chain2.hpp"
