// tight_rectangle covering criteria for the gfx950 kernels.
//
// Same rounded-operation sequence as the reference (cited per function), so
// verdicts are bit-identical; GPU-side changes are purely structural:
//   * required_weight_for(la), which every reference criterion recomputes,
//     is computed once per cell and passed in as `W`;
//   * the break-even bounds of critical_ratio (reference recomputes them per
//     call) come precomputed, as uniform constants, in `Konst`;
//   * traps become a per-lane flag.
#pragma once
#include "iv.hpp"

// The can_recurse helpers are called from many criteria.  Kept out of line
// by default (smaller code, faster builds); a translation unit that defines
// CC_INLINE_RECURSE 1 before including this header inlines them, which
// removes their call stack (scratch traffic) at the cost of code size.
#ifndef CC_INLINE_RECURSE
#define CC_INLINE_RECURSE 0
#endif
// CC_EXACT_SHORTCUTS 1 replaces parts of loops by computations that provably
// give the loop's own result (see r123_cover_strip); off by default.
#ifndef CC_EXACT_SHORTCUTS
#define CC_EXACT_SHORTCUTS 0
#endif
#if CC_INLINE_RECURSE
#define CC_RECURSE_FN CCD
#else
#define CC_RECURSE_FN static __device__ __noinline__
#endif

namespace ccd {

static constexpr double kRsbRatio = 0.61;
static constexpr double kDiskRadiusMax = 3.0 / 8.0;
static constexpr double kDiskWeightMax = 9.0 / 64.0;
static constexpr double kWorstRatio = 195.0 / 256.0;
static constexpr double kRsbToleratedLambda = 2.0898841580413813901;
// Bisections over doubles end within ~2300 halvings for finite bounds; the
// cap only matters for NaN bounds, on which the reference would spin forever.
static constexpr int kBisectCap = 4096;

// Cell-independent constants, computed once by ccrd_init_konst and read by
// every lane through scalar loads (uniform, so they live in SGPRs).
struct Konst {
  double be_lo, be_hi;  // bounds on lambda_2 = sqrt(sqrt(7)/2 - 1/4)
  // rectangle_size_bound Intermediate_values constants (variables.cuh:30-52)
  iv critical_ratio_sq;
  iv min_width_factor, max_width_factor;
  double weight_bound_wall_building;
  double width_fraction_incomplete;
  iv phase_factor_wall_building;
  double lb_w_infty_denom;
};

CCD bool disk_fits(double r_sq_ub, double h_lb) {  // disk_satisfies_size_bound.cu:25-28
  double s = rcp_ru(h_lb);
  return mul_ru(r_sq_ub, mul_ru(s, s)) <= kDiskWeightMax;
}

namespace trc {

// critical_ratio(IV la): critical_ratio.cu:58-78
CCD iv crit_ratio(const Konst& K, iv la) {
  iv a, b;
  bool lo_side = definitely(la <= K.be_lo);
  bool hi_side = !lo_side && definitely(la >= K.be_hi);
  if (!hi_side) {
    iv sq = square(la);
    a = ((3.0 / 256.0) / la) * (16.0 * sq + 40.0 + 9.0 / sq);
    if (lo_side) return a;
  }
  b = 0.25 * la + 0.5 * reciprocal(la);
  if (hi_side) return b;
  return iv{a.lo < b.lo ? a.lo : b.lo, a.hi < b.hi ? b.hi : a.hi};
}
// critical_ratio(IV w, IV h): critical_ratio.cu:32-52
CCD iv crit_ratio(const Konst& K, iv w, iv h) {
  if (w.hi <= h.lo) return crit_ratio(K, h / w);
  if (h.hi <= w.lo) return crit_ratio(K, w / h);
  iv hw = h / w;
  tighten_lo(hw, 1.0);
  iv wh = w / h;
  tighten_lo(wh, 1.0);
  iv c1 = crit_ratio(K, hw);
  iv c2 = crit_ratio(K, wh);
  return iv{c1.lo < c2.lo ? c1.lo : c2.lo, c1.hi < c2.hi ? c2.hi : c1.hi};
}
// RD and RU of n/d from one quotient: RU(n/d) is RD(n/d), or the next double
// when the residual n - q*d (exact for operands in div_rd's fast range) is
// nonzero.  Same pair as pt(n) / pt(d) for such operands.
CCD iv div_pt_fast(double n, double d) {  // n, d nonzero, div_fast_ok(n, d)
  const double q = d > 0.0 ? div_rd_fast_pos(n, d) : div_rd_fast_pos(-n, -d);
  return iv{q, fnma_rd(q, d, n) == 0.0 ? q : next_up(q)};
}

// crit_ratio(K, la).hi forming only the upper bounds that reach it (three
// quotients instead of six), for la = [lo, hi] of one sign with both ends
// of magnitude in [2^-500, 2^500] (checked by crit_hi_ok).  Each bound is
// the one the interval operations of crit_ratio(K, la) select: with
// x = (3/256)/la > 0 or < 0 and y = 16 la^2 + 40 + 9/la^2 > 0, the largest
// corner of x*y is RU(x.hi*y.hi) or RU(x.hi*y.lo); b = la/4 + 1/(2 la) only
// occurs for la > be_lo > 0.
CCD bool crit_hi_ok(iv la) {
  // |lo|, |hi| in [2^-500, 2^500) and one sign: compares with absolute-value
  // operands (false for zeros, subnormals, infinities and NaN)
  const double fl = __builtin_fabs(la.lo), fh = __builtin_fabs(la.hi);
  return (fl >= 0x1p-500) & (fl < 0x1p500) & (fh >= 0x1p-500) & (fh < 0x1p500) & ((la.lo > 0.0) == (la.hi > 0.0));
}
CCD double crit_ratio_hi_fast(const Konst& K, iv la) {
  const bool lo_side = la.hi <= K.be_lo;
  const bool hi_side = !lo_side && la.lo >= K.be_hi;
  double ahi = 0.0;
  if (!hi_side) {
    // la > 0: y.hi = (RU(16 sq.hi + 40)) (+RU) RU(9 / sq.lo), a.hi = RU(x.hi * y.hi);
    // la < 0: y.lo = (RD(16 sq.lo + 40)) (+RD) RD(9 / sq.hi), a.hi = RU(x.hi * y.lo).
    // Written once with sg = -1 (la > 0) or +1: RU(u) = -RD(-u) on exact negations.
    const bool pos = la.lo > 0.0;
    const double sg = pos ? -1.0 : 1.0;
    const double sql = pos ? mul_rd(la.lo, la.lo) : mul_rd(la.hi, la.hi);
    const double sqh = pos ? mul_ru(la.hi, la.hi) : mul_ru(la.lo, la.lo);
    const double t9 = div_rd(sg * 9.0, pos ? sql : sqh);
    const double y = sg * add_rd(add_rd(sg * mul_rd(pos ? sqh : sql, 16.0), sg * 40.0), t9);
    ahi = mul_ru(div_ru(3.0 / 256.0, la.lo), y);
    if (lo_side) return ahi;
  }
  const double bhi = add_ru(mul_rd(la.hi, 0.25), mul_rd(rcp_ru(la.lo), 0.5));
  if (hi_side) return bhi;
  return ahi < bhi ? bhi : ahi;
}
CCD double crit_ratio_hi(const Konst& K, iv la) {
  return crit_hi_ok(la) ? crit_ratio_hi_fast(K, la) : crit_ratio(K, la).hi;
}
// crit_ratio(K, pt(w), pt(h)).hi: for ordered points the reference divides
// the larger by the smaller.
CCD double crit_ratio_pt_hi(const Konst& K, double w, double h) {
  const double n = w <= h ? h : w, d = w <= h ? w : h;
  if ((w <= h || h <= w) && n != 0.0 && div_fast_ok(n, d)) return crit_ratio_hi(K, div_pt_fast(n, d));
  return crit_ratio(K, pt(w), pt(h)).hi;
}

CCD iv weight_for(const Konst& K, iv w, iv h) { return crit_ratio(K, w, h) * w * h; }
// weight_for(K, pt(w), pt(h)).hi: for w, h > 0 the products' upper corners
// are RU(RU(crit.hi * w) * h).
CCD double weight_for_pt_hi(const Konst& K, double w, double h) {
  if ((w >= 0x1p-200) & (w < 0x1p200) & (h >= 0x1p-200) & (h < 0x1p200))  // w, h in [2^-200, 2^200)
    return mul_ru(mul_ru(crit_ratio_pt_hi(K, w, h), w), h);
  return weight_for(K, pt(w), pt(h)).hi;
}
CCD iv weight_for(const Konst& K, iv la) { return crit_ratio(K, la) * la; }

// can_recurse(ub_w, ub_h, ub_largest, lb_rem): can_recurse.cu:32-62
CC_RECURSE_FN bool can_recurse(const Konst& K, double ub_w, double ub_h, double ub_largest, double lb_rem) {
  double cr = crit_ratio_pt_hi(K, ub_w, ub_h);
  double area = mul_ru(ub_w, ub_h);
  double side = ub_w < ub_h ? ub_w : ub_h;
  double scale = rcp_ru(side);
  double norm = mul_ru(mul_ru(scale, scale), ub_largest);
  double cor = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(mul_ru(norm, norm), 1.0)), 1.0)));
  if (cor < kWorstRatio) cor = kWorstRatio;
  if (cr > cor) cr = cor;
  if (norm <= kDiskWeightMax) cr = kRsbRatio;
  return lb_rem >= mul_ru(cr, area);
}

// two_disks_maximize_height (rectangle_size_bound/maximize_covered.cu:106-130)
CCD double two_disk_height(double r1, double r2, double ub_w) {
  double dlo = add_rd(r1, -r2);
  double dhi = add_ru(r1, -r2);
  double dsq = mul_ru(dhi, dhi);
  if (ub_w == 1.0) {
    // The same operations with every product and quotient by ub_w = 1
    // (exact in any rounding) left out; call sites passing the literal fold
    // this branch at compile time.
    double w1 = add_ru(0.5, mul_ru(2.0, dhi));
    double w2 = add_ru(0.5, -mul_rd(2.0, dlo));
    w1 = mul_ru(w1, w1);
    w2 = mul_ru(w2, w2);
    if (r1 <= mul_ru(0.25, w1) || r2 <= mul_ru(0.25, w2)) return 0.0;
    double h = add_rd(mul_rd(2.0, add_rd(r1, r2)), -add_ru(0.25, mul_ru(4.0, dsq)));
    if (h <= 0) return 0.0;
    return sqrt_rd(h);
  }
  double wsq = mul_ru(ub_w, ub_w);
  const iv two_w = ub_w != 0.0 && div_fast_ok(2.0, ub_w) ? div_pt_fast(2.0, ub_w) : iv{div_rd(2.0, ub_w), div_ru(2.0, ub_w)};
  double w1 = add_ru(mul_ru(0.5, ub_w), mul_ru(two_w.hi, dhi));
  double w2 = add_ru(mul_ru(0.5, ub_w), -mul_rd(two_w.lo, dlo));
  w1 = mul_ru(w1, w1);
  w2 = mul_ru(w2, w2);
  if (r1 <= mul_ru(0.25, w1) || r2 <= mul_ru(0.25, w2)) return 0.0;
  double h = add_rd(mul_rd(2.0, add_rd(r1, r2)), -add_ru(mul_ru(0.25, wsq), mul_ru(4.0, div_ru(dsq, wsq))));
  if (h <= 0) return 0.0;
  return sqrt_rd(h);
}

// r1_strategies: r1_strategies.cu:27-87.  W = required_weight_for(la).
// r1_strategies = r1_strip || r1_rest: the first strategy (a full-height
// strip) proves most boxes, so the batch kernel runs it alone and leaves the
// others to a second pass over the boxes it does not prove.
CCD bool r1_strip(const Konst& K, iv la, iv W, iv r1, double ub_r2) {
  double strip_w = add_rd(mul_rd(4.0, r1.lo), -1.0);
  if (!(strip_w > 0.0)) return false;
  strip_w = sqrt_rd(strip_w);
  return can_recurse(K, add_ru(la.hi, -strip_w), 1.0, ub_r2, (W - r1).lo);
}
CCD bool r1_rest(const Konst& K, iv la, iv W, iv r1, double ub_r2) {
  double sq_w = sqrt_rd(mul_rd(2.0, r1.lo));
  const double rem = (W - r1).lo;
  if (sq_w < 1.0) {
    double wright = add_ru(la.hi, -sq_w);
    double minright = wright < 1.0 ? wright : 1.0;
    double cost = disk_fits(ub_r2, minright) ? mul_ru(wright, kRsbRatio) : weight_for_pt_hi(K, wright, 1.0);
    cost = add_ru(cost, ub_r2);
    double rem_top = add_rd(rem, -cost);
    if (rem_top > 0 && can_recurse(K, sq_w, add_ru(1.0, -sq_w), ub_r2, rem_top)) return true;
  }
  double s1 = add_rd(mul_rd(4.0, r1.lo), -1.0);
  if (s1 > 0.0) {
    s1 = sqrt_rd(s1);
    double t = add_ru(la.hi, -mul_rd(0.5, s1));
    double h1 = add_rd(r1.lo, -mul_ru(t, t));
    if (h1 > 0.0) {
      h1 = mul_rd(2.0, sqrt_rd(h1));
      double ph = mul_ru(0.5, add_ru(1.0, -h1));
      double pw = add_ru(la.hi, -s1);
      double ru = mul_rd(0.5, add_rd(rem, -ub_r2));
      if (ru > 0 && can_recurse(K, pw, ph, ub_r2, ru)) return true;
    }
  }
  return false;
}
CCD bool r1_strategies(const Konst& K, iv la, iv W, iv r1, double ub_r2) {
  return r1_strip(K, la, W, r1, ub_r2) || r1_rest(K, la, W, r1, ub_r2);
}

// manually_proved_two_pocket: r1_r2_strategies.cu:29-44
CCD bool two_pocket(iv la, iv r1, iv r2, Trap& tp) {
  if (r2.lo < 0.2400000000000000188 || r1.lo < 0.4900000000000000466) return false;
  iv lasq = square(la);
  double r1req = (0.25 * (lasq + 1.0) - 0.00999999999999999847).hi;
  if (r1.lo < r1req) return false;
  iv S1 = 2.0 * sqrt(r1 - 0.25, tp);
  iv pw = la - S1;
  tighten_lo(pw, 0.0);
  double r2req = (square(pw) + 0.25 - 0.00999999999999999847).hi;
  return r2.lo >= r2req;
}

// r1_r2_strategies_simple: r1_r2_strategies.cu:46-72
CCD bool r1_r2_simple(const Konst& K, iv la, iv W, iv r1, iv r2, double ub_r3, Trap& tp) {
  if (two_pocket(la, r1, r2, tp)) return true;
  double width = two_disk_height(r1.lo, r2.lo, 1.0);
  if (width <= 0.0) return false;
  double w2 = add_rd(mul_rd(4.0, r2.lo), -1.0);
  if (w2 > 0.0) {
    w2 = sqrt_rd(w2);
    double w = add_rd(sqrt_rd(add_rd(mul_rd(4.0, r1.lo), -1.0)), w2);
    if (w > width) {
      width = w;
      if (w > la.hi) return true;
    }
  }
  // the reference adds the covered width here (r1_r2_strategies.cu:69)
  return can_recurse(K, add_rd(la.hi, width), 1.0, ub_r3, (W - r1 - r2).lo);
}

// r1_r2_strategies_stacked: r1_r2_strategies.cu:74-130
CCD bool r1_r2_stacked(const Konst& K, iv la, iv W, iv r1, iv r2, double ub_r3) {
  double side = sqrt_rd(mul_rd(2.0, r1.lo));
  if (side > 1.0) return false;
  double top_h = add_ru(1.0, -side);
  double w2 = add_rd(mul_rd(4.0, r2.lo), -mul_ru(top_h, top_h));
  if (w2 < 0.0) return false;
  w2 = sqrt_rd(w2);
  double w_top = add_ru(la.hi, -w2);
  double w_bot = add_ru(la.hi, -side);
  double top_min = w_top < top_h ? w_top : top_h;
  double bot_min = w_bot < side ? w_bot : side;
  double R3 = (W - r1 - r2).lo;
  // The three cases differ only in what the recursion test gets; one call
  // site keeps a wave from running the test once per case.
  double need, cw = w_bot, ch = side;
  if (disk_fits(ub_r3, top_min)) {
    need = add_ru(mul_ru(kRsbRatio, mul_ru(w_top, top_h)), ub_r3);
  } else if (disk_fits(ub_r3, bot_min)) {
    need = add_ru(mul_ru(kRsbRatio, mul_ru(w_bot, side)), ub_r3);
    cw = w_top;
    ch = top_h;
  } else {
    need = add_ru(weight_for_pt_hi(K, w_top, top_h), ub_r3);
  }
  double left = add_rd(R3, -need);
  return left > 0.0 && can_recurse(K, cw, ch, ub_r3, left);
}

// r1_r2_r3_strategy_horizontal: r1_r2_r3_strategies.cu:28-46
CCD bool r123_horizontal(const Konst& K, iv la, iv W, iv r1, iv r2, iv r3, double ub_r4) {
  double w3 = add_rd(mul_rd(4.0, r3.lo), -1.0);
  if (w3 <= 0.0) return false;
  w3 = sqrt_rd(w3);
  double w2 = sqrt_rd(add_rd(mul_rd(4.0, r2.lo), -1.0));
  double w1 = sqrt_rd(add_rd(mul_rd(4.0, r1.lo), -1.0));
  double cov = add_rd(w1, add_rd(w2, w3));
  double wrem = add_ru(la.hi, -cov);
  if (wrem <= 0.0) return true;
  return can_recurse(K, wrem, 1.0, ub_r4, (W - r1 - r2 - r3).lo);
}

// can_realize_strip_width: r1_r2_r3_strategies.cu:48-67 (4*r_i precomputed)
CCD bool strip_width_ok(double q1, double q2, double q3, double w) {
  double h1 = add_rd(q1, -mul_ru(w, w));
  if (h1 <= 0.0) return false;
  double hrem = add_ru(1.0, -sqrt_rd(h1));
  double hsq = mul_ru(hrem, hrem);
  double w2 = add_rd(q2, -hsq);
  double w3 = add_rd(q3, -hsq);
  if (w3 <= 0.0 || w2 <= 0.0) return false;
  return add_rd(sqrt_rd(w2), sqrt_rd(w3)) >= w;
}

#if CC_EXACT_SHORTCUTS
// Estimate of the largest strip width w in [a, b] that strip_width_ok
// accepts, from the exact real-number form of the test.  With
// u = sqrt(q1 - w^2) (decreasing in w) and v = 1 - u the test reads
//   u > 0,  w2 = q2 - v^2 > 0,  w3 = q3 - v^2 > 0,  sqrt(w2) + sqrt(w3) >= w,
// and the last condition is R = w^2 - w2 - w3 <= 0 or F = 4 w2 w3 - R^2 >= 0:
// every term is a polynomial in u, so F has no square-root singularity at
// the height limits and Newton's method converges quadratically on it.
// Safeguarded Newton (bisection where R <= 0 or a step leaves the bracket)
// finds the threshold u*, and w = sqrt(q1 - u*^2) is returned.  Only a
// starting point: arithmetic rounding here does not affect any result.
// Approximate square roots and quotient for the estimate: a Goldschmidt step
// from the rsq seed (relative error about 2^-49), plus a Newton step where
// the result must be good to an ulp or two; a Newton step from the rcp seed.
// est_sqrt(0) = 0; a zero divisor gives inf or NaN, which the caller rejects.
CCD double est_sqrt(double x, bool refine = false) {
  const double y = rsq_seed(x);
  double g = mul_rd(x, y), h = mul_rd(y, 0.5);
  const double r = fnma_rd(g, h, 0.5);
  g = fma_rd(g, r, g);
  if (refine) {
    h = fma_rd(h, r, h);
    g = fma_rd(fnma_rd(g, g, x), h, g);
  }
  return x > 0.0 ? g : 0.0;
}
CCD double est_div(double n, double d) {
  double y = rcp_seed(d);
  y = fma_rd(y, fnma_rd(d, y, 1.0), y);
  return mul_rd(n, y);
}
CCD double strip_width_estimate(double q1, double q2, double q3, double a, double b) {
  const double u_hi = est_sqrt(fmax(add_rd(q1, -mul_rd(a, a)), 0.0));
  const double h_b = fmax(add_rd(q1, -mul_rd(b, b)), 0.0);
  const double qm = fmin(q2, q3);
  const double u_w = qm < 1.0 ? add_rd(1.0, -est_sqrt(qm)) : 0.0;
  double x0 = fmax(est_sqrt(h_b), u_w), x1 = u_hi, x = x0;
  bool first = true;
  for (int it = 0; it < 32; ++it) {
    const double v = add_rd(1.0, -x);
    const double vv = mul_rd(v, v);
    const double w2 = add_rd(q2, -vv), w3 = add_rd(q3, -vv);
    const double R = add_rd(add_rd(q1, -mul_rd(x, x)), -add_rd(w2, w3));
    const double F = fma_rd(4.0 * w2, w3, -mul_rd(R, R));
    if (R <= 0.0 || F >= 0.0) {
      if (first) break;  // accepted at the narrowest end of the range
      x1 = x;
    } else {
      x0 = x;
    }
    first = false;
    const double dF = fma_rd(8.0 * v, add_rd(w2, w3), mul_rd(4.0 * R, fma_rd(2.0, v, x)));
    double nx = add_rd(x, -est_div(F, dF));
    if (R <= 0.0 || !(nx >= x0 && nx <= x1)) {
      nx = mul_rd(0.5, add_rd(x0, x1));
    } else if (fabs(add_rd(nx, -x)) <= 0x1p-51) {  // u is O(1): stop at a few ulps
      x = nx;
      break;
    }
    x = nx;
    if (add_rd(x1, -x0) <= 0x1p-51) break;
  }
  return est_sqrt(fmax(add_rd(q1, -mul_rd(x, x)), 0.0), true);
}

// strip_width_ok, also returning the covered width s = sqrt(w2) + sqrt(w3)
// (0 when a height test fails).
CCD bool strip_eval(double q1, double q2, double q3, double w, double& s) {
  s = 0.0;
  double h1 = add_rd(q1, -mul_ru(w, w));
  if (h1 <= 0.0) return false;
  double hrem = add_ru(1.0, -sqrt_rd(h1));
  double hsq = mul_ru(hrem, hrem);
  double w2 = add_rd(q2, -hsq);
  double w3 = add_rd(q3, -hsq);
  if (w3 <= 0.0 || w2 <= 0.0) return false;
  s = add_rd(sqrt_rd(w2), sqrt_rd(w3));
  return s >= w;
}

// Is the test true-then-false on [lo, b], b the largest double below hi,
// given that it holds at lo with covered width s_lo?  Every rounded step is
// monotone in w; with h1(w) = q1 (-) w*w (rounded down, non-increasing):
//  * where h1(w) <= 1 the remaining height 1 - sqrt(h1) is >= 0 and grows
//    with w, so w2, w3 and s shrink while w grows: true, then false;
//  * where h1(w) >= 1 the remaining height is <= 0 and its square shrinks as
//    w grows, so w2, w3 and s grow: true at lo means the height tests pass
//    on the rest of that range, and s(w) >= s_lo >= w there as long as
//    w <= s_lo.  That range ends at or below sqrt(q1 - 1) <= U (computed
//    rounding up), so s_lo >= min(b, U) makes the test true on all of it.
// The h1 >= 1 range precedes the h1 <= 1 range, hence true-then-false.
CCD bool strip_monotone_from(double q1, double U, double lo, double hi, double s_lo) {
  if (!(hi > lo) || !(hi < __builtin_inf())) return false;
  if (add_rd(q1, -mul_ru(lo, lo)) <= 1.0) return true;
  const double b = next_down(hi);
  return s_lo >= (b < U ? b : U);
}

// The cover-strip bisection from [lo, hi] (test true at lo, covered width
// s_lo).  Bisection steps run until strip_monotone_from holds; from then on
// the loop's final lo is the largest double T below hi that the test
// accepts, found by one test at the estimate, a gallop outward by 1, 2, 4,
// ... ulps to the first change of outcome, and a bisection of the bit
// patterns (positive doubles: bit order is numeric order).  Each lane walks
// these phases at its own pace, but every iteration of the one loop runs
// exactly one test, so a wave runs the test as often as its slowest lane
// needs, not once per phase and lane group.  Returns true with the final lo,
// or false with (lo, hi) at the state reached when more than `cap`
// bisection steps, or 48 tests in all, would be needed.
CCD bool strip_search(double q1, double q2, double q3, double& lo, double& hi, double s_lo, int cap) {
  const double U = q1 > 1.0 ? sqrt_ru(add_ru(q1, -1.0)) : 0.0;
  const double g = strip_width_estimate(q1, q2, q3, lo, next_down(hi));
  enum { kBisect, kFirst, kUp, kDown, kBits, kDone };
  int phase = kBisect, steps = 0;
  double b = 0.0, t = 0.0, f = 0.0, c = 0.0;
  uint64_t d = 1;
  bool ok = true;
  for (int n = 0;; ++n) {
    if (phase == kBisect) {
      if (strip_monotone_from(q1, U, lo, hi, s_lo)) {
        b = next_down(hi);
        if (b > lo) {
          c = g > lo ? (g < b ? g : b) : lo;
          phase = kFirst;
        } else {
          phase = kDone;
        }
      } else if (steps >= cap) {
        ok = false;
        phase = kDone;
      } else {
        c = rn_mid(lo, hi);
        ++steps;
        if (c <= lo || c >= hi) phase = kDone;
      }
    } else if (phase == kUp) {
      if (bits(t) == bits(b)) {
        lo = b;
        phase = kDone;
      } else {
        c = from_bits(bits(b) - bits(t) > d ? bits(t) + d : bits(b));
      }
    } else if (phase == kDown) {
      if (bits(f) - bits(lo) <= d) {
        t = lo;
        phase = kBits;
      } else {
        c = from_bits(bits(f) - d);
      }
    }
    if (phase == kBits) {
      if (bits(f) - bits(t) <= 1) {
        lo = t;
        phase = kDone;
      } else {
        c = from_bits(bits(t) + (bits(f) - bits(t)) / 2);
      }
    }
    if (phase == kDone) break;
    if (n >= 48) {
      ok = false;
      break;
    }
    double s;
    const bool r = strip_eval(q1, q2, q3, c, s);
    switch (phase) {
      case kBisect:
        if (r) {
          lo = c;
          s_lo = s;
        } else {
          hi = c;
        }
        break;
      case kFirst:
        if (r) t = c;
        else f = c;
        phase = r ? kUp : kDown;
        d = 1;
        break;
      case kUp:
        if (r) {
          t = c;
          d *= 2;
        } else {
          f = c;
          phase = kBits;
        }
        break;
      case kDown:
        if (r) {
          t = c;
          phase = kBits;
        } else {
          f = c;
          d *= 2;
        }
        break;
      default:  // kBits
        if (r) t = c;
        else f = c;
        break;
    }
  }
  return ok;
}
#endif

// r1_r2_r3_strategy_cover_strip: r1_r2_r3_strategies.cu:69-94
CCD bool r123_cover_strip(const Konst& K, iv la, iv W, iv r1, iv r2, iv r3, double ub_r4) {
  const double q1 = mul_rd(4.0, r1.lo), q2 = mul_rd(4.0, r2.lo), q3 = mul_rd(4.0, r3.lo);
  double lo = 0.3, hi = la.hi;
#if CC_EXACT_SHORTCUTS
  // The bisection keeps lo accepted and hi rejected (or the initial bound)
  // and stops when no double lies strictly between them; on a range where
  // the test is true-then-false it therefore ends at lo = the last accepted
  // double below hi, which strip_search finds with a handful of tests.
  double s_lo;
  if (!strip_eval(q1, q2, q3, lo, s_lo)) return false;
  if (!strip_search(q1, q2, q3, lo, hi, s_lo, kBisectCap)) {
    for (int it = 0; it < kBisectCap; ++it) {
      double mid = rn_mid(lo, hi);
      if (mid <= lo || mid >= hi) break;
      if (strip_width_ok(q1, q2, q3, mid)) lo = mid;
      else hi = mid;
    }
  }
  return can_recurse(K, add_ru(la.hi, -lo), 1.0, ub_r4, (W - r1 - r2 - r3).lo);
#else
  if (!strip_width_ok(q1, q2, q3, lo)) return false;
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) break;
    if (strip_width_ok(q1, q2, q3, mid)) lo = mid;
    else hi = mid;
  }
  return can_recurse(K, add_ru(la.hi, -lo), 1.0, ub_r4, (W - r1 - r2 - r3).lo);
#endif
}

#if CC_EXACT_SHORTCUTS
// r123_cover_strip for the batch kernel: at most `cap` bisection steps before
// the shortcut applies.  A box that needs more is left undecided
// (`deferred` = true, result false) for a second pass that runs the plain
// loop on the few such boxes together, instead of holding up the 63 other
// lanes of its wave.
CCD bool r123_cover_strip_capped(const Konst& K, iv la, iv W, iv r1, iv r2, iv r3, double ub_r4, int cap,
                                 bool& deferred) {
  const double q1 = mul_rd(4.0, r1.lo), q2 = mul_rd(4.0, r2.lo), q3 = mul_rd(4.0, r3.lo);
  double lo = 0.3, hi = la.hi, s_lo;
  deferred = false;
  if (!strip_eval(q1, q2, q3, lo, s_lo)) return false;
  if (!strip_search(q1, q2, q3, lo, hi, s_lo, cap)) {
    deferred = true;
    return false;
  }
  return can_recurse(K, add_ru(la.hi, -lo), 1.0, ub_r4, (W - r1 - r2 - r3).lo);
}
#endif

CCD bool r123_strategies(const Konst& K, iv la, iv W, iv r1, iv r2, iv r3, double ub_r4) {  // :96-99
  return r123_horizontal(K, la, W, r1, r2, r3, ub_r4) || r123_cover_strip(K, la, W, r1, r2, r3, ub_r4);
}
CCD bool r1_r2_strategies(const Konst& K, iv la, iv W, iv r1, iv r2, double ub_r3, Trap& tp) {  // :132-134
  return r1_r2_simple(K, la, W, r1, r2, ub_r3, tp) || r1_r2_stacked(K, la, W, r1, r2, ub_r3);
}

// ---------------------------------------------------------------- leaf chain
// A leaf: lambda and the four squared radii (tightened), as in
// kernel_find_critical_intervals_in (tight_rectangle.cu:238-244).
struct cell5 {
  iv la;
  iv r[4];
};

// three_disk_recursion_big_strip: r1_r2_r3_strategies.cu:101-121
CCD bool three_big_strip(const Konst& K, double la, double r1, double r2, double r3, double ub_r4, double R4) {
  double ssq = add_rd(add_rd(mul_rd(2.0, r2), add_rd(mul_rd(2.0, r3), -0.25)),
                      add_rd(add_rd(-mul_ru(4.0, mul_ru(r2, r2)), mul_rd(8.0, mul_rd(r2, r3))),
                             -mul_ru(4.0, mul_ru(r3, r3))));
  if (ssq <= 0.0) return false;
  double la_rem = add_ru(la, -sqrt_rd(ssq));
  double h1sq = add_rd(mul_rd(4.0, r1), -mul_ru(la_rem, la_rem));
  if (h1sq <= 0.0) return false;
  if (h1sq >= 1.0) return true;
  return can_recurse(K, la_rem, add_ru(1.0, -sqrt_rd(h1sq)), ub_r4, R4);
}

// three_disk_recursion_small_corner: r1_r2_r3_strategies.cu:123-165
CCD bool three_small_corner(const Konst& K, double la, double r1, double r2, double r3, double ub_r4, double R4) {
  double c3 = sqrt_rd(mul_rd(2.0, r3));
  if (c3 >= 1.0) return false;
  double la_rem = add_ru(la, -c3);
  double la_half = mul_ru(la_rem, 0.5);
  double one_rem = add_ru(1.0, -c3);
  if (la_half <= sqrt_rd(r2)) {
    double c1 = sqrt_rd(add_rd(mul_rd(4.0, r1), -mul_ru(one_rem, one_rem)));
    double c2 = sqrt_rd(add_rd(mul_rd(4.0, r2), -mul_ru(la_rem, la_rem)));
    return can_recurse(K, add_ru(la, -c1), add_ru(1.0, -c2), ub_r4, R4);
  }
  if (la_half >= sqrt_rd(r1)) return false;
  if (one_rem >= mul_rd(2.0, sqrt_rd(r2))) return false;
  double c2 = sqrt_rd(add_rd(mul_rd(4.0, r2), -mul_ru(one_rem, one_rem)));
  double c1 = sqrt_rd(add_rd(mul_rd(4.0, r1), -mul_ru(la_rem, la_rem)));
  return can_recurse(K, add_ru(la, -c2), add_ru(1.0, -c1), ub_r4, R4);
}

// Shared prefix of the two pocket placements (r1_r2_r3_strategies.cu:168-183
// and :222-235): S1, then h2, then S3; false when a radius is too small.
CCD bool pocket_chain(iv la, iv r1, iv r2, iv r3, iv& S1, iv& S3, Trap& tp) {
  S1 = 2.0 * sqrt(r1 - 0.25, tp);
  iv a = la - S1;
  iv h2_rest = 0.25 * a * a;
  if (r2.lo <= h2_rest.hi) return false;
  iv h2 = 2.0 * sqrt(r2 - h2_rest, tp);
  iv b = 1.0 - h2;
  iv s3_rest = 0.25 * b * b;
  if (r3.lo <= s3_rest.hi) return false;
  S3 = 2.0 * sqrt(r3 - s3_rest, tp);
  return true;
}

// three_disk_recursion_r1_r3_pocket: r1_r2_r3_strategies.cu:167-206
CCD bool three_r1_r3_pocket(const Konst& K, iv la, iv r1, iv r2, iv r3, double ub_r4, double R4, Trap& tp) {
  if (r1.lo <= 0.25) return false;
  iv S1, S3;
  if (!pocket_chain(la, r1, r2, r3, S1, S3, tp)) return false;
  iv X = la - S1 - S3;
  if (X.hi <= 0.0) return true;
  iv c = la - 0.5 * S1 - S3;
  iv y_rest = c * c;
  if (r1.lo <= y_rest.hi) return false;
  iv Y = 0.5 - sqrt(r1 - y_rest, tp);
  if (Y.hi <= 0.0) return true;
  return can_recurse(K, X.hi, Y.hi, ub_r4, R4);
}

// manually_proved_three_disk_pocket: r1_r2_r3_strategies.cu:208-247
CCD bool three_manual_pocket(iv la, iv r1, iv r2, iv r3, Trap& tp) {
  if (la.hi > 1.0499999999999998) return false;
  iv lasq = square(la);
  iv rstar = 0.0625 * lasq + 0.15625 + 0.03515625 / lasq;
  iv d1 = rstar - r1;
  if (d1.lo < -0.0009999999999999998 || d1.hi > 0.0009999999999999998) return false;
  iv d2 = rstar - r2;
  if (d2.lo < -0.0019999999999999996 || d2.hi > 0.0019999999999999996) return false;
  iv d3 = rstar - r3;
  if (d3.lo < -0.0029999999999999996 || d3.hi > 0.0029999999999999996) return false;
  iv S1, S3;
  if (!pocket_chain(la, r1, r2, r3, S1, S3, tp)) return false;
  iv X = la - S1 - S3;
  return !(X.hi > 0.046999999999999993);
}

// r1_r2_r3_strategies(vars, R4): r1_r2_r3_strategies.cu:249-260
CCD bool three_disk_leaf(const Konst& K, const cell5& c, iv R4, Trap& tp) {
  const double la = c.la.hi, r1 = c.r[0].lo, r2 = c.r[1].lo, r3 = c.r[2].lo;
  const double ub_r4 = c.r[3].hi, R4lo = R4.lo;
  return three_big_strip(K, la, r1, r2, r3, ub_r4, R4lo) || three_small_corner(K, la, r1, r2, r3, ub_r4, R4lo) ||
         three_r1_r3_pocket(K, c.la, c.r[0], c.r[1], c.r[2], ub_r4, R4lo, tp) ||
         three_manual_pocket(c.la, c.r[0], c.r[1], c.r[2], tp);
}

// two_by_two_strategies.cu:29-51
CCD bool two_by_two_r1_large(const Konst& K, const cell5& c, iv R) {
  double w1 = sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lo), -1.0));
  double wrem = add_ru(c.la.hi, -w1);
  double wrem_sq = mul_ru(wrem, wrem);
  double h4 = add_rd(mul_rd(4.0, c.r[3].lo), -wrem_sq);
  if (h4 <= 0.0) return false;
  h4 = sqrt_rd(h4);
  double h3 = sqrt_rd(add_rd(mul_rd(4.0, c.r[2].lo), -wrem_sq));
  double h2 = sqrt_rd(add_rd(mul_rd(4.0, c.r[1].lo), -wrem_sq));
  double htot = add_rd(h2, add_rd(h3, h4));
  if (htot >= 1.0) return true;
  double ub_r5 = c.r[3].hi < R.hi ? c.r[3].hi : R.hi;
  return can_recurse(K, wrem, add_ru(1.0, -htot), ub_r5, R.lo);
}

// two_by_two_strategies.cu:53-83
CCD bool two_by_two_small(const Konst& K, const cell5& c, iv R) {
  double side = sqrt_rd(mul_rd(2.0, c.r[0].lo));
  if (side >= 1.0) return two_by_two_r1_large(K, c, R);
  double wrem = add_ru(c.la.hi, -side);
  double hrem = add_ru(1.0, -side);
  double h2 = add_rd(mul_rd(4.0, c.r[1].lo), -mul_ru(wrem, wrem));
  double w4 = add_rd(mul_rd(4.0, c.r[3].lo), -mul_ru(hrem, hrem));
  if (h2 <= 0.0 || w4 <= 0.0) return false;
  h2 = sqrt_rd(h2);
  w4 = sqrt_rd(w4);
  if (w4 > side) w4 = side;
  if (h2 > side) h2 = side;
  double h = add_ru(1.0, -h2);
  double w = add_ru(c.la.hi, -w4);
  return mul_rd(4.0, c.r[2].lo) >= add_ru(mul_ru(h, h), mul_ru(w, w));
}

// two_by_two_strategies.cu:85-111: best of the three pairings into two
// columns of two disks
CCD bool two_by_two_big(const Konst& K, const cell5& c, iv R) {
  const double q0 = c.r[0].lo, q1 = c.r[1].lo, q2 = c.r[2].lo, q3 = c.r[3].lo;
  double w = add_rd(two_disk_height(q0, q1, 1.0), two_disk_height(q2, q3, 1.0));
  double s = add_rd(two_disk_height(q0, q3, 1.0), two_disk_height(q1, q2, 1.0));
  if (s > w) w = s;
  s = add_rd(two_disk_height(q0, q2, 1.0), two_disk_height(q1, q3, 1.0));
  if (s > w) w = s;
  if (w <= 0.0) return false;
  double ub_r5 = R.hi < c.r[3].hi ? R.hi : c.r[3].hi;
  return can_recurse(K, add_ru(c.la.hi, -w), 1.0, ub_r5, R.lo);
}

CCD bool two_by_two_leaf(const Konst& K, const cell5& c, iv R) {  // :113-115
  return two_by_two_small(K, c, R) || two_by_two_big(K, c, R);
}

// can_realize_gap_width: r1_r2_large_r3_r4_gaps_strategy.cu:29-57
CCD bool gap_width_ok(double gap, double r1, double w1, double r2, double w2, double r4, Trap& tp) {
  iv y4 = r4 - 0.25 * pt(gap) * gap;
  if (possibly(y4 <= 0.0)) return false;
  iv x1 = pt(0.5) * w1;
  iv x2 = pt(0.5) * w2 + gap + w1;
  circle c1{{x1, pt(0.5)}, pt(r1)};
  circle c2{{x2, pt(0.5)}, pt(r2)};
  crossing x12 = intersect_circles(c1, c2, tp);
  if (!x12.sure) return false;
  UB lower = x12.p[0].y <= x12.p[1].y;
  if (!is_certain(lower)) return false;
  point p4{pt(0.5) * gap + w1, sqrt(y4, tp)};
  const point& px = x12.p[lower.lo ? 0 : 1];
  return definitely(dist2(p4, px) <= r4);
}

// r1_r2_large_r3_r4_gaps_strategy.cu:59-98
CCD bool gaps_leaf(const Konst& K, const cell5& c, iv R, Trap& tp) {
  const double r1 = c.r[0].lo, r2 = c.r[1].lo, r4 = c.r[3].lo;
  double w1 = add_rd(mul_rd(4.0, r1), -1.0);
  double w2 = add_rd(mul_rd(4.0, r2), -1.0);
  if (w1 <= 0.0 || w2 <= 0.0) return false;
  w1 = sqrt_rd(w1);
  w2 = sqrt_rd(w2);
  double lo = 0.0, hi = add_rd(sqrt_rd(r1), sqrt_rd(r2));
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) break;
    if (gap_width_ok(mid, r1, w1, r2, w2, r4, tp)) lo = mid;
    else hi = mid;
  }
  double wrem = add_ru(c.la.hi, -add_rd(w1, add_rd(lo, w2)));
  if (wrem <= 0.0) return true;
  double ub_r5 = R.hi;
  if (ub_r5 > c.r[3].hi) ub_r5 = c.r[3].hi;
  return can_recurse(K, wrem, 1.0, ub_r5, R.lo);
}

// Valid squared-radius ranges: tight_rectangle.cu:72-102 (W = required_weight_for(la)).
CCD iv r1_range(const Konst& K, iv la) {
  double f = div_rd(crit_ratio(K, la).lo, kRsbRatio);
  if (f <= 1.0) f = 1.0;
  return iv{mul_rd(kDiskWeightMax, f), mul_ru(0.25, add_ru(1.0, mul_ru(la.hi, la.hi)))};
}
CCD iv r2_range(iv W, iv r1) {
  double b = (W - r1).hi;
  return iv{0.0, r1.hi < b ? r1.hi : b};
}
CCD iv r3_range(iv W, iv r1, iv r2) {
  double b = (W - r1 - r2).hi;
  return iv{0.0, r2.hi < b ? r2.hi : b};
}
CCD iv r4_range(iv W, iv r1, iv r2, iv r3) {
  double b = (W - r1 - r2 - r3).hi;
  return iv{0.0, r3.hi < b ? r3.hi : b};
}

}  // namespace trc
}  // namespace ccd
