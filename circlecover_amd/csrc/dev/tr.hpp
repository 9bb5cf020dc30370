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
CCD iv weight_for(const Konst& K, iv w, iv h) { return crit_ratio(K, w, h) * w * h; }
CCD iv weight_for(const Konst& K, iv la) { return crit_ratio(K, la) * la; }

// can_recurse(ub_w, ub_h, ub_largest, lb_rem): can_recurse.cu:32-62
__device__ __noinline__ bool can_recurse(const Konst& K, double ub_w, double ub_h, double ub_largest, double lb_rem) {
  double cr = crit_ratio(K, pt(ub_w), pt(ub_h)).hi;
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
  double wsq = mul_ru(ub_w, ub_w);
  double dlo = add_rd(r1, -r2);
  double dhi = add_ru(r1, -r2);
  double dsq = mul_ru(dhi, dhi);
  double w1 = add_ru(mul_ru(0.5, ub_w), mul_ru(div_ru(2.0, ub_w), dhi));
  double w2 = add_ru(mul_ru(0.5, ub_w), -mul_rd(div_rd(2.0, ub_w), dlo));
  w1 = mul_ru(w1, w1);
  w2 = mul_ru(w2, w2);
  if (r1 <= mul_ru(0.25, w1) || r2 <= mul_ru(0.25, w2)) return 0.0;
  double h = add_rd(mul_rd(2.0, add_rd(r1, r2)), -add_ru(mul_ru(0.25, wsq), mul_ru(4.0, div_ru(dsq, wsq))));
  if (h <= 0) return 0.0;
  return sqrt_rd(h);
}

// r1_strategies: r1_strategies.cu:27-87.  W = required_weight_for(la).
CCD bool r1_strategies(const Konst& K, iv la, iv W, iv r1, double ub_r2) {
  double sq_w = sqrt_rd(mul_rd(2.0, r1.lo));
  double strip_w = add_rd(mul_rd(4.0, r1.lo), -1.0);
  const double rem = (W - r1).lo;
  if (strip_w > 0.0) {
    strip_w = sqrt_rd(strip_w);
    if (can_recurse(K, add_ru(la.hi, -strip_w), 1.0, ub_r2, rem)) return true;
  }
  if (sq_w < 1.0) {
    double wright = add_ru(la.hi, -sq_w);
    double minright = wright < 1.0 ? wright : 1.0;
    double cost = disk_fits(ub_r2, minright) ? mul_ru(wright, kRsbRatio) : weight_for(K, pt(wright), pt(1.0)).hi;
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
  if (disk_fits(ub_r3, top_min)) {
    double need = add_ru(mul_ru(kRsbRatio, mul_ru(w_top, top_h)), ub_r3);
    double left = add_rd(R3, -need);
    return left > 0.0 && can_recurse(K, w_bot, side, ub_r3, left);
  }
  if (disk_fits(ub_r3, bot_min)) {
    double need = add_ru(mul_ru(kRsbRatio, mul_ru(w_bot, side)), ub_r3);
    double left = add_rd(R3, -need);
    return left > 0.0 && can_recurse(K, w_top, top_h, ub_r3, left);
  }
  double need = add_ru(weight_for(K, pt(w_top), pt(top_h)).hi, ub_r3);
  double left = add_rd(R3, -need);
  return left > 0.0 && can_recurse(K, w_bot, side, ub_r3, left);
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

// r1_r2_r3_strategy_cover_strip: r1_r2_r3_strategies.cu:69-94
// `trace` (diagnostics only, null in production): per iteration lo, hi, mid.
CCD bool r123_cover_strip(const Konst& K, iv la, iv W, iv r1, iv r2, iv r3, double ub_r4,
                          double* trace = nullptr, int trace_cap = 0) {
  const double q1 = mul_rd(4.0, r1.lo), q2 = mul_rd(4.0, r2.lo), q3 = mul_rd(4.0, r3.lo);
  double lo = 0.3, hi = la.hi;
  if (!strip_width_ok(q1, q2, q3, lo)) return false;
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (trace && it < trace_cap) {
      trace[3 * it + 0] = lo;
      trace[3 * it + 1] = hi;
      trace[3 * it + 2] = mid;
    }
    if (mid <= lo || mid >= hi) break;
    if (strip_width_ok(q1, q2, q3, mid)) lo = mid;
    else hi = mid;
  }
  return can_recurse(K, add_ru(la.hi, -lo), 1.0, ub_r4, (W - r1 - r2 - r3).lo);
}

}  // namespace trc
}  // namespace ccd
