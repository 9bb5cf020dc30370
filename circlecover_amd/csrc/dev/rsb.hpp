// rectangle_size_bound helpers and cell-level shortcuts for the gfx950
// kernels; same rounded-operation sequence as the cited reference lines.
#pragma once
#include "tr.hpp"

namespace ccd {
namespace rsbc {

CCD double required_height(double r_sq_ub) { return div_ru(sqrt_ru(r_sq_ub), kDiskRadiusMax); }  // :30-32

// can_recurse (size-bound version): rectangle_size_bound/can_recurse.cu:26-88
CC_RECURSE_FN bool can_recurse(const Konst& K, double lb_rem, double ub_largest, double ub_w, double ub_h) {
  if (lb_rem <= 0.0) {
    if (lb_rem < 0.0) return false;
    return (ub_w <= 0.0) | (ub_h <= 0.0);
  }
  if (ub_w <= 0.0 || ub_h <= 0.0) return true;
  double smin = ub_w < ub_h ? ub_w : ub_h;
  double smax = ub_w < ub_h ? ub_h : ub_w;
  double area = div_rd(lb_rem, kRsbRatio);
  double side = div_rd(area, smax);
  if (side < smin) return false;
  if (side > smax) side = sqrt_rd(area);
  if (mul_rd(kDiskWeightMax, mul_rd(side, side)) >= ub_largest) return true;
  double skew = div_ru(smax, smin);
  if (skew <= kRsbToleratedLambda) {
    double cr = trc::crit_ratio_hi(K, iv{div_rd(smax, smin), skew});
    return mul_ru(cr, mul_ru(ub_w, ub_h)) <= lb_rem;
  }
  double sig = div_ru(ub_largest, mul_rd(smin, smin));
  sig = mul_ru(sig, sig);
  double cr = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(sig, 1.0)), 1.0)));
  if (cr < kWorstRatio) cr = kWorstRatio;
  if (mul_ru(cr, mul_ru(ub_w, ub_h)) <= lb_rem) return true;
  return add_ru(mul_ru(0.25, mul_ru(smax, smax)), mul_ru(0.5, mul_ru(smin, smin))) <= lb_rem;
}

// bound_worst_case_ratio: can_recurse.cu:221-235
CCD double worst_case_ratio(iv w, iv h) {
  if (w.lo <= 0.0 || h.lo <= 0.0) return __builtin_inf();
  double a = div_ru(w.hi, h.lo);
  double b = div_ru(h.hi, w.lo);
  double skew = a < b ? b : a;
  double r = add_ru(mul_ru(0.25, skew), mul_ru(0.5, rcp_ru(skew)));
  return r < kWorstRatio ? kWorstRatio : r;
}

// S1 shortcut_even_split_recursion: even_split_recursion.cu:42-57
CCD bool s1_even_split(iv la, iv r1, iv r2) {
  double big_w = mul_rd(0.5, la.lo);
  if (big_w < 1.0 && !disk_fits(r1.hi, big_w)) return false;
  double small_w = add_rd(big_w, -div_ru(r2.hi, kRsbRatio));
  return small_w >= 1.0 || disk_fits(r2.hi, small_w);
}

// S2/S3 shortcut_uneven_split_recursion: uneven_split_recursion.cu:78-125.
// S2 passes rk = r2, S3 passes rk = r3; the two bodies are otherwise equal.
CCD bool s23_uneven_split(iv la, iv r1, iv rk) {
  if (r1.lo <= 0.0) return false;
  double need_w = required_height(r1.hi);
  if (need_w >= la.lo) return false;
  double g1 = add_ru(rk.hi, div_ru(need_w, kRsbRatio));
  double rem = add_rd(mul_rd(la.lo, kRsbRatio), -g1);
  if (rem <= 0.0) return false;
  double rem_w = div_rd(rem, kRsbRatio);
  return rem_w >= 1.0 || disk_fits(rk.hi, rem_w);
}

// S4 shortcut_r1_r2_opposite_corners_recursion: r1_r2_opposite_corners_recursion.cu:28-70
CCD bool s4_opposite_corners(iv la, iv r1, iv r2, iv r3) {
  iv w1{sqrt_rd(mul_rd(2.0, r1.lo)), sqrt_ru(mul_ru(2.0, r1.hi))};
  iv w2{sqrt_rd(mul_rd(2.0, r2.lo)), sqrt_ru(mul_ru(2.0, r2.hi))};
  iv both = w1 + w2;
  if (!(both.hi < 1.0)) return false;
  iv wA = la - w2, wB = la - w1, hA = 1.0 - w1;
  double minA = hA.lo < wA.lo ? hA.lo : wA.lo;
  double minB = w1.lo < wB.lo ? w1.lo : wB.lo;
  if (!disk_fits(r3.hi, minA < minB ? minA : minB)) return false;
  iv hC = 1.0 - both;
  double minC = hC.lo < w2.lo ? hC.lo : w2.lo;
  double extra12 = mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), add_rd(r1.lo, r2.lo));
  if (disk_fits(r3.hi, minC)) return !(mul_ru(2.0, r3.hi) > extra12);
  double before = add_rd(extra12, -mul_ru(2.0, r3.hi));
  double extra = mul_ru(add_ru(worst_case_ratio(w2, hC), -kRsbRatio), mul_ru(w2.hi, hC.hi));
  return !(before < extra);
}

// S5 shortcut_r1_r2_vertical_strip: r1_r2_vertical_strip.cu:28-41
CCD bool s5_vertical_strip(const Konst& K, iv la, iv r1, iv r2, iv r3) {
  double w = trc::two_disk_height(r1.lo, r2.lo, 1.0);
  if (w <= 0.0) return false;
  double R3 = add_rd(mul_rd(la.lo, kRsbRatio), -add_ru(r1.hi, r2.hi));
  iv wrem = la - w;
  if (wrem.lo >= 1.0 || disk_fits(r3.hi, wrem.lo)) return R3 >= mul_ru(kRsbRatio, wrem.hi);
  return can_recurse(K, R3, r3.hi, wrem.hi, 1.0);
}

}  // namespace rsbc
}  // namespace ccd
