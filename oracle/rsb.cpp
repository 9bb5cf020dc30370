// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// rectangle_size_bound helpers and cell-level shortcuts, restated step for
// step from the reference (Prover/rectangle_size_bound/*.cu).
#include "criteria.hpp"

namespace ccor {
namespace rsb {

// disk_satisfies_size_bound.cu:25-37
bool disk_satisfies_size_bound(double r_sq_ub, double h_lb) {
  double s = rcp_ru(h_lb);
  return mul_ru(r_sq_ub, mul_ru(s, s)) <= kDiskWeightMax;
}
double bound_required_height(double r_sq_ub) { return div_ru(sqrt_ru(r_sq_ub), kDiskRadiusMax); }
double bound_allowed_weight(double h_lb) { return mul_rd(kDiskWeightMax, mul_rd(h_lb, h_lb)); }

// compute_efficient_rectangle_cover_weight_range.cu:25-35
IV compute_efficient_rectangle_cover_weight_range(double lb_w, const IV& h) {
  double wmax = mul_rd(kRsbRatio, mul_rd(h.lb, lb_w));
  double lo_esq_hsq = mul_ru(mul_ru(kRsbRatio, kRsbRatio), mul_ru(h.ub, h.ub));
  double hi_esq_hsq = mul_rd(mul_rd(kRsbRatio, kRsbRatio), mul_rd(h.lb, h.lb));
  double root = sqrt_rd(add_rd(4.0, -rcp_ru(mul_rd(kRsbRatio, kRsbRatio))));
  double lo = mul_ru(lo_esq_hsq, add_ru(2.0, -root));
  double hi = mul_rd(hi_esq_hsq, add_rd(2.0, root));
  return IV(lo, wmax < hi ? wmax : hi);
}

// can_recurse.cu:26-88
bool can_recurse(double lb_rem, double ub_largest, double ub_w, double ub_h) {
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
    double cr = tr::critical_ratio(IV(div_rd(smax, smin), skew)).ub;
    return mul_ru(cr, mul_ru(ub_w, ub_h)) <= lb_rem;
  }
  double sig = div_ru(ub_largest, mul_rd(smin, smin));
  sig = mul_ru(sig, sig);
  double cr = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(sig, 1.0)), 1.0)));
  if (cr < kWorstRatio) cr = kWorstRatio;
  if (mul_ru(cr, mul_ru(ub_w, ub_h)) <= lb_rem) return true;
  double need = add_ru(mul_ru(0.25, mul_ru(smax, smax)), mul_ru(0.5, mul_ru(smin, smin)));
  return need <= lb_rem;
}

// can_recurse.cu:90-161
double bound_critical_ratio(double lb_rem, double ub_largest, double ub_w, double ub_h) {
  if (lb_rem < 0.0) return kInf;
  if (ub_w <= 0.0 || ub_h <= 0.0) return 0.0;
  double smin = ub_w < ub_h ? ub_w : ub_h;
  double smax = ub_w < ub_h ? ub_h : ub_w;
  double best = kInf;
  double area = mul_ru(ub_w, ub_h);
  double req_side = sqrt_ru(div_ru(ub_largest, kDiskWeightMax));
  if (smin >= req_side) {
    if (mul_ru(kRsbRatio, area) <= lb_rem) return kRsbRatio;
    return kInf;
  }
  if (req_side > smax) {
    double sq_area = div_rd(lb_rem, kRsbRatio);
    double sq_side = sqrt_rd(sq_area);
    if (sq_side >= req_side) {
      double a = mul_ru(req_side, req_side);
      double wgt = mul_ru(a, kRsbRatio);
      double orig = mul_rd(ub_w, ub_h);
      best = div_ru(wgt, orig);
    }
  } else {
    double req = mul_ru(kRsbRatio, mul_ru(req_side, smax));
    if (req <= lb_rem) {
      double orig = mul_rd(ub_w, ub_h);
      best = div_ru(req, orig);
    }
  }
  double sig = div_ru(ub_largest, mul_rd(smin, smin));
  sig = mul_ru(sig, sig);
  double eff = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(sig, 1.0)), 1.0)));
  if (eff < kWorstRatio) eff = kWorstRatio;
  if (eff < best) {
    if (mul_ru(eff, area) <= lb_rem) best = eff;
  }
  IV skew(div_rd(smax, smin), div_ru(smax, smin));
  double cr = tr::critical_ratio(skew).ub;
  if (cr < best) {
    if (mul_ru(cr, area) <= lb_rem) best = cr;
  }
  return best;
}

// can_recurse.cu:163-218
double recursion_bound_size(double lb_rem, double ub_w, double ub_h) {
  if (lb_rem <= 0.0) {
    if (lb_rem < 0.0) return 0.0;
    return (ub_w <= 0.0 || ub_h <= 0.0) ? kInf : 0.0;
  }
  if (ub_w <= 0.0 || ub_h <= 0.0) return kInf;
  double area = mul_ru(ub_w, ub_h);
  double max_cr = div_rd(lb_rem, area);
  double smin = ub_w < ub_h ? ub_w : ub_h;
  double smax = ub_w < ub_h ? ub_h : ub_w;
  IV skew(div_rd(smax, smin), div_ru(smax, smin));
  double cr_wc = tr::critical_ratio(skew).ub;
  if (cr_wc <= max_cr) return kInf;
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

// can_recurse.cu:221-235
double bound_worst_case_ratio(const IV& w, const IV& h) {
  if (w.lb <= 0.0 || h.lb <= 0.0) return kInf;
  double a = div_ru(w.ub, h.lb);
  double b = div_ru(h.ub, w.lb);
  double skew = a < b ? b : a;
  double r = add_ru(mul_ru(0.25, skew), mul_ru(0.5, rcp_ru(skew)));
  if (r < kWorstRatio) r = kWorstRatio;
  return r;
}

// can_recurse.cu:237-261
bool can_recurse_worst_case(double lb_rem, double ub_w, double ub_h) {
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

// maximize_covered.cu:28-51
static bool nd_can_cover_no_small(double ub_w, double ub_h, const IV* d, int n) {
  double htot = 0.0, wtot = 0.0;
  double wsq = mul_ru(ub_w, ub_w);
  double hsq = mul_ru(ub_h, ub_h);
  for (int i = 0; i < n; ++i) {
    double wc = add_rd(mul_rd(4.0, d[i].lb), -hsq);
    double hc = add_rd(mul_rd(4.0, d[i].lb), -wsq);
    wc = wc <= 0 ? 0 : sqrt_rd(wc);
    hc = hc <= 0 ? 0 : sqrt_rd(hc);
    wtot = add_rd(wtot, wc);
    htot = add_rd(htot, hc);
  }
  return htot >= ub_h || wtot >= ub_w;
}

// maximize_covered.cu:53-71
bool nd_can_cover(double ub_w, double ub_h, const IV* d, int n) {
  if (n == 1) {
    double wsq = mul_ru(ub_w, ub_w);
    double hsq = mul_ru(ub_h, ub_h);
    return mul_rd(4.0, d[0].lb) >= add_ru(wsq, hsq);
  }
  if (n == 2) {
    double h = two_disks_maximize_height(d[0], d[1], ub_w);
    double w = two_disks_maximize_height(d[0], d[1], ub_h);
    return h >= ub_h || w >= ub_w;
  }
  if (n == 3) return three_disks_can_cover(d[0], d[1], d[2], ub_w, ub_h);
  return nd_can_cover_no_small(ub_w, ub_h, d, n);
}

// maximize_covered.cu:73-104
double nd_maximize_covered(double ub_w, const IV* d, int n) {
  if (n == 1) {
    double wsq = mul_ru(ub_w, ub_w);
    double h = add_rd(mul_rd(4.0, d[0].lb), -wsq);
    return h <= 0 ? 0.0 : sqrt_rd(h);
  }
  if (n == 2) return two_disks_maximize_height(d[0], d[1], ub_w);
  if (n == 3) return three_disks_maximize_height(d[0], d[1], d[2], ub_w);
  double hi = mul_rd(2.0, sqrt_rd(d[n - 1].lb));
  double lo = 0.0;
  for (int it = 0; it < kBisectCap; ++it) {
    double h = rn_mid(lo, hi);
    if (h <= lo || h >= hi) return lo;
    if (nd_can_cover_no_small(ub_w, h, d, n)) lo = h;
    else hi = h;
  }
  return lo;
}

// maximize_covered.cu:106-130 (uses the lower bounds only)
double two_disks_maximize_height(const IV& r1_, const IV& r2_, double ub_w) {
  const double r1 = r1_.lb, r2 = r2_.lb;
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

// maximize_covered.cu:171-196
bool three_disks_can_cover(const IV& r1, const IV& r2, const IV& r3, double w, double h) {
  double hsq = mul_ru(h, h);
  double w3 = add_rd(mul_rd(4.0, r3.lb), -hsq);
  w3 = w3 <= 0 ? 0 : sqrt_rd(w3);
  double w2 = add_rd(mul_rd(4.0, r2.lb), -hsq);
  w2 = w2 <= 0 ? 0 : sqrt_rd(w2);
  double w1 = add_rd(mul_rd(4.0, r1.lb), -hsq);
  w1 = w1 <= 0 ? 0 : sqrt_rd(w1);
  return add_rd(w1, add_rd(w2, w3)) >= w;
}

// maximize_covered.cu:198-237
double three_disks_maximize_height(const IV& r1, const IV& r2, const IV& r3, double ub_w) {
  double lo = 0.0, hi = mul_rd(2.0, sqrt_rd(r3.lb));
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) return lo;
    double hsq = mul_ru(mid, mid);
    double w3 = add_rd(mul_rd(4.0, r3.lb), -hsq);
    w3 = w3 <= 0 ? 0 : sqrt_rd(w3);
    double w2 = add_rd(mul_rd(4.0, r2.lb), -hsq);
    w2 = w2 <= 0 ? 0 : sqrt_rd(w2);
    double w1 = add_rd(mul_rd(4.0, r1.lb), -hsq);
    w1 = w1 <= 0 ? 0 : sqrt_rd(w1);
    if (add_rd(w1, add_rd(w2, w3)) < ub_w) hi = mid;
    else lo = mid;
  }
  return lo;
}

// even_split_recursion.cu:42-57
bool shortcut_even_split_recursion(const IV& la, const IV& r1, const IV& r2) {
  double big_w = mul_rd(0.5, la.lb);
  if (big_w < 1.0 && !disk_satisfies_size_bound(r1.ub, big_w)) return false;
  double small_w = add_rd(big_w, -div_ru(r2.ub, kRsbRatio));
  return small_w >= 1.0 || disk_satisfies_size_bound(r2.ub, small_w);
}

// uneven_split_recursion.cu:78-100
bool shortcut_uneven_split_recursion(const IV& la, const IV& r1, const IV& r2) {
  if (r1.lb <= 0.0) return false;
  double need_w = bound_required_height(r1.ub);
  if (need_w >= la.lb) return false;
  double g1 = add_ru(r2.ub, div_ru(need_w, kRsbRatio));
  double rem = add_rd(mul_rd(la.lb, kRsbRatio), -g1);
  if (rem <= 0.0) return false;
  double rem_w = div_rd(rem, kRsbRatio);
  return rem_w >= 1.0 || disk_satisfies_size_bound(r2.ub, rem_w);
}

// uneven_split_recursion.cu:102-125
bool shortcut_uneven_split_recursion(const IV& la, const IV& r1, const IV& r2, const IV& r3) {
  (void)r2;
  if (r1.lb <= 0.0) return false;
  double need_w = bound_required_height(r1.ub);
  if (need_w >= la.lb) return false;
  double g1 = add_ru(r3.ub, div_ru(need_w, kRsbRatio));
  double rem = add_rd(mul_rd(la.lb, kRsbRatio), -g1);
  if (rem <= 0.0) return false;
  double rem_w = div_rd(rem, kRsbRatio);
  return rem_w >= 1.0 || disk_satisfies_size_bound(r3.ub, rem_w);
}

// r1_r2_opposite_corners_recursion.cu:28-138 (only the live first case)
bool shortcut_r1_r2_opposite_corners_recursion(const IV& la, const IV& r1, const IV& r2, const IV& r3) {
  IV w1(sqrt_rd(mul_rd(2.0, r1.lb)), sqrt_ru(mul_ru(2.0, r1.ub)));
  IV w2(sqrt_rd(mul_rd(2.0, r2.lb)), sqrt_ru(mul_ru(2.0, r2.ub)));
  IV both = w1 + w2;
  if (!(both.ub < 1.0)) return false;
  IV wA = la - w2;
  IV wB = la - w1;
  IV hA = 1.0 - w1;
  IV hB = w1;
  double minA = wA.lb;
  if (hA.lb < minA) minA = hA.lb;
  double minB = wB.lb;
  if (hB.lb < minB) minB = hB.lb;
  double minAB = minA < minB ? minA : minB;
  if (!disk_satisfies_size_bound(r3.ub, minAB)) return false;
  IV wC = w2;
  IV hC = 1.0 - both;
  double minC = wC.lb;
  if (hC.lb < minC) minC = hC.lb;
  if (disk_satisfies_size_bound(r3.ub, minC)) {
    if (mul_ru(2.0, r3.ub) > mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), add_rd(r1.lb, r2.lb))) return false;
  } else {
    double before = mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), add_rd(r1.lb, r2.lb));
    before = add_rd(before, -mul_ru(2.0, r3.ub));
    double crC = bound_worst_case_ratio(wC, hC);
    double areaC = mul_ru(wC.ub, hC.ub);
    double extra = mul_ru(add_ru(crC, -kRsbRatio), areaC);
    if (before < extra) return false;
  }
  return true;
}

// r1_r2_vertical_strip.cu:28-41
bool shortcut_r1_r2_vertical_strip(const IV& la, const IV& r1, const IV& r2, const IV& r3) {
  double w = two_disks_maximize_height(r1, r2, 1.0);
  if (w <= 0.0) return false;
  double R3 = add_rd(mul_rd(la.lb, kRsbRatio), -add_ru(r1.ub, r2.ub));
  IV wrem = la - w;
  if (wrem.lb >= 1.0 || disk_satisfies_size_bound(r3.ub, wrem.lb)) return R3 >= mul_ru(kRsbRatio, wrem.ub);
  return can_recurse(R3, r3.ub, wrem.ub, 1.0);
}

}  // namespace rsb
}  // namespace ccor
