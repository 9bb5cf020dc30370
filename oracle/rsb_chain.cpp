// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// rectangle_size_bound leaf-criterion chain C1..C14, restated step for step
// from the reference (Prover/rectangle_size_bound/*.cu, cited per function).
// A device trap of the reference (algcuda::trap) becomes: set g_trap and
// return false (or 0 / no change) at that site.
#include "rsb_chain.hpp"

namespace ccor {
namespace rsb {

static inline bool trap_false() {
  g_trap = 1;
  return false;
}

// variables.hpp:53-66: radii[i].lb = max_{j>=i} radii[j].lb for i >= 1.
Cell tighten(const Cell& c) {
  Cell r = c;
  double lb = 0.0;
  for (int i = 5; i > 0; --i) {
    if (lb < c.r[i].lb) lb = c.r[i].lb;
    else r.r[i].lb = lb;
  }
  return r;
}
// variables.hpp:68-80 (lb == ub counts as empty)
bool is_empty(const Cell& c) {
  if (c.la.empty()) return true;
  for (int i = 0; i < 6; ++i)
    if (c.r[i].lb >= c.r[i].ub) return true;
  return false;
}

// variables.cuh:30-52
Vals make_vals(const Cell& c) {
  Vals v;
  const double cr = kRsbRatio;
  v.critical_ratio_sq = IV(mul_rd(cr, cr), mul_ru(cr, cr));
  v.R = c.la * cr;
  for (int i = 0; i < 6; ++i) v.R -= c.r[i];
  if (v.R.lb < 0.0) v.R.lb = 0.0;
  IV additive = iv_reciprocal(IV::pt(cr)) * iv_sqrt(4.0 * v.critical_ratio_sq - 1.0);
  v.min_width_factor = iv_sqrt(2.0 - additive);
  v.max_width_factor = iv_sqrt(2.0 + additive);
  const double den = sqrt_rd(add_rd(1.0, sqrt_rd(add_rd(1.0, -div_ru(0.25, v.critical_ratio_sq.lb)))));
  const double fac = add_rd(1.0, -rcp_ru(den));
  v.weight_bound_wall_building = mul_rd(0.5, mul_rd(fac, fac));
  v.phase_factor_wall_building = 4.0 * v.critical_ratio_sq - cr * iv_sqrt(16.0 * v.critical_ratio_sq - 4.0);
  v.lb_w_infty_denom = add_rd(1.0, -sqrt_ru(v.phase_factor_wall_building.ub));
  v.width_fraction_incomplete =
      rcp_ru(sqrt_rd(add_rd(1.0, sqrt_rd(add_rd(1.0, -div_ru(0.25, v.critical_ratio_sq.lb))))));
  v.cover_2x2.width = 0.0;
  return v;
}

// ------------------------------------------------------------------ C1
// compute_even_split_distance_bound.cu:25-76
static double even_split_distance_bound(const Cell& c, const Vals& v) {
  IV g1 = c.r[0];
  IV g2 = c.r[1] + c.r[2];
  for (int i = 3; i < 6; ++i) {
    if (g1.ub <= g2.ub) g1 += c.r[i];
    else g2 += c.r[i];
  }
  double d = add_ru(g1.ub, -g2.lb);
  double d2 = add_ru(g2.ub, -g1.lb);
  if (d < d2) d = d2;
  double d_minus = add_ru(d, -c.r[5].lb);
  if (d_minus <= v.R.lb) return c.r[5].ub;
  if (g1.ub <= g2.ub) g1 += v.R;
  else g2 += v.R;
  d = add_ru(g1.ub, -g2.lb);
  d2 = add_ru(g2.ub, -g1.lb);
  if (d < d2) d = d2;
  double res = c.r[0].ub;
  if (d <= res) res = d;
  return res;
}
// even_split_recursion.cu:25-40
bool even_split_recursion(const Cell& c, const Vals& v) {
  double dist = even_split_distance_bound(c, v);
  double frac = add_rd(0.5, -div_ru(dist, mul_rd(2.0, mul_rd(c.la.lb, kRsbRatio))));
  double w = mul_rd(frac, c.la.lb);
  if (w >= 1.0) return true;
  return disk_satisfies_size_bound(c.r[0].ub, w);
}

// ------------------------------------------------------------------ C2
// uneven_split_recursion.cu:25-76
bool uneven_split_recursion(const Cell& c, const Vals&) {
  if (c.r[0].lb <= 0.0) return false;
  double need_w = bound_required_height(c.r[0].ub);
  if (need_w >= c.la.lb) return false;
  double need = mul_ru(need_w, kRsbRatio);
  int possibly_at = -1, definitely_at = -1;
  IV wl(0.0, 0.0);
  for (int i = 0; i < 6; ++i) {
    wl += c.r[i];
    if (possibly_at < 0 && wl.ub >= need) possibly_at = i;
    if (wl.lb >= need) {
      definitely_at = i;
      break;
    }
  }
  double largest_rem = (possibly_at < 0 || possibly_at == 5) ? c.r[5].ub : c.r[possibly_at + 1].ub;
  double D1 = definitely_at < 0 ? add_ru(need, c.r[5].ub) : wl.ub;
  if (D1 >= mul_rd(c.la.lb, kRsbRatio)) return false;
  double w1 = div_ru(D1, kRsbRatio);
  double wrem = add_rd(c.la.lb, -w1);
  return wrem >= 1.0 || disk_satisfies_size_bound(largest_rem, wrem);
}

// ------------------------------------------------------------------ C3
// multi_disk_strip.cu:31-59
static bool strip_vertical_nd(const Cell& c, int k) {
  double wgt = 0.0;
  for (int j = 0; j < k; ++j) wgt = add_ru(wgt, c.r[j].ub);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h = 0.0;
  for (int j = 0; j < k; ++j) {
    double hc = add_rd(mul_rd(4.0, c.r[j].lb), -wsq);
    if (hc <= 0.0) break;
    h = add_rd(h, sqrt_rd(hc));
  }
  if (h < 1.0) return false;
  double largest = k == 6 ? c.r[5].ub : c.r[k].ub;
  double rem = add_rd(c.la.lb, -w);
  return rem >= 1.0 || disk_satisfies_size_bound(largest, rem);
}
// multi_disk_strip.cu:61-93
static bool stacked_r3_r4_horizontal(const Cell& c) {
  double wgt = 0.0;
  for (int j = 0; j < 4; ++j) wgt = add_ru(wgt, c.r[j].ub);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h1 = add_rd(mul_rd(4.0, c.r[0].lb), -wsq);
  double h2 = add_rd(mul_rd(4.0, c.r[1].lb), -wsq);
  if (h1 <= 0.0 || h2 <= 0.0) return false;
  double h12 = add_rd(sqrt_rd(h1), sqrt_rd(h2));
  double hrem = add_ru(1.0, -h12);
  double hrem_sq = mul_ru(hrem, hrem);
  double w4 = add_rd(mul_rd(4.0, c.r[3].lb), -hrem_sq);
  double w3 = add_rd(mul_rd(4.0, c.r[2].lb), -hrem_sq);
  if (w4 <= 0.0 || w3 <= 0.0) return false;
  if (add_rd(sqrt_rd(w3), sqrt_rd(w4)) < w) return false;
  double rem = add_rd(c.la.lb, -w);
  return rem >= 1.0 || disk_satisfies_size_bound(c.r[4].ub, rem);
}
// multi_disk_strip.cu:95-145
static bool stacked_r34_r56_horizontal(const Cell& c) {
  double wgt = 0.0;
  for (int j = 0; j < 6; ++j) wgt = add_ru(wgt, c.r[j].ub);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h1 = add_rd(mul_rd(4.0, c.r[0].lb), -wsq);
  double h2 = add_rd(mul_rd(4.0, c.r[1].lb), -wsq);
  if (h1 <= 0.0 || h2 <= 0.0) return false;
  IV d34 = c.r[2] - c.r[3];
  d34.tighten_lb(0.0);
  IV w3 = (2.0 * d34) / w + 0.5 * IV::pt(w);
  IV w3sq = w3 * w3;
  if (!definitely(4.0 * c.r[2] > w3sq) || !definitely(4.0 * c.r[3] > (w - w3) * (w - w3))) return false;
  double h34 = sqrt_rd(add_rd(mul_rd(4.0, c.r[2].lb), -w3sq.ub));
  IV d56 = c.r[4] - c.r[5];
  d56.tighten_lb(0.0);
  IV w5 = (2.0 * d56) / w + 0.5 * IV::pt(w);
  IV w5sq = w5 * w5;
  if (!definitely(4.0 * c.r[4] > w5sq) || !definitely(4.0 * c.r[5] > (w - w5) * (w - w5))) return false;
  double h56 = sqrt_rd(add_rd(mul_rd(4.0, c.r[4].lb), -w5sq.ub));
  double h12 = add_rd(sqrt_rd(h1), sqrt_rd(h2));
  if (add_rd(h12, add_rd(h34, h56)) < 1.0) return false;
  double rem = add_rd(c.la.lb, -w);
  return rem >= 1.0 || disk_satisfies_size_bound(c.r[5].ub, rem);
}
// multi_disk_strip.cu:147-240 (lb_h14/lb_h23 take sqrt twice: reference quirk)
static bool six_disk_strip_three_rows(const Cell& c, const Vals& v) {
  double wgt = 0.0;
  for (int i = 0; i < 6; ++i) wgt = add_ru(wgt, c.r[i].ub);
  double wn = div_ru(wgt, kRsbRatio);
  double ub_r7 = c.r[5].ub;
  if (v.R.ub < ub_r7) ub_r7 = v.R.ub;
  double wrem = add_rd(c.la.lb, -wn);
  if (wrem < 1.0 && !disk_satisfies_size_bound(ub_r7, wrem)) return false;
  IV w = IV::pt(wn);
  IV wsq = w * wn;
  IV w1 = 2.0 / w * (c.r[0] - c.r[3]) + 0.5 * w;
  IV w2 = 2.0 / w * (c.r[1] - c.r[2]) + 0.5 * w;
  double d14 = add_ru(c.r[0].ub, -c.r[3].lb);
  double d14sq = mul_ru(d14, d14);
  double h14 = sqrt_rd(add_rd(add_rd(mul_rd(2.0, add_rd(c.r[0].lb, c.r[3].lb)), -mul_ru(0.25, wsq.ub)),
                              -div_ru(mul_ru(4.0, d14sq), wsq.lb)));
  if (h14 < 0) return false;
  h14 = sqrt_rd(h14);
  double d23 = add_ru(c.r[1].ub, -c.r[2].lb);
  double d23sq = mul_ru(d23, d23);
  double h23 = sqrt_rd(add_rd(add_rd(mul_rd(2.0, add_rd(c.r[1].lb, c.r[2].lb)), -mul_ru(0.25, wsq.ub)),
                              -div_ru(mul_ru(4.0, d23sq), wsq.lb)));
  if (h23 < 0) return false;
  h23 = sqrt_rd(h23);
  IV hrem = 1.0 - IV::pt(h23) - h14;
  if (hrem.ub <= 0.0) return true;
  double x5 = add_rd(c.r[4].lb, -mul_ru(0.25, mul_ru(hrem.ub, hrem.ub)));
  double w6 = add_rd(c.r[5].lb, -mul_ru(0.25, mul_ru(hrem.ub, hrem.ub)));
  if (x5 < 0 || w6 < 0) return false;
  x5 = sqrt_rd(x5);
  w6 = sqrt_rd(w6);
  double x6 = add_ru(wn, -w6);
  IV y5 = 0.5 * hrem + h23;
  double y1 = add_ru(1.0, -mul_rd(0.5, h14));
  double y2 = mul_rd(0.5, h23);
  Point p5{IV::pt(x5), y5};
  Point p1{w - 0.5 * w1, IV::pt(y1)};
  Point p2{w - 0.5 * w2, IV::pt(y2)};
  Point p3{0.5 * (w - w2), IV::pt(y2)};
  Point p4{0.5 * (w - w1), IV::pt(y1)};
  Point p6{IV::pt(x6), y5};
  Circle c1{p1, IV::pt(c.r[0].lb)};
  Circle c2{p2, IV::pt(c.r[1].lb)};
  Circle c5{p5, IV::pt(c.r[4].lb)};
  Intersection x15 = intersection(c5, c1);
  Intersection x25 = intersection(c5, c2);
  if (!x15.definitely_intersecting || !x25.definitely_intersecting) return false;
  if (squared_distance(p4, x15.p[0]).ub > c.r[3].lb) return false;
  if (squared_distance(p3, x25.p[0]).ub > c.r[2].lb) return false;
  return squared_distance(p6, x15.p[1]).ub <= c.r[5].lb && squared_distance(p6, x25.p[1]).ub <= c.r[5].lb;
}
// multi_disk_strip.cu:242-262
bool multi_disk_strip_vertical(const Cell& c, const Vals& v) {
  for (int i = 2; i <= 6; ++i)
    if (strip_vertical_nd(c, i)) return true;
  if (stacked_r3_r4_horizontal(c)) return true;
  if (stacked_r34_r56_horizontal(c)) return true;
  return six_disk_strip_three_rows(c, v);
}

// ------------------------------------------------------------------ C11
// multi_disk_strip.cu:264-307
static bool adv_strip_one(const Cell& c, unsigned set) {
  double wgt = 0.0;
  for (int j = 0; j < 6; ++j)
    if (set & (1u << j)) wgt = add_ru(wgt, c.r[j].ub);
  double w = div_ru(wgt, kRsbRatio);
  double wsq = mul_ru(w, w);
  double h = 0.0;
  int largest = -1;
  for (int j = 0; j < 6; ++j) {
    if (!(set & (1u << j))) {
      if (largest < 0) largest = j;
      continue;
    }
    double hc = add_rd(mul_rd(4.0, c.r[j].lb), -wsq);
    if (hc <= 0.0) return false;
    h = add_rd(h, sqrt_rd(hc));
  }
  if (h < 1.0) return false;
  double rem = add_rd(c.la.lb, -w);
  return rem >= 1.0 || disk_satisfies_size_bound(c.r[largest].ub, rem);
}
// multi_disk_strip.cu:309-317
bool advanced_multi_disk_strip_vertical(const Cell& c, const Vals&) {
  for (unsigned s = 1; s < 63; ++s)
    if (adv_strip_one(c, s)) return true;
  return false;
}

// ------------------------------------------------------------------ C4
// vertical_wall_building_recursion.cu:25-78 ("remaining width" adds: quirk)
bool vertical_wall_building_recursion(const Cell& c, const Vals& v) {
  double col_w = sqrt_ru(mul_ru(c.r[5].ub, 2.0));
  double rem_w = add_rd(c.la.lb, col_w);
  if (c.r[5].lb <= 0 || v.R.lb <= 0) return false;
  if (rem_w < 1.0 && !disk_satisfies_size_bound(c.r[0].ub, rem_w)) return false;
  if (c.r[5].ub > v.weight_bound_wall_building) return false;
  double R6 = add_rd(c.r[5].lb, v.R.lb);
  double below = mul_ru(c.r[5].ub, v.phase_factor_wall_building.ub);
  double in_phases = mul_ru(kRsbRatio, sqrt_ru(mul_ru(2.0, c.r[5].ub)));
  for (int k = 0; k < 32; ++k) {
    double total_below = add_rd(R6, -in_phases);
    if (total_below <= 0) break;
    double w_below = bound_required_height(below);
    double req = mul_ru(kRsbRatio, w_below);
    if (req <= total_below) {
      double used = add_ru(req, below);
      double strip = div_ru(used, kRsbRatio);
      double rem = add_rd(c.la.lb, strip);
      if (rem >= 1.0 || disk_satisfies_size_bound(c.r[0].ub, rem)) return true;
    }
    in_phases = add_ru(in_phases, mul_ru(kRsbRatio, sqrt_ru(mul_ru(2.0, below))));
    below = mul_ru(below, v.phase_factor_wall_building.ub);
  }
  return false;
}

// ------------------------------------------------------------------ C5
// r1_in_corner_explicit_recursion.cu:36-111 (above) and :113-187 (right)
static bool corner_explicit_nd(const Cell& c, const Vals& v, int nd, bool right) {
  double wgt = 0.0;
  for (int j = nd; j >= 0; --j) wgt = add_ru(wgt, c.r[j].ub);
  double span;  // the width (above) or height (right) the strip must cover
  double max_rem = (nd == 5) ? (c.r[5].ub > v.R.ub ? v.R.ub : c.r[5].ub) : c.r[nd + 1].ub;
  double req_rem = bound_required_height(max_rem);
  if (!right) {
    span = div_ru(wgt, kRsbRatio);
    if (span > add_rd(c.la.lb, -req_rem)) return false;
  } else {
    span = div_ru(div_ru(wgt, kRsbRatio), c.la.lb);
    if (span > add_rd(1.0, -req_rem)) return false;
  }
  double start = add_rd(mul_rd(4.0, c.r[0].lb), -mul_ru(span, span));
  if (start <= 0) return false;
  start = sqrt_rd(start);
  struct Entry {
    double reached;
    int used;
    int choice;
  } st[6];
  int sh = 1;
  st[0] = {start, 0, 0};
  const double goal = right ? c.la.ub : 1.0;
  while (sh > 0) {
    int s = sh - 1;
    int remaining = nd - st[s].used;
    int next = ++st[s].choice;
    if (next > remaining) {
      --sh;
      continue;
    }
    int first = 1 + st[s].used;
    if (next == remaining) {
      if (nd_can_cover(span, add_ru(goal, -st[s].reached), &c.r[first], next)) return true;
      --sh;
      continue;
    }
    double got = nd_maximize_covered(span, &c.r[first], next);
    if (got == 0) continue;
    double reach = add_rd(st[s].reached, got);
    if (reach >= goal) return true;
    st[sh] = {reach, st[s].used + next, 0};
    ++sh;
  }
  return false;
}
// r1_in_corner_explicit_recursion.cu:189-251
static bool corner_right_34_and_2(const Cell& c) {
  IV w1 = iv_sqrt(2.0 * c.r[0]);
  IV wrem = c.la - w1;
  double hA = add_rd(1.0, -w1.ub);
  if (!disk_satisfies_size_bound(c.r[4].ub, hA)) return false;
  double h34 = two_disks_maximize_height(c.r[2], c.r[3], wrem.ub);
  if (h34 <= 0.0 || h34 >= w1.lb) return false;
  double hrem = add_ru(w1.ub, -h34);
  double w2 = add_rd(mul_rd(4.0, c.r[1].lb), -mul_ru(hrem, hrem));
  if (w2 <= 0.0) return false;
  w2 = sqrt_rd(w2);
  bool no_B = false;
  if (w2 >= wrem.ub) {
    no_B = true;
    w2 = wrem.lb;
  }
  if (w2 > wrem.lb) w2 = wrem.lb;
  double a234 = add_rd(mul_rd(w2, w1.lb), mul_rd(h34, add_rd(wrem.lb, -w2)));
  double wt234 = add_ru(c.r[1].ub, add_ru(c.r[2].ub, c.r[3].ub));
  double extra = add_rd(mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), c.r[0].lb),
                        add_rd(mul_rd(kRsbRatio, a234), -wt234));
  if (extra <= 0.0) return false;
  if (no_B) return true;
  IV hB = w1 - h34;
  IV wB = wrem - w2;
  double minB = hB.lb;
  if (minB > wB.lb) minB = wB.lb;
  double cost = c.r[5].ub;
  if (!disk_satisfies_size_bound(c.r[5].ub, minB)) {
    double crB = bound_worst_case_ratio(wB, hB);
    double aB = mul_ru(hB.ub, wB.ub);
    cost = add_ru(cost, mul_ru(add_ru(crB, -kRsbRatio), aB));
  }
  return cost <= extra;
}
// r1_in_corner_explicit_recursion.cu:253-261
bool r1_in_corner_explicit_recursion(const Cell& c, const Vals& v) {
  for (int i = 2; i <= 5; ++i)
    if (corner_explicit_nd(c, v, i, false) || corner_explicit_nd(c, v, i, true)) return true;
  return corner_right_34_and_2(c);
}

// ------------------------------------------------------------------ C6
namespace {
struct ShortLong {
  IV shorter, longer;
};
struct SizeCheck {
  double cost_bigger, cost_smaller;
  bool works;
};
}  // namespace
// r1_r2_opposite_corners_recursion.cu:146-179
static ShortLong shorter_longer(const IV& w, const IV& h) {
  UB cmp = w <= h;
  if (is_certain(cmp)) return cmp.lo ? ShortLong{w, h} : ShortLong{h, w};
  ShortLong r;
  if (w.lb < h.lb) {
    r.shorter.lb = w.lb;
    r.longer.lb = h.lb;
  } else {
    r.shorter.lb = h.lb;
    r.longer.lb = w.lb;
  }
  if (w.ub < h.ub) {
    r.shorter.ub = w.ub;
    r.longer.ub = h.ub;
  } else {
    r.shorter.ub = h.ub;
    r.longer.ub = w.ub;
  }
  return r;
}
// r1_r2_opposite_corners_recursion.cu:188-252
static SizeCheck opposite_check_sizes(const Cell& c, const Vals& v, const IV& wA, const IV& hA, const IV& wB,
                                      const IV& hB, int k) {
  double minA = wA.lb;
  if (hA.lb < minA) minA = hA.lb;
  double minB = wB.lb;
  if (hB.lb < minB) minB = hB.lb;
  double areaA = mul_rd(wA.lb, hA.lb);
  double areaB = mul_rd(wB.lb, hB.lb);
  double big = minA < minB ? minB : minA;
  double small = minA < minB ? minA : minB;
  if (k < 6) {
    if (!disk_satisfies_size_bound(c.r[k].ub, big)) return {0.0, 0.0, false};
    double big_area = minA < minB ? areaB : areaA;
    double big_weight = mul_rd(kRsbRatio, big_area);
    double in_big = 0.0;
    for (; k < 6; ++k) {
      in_big = add_ru(in_big, c.r[k].ub);
      if (in_big >= big_weight) break;
    }
  }
  if (k < 6) {
    if (!disk_satisfies_size_bound(c.r[k].ub, small)) return {0.0, 0.0, false};
    double cost_big = c.r[k].ub;
    double small_area = minA < minB ? areaA : areaB;
    double small_weight = mul_rd(kRsbRatio, small_area);
    double in_small = 0.0;
    for (; k < 6; ++k) {
      in_small = add_ru(in_small, c.r[k].ub);
      if (in_small >= small_weight) break;
    }
    if (k < 6) return {cost_big, c.r[k].ub, true};
    double r7 = c.r[5].ub;
    if (v.R.ub < r7) r7 = v.R.ub;
    return {cost_big, r7, true};
  }
  double r7 = c.r[5].ub;
  if (v.R.ub < r7) r7 = v.R.ub;
  if (!disk_satisfies_size_bound(r7, small)) return {0.0, 0.0, false};
  return {r7, r7, true};
}
// r1_r2_opposite_corners_recursion.cu:254-340
bool r1_r2_opposite_corners_recursion(const Cell& c, const Vals& v) {
  const IV& r1 = c.r[0];
  const IV& r2 = c.r[1];
  IV w1(sqrt_rd(mul_rd(2.0, r1.lb)), sqrt_ru(mul_ru(2.0, r1.ub)));
  IV w2(sqrt_rd(mul_rd(2.0, r2.lb)), sqrt_ru(mul_ru(2.0, r2.ub)));
  IV both = w1 + w2;
  double extra12 = mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), add_rd(r1.lb, r2.lb));
  if (!(both.ub < 1.0)) return false;
  IV wA = c.la - w2, wB = c.la - w1, hA = 1.0 - w1, hB = w1, wC = w2, hC = 1.0 - both;
  ShortLong sC = shorter_longer(wC, hC);
  IV range = compute_efficient_rectangle_cover_weight_range(sC.longer.lb, sC.shorter);
  int k;
  for (k = 2; k < 6; ++k) {
    if (range.lb > c.r[k].lb || c.r[k].ub > range.ub) break;
    IV len = 4.0 * c.r[k] - iv_square(sC.shorter);
    if (possibly(len <= 0.0)) break;
    len = iv_sqrt(len);
    sC.longer -= len;
    sC.longer.tighten_lb(0.0);
    if (sC.longer.empty()) {
      SizeCheck sb = opposite_check_sizes(c, v, wA, hA, wB, hB, k + 1);
      if (!sb.works) return false;
      return sb.cost_bigger <= extra12;
    }
    sC = shorter_longer(sC.shorter, sC.longer);
    range = compute_efficient_rectangle_cover_weight_range(sC.longer.lb, sC.shorter);
  }
  SizeCheck first = opposite_check_sizes(c, v, wA, hA, wB, hB, k);
  if (first.works) {
    if (disk_satisfies_size_bound(first.cost_smaller, sC.shorter.lb))
      return add_ru(first.cost_smaller, first.cost_bigger) <= extra12;
  }
  double minareaC = mul_rd(wC.lb, hC.lb);
  double crC = bound_worst_case_ratio(wC, hC);
  double minweightC = mul_rd(crC, minareaC);
  double big_in_C = 0.0;
  for (; k < 6; ++k) {
    big_in_C = add_ru(big_in_C, c.r[k].ub);
    if (big_in_C >= minweightC) break;
  }
  SizeCheck second = opposite_check_sizes(c, v, wA, hA, wB, hB, k);
  if (second.works) {
    double split = add_ru(second.cost_smaller, second.cost_bigger);
    double extraC = mul_ru(add_ru(crC, -kRsbRatio), mul_ru(wC.ub, hC.ub));
    return add_ru(split, extraC) <= extra12;
  }
  return false;
}

// ------------------------------------------------------------------ C7
// r1_r2_vertical_strip.cu:43-95
bool r1_r2_vertical_strip(const Cell& c, const Vals& v) {
  const IV r1 = c.r[0], r2 = c.r[1];
  double w = two_disks_maximize_height(r1, r2, 1.0);
  if (w <= 0.0) return false;
  double R3 = add_rd(mul_rd(c.la.lb, kRsbRatio), -add_ru(r1.ub, r2.ub));
  IV wrem = c.la - w;
  double ub_w = wrem.ub, ub_h = 1.0, R = R3;
  for (int k = 2; k < 6; ++k) {
    if (can_recurse(R, c.r[k].ub, ub_w, ub_h)) return true;
    double m = ub_w < ub_h ? ub_w : ub_h;
    double cut = add_rd(mul_rd(4.0, c.r[k].lb), -mul_ru(m, m));
    if (cut <= 0) return false;
    cut = sqrt_rd(cut);
    if (ub_w < ub_h) {
      ub_h = add_ru(ub_h, -cut);
      if (ub_h <= 0.0) return true;
    } else {
      ub_w = add_ru(ub_w, -cut);
      if (ub_w <= 0.0) return true;
    }
    R = add_rd(R, -c.r[k].ub);
  }
  double r7 = c.r[5].ub < v.R.ub ? c.r[5].ub : v.R.ub;
  return can_recurse(R, r7, ub_w, ub_h);
}

// ------------------------------------------------------------------ 2x2 cover + C8
// two_by_two_cover.cu:29-76
static TwoByTwo two_by_two_concrete(double r1, double r2, double r3, double r4) {
  TwoByTwo res;
  res.width = 0.0;
  IV w1;
  {
    IV d = IV::pt(r1) - r2;
    IV dsq = iv_square(d);
    IV h1 = 0.5 + 2.0 * d;
    IV h2 = 0.5 - 2.0 * d;
    w1 = 2.0 * (IV::pt(r1) + r2) - 0.25 - 4.0 * dsq;
    if (possibly(r1 - 0.25 * iv_square(h1) <= 0.0) || possibly(r2 - 0.25 * iv_square(h2) <= 0.0) ||
        possibly(w1 <= 0.0)) {
      w1 = IV(0.0, 0.0);
      res.circles[0] = Circle{{IV(0, 0), IV(0, 0)}, IV::pt(r1)};
      res.circles[1] = Circle{{IV(0, 0), IV(0, 0)}, IV::pt(r2)};
    } else {
      w1 = iv_sqrt(w1);
      IV x = 0.5 * w1;
      res.circles[0] = Circle{{x, 0.5 * h1}, IV::pt(r1)};
      res.circles[1] = Circle{{x, 1.0 - 0.5 * h2}, IV::pt(r2)};
      res.width = w1.lb;
    }
  }
  {
    IV d = IV::pt(r3) - r4;
    IV dsq = iv_square(d);
    IV h1 = 0.5 + 2.0 * d;
    IV h2 = 0.5 - 2.0 * d;
    IV w2 = 2.0 * (IV::pt(r3) + r4) - 0.25 - 4.0 * dsq;
    if (possibly(r3 - 0.25 * iv_square(h1) <= 0.0) || possibly(r4 - 0.25 * iv_square(h2) <= 0.0) ||
        possibly(w2 <= 0.0)) {
      res.circles[3] = Circle{{IV(0, 0), IV(0, 0)}, IV::pt(r3)};
      res.circles[2] = Circle{{IV(0, 0), IV(0, 0)}, IV::pt(r4)};
    } else {
      w2 = iv_sqrt(w2);
      IV x = w1 + 0.5 * w2;
      res.circles[3] = Circle{{x, 1.0 - 0.5 * h1}, IV::pt(r3)};
      res.circles[2] = Circle{{x, 0.5 * h2}, IV::pt(r4)};
      res.width = add_rd(res.width, w2.lb);
    }
  }
  return res;
}
// two_by_two_cover.cu:78-88
TwoByTwo compute_two_by_two_cover(const IV* d) {
  TwoByTwo c1 = two_by_two_concrete(d[0].lb, d[1].lb, d[2].lb, d[3].lb);
  TwoByTwo c2 = two_by_two_concrete(d[0].lb, d[2].lb, d[1].lb, d[3].lb);
  TwoByTwo c3 = two_by_two_concrete(d[0].lb, d[3].lb, d[1].lb, d[2].lb);
  if (c1.width > c2.width) return c1.width > c3.width ? c1 : c3;
  return c2.width > c3.width ? c2 : c3;
}
// two_by_two_cover_with_strip_and_recursion.cu:26-105
bool two_by_two_cover_with_strip_and_recursion(const Cell& c, const Vals& v) {
  const double r5 = c.r[4].lb, r6 = c.r[5].lb, R = v.R.lb;
  double ub_w = add_ru(c.la.ub, -v.cover_2x2.width);
  double R5 = mul_rd(kRsbRatio, c.la.lb);
  double R6 = add_rd(R5, -c.r[4].ub);
  for (int i = 0; i < 4; ++i) {
    R5 = add_rd(R5, -c.r[i].ub);
    R6 = add_rd(R6, -c.r[i].ub);
  }
  if (can_recurse(R5, c.r[4].ub, ub_w, 1.0)) return true;
  double wsq = mul_ru(ub_w, ub_w);
  double h5 = add_rd(mul_rd(4.0, r5), -wsq);
  if (h5 <= 0.0) return false;
  h5 = sqrt_rd(h5);
  double hrem = add_ru(1.0, -h5);
  if (h5 >= 1.0 || can_recurse(R6, c.r[5].ub, ub_w, hrem)) return true;
  double h6 = add_rd(mul_rd(4.0, r6), -wsq);
  if (h6 > 0.0) {
    h6 = sqrt_rd(h6);
    if (h6 >= hrem) return true;
    if (can_recurse(R, c.r[5].ub, ub_w, add_ru(hrem, -h6))) return true;
  }
  Circle c5{{c.la - IV::pt(0.5) * ub_w, IV::pt(0.5) * h5}, IV::pt(r5)};
  double hremsq = mul_ru(hrem, hrem);
  double w6 = add_rd(mul_rd(4.0, r6), -hremsq);
  if (w6 > 0.0) {
    w6 = sqrt_rd(w6);
    if (w6 >= ub_w) return true;
    Circle c6{{c.la - IV::pt(0.5) * w6, 1.0 - IV::pt(0.5) * hrem}, IV::pt(r6)};
    Intersection x56 = intersection(c5, c6);
    if (x56.definitely_intersecting && x56.p[0].x.ub < x56.p[1].x.lb) {
      const Circle& c4 = v.cover_2x2.circles[3];
      if (definitely(squared_distance(x56.p[0], c4.center) <= c4.squared_radius)) {
        Intersection x46 = intersection(c6, c4);
        if (x46.definitely_intersecting) {
          UB first_higher = x46.p[0].y > x46.p[1].y;
          if (is_certain(first_higher)) {
            double y = x46.p[first_higher.lo ? 0 : 1].y.lb;
            double hrect = add_ru(1.0, -y);
            if (can_recurse(v.R.lb, c.r[5].ub, add_ru(ub_w, -w6), hrect)) return true;
          }
        }
      }
    }
  }
  return false;
}

// ------------------------------------------------------------------ C9
// l_shaped_recursion.cu:28-120
static bool l_shaped_one_horizontal(const Cell& c, double ub_w, double lb_weight) {
  double R4 = add_rd(lb_weight, -c.r[2].ub);
  double side3 = sqrt_rd(mul_rd(2.0, c.r[2].lb));
  IV h_top = IV(1.0, 1.0) - side3;
  double min_top = h_top.lb;
  if (ub_w < min_top) min_top = ub_w;
  if (disk_satisfies_size_bound(c.r[3].ub, min_top)) {
    double w_top = add_ru(mul_ru(kRsbRatio, mul_ru(h_top.ub, ub_w)), c.r[3].ub);
    double rem_br = add_rd(R4, -w_top);
    if (rem_br > 0) {
      double w_br = add_ru(ub_w, -side3);
      double m = w_br;
      if (side3 < m) m = side3;
      if (disk_satisfies_size_bound(c.r[4].ub, m)) {
        if (mul_ru(kRsbRatio, mul_ru(side3, w_br)) <= rem_br) return true;
      } else if (can_recurse(rem_br, c.r[4].ub, w_br, side3)) {
        return true;
      }
    }
  }
  IV w_right = c.la - side3;
  double min_right = w_right.lb;
  if (min_right > 1.0) min_right = 1.0;
  if (min_right >= 1.0 || disk_satisfies_size_bound(c.r[3].ub, min_right)) {
    double w_r = add_ru(mul_ru(kRsbRatio, w_right.ub), c.r[3].ub);
    double rem_tl = add_rd(R4, -w_r);
    if (rem_tl > 0) {
      double h_tl = add_ru(1.0, -side3);
      double m = h_tl;
      if (side3 < m) m = side3;
      if (disk_satisfies_size_bound(c.r[4].ub, m)) {
        if (mul_ru(kRsbRatio, mul_ru(side3, h_tl)) <= rem_tl) return true;
      } else if (can_recurse(rem_tl, c.r[4].ub, side3, h_tl)) {
        return true;
      }
    }
  }
  {
    double wsq = mul_ru(ub_w, ub_w);
    double h3 = add_rd(mul_rd(4.0, c.r[2].lb), -wsq);
    if (h3 > 0.0) {
      h3 = sqrt_rd(h3);
      double dw = c.r[3].ub;
      double hrem = add_ru(1.0, -h3);
      double m = hrem;
      double area = mul_ru(ub_w, hrem);
      if (ub_w < m) m = ub_w;
      if (disk_satisfies_size_bound(dw, m)) {
        if (mul_ru(area, kRsbRatio) <= R4) return true;
      } else if (can_recurse(R4, dw, ub_w, hrem)) {
        return true;
      }
    }
  }
  return false;
}
// l_shaped_recursion.cu:122-154
bool l_shaped_recursion(const Cell& c, const Vals&) {
  {
    double wv = add_ru(c.r[0].ub, c.r[1].ub);
    double rem = add_rd(mul_rd(c.la.lb, kRsbRatio), -wv);
    double wmax = two_disks_maximize_height(c.r[0], c.r[1], 1.0);
    double wrem = add_ru(c.la.ub, -wmax);
    if (!(mul_rd(wrem, kRsbRatio) > rem) && l_shaped_one_horizontal(c, wrem, rem)) return true;
  }
  double w12 = two_disks_maximize_height(c.r[0], c.r[1], 1.0);
  if (w12 <= 0.0) return false;
  double wrem = add_ru(c.la.ub, -w12);
  double h34 = two_disks_maximize_height(c.r[2], c.r[3], wrem);
  if (h34 <= 0.0) return false;
  double hrem = add_ru(1.0, -h34);
  double R5 = add_rd(mul_rd(c.la.lb, kRsbRatio),
                     -add_ru(c.r[0].ub, add_ru(c.r[1].ub, add_ru(c.r[2].ub, c.r[3].ub))));
  return can_recurse(R5, c.r[5].ub, wrem, hrem);
}

// ------------------------------------------------------------------ C10
// r1_in_corner_wall_building_recursion.cu:28-149
bool r1_in_corner_wall_building_recursion(const Cell& c, const Vals& v) {
  IV w1 = iv_sqrt(2.0 * c.r[0]);
  double extra_r1 = mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), c.r[0].lb);
  IV wB = c.la - w1;
  IV hA = 1.0 - w1;
  IV minA(hA.lb < w1.lb ? hA.lb : w1.lb, hA.ub < w1.ub ? hA.ub : w1.ub);
  IV maxA(hA.lb < w1.lb ? w1.lb : hA.lb, hA.ub < w1.ub ? w1.ub : hA.ub);
  IV minA_sq = iv_square(minA);
  IV maxA_sq = iv_square(maxA);
  IV boundA = kDiskWeightMax * minA_sq;
  double costA = boundA.ub;
  if (costA >= c.r[5].ub) costA = c.r[5].ub;
  if (boundA.lb <= 0.0 || costA > extra_r1) return false;
  if (possibly(wB < 1.0) && !disk_satisfies_size_bound(c.r[1].ub, wB.lb)) {
    double e23 = -kInf;
    {
      double h23 = two_disks_maximize_height(c.r[1], c.r[2], wB.ub);
      double hremB = add_rd(1.0, -h23);
      double mB = hremB < wB.lb ? hremB : wB.lb;
      if (disk_satisfies_size_bound(c.r[3].ub, mB)) {
        double wt = add_ru(c.r[1].ub, c.r[2].ub);
        e23 = add_rd(mul_rd(kRsbRatio, mul_rd(h23, wB.lb)), -wt);
      }
    }
    double e234 = -kInf;
    {
      double h234 = three_disks_maximize_height(c.r[1], c.r[2], c.r[3], wB.ub);
      double hremB = add_rd(1.0, -h234);
      double mB = hremB < wB.lb ? hremB : wB.lb;
      if (disk_satisfies_size_bound(c.r[4].ub, mB)) {
        double wt = add_ru(c.r[1].ub, add_ru(c.r[2].ub, c.r[3].ub));
        e234 = add_rd(mul_rd(kRsbRatio, mul_rd(h234, wB.lb)), -wt);
      }
    }
    double better = e23 > e234 ? e23 : e234;
    extra_r1 = add_rd(extra_r1, better);
    if (extra_r1 <= 0.0) return false;
  }
  double wb_short = mul_rd(v.weight_bound_wall_building, minA_sq.lb);
  double wb_long = mul_rd(v.weight_bound_wall_building, maxA_sq.lb);
  if (wb_long < c.r[5].ub) return false;
  bool use_short = wb_short >= c.r[5].ub;
  double costA_wb = mul_ru(mul_ru(kRsbRatio, use_short ? minA.ub : maxA.ub), sqrt_ru(mul_ru(2.0, c.r[5].ub)));
  if (costA_wb > costA) costA = costA_wb;
  if (costA > extra_r1) return false;
  if (c.r[5].ub <= boundA.lb) return true;
  double R6 = mul_rd(kRsbRatio, c.la.lb);
  for (int i = 0; i < 5; ++i) R6 = add_rd(R6, -c.r[i].ub);
  double max_incomplete = mul_ru(use_short ? minA.ub : maxA.ub, v.width_fraction_incomplete);
  double num = mul_ru(kRsbRatio, mul_ru(max_incomplete, sqrt_ru(mul_ru(2.0, c.r[5].ub))));
  double dropped_inf = div_ru(num, v.lb_w_infty_denom);
  double kept_inf = add_rd(R6, -dropped_inf);
  double areaA = mul_ru(w1.ub, hA.ub);
  double needA = mul_ru(kRsbRatio, areaA);
  if (kept_inf >= needA) return true;
  double dropped = needA;
  double cur = c.r[5].ub;
  for (int it = 0; it < kBisectCap && cur > boundA.lb; ++it) {
    dropped = add_ru(dropped, mul_ru(kRsbRatio, mul_ru(max_incomplete, sqrt_ru(mul_ru(2.0, cur)))));
    cur = mul_ru(cur, v.phase_factor_wall_building.ub);
  }
  return add_rd(R6, -dropped) >= needA;
}

}  // namespace rsb
}  // namespace ccor
