// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// tight_rectangle criteria, restated step for step from the reference
// (Prover/tight_rectangle/*.cu).  Every rounded operation, branch and
// comparison follows the cited reference lines in the same order.
#include "criteria.hpp"

namespace ccor {

thread_local int g_trap = 0;
#ifdef CC_ORACLE_COUNT
thread_local uint64_t g_op_count = 0;
#endif

// Prover/operations.cu:30-83
Intersection intersection(const Circle& c1, const Circle& c2) {
  Intersection res;
  res.definitely_intersecting = false;
  res.p[0] = res.p[1] = Point{IV(0, 0), IV(0, 0)};
  IV d2 = squared_distance(c1.center, c2.center);
  if (d2.lb == 0) return res;
  IV rdiff = c1.squared_radius - c2.squared_radius;
  IV rdiff_sq = iv_square(rdiff);
  IV f1 = (2.0 * (c1.squared_radius + c2.squared_radius)) / d2 - 1.0 - rdiff_sq / iv_square(d2);
  if (f1.lb < 0) return res;
  f1 = 0.5 * iv_sqrt(f1);
  res.definitely_intersecting = true;
  IV f2 = (0.5 * rdiff) / d2;
  const IV& x1 = c1.center.x;
  const IV& x2 = c2.center.x;
  const IV& y1 = c1.center.y;
  const IV& y2 = c2.center.y;
  IV cx = 0.5 * (x1 + x2);
  IV cy = 0.5 * (y1 + y2);
  IV dx = x2 - x1;
  IV dy = y2 - y1;
  cx += f2 * dx;
  cy += f2 * dy;
  Point a{cx + f1 * dy, cy - f1 * dx};
  Point b{cx - f1 * dy, cy + f1 * dx};
  // Point order is chosen from the bounds without a certainty check.
  bool a_first = dy.ub < 0 || (dy.ub == 0 && dx.lb >= 0);
  res.p[0] = a_first ? a : b;
  res.p[1] = a_first ? b : a;
  return res;
}

namespace tr {

// critical_ratio.cu:58-78.  The break-even bounds are recomputed on every
// call exactly as the reference does.
IV critical_ratio(const IV& la) {
  const double be_lo = sqrt_rd(add_rd(mul_rd(0.5, sqrt_rd(7.0)), -0.25));
  const double be_hi = sqrt_ru(add_ru(mul_ru(0.5, sqrt_ru(7.0)), -0.25));
  auto small_branch = [&]() {
    IV sq = iv_square(la);
    return ((3.0 / 256.0) / la) * (16.0 * sq + 40.0 + 9.0 / sq);
  };
  auto large_branch = [&]() { return 0.25 * la + 0.5 * iv_reciprocal(la); };
  if (definitely(la <= be_lo)) return small_branch();
  if (definitely(la >= be_hi)) return large_branch();
  IV a = small_branch();
  IV b = large_branch();
  return IV(a.lb < b.lb ? a.lb : b.lb, a.ub < b.ub ? b.ub : a.ub);
}

// critical_ratio.cu:32-52
IV critical_ratio(const IV& w, const IV& h) {
  if (w.ub <= h.lb) return critical_ratio(h / w);
  if (h.ub <= w.lb) return critical_ratio(w / h);
  IV hw = h / w;
  hw.tighten_lb(1.0);
  IV wh = w / h;
  wh.tighten_lb(1.0);
  IV c1 = critical_ratio(hw);
  IV c2 = critical_ratio(wh);
  return IV(c1.lb < c2.lb ? c1.lb : c2.lb, c1.ub < c2.ub ? c2.ub : c1.ub);
}

IV required_weight_for(const IV& w, const IV& h) { return critical_ratio(w, h) * w * h; }  // :80-82
IV required_weight_for(const IV& la) { return critical_ratio(la) * la; }                   // :84-86

// can_apply_size_bounded_covering.cu:49-52
double required_weight_for_bounded_size(double ub_w, double ub_h) {
  return mul_ru(kRsbRatio, mul_ru(ub_w, ub_h));
}

// can_recurse.cu:32-62
bool can_recurse(double ub_w, double ub_h, double ub_largest, double lb_rem) {
  double cr = critical_ratio(IV::pt(ub_w), IV::pt(ub_h)).ub;
  double area = mul_ru(ub_w, ub_h);
  double side = ub_w < ub_h ? ub_w : ub_h;
  double scale = rcp_ru(side);
  double scale_sq = mul_ru(scale, scale);
  double norm = mul_ru(scale_sq, ub_largest);
  double cor = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(mul_ru(norm, norm), 1.0)), 1.0)));
  if (cor < kWorstRatio) cor = kWorstRatio;
  if (cr > cor) cr = cor;
  if (norm <= kDiskWeightMax) cr = kRsbRatio;
  double need = mul_ru(cr, area);
  return lb_rem >= need;
}

// r1_strategies.cu:27-87
bool r1_strategies(const IV& la, const IV& r1, double ub_r2) {
  double sq_w = sqrt_rd(mul_rd(2.0, r1.lb));
  double strip_w = add_rd(mul_rd(4.0, r1.lb), -1.0);
  const double rem = (required_weight_for(la) - r1).lb;
  if (strip_w > 0.0) {
    strip_w = sqrt_rd(strip_w);
    IV wr = la - strip_w;
    if (can_recurse(wr.ub, 1.0, ub_r2, rem)) return true;
  }
  if (sq_w < 1.0) {
    double wright = add_ru(la.ub, -sq_w);
    double minright = wright < 1.0 ? wright : 1.0;
    double cost_right;
    if (rsb::disk_satisfies_size_bound(ub_r2, minright)) cost_right = mul_ru(wright, kRsbRatio);
    else cost_right = required_weight_for(IV::pt(wright), IV::pt(1.0)).ub;
    cost_right = add_ru(cost_right, ub_r2);
    double rem_top = add_rd(rem, -cost_right);
    if (rem_top > 0) {
      double htop = add_ru(1.0, -sq_w);
      if (can_recurse(sq_w, htop, ub_r2, rem_top)) return true;
    }
  }
  double s1 = add_rd(mul_rd(4.0, r1.lb), -1.0);
  if (s1 > 0.0) {
    s1 = sqrt_rd(s1);
    double t = add_ru(la.ub, -mul_rd(0.5, s1));
    t = mul_ru(t, t);
    double h1 = add_rd(r1.lb, -t);
    if (h1 > 0.0) {
      h1 = mul_rd(2.0, sqrt_rd(h1));
      double ph = mul_ru(0.5, add_ru(1.0, -h1));
      double pw = add_ru(la.ub, -s1);
      double rem_uneven = mul_rd(0.5, add_rd(rem, -ub_r2));
      if (rem_uneven > 0 && can_recurse(pw, ph, ub_r2, rem_uneven)) return true;
    }
  }
  return false;
}

// r1_r2_strategies.cu:29-44
static bool manually_proved_two_pocket(const IV& la, const IV& r1, const IV& r2) {
  if (r2.lb < 0.2400000000000000188 || r1.lb < 0.4900000000000000466) return false;
  IV lasq = iv_square(la);
  double r1req = (0.25 * (lasq + 1.0) - 0.00999999999999999847).ub;
  if (r1.lb < r1req) return false;
  IV S1 = 2.0 * iv_sqrt(r1 - 0.25);
  IV pw = la - S1;
  pw.tighten_lb(0.0);
  double r2req = (0.25 + iv_square(pw) - 0.00999999999999999847).ub;
  return r2.lb >= r2req;
}

// r1_r2_strategies.cu:46-72
bool r1_r2_strategies_simple(const IV& la, const IV& r1, const IV& r2, double ub_r3) {
  if (manually_proved_two_pocket(la, r1, r2)) return true;
  double width = rsb::two_disks_maximize_height(r1, r2, 1.0);
  if (width <= 0.0) return false;
  double w2 = add_rd(mul_rd(4.0, r2.lb), -1.0);
  if (w2 > 0.0) {
    w2 = sqrt_rd(w2);
    double w1 = sqrt_rd(add_rd(mul_rd(4.0, r1.lb), -1.0));
    double w = add_rd(w1, w2);
    if (w > width) {
      width = w;
      if (w > la.ub) return true;
    }
  }
  double wrem = add_rd(la.ub, width);  // the reference adds here
  double R3 = (required_weight_for(la) - r1 - r2).lb;
  return can_recurse(wrem, 1.0, ub_r3, R3);
}

// r1_r2_strategies.cu:74-130
bool r1_r2_strategies_stacked(const IV& la, const IV& r1, const IV& r2, double ub_r3) {
  double side = sqrt_rd(mul_rd(2.0, r1.lb));
  if (side > 1.0) return false;
  double top_h = add_ru(1.0, -side);
  double w2 = add_rd(mul_rd(4.0, r2.lb), -mul_ru(top_h, top_h));
  if (w2 < 0.0) return false;
  w2 = sqrt_rd(w2);
  double w_top = add_ru(la.ub, -w2);
  double w_bot = add_ru(la.ub, -side);
  double top_min = w_top < top_h ? w_top : top_h;
  double bot_min = w_bot < side ? w_bot : side;
  double R3 = (required_weight_for(la) - r1 - r2).lb;
  if (rsb::disk_satisfies_size_bound(ub_r3, top_min)) {
    double need_top = add_ru(required_weight_for_bounded_size(w_top, top_h), ub_r3);
    double left = add_rd(R3, -need_top);
    if (left > 0.0 && can_recurse(w_bot, side, ub_r3, left)) return true;
  } else if (rsb::disk_satisfies_size_bound(ub_r3, bot_min)) {
    double need_bot = add_ru(required_weight_for_bounded_size(w_bot, side), ub_r3);
    double left = add_rd(R3, -need_bot);
    if (left > 0.0 && can_recurse(w_top, top_h, ub_r3, left)) return true;
  } else {
    double need_top = required_weight_for(IV::pt(w_top), IV::pt(top_h)).ub;
    need_top = add_ru(need_top, ub_r3);
    double left = add_rd(R3, -need_top);
    if (left > 0.0) return can_recurse(w_bot, side, ub_r3, left);
  }
  return false;
}

bool r1_r2_strategies(const IV& la, const IV& r1, const IV& r2, double ub_r3) {  // :132-134
  return r1_r2_strategies_simple(la, r1, r2, ub_r3) || r1_r2_strategies_stacked(la, r1, r2, ub_r3);
}

// r1_r2_r3_strategies.cu:28-46
bool r1_r2_r3_strategy_horizontal(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4) {
  double w3 = add_rd(mul_rd(4.0, r3.lb), -1.0);
  if (w3 <= 0.0) return false;
  w3 = sqrt_rd(w3);
  double w2 = sqrt_rd(add_rd(mul_rd(4.0, r2.lb), -1.0));
  double w1 = sqrt_rd(add_rd(mul_rd(4.0, r1.lb), -1.0));
  double cov = add_rd(w1, add_rd(w2, w3));
  double R4 = (required_weight_for(la) - r1 - r2 - r3).lb;
  double wrem = add_ru(la.ub, -cov);
  if (wrem <= 0.0) return true;
  return can_recurse(wrem, 1.0, ub_r4, R4);
}

// r1_r2_r3_strategies.cu:48-67
static bool can_realize_strip_width(double r1, double r2, double r3, double w) {
  double h1 = add_rd(mul_rd(4.0, r1), -mul_ru(w, w));
  if (h1 <= 0.0) return false;
  h1 = sqrt_rd(h1);
  double hrem = add_ru(1.0, -h1);
  double hsq = mul_ru(hrem, hrem);
  double w2 = add_rd(mul_rd(4.0, r2), -hsq);
  double w3 = add_rd(mul_rd(4.0, r3), -hsq);
  if (w3 <= 0.0 || w2 <= 0.0) return false;
  w2 = sqrt_rd(w2);
  w3 = sqrt_rd(w3);
  return add_rd(w2, w3) >= w;
}

// r1_r2_r3_strategies.cu:69-94
bool r1_r2_r3_strategy_cover_strip(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4) {
  double lo = 0.3, hi = la.ub;
  if (!can_realize_strip_width(r1.lb, r2.lb, r3.lb, lo)) return false;
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) break;
    if (can_realize_strip_width(r1.lb, r2.lb, r3.lb, mid)) lo = mid;
    else hi = mid;
  }
  double wrem = add_ru(la.ub, -lo);
  double R4 = (required_weight_for(la) - r1 - r2 - r3).lb;
  return can_recurse(wrem, 1.0, ub_r4, R4);
}

bool r1_r2_r3_strategies(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4) {  // :96-99
  return r1_r2_r3_strategy_horizontal(la, r1, r2, r3, ub_r4) || r1_r2_r3_strategy_cover_strip(la, r1, r2, r3, ub_r4);
}

// r1_r2_r3_strategies.cu:101-121
static bool three_disk_recursion_big_strip(double la, double r1, double r2, double r3, double ub_r4, double R4) {
  double ssq = add_rd(add_rd(mul_rd(2.0, r2), add_rd(mul_rd(2.0, r3), -0.25)),
                      add_rd(add_rd(-mul_ru(4.0, mul_ru(r2, r2)), mul_rd(8.0, mul_rd(r2, r3))),
                             -mul_ru(4.0, mul_ru(r3, r3))));
  if (ssq <= 0.0) return false;
  double s = sqrt_rd(ssq);
  double la_rem = add_ru(la, -s);
  double h1sq = add_rd(mul_rd(4.0, r1), -mul_ru(la_rem, la_rem));
  if (h1sq <= 0.0) return false;
  if (h1sq >= 1.0) return true;
  double h1_rem = add_ru(1.0, -sqrt_rd(h1sq));
  return can_recurse(la_rem, h1_rem, ub_r4, R4);
}

// r1_r2_r3_strategies.cu:123-165
static bool three_disk_recursion_small_corner(double la, double r1, double r2, double r3, double ub_r4, double R4) {
  double c3 = sqrt_rd(mul_rd(2.0, r3));
  if (c3 >= 1.0) return false;
  double la_rem = add_ru(la, -c3);
  double la_half = mul_ru(la_rem, 0.5);
  double one_rem = add_ru(1.0, -c3);
  if (la_half <= sqrt_rd(r2)) {
    double c1 = sqrt_rd(add_rd(mul_rd(4.0, r1), -mul_ru(one_rem, one_rem)));
    double la_rem1 = add_ru(la, -c1);
    double c2 = sqrt_rd(add_rd(mul_rd(4.0, r2), -mul_ru(la_rem, la_rem)));
    double one_rem2 = add_ru(1.0, -c2);
    return can_recurse(la_rem1, one_rem2, ub_r4, R4);
  }
  if (la_half >= sqrt_rd(r1)) return false;
  if (one_rem >= mul_rd(2.0, sqrt_rd(r2))) return false;
  double c2 = sqrt_rd(add_rd(mul_rd(4.0, r2), -mul_ru(one_rem, one_rem)));
  double la_rem2 = add_ru(la, -c2);
  double c1 = sqrt_rd(add_rd(mul_rd(4.0, r1), -mul_ru(la_rem, la_rem)));
  double one_rem1 = add_ru(1.0, -c1);
  return can_recurse(la_rem2, one_rem1, ub_r4, R4);
}

// r1_r2_r3_strategies.cu:167-206
static bool three_disk_recursion_r1_r3_pocket(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4,
                                              double R4) {
  if (r1.lb <= 0.25) return false;
  IV S1 = 2.0 * iv_sqrt(r1 - 0.25);
  IV a = la - S1;
  IV h2_rest = 0.25 * a * a;
  if (r2.lb <= h2_rest.ub) return false;
  IV h2 = 2.0 * iv_sqrt(r2 - h2_rest);
  IV b = 1.0 - h2;
  IV s3_rest = 0.25 * b * b;
  if (r3.lb <= s3_rest.ub) return false;
  IV S3 = 2.0 * iv_sqrt(r3 - s3_rest);
  IV X = la - S1 - S3;
  if (X.ub <= 0.0) return true;
  IV c = la - 0.5 * S1 - S3;
  IV y_rest = c * c;
  if (r1.lb <= y_rest.ub) return false;
  IV Y = 0.5 - iv_sqrt(r1 - y_rest);
  if (Y.ub <= 0.0) return true;
  return can_recurse(X.ub, Y.ub, ub_r4, R4);
}

// r1_r2_r3_strategies.cu:208-247
static bool manually_proved_three_disk_pocket(const IV& la, const IV& r1, const IV& r2, const IV& r3) {
  if (la.ub > 1.0499999999999998) return false;
  IV lasq = iv_square(la);
  IV rstar = 0.0625 * lasq + 0.15625 + 0.03515625 / lasq;
  IV d1 = rstar - r1;
  if (d1.lb < -0.0009999999999999998 || d1.ub > 0.0009999999999999998) return false;
  IV d2 = rstar - r2;
  if (d2.lb < -0.0019999999999999996 || d2.ub > 0.0019999999999999996) return false;
  IV d3 = rstar - r3;
  if (d3.lb < -0.0029999999999999996 || d3.ub > 0.0029999999999999996) return false;
  IV S1 = 2.0 * iv_sqrt(r1 - 0.25);
  IV a = la - S1;
  IV h2_rest = 0.25 * a * a;
  if (r2.lb <= h2_rest.ub) return false;
  IV h2 = 2.0 * iv_sqrt(r2 - h2_rest);
  IV b = 1.0 - h2;
  IV s3_rest = 0.25 * b * b;
  if (r3.lb <= s3_rest.ub) return false;
  IV S3 = 2.0 * iv_sqrt(r3 - s3_rest);
  IV X = la - S1 - S3;
  return !(X.ub > 0.046999999999999993);
}

// r1_r2_r3_strategies.cu:249-260
bool r1_r2_r3_strategies(const Cell& c, const IV& R4) {
  const double la = c.la.ub, r1 = c.r[0].lb, r2 = c.r[1].lb, r3 = c.r[2].lb;
  const double ub_r4 = c.r[3].ub, lb_R4 = R4.lb;
  return three_disk_recursion_big_strip(la, r1, r2, r3, ub_r4, lb_R4) ||
         three_disk_recursion_small_corner(la, r1, r2, r3, ub_r4, lb_R4) ||
         three_disk_recursion_r1_r3_pocket(c.la, c.r[0], c.r[1], c.r[2], ub_r4, lb_R4) ||
         manually_proved_three_disk_pocket(c.la, c.r[0], c.r[1], c.r[2]);
}

// two_by_two_strategies.cu:29-51
static bool two_by_two_r1_large(const Cell& c, const IV& R) {
  double w1 = sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lb), -1.0));
  double wrem = add_ru(c.la.ub, -w1);
  double wrem_sq = mul_ru(wrem, wrem);
  double h4 = add_rd(mul_rd(4.0, c.r[3].lb), -wrem_sq);
  if (h4 <= 0.0) return false;
  h4 = sqrt_rd(h4);
  double h3 = sqrt_rd(add_rd(mul_rd(4.0, c.r[2].lb), -wrem_sq));
  double h2 = sqrt_rd(add_rd(mul_rd(4.0, c.r[1].lb), -wrem_sq));
  double htot = add_rd(h2, add_rd(h3, h4));
  if (htot >= 1.0) return true;
  double hrem = add_ru(1.0, -htot);
  double ub_r5 = c.r[3].ub < R.ub ? c.r[3].ub : R.ub;
  return can_recurse(wrem, hrem, ub_r5, R.lb);
}

// two_by_two_strategies.cu:53-83
static bool two_by_two_lambda_small(const Cell& c, const IV& R) {
  double side = sqrt_rd(mul_rd(2.0, c.r[0].lb));
  if (side >= 1.0) return two_by_two_r1_large(c, R);
  double wrem = add_ru(c.la.ub, -side);
  double hrem = add_ru(1.0, -side);
  double h2 = add_rd(mul_rd(4.0, c.r[1].lb), -mul_ru(wrem, wrem));
  double w4 = add_rd(mul_rd(4.0, c.r[3].lb), -mul_ru(hrem, hrem));
  if (h2 <= 0.0 || w4 <= 0.0) return false;
  h2 = sqrt_rd(h2);
  w4 = sqrt_rd(w4);
  if (w4 > side) w4 = side;
  if (h2 > side) h2 = side;
  double h = add_ru(1.0, -h2);
  double w = add_ru(c.la.ub, -w4);
  double a = add_ru(mul_ru(h, h), mul_ru(w, w));
  return mul_rd(4.0, c.r[2].lb) >= a;
}

// two_by_two_strategies.cu:85-111
static bool two_by_two_lambda_big(const Cell& c, const IV& R) {
  double w = add_rd(rsb::two_disks_maximize_height(c.r[0], c.r[1], 1.0),
                    rsb::two_disks_maximize_height(c.r[2], c.r[3], 1.0));
  {
    double a = rsb::two_disks_maximize_height(c.r[0], c.r[3], 1.0);
    double b = rsb::two_disks_maximize_height(c.r[1], c.r[2], 1.0);
    double s = add_rd(a, b);
    if (s > w) w = s;
  }
  {
    double a = rsb::two_disks_maximize_height(c.r[0], c.r[2], 1.0);
    double b = rsb::two_disks_maximize_height(c.r[1], c.r[3], 1.0);
    double s = add_rd(a, b);
    if (s > w) w = s;
  }
  if (w <= 0.0) return false;
  double ub_r5 = R.ub < c.r[3].ub ? R.ub : c.r[3].ub;
  return can_recurse(add_ru(c.la.ub, -w), 1.0, ub_r5, R.lb);
}

bool two_by_two_strategies(const Cell& c, const IV& R) {  // :113-115
  return two_by_two_lambda_small(c, R) || two_by_two_lambda_big(c, R);
}

// r1_r2_large_r3_r4_gaps_strategy.cu:29-57
static bool can_realize_gap_width(double gap, double r1, double w1, double r2, double w2, double r4) {
  IV y4 = r4 - 0.25 * IV::pt(gap) * gap;
  if (possibly(y4 <= 0.0)) return false;
  IV x1 = IV::pt(0.5) * w1;
  IV x2 = IV::pt(0.5) * w2 + gap + w1;
  Circle c1{{x1, IV::pt(0.5)}, IV::pt(r1)};
  Circle c2{{x2, IV::pt(0.5)}, IV::pt(r2)};
  Intersection x12 = intersection(c1, c2);
  if (!x12.definitely_intersecting) return false;
  UB lower = x12.p[0].y <= x12.p[1].y;
  if (!is_certain(lower)) return false;
  IV x4 = IV::pt(0.5) * gap + w1;
  y4 = iv_sqrt(y4);
  Point p4{x4, y4};
  const Point& px = x12.p[lower.lo ? 0 : 1];
  return definitely(squared_distance(p4, px) <= r4);
}

// r1_r2_large_r3_r4_gaps_strategy.cu:59-98
bool r1_r2_large_r3_r4_gaps_strategy(const Cell& c, const IV& R) {
  const double r1 = c.r[0].lb, r2 = c.r[1].lb, r4 = c.r[3].lb;
  double w1 = add_rd(mul_rd(4.0, r1), -1.0);
  double w2 = add_rd(mul_rd(4.0, r2), -1.0);
  if (w1 <= 0.0 || w2 <= 0.0) return false;
  w1 = sqrt_rd(w1);
  w2 = sqrt_rd(w2);
  double lo = 0.0, hi = add_rd(sqrt_rd(r1), sqrt_rd(r2));
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) break;
    if (can_realize_gap_width(mid, r1, w1, r2, w2, r4)) lo = mid;
    else hi = mid;
  }
  double wtot = add_rd(w1, add_rd(lo, w2));
  double wrem = add_ru(c.la.ub, -wtot);
  if (wrem <= 0.0) return true;
  double ub_r5 = R.ub;
  if (ub_r5 > c.r[3].ub) ub_r5 = c.r[3].ub;
  return can_recurse(wrem, 1.0, ub_r5, R.lb);
}

// The leaf criteria of one tight_rectangle leaf, each evaluated on its own,
// with the strategy inside each that decides (test support, no reference
// counterpart): bit 0 two_by_two_strategies, 1 r1_r2_r3_strategies(vars, R4),
// 2 r1_r2_large_r3_r4_gaps_strategy (tight_rectangle.cu:247-249); bit 8 the
// two-by-two lambda_small placement, 9 lambda_big; 10 the three-disk big
// strip, 11 small corner, 12 r1-r3 pocket, 13 manual three-disk pocket.
uint32_t leaf_strategy_bits(const Cell& c, const IV& R4, const IV& R) {
  const double la = c.la.ub, r1 = c.r[0].lb, r2 = c.r[1].lb, r3 = c.r[2].lb;
  const double ub_r4 = c.r[3].ub, lb_R4 = R4.lb;
  uint32_t m = 0;
  if (two_by_two_lambda_small(c, R)) m |= 1u << 8;
  if (two_by_two_lambda_big(c, R)) m |= 1u << 9;
  if (three_disk_recursion_big_strip(la, r1, r2, r3, ub_r4, lb_R4)) m |= 1u << 10;
  if (three_disk_recursion_small_corner(la, r1, r2, r3, ub_r4, lb_R4)) m |= 1u << 11;
  if (three_disk_recursion_r1_r3_pocket(c.la, c.r[0], c.r[1], c.r[2], ub_r4, lb_R4)) m |= 1u << 12;
  if (manually_proved_three_disk_pocket(c.la, c.r[0], c.r[1], c.r[2])) m |= 1u << 13;
  if (two_by_two_strategies(c, R)) m |= 1u;
  if (r1_r2_r3_strategies(c, R4)) m |= 2u;
  if (r1_r2_large_r3_r4_gaps_strategy(c, R)) m |= 4u;
  return m;
}

// Valid squared-radius ranges (tight_rectangle.cu:72-102).
IV r1_range(const IV& la) {
  double f = div_rd(critical_ratio(la).lb, kRsbRatio);
  if (f <= 1.0) f = 1.0;
  return IV(mul_rd(kDiskWeightMax, f), mul_ru(0.25, add_ru(1.0, mul_ru(la.ub, la.ub))));
}
IV r2_range(const IV& la, const IV& r1) {
  double a = r1.ub, b = (required_weight_for(la) - r1).ub;
  return IV(0.0, a < b ? a : b);
}
IV r3_range(const IV& la, const IV& r1, const IV& r2) {
  double a = r2.ub, b = (required_weight_for(la) - r1 - r2).ub;
  return IV(0.0, a < b ? a : b);
}
IV r4_range(const IV& la, const IV& r1, const IV& r2, const IV& r3) {
  double a = r3.ub, b = (required_weight_for(la) - r1 - r2 - r3).ub;
  return IV(0.0, a < b ? a : b);
}

}  // namespace tr
}  // namespace ccor
