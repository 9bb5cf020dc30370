// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// rectangle_size_bound C12..C14 and their helpers, restated step for step
// from the reference (cited per function).  Traps: set g_trap, return false
// (or 0 / leave r7 unchanged) at the trap site.
#include "rsb_chain.hpp"

#include <cfloat>
#include <cmath>

namespace ccor {
namespace rsb {

static inline bool trap_false() {
  g_trap = 1;
  return false;
}

// maximize_covered.cu:132-150
Strip2 two_disks_maximal_height_strip(const IV& la_, const IV& r1_, const IV& r2_) {
  const double la = la_.ub, r1 = r1_.lb, r2 = r2_.lb;
  IV rdiff(add_rd(r1, -r2), add_ru(r1, -r2));
  IV wdiff = (2.0 * rdiff) / la;
  IV w1 = IV::pt(0.5) * la + wdiff;
  IV w2 = IV::pt(0.5) * la - wdiff;
  IV h1 = r1 - 0.25 * iv_square(w1);
  IV h2 = r2 - 0.25 * iv_square(w2);
  if (possibly(h1 <= 0) || possibly(h2 <= 0))
    return Strip2{Circle{{IV(0, 0), IV(0, 0)}, IV::pt(r1)}, Circle{{IV(0, 0), IV(0, 0)}, IV::pt(r2)}, 0.0};
  IV h = iv_sqrt(h1);
  return Strip2{Circle{{0.5 * w1, h}, IV::pt(r1)}, Circle{{la - 0.5 * w2, h}, IV::pt(r2)}, mul_rd(2.0, h.lb)};
}

// maximize_covered.cu:239-278
Strip2 two_disks_maximal_height_strip_wc_recursion(const IV& r1, const IV& r2, const IV& R, double ub_w) {
  double lo = 0.0, hi = mul_rd(2.0, r2.lb);
  const double lb_R = R.lb, lr1 = r1.lb, lr2 = r2.lb;
  IV c1x(0, 0), c2x(0, 0);
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) break;
    double hsq = mul_ru(mid, mid);
    double w1 = add_rd(mul_rd(4.0, lr1), -hsq);
    double w2 = add_rd(mul_rd(4.0, lr2), -hsq);
    if (w1 < 0 || w2 < 0) {
      hi = mid;
      continue;
    }
    w1 = sqrt_rd(w1);
    w2 = sqrt_rd(w2);
    double wrem = add_ru(ub_w, -add_rd(w1, w2));
    if (wrem <= 0.0 || can_recurse_worst_case(lb_R, wrem, mid)) {
      lo = mid;
      c1x = IV(mul_rd(0.5, w1), mul_ru(0.5, w1));
      c2x = IV(add_rd(ub_w, -mul_ru(0.5, w2)), add_ru(ub_w, -mul_rd(0.5, w2)));
    } else {
      hi = mid;
    }
  }
  if (lo <= 0.0) return Strip2{Circle{{IV(0, 0), IV(0, 0)}, IV(0, 0)}, Circle{{IV(0, 0), IV(0, 0)}, IV(0, 0)}, 0.0};
  IV y(mul_rd(0.5, lo), mul_ru(0.5, lo));
  return Strip2{Circle{{c1x, y}, IV::pt(lr1)}, Circle{{c2x, y}, IV::pt(lr2)}, lo};
}

// ------------------------------------------------------------------ C12
// five_disk_full_cover.cu:91-177
static bool five_3v2_recursion_center(const Cell& c, const Vals& v) {
  IV lac = IV::pt(c.la.ub);
  IV R6 = c.r[5] + v.R;
  Strip2 bot = two_disks_maximal_height_strip(lac, c.r[0], c.r[1]);
  if (bot.lb_height <= 0.0) return false;
  const double r3 = c.r[2].lb, r4 = c.r[3].lb, r5 = c.r[4].lb;
  double hrem = add_ru(1.0, -bot.lb_height);
  double hrem_sq = mul_ru(hrem, hrem);
  double w4 = add_rd(mul_rd(4.0, r4), -hrem_sq);
  double w5 = add_rd(mul_rd(4.0, r5), -hrem_sq);
  if (w4 <= 0.0 || w5 <= 0.0) return false;
  w4 = sqrt_rd(w4);
  w5 = sqrt_rd(w5);
  IV y45 = 1.0 - IV::pt(0.5) * hrem;
  Circle c4{{lac - IV::pt(0.5) * w4, y45}, IV::pt(r4)};
  Circle c5{{IV::pt(0.5) * w5, y45}, IV::pt(r5)};
  double wrem = add_ru(lac.ub, -add_rd(w4, w5));
  double wrem_sq = mul_ru(wrem, wrem);
  double h3 = add_rd(mul_rd(4.0, r3), -wrem_sq);
  if (h3 <= 0.0) return false;
  h3 = sqrt_rd(h3);
  if (h3 >= hrem) return true;
  Circle c3{{w5 + IV::pt(0.5) * wrem, 1.0 - IV::pt(0.5) * h3}, IV::pt(r3)};
  Intersection x15 = intersection(bot.c1, c5);
  Intersection x12 = intersection(bot.c1, bot.c2);
  Intersection x24 = intersection(bot.c2, c4);
  if (!x12.definitely_intersecting || !x15.definitely_intersecting || !x24.definitely_intersecting)
    return trap_false();
  if (x15.p[0].x.ub >= x15.p[1].x.lb || x24.p[0].x.ub >= x24.p[1].x.lb) return false;
  UB higher = x12.p[0].y > x12.p[1].y;
  if (!is_certain(higher)) return false;
  const Point& top12 = x12.p[higher.lo ? 0 : 1];
  bool in12 = definitely(squared_distance(top12, c3.center) <= c3.squared_radius);
  bool in15 = definitely(squared_distance(x15.p[1], c3.center) <= c3.squared_radius);
  bool in24 = definitely(squared_distance(x24.p[0], c3.center) <= c3.squared_radius);
  if (in12 && in15 && in24) return true;
  if (in15 && in24) {
    Intersection x31 = intersection(bot.c1, c3);
    Intersection x32 = intersection(bot.c2, c3);
    if (x31.definitely_intersecting && x32.definitely_intersecting && definitely(x31.p[0].x < x31.p[1].x) &&
        definitely(x32.p[0].x < x32.p[1].x)) {
      double pw = (x32.p[0].x - x31.p[1].x).ub;
      double y31 = x31.p[1].y.ub, y32 = x32.p[0].y.ub;
      double y = y31 > y32 ? y31 : y32;
      double ph = add_ru(y, -top12.y.lb);
      return can_recurse(R6.lb, c.r[5].ub, pw, ph);
    }
  }
  return false;
}
// five_disk_full_cover.cu:179-203
static bool five_border_cover32(const Cell& c, int t1, int t2, int t3, int a, int b) {
  double htop = three_disks_maximize_height(c.r[t1], c.r[t2], c.r[t3], c.la.ub);
  if (htop <= 0.0) return false;
  double hrem = add_ru(1.0, -htop);
  double wright = two_disks_maximize_height(c.r[a], c.r[b], hrem);
  if (wright <= 0.0) return false;
  double wrem = add_ru(c.la.ub, -wright);
  double R6 = mul_rd(c.la.lb, kRsbRatio);
  for (int i = 0; i < 5; ++i) R6 = add_rd(R6, -c.r[i].ub);
  return can_recurse(R6, c.r[5].ub, wrem, hrem);
}
// five_disk_full_cover.cu:205-210
bool five_disk_full_cover(const Cell& c, const Vals& v) {
  return five_3v2_recursion_center(c, v) || five_border_cover32(c, 0, 1, 2, 3, 4) ||
         five_border_cover32(c, 0, 1, 4, 2, 3);
}

// ------------------------------------------------------------------ C13
// six_disk_full_cover.cu:29-108
static bool six_4_plus_2(const Cell& c) {
  IV la = IV::pt(c.la.ub);
  Strip2 bot = two_disks_maximal_height_strip(la, c.r[0], c.r[1]);
  if (bot.lb_height <= 0.0) return false;
  double hrem = add_ru(1.0, -bot.lb_height);
  double hrem_sq = mul_ru(hrem, hrem);
  IV w4 = IV::pt(c.r[3].lb) - IV::pt(0.25) * hrem_sq;
  IV w3 = IV::pt(c.r[2].lb) - IV::pt(0.25) * hrem_sq;
  if (possibly(w4 <= 0.0) || possibly(w3 <= 0.0)) return false;
  w4 = iv_sqrt(w4);
  w3 = iv_sqrt(w3);
  IV y = 1.0 - IV::pt(0.5) * hrem;
  Circle c4{{w4, y}, IV::pt(c.r[3].lb)};
  Circle c3{{la - w3, y}, IV::pt(c.r[2].lb)};
  Intersection x34 = intersection(c3, c4);
  if (!x34.definitely_intersecting) return false;
  bool p1_up = x34.p[0].y.lb < x34.p[1].y.lb;
  Point up_top = p1_up ? x34.p[1] : x34.p[0];
  Point low_top = p1_up ? x34.p[0] : x34.p[1];
  Intersection x12 = intersection(bot.c1, bot.c2);
  if (!x12.definitely_intersecting) return trap_false();
  Point up_bot = x12.p[0].y.lb < x12.p[1].y.lb ? x12.p[1] : x12.p[0];
  Intersection x14 = intersection(bot.c1, c4);
  Intersection x23 = intersection(bot.c2, c3);
  if (!x14.definitely_intersecting || !x23.definitely_intersecting) return false;
  Point top_c{up_top.x, 0.5 * (1.0 + up_top.y)};
  IV t1 = squared_distance(top_c, Point{2.0 * w4, IV(1.0, 1.0)});
  IV t2 = squared_distance(top_c, Point{la - 2.0 * w3, IV(1.0, 1.0)});
  IV t3 = squared_distance(top_c, up_top);
  Point bc = center(low_top, up_bot);
  IV b1 = squared_distance(bc, x14.p[1]);
  IV b2 = squared_distance(bc, x23.p[0]);
  IV b3 = squared_distance(bc, low_top);
  IV b4 = squared_distance(bc, up_bot);
  const double r5 = c.r[4].lb, r6 = c.r[5].lb;
  if (definitely(t1 <= r6) && definitely(t2 <= r6) && definitely(t3 <= r6))
    return definitely(b1 <= r5) && definitely(b2 <= r5) && definitely(b3 <= r5) && definitely(b4 <= r5);
  if (definitely(t1 <= r5) && definitely(t2 <= r5) && definitely(t3 <= r5))
    return definitely(b1 <= r6) && definitely(b2 <= r6) && definitely(b3 <= r6) && definitely(b4 <= r6);
  return false;
}
// six_disk_full_cover.cu:110-167
static bool six_2v3_r6_center(const Cell& c, const Vals& v) {
  Strip2 bot = two_disks_maximal_height_strip_wc_recursion(c.r[0], c.r[1], v.R, c.la.ub);
  if (bot.lb_height <= 0.0) return false;
  double hrem = add_ru(1.0, -bot.lb_height);
  double hrem_sq = mul_ru(hrem, hrem);
  double w3 = add_rd(mul_rd(4.0, c.r[2].lb), -hrem_sq);
  double w4 = add_rd(mul_rd(4.0, c.r[3].lb), -hrem_sq);
  if (w3 <= 0 || w4 <= 0) return false;
  w3 = sqrt_rd(w3);
  w4 = sqrt_rd(w4);
  double wrem = add_ru(c.la.ub, -add_rd(w3, w4));
  if (wrem <= 0.0) return true;
  double wrem_sq = mul_ru(wrem, wrem);
  const double r5 = c.r[4].lb;
  double dy5 = add_rd(r5, -mul_ru(0.25, wrem_sq));
  if (dy5 <= 0.0) return false;
  dy5 = sqrt_rd(dy5);
  if (add_rd(mul_rd(2.0, dy5), bot.lb_height) >= 1.0) return true;
  Point p5{w4 + IV::pt(0.5) * wrem, 1.0 - IV::pt(dy5)};
  Circle c5{p5, IV::pt(r5)};
  Intersection x15 = intersection(bot.c1, c5), x25 = intersection(bot.c2, c5);
  if (!x15.definitely_intersecting || !x25.definitely_intersecting) return false;
  const Point& p15 = x15.p[1];
  const Point& p25 = x25.p[0];
  if (definitely(p15.y <= bot.lb_height) && definitely(p25.y <= bot.lb_height)) return true;
  double w6 = add_ru(p25.x.ub, -p15.x.lb);
  double ymax = p15.y.ub < p25.y.ub ? p25.y.ub : p15.y.ub;
  double h6 = add_ru(ymax, -bot.lb_height);
  return add_ru(mul_ru(w6, w6), mul_ru(h6, h6)) <= mul_rd(4.0, c.r[5].lb);
}
// six_disk_full_cover.cu:169-193
static bool six_r12_r456_r3(const Cell& c, const Vals& v) {
  double w12 = two_disks_maximize_height(c.r[0], c.r[1], 1.0);
  if (w12 <= 0.0) return false;
  double w456 = three_disks_maximize_height(c.r[3], c.r[4], c.r[5], 1.0);
  if (w456 <= 0.0) return false;
  double wrem = add_ru(c.la.ub, -add_rd(w12, w456));
  if (wrem <= 0.0) return true;
  double h3 = add_rd(mul_rd(4.0, c.r[2].lb), -mul_ru(wrem, wrem));
  if (h3 <= 0.0) return false;
  h3 = sqrt_rd(h3);
  return can_recurse(v.R.lb, c.r[5].ub, wrem, add_ru(1.0, -h3));
}
// six_disk_full_cover.cu:195-199
bool six_disk_full_cover(const Cell& c, const Vals& v) {
  return six_r12_r456_r3(c, v) || six_4_plus_2(c) || six_2v3_r6_center(c, v);
}

// ------------------------------------------------------------------ six-disk width
// six_disks_cover_width.cu:30-93
static bool six_r56_center(double w, double r1, double r2, double r3, double r4, double r5, double r6) {
  double h14 = two_disks_maximize_height(IV::pt(r1), IV::pt(r4), w);
  double h23 = two_disks_maximize_height(IV::pt(r2), IV::pt(r3), w);
  if (h14 <= 0.0 || h23 <= 0.0) return false;
  double hrem = add_ru(1.0, -add_rd(h14, h23));
  double q = mul_ru(0.25, mul_ru(hrem, hrem));
  double x5 = add_rd(r5, -q);
  double x6 = add_rd(r6, -q);
  if (x5 <= 0.0 || x6 <= 0.0) return false;
  x5 = sqrt_rd(x5);
  x6 = sqrt_rd(x6);
  if (mul_rd(2.0, add_rd(x5, x6)) >= w) return true;
  x6 = add_ru(w, -x6);
  IV y56(add_rd(h14, mul_rd(0.5, hrem)), add_ru(h14, mul_ru(0.5, hrem)));
  IV y14(mul_rd(0.5, h14), mul_ru(0.5, h14));
  IV y23(add_rd(1.0, -mul_ru(0.5, h23)), add_ru(1.0, -mul_rd(0.5, h23)));
  IV h14sq(mul_rd(h14, h14), mul_ru(h14, h14));
  IV h23sq(mul_rd(h23, h23), mul_ru(h23, h23));
  h14sq = h14sq * 0.25;
  h23sq = h23sq * 0.25;
  IV w1 = iv_sqrt(r1 - h14sq);
  IV w2 = iv_sqrt(r2 - h23sq);
  IV x1 = w - w1;
  IV x2 = w - w2;
  w1 = w1 * 2.0;
  w2 = w2 * 2.0;
  Point p5{IV::pt(x5), y56};
  Point p6{IV::pt(x6), y56};
  Point ps1{w - w1, IV::pt(h14)};
  Point ps2{w - w2, 1.0 - IV::pt(h23)};
  Circle c1{{x1, y14}, IV::pt(r1)};
  Circle c2{{x2, y23}, IV::pt(r2)};
  Intersection x12 = intersection(c1, c2);
  if (!x12.definitely_intersecting || !definitely(x12.p[0].x < x12.p[1].x)) return false;
  return squared_distance(p5, ps1).ub <= r5 && squared_distance(p5, ps2).ub <= r5 &&
         squared_distance(p5, x12.p[0]).ub <= r5 && squared_distance(p6, x12.p[1]).ub <= r6;
}
static bool six_r56_center(const Cell& c, double w) {
  return six_r56_center(w, c.r[0].lb, c.r[1].lb, c.r[2].lb, c.r[3].lb, c.r[4].lb, c.r[5].lb);
}
// Pairings of the six disks (six_disks_cover_width.cu:111-140).
static const int kPairs23[10][6] = {{0, 1, 2, 3, 4, 5}, {0, 1, 3, 2, 4, 5}, {0, 1, 4, 2, 3, 5}, {0, 1, 5, 2, 3, 4},
                                    {0, 2, 3, 1, 4, 5}, {0, 2, 4, 1, 3, 5}, {0, 2, 5, 1, 3, 4}, {0, 3, 4, 1, 2, 5},
                                    {0, 3, 5, 1, 2, 4}, {0, 4, 5, 1, 2, 3}};
static const int kPairs32[15][6] = {{0, 1, 2, 3, 4, 5}, {0, 1, 2, 4, 3, 5}, {0, 1, 2, 5, 3, 4}, {0, 2, 1, 3, 4, 5},
                                    {0, 2, 1, 4, 3, 5}, {0, 2, 1, 5, 3, 4}, {0, 3, 1, 2, 4, 5}, {0, 3, 1, 4, 2, 5},
                                    {0, 3, 1, 5, 2, 4}, {0, 4, 1, 2, 3, 5}, {0, 4, 1, 3, 2, 5}, {0, 4, 1, 5, 2, 3},
                                    {0, 5, 1, 2, 3, 4}, {0, 5, 1, 3, 2, 4}, {0, 5, 1, 4, 2, 3}};
// six_disks_cover_width.cu:156-233
double six_disks_maximize_covered_width(const Cell& c) {
  double w = 0.0;
  for (const auto& p : kPairs32) {
    double a = two_disks_maximize_height(c.r[p[0]], c.r[p[1]], 1.0);
    double b = two_disks_maximize_height(c.r[p[2]], c.r[p[3]], 1.0);
    double d = two_disks_maximize_height(c.r[p[4]], c.r[p[5]], 1.0);
    double t = add_rd(a, add_rd(b, d));
    if (t > w) w = t;
  }
  double w2 = 0.0;
  for (const auto& p : kPairs23) {
    double a = three_disks_maximize_height(c.r[p[0]], c.r[p[1]], c.r[p[2]], 1.0);
    double b = three_disks_maximize_height(c.r[p[3]], c.r[p[4]], c.r[p[5]], 1.0);
    double t = add_rd(a, b);
    if (t > w2) w2 = t;
  }
  if (w2 > w) w = w2;
  if (six_r56_center(c, w)) {
    double lo = w, hi = std::nextafter(c.la.ub, DBL_MAX);
    for (int it = 0; it < kBisectCap; ++it) {
      double mid = rn_mid(lo, hi);
      if (mid <= lo || mid >= hi) {
        w = lo;
        break;
      }
      if (six_r56_center(c, mid)) lo = mid;
      else hi = mid;
    }
  }
  return w;
}

// ------------------------------------------------------------------ C14 helpers
// seven_disk_strategies.cu:38-47
static bool r7_in_corner_and_recurse(const IV& R, const IV& r7, double ub_w) {
  double R_no7 = add_rd(R.lb, -r7.ub);
  double h7 = add_rd(mul_rd(4.0, r7.lb), -mul_ru(ub_w, ub_w));
  if (h7 < 0) return false;
  h7 = sqrt_rd(h7);
  return can_recurse(R_no7, r7.ub, ub_w, add_ru(1.0, -h7));
}
// seven_disk_strategies.cu:49-115
static bool three_disk_top_strip_can_cover(double la, double t1, double t2, double t3, const double* rem4,
                                           double ub_max_disk, double R_no7) {
  double htop = three_disks_maximize_height(IV::pt(t1), IV::pt(t2), IV::pt(t3), la);
  if (htop <= 0) return false;
  double hrem = add_rd(1.0, -htop);  // rounded down in the reference
  if (hrem <= 0) return true;
  double hrem_sq = mul_ru(hrem, hrem);
  double wrem;
  double area_sb = div_rd(R_no7, kRsbRatio);
  double width_sb = div_rd(area_sb, hrem);
  double min_sb = width_sb < hrem ? width_sb : hrem;
  if (disk_satisfies_size_bound(ub_max_disk, min_sb)) {
    wrem = add_ru(la, -width_sb);
  } else {
    double area_wc = div_rd(R_no7, kWorstRatio);
    double width_wc = div_rd(area_wc, hrem);
    double long_wc = width_wc < hrem ? hrem : width_wc;
    double short_wc = width_wc < hrem ? width_wc : hrem;
    double skew = div_ru(long_wc, short_wc);
    if (skew > kRsbToleratedLambda) {
      if (width_wc > hrem) {
        width_wc = mul_rd(hrem, kRsbToleratedLambda);
      } else {
        width_wc = add_rd(mul_rd(2.0, R_no7), -mul_ru(0.5, hrem_sq));
        width_wc = width_wc <= 0 ? 0.0 : sqrt_rd(width_wc);
      }
    }
    wrem = add_ru(la, -width_wc);
  }
  double hbot = two_disks_maximize_height(IV::pt(rem4[1]), IV::pt(rem4[2]), wrem);
  if (hbot <= 0) return false;
  if (hbot >= hrem) return true;
  double hrem2 = add_ru(hrem, -hbot);
  double hl2 = two_disks_maximize_height(IV::pt(rem4[0]), IV::pt(rem4[3]), wrem);
  if (hl2 <= 0.0) return false;
  return hl2 >= hrem2;
}
// seven_disk_strategies.cu:117-129
static bool split_largest_r7_works(double ub_r7, double wrem, double wf, double hf, double ws, double hs) {
  double cr = bound_critical_ratio(wrem, ub_r7, wf, hf);
  if (cr >= kInf) return false;
  double need = mul_ru(cr, mul_ru(wf, hf));
  double need_split = add_ru(need, ub_r7);
  double rem2 = add_rd(wrem, -need_split);
  return can_recurse(rem2, ub_r7, ws, hs);
}
// seven_disk_strategies.cu:131-154
static void split_largest_r7(IV& r7, double wrem, double wf, double hf, double ws, double hs) {
  if (split_largest_r7_works(r7.ub, wrem, wf, hf, ws, hs)) {
    r7.lb = kInf;
    return;
  }
  double lo = r7.lb, hi = r7.ub;
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) {
      r7.tighten_lb(lo);
      break;
    }
    if (split_largest_r7_works(mid, wrem, wf, hf, ws, hs)) lo = mid;
    else hi = mid;
  }
}
// seven_disk_strategies.cu:156-222
static void split_recursion(const Cell& c, IV& r7, double wrem, int largest, double wf, double hf, double ws,
                            double hs) {
  if (largest > 5) {
    split_largest_r7(r7, wrem, wf, hf, ws, hs);
    return;
  }
  double ub_largest = c.r[largest].ub;
  double cr = bound_critical_ratio(wrem, ub_largest, wf, hf);
  if (!(cr < kInf)) return;
  double need = mul_ru(cr, mul_ru(wf, hf));
  double in_disks = 0.0;
  for (int i = largest; i <= 5; ++i) in_disks = add_ru(in_disks, c.r[i].ub);
  if (in_disks > need) {
    double lb_in_disks = 0.0;
    for (int i = largest; i <= 5; ++i) lb_in_disks = add_rd(lb_in_disks, c.r[i].lb);
    if (lb_in_disks >= need) r7.tighten_lb(recursion_bound_size(add_rd(wrem, -in_disks), ws, hs));
    return;
  }
  double need_split = add_ru(need, r7.ub);
  double rem2 = add_rd(wrem, -need_split);
  r7.tighten_lb(recursion_bound_size(rem2, ws, hs));
  if (!r7.empty()) {
    double hi = r7.ub, lo = r7.lb;
    for (int it = 0; it < kBisectCap; ++it) {
      double mid = rn_mid(lo, hi);
      if (mid <= lo || mid >= hi) {
        r7.lb = lo;
        return;
      }
      double ns = add_ru(need, mid);
      double r2 = add_rd(wrem, -ns);
      if (recursion_bound_size(r2, ws, hs) < mid) hi = mid;
      else lo = mid;
    }
  }
}
namespace {
struct R6Pos {
  Point p6;
  double left;
};
}  // namespace
// seven_disk_strategies.cu:233-262
static R6Pos compute_r6(const IV& r6, const Point& x1, const Point& x2, const IV& lower, const IV& upper) {
  R6Pos pos;
  pos.left = 0.0;
  pos.p6 = Point{IV(0, 0), IV(0, 0)};
  double left = x1.x.lb < x2.x.lb ? x1.x.lb : x2.x.lb;
  double top = x1.y.ub < x2.y.ub ? x2.y.ub : x1.y.ub;
  double bot = x1.y.lb < x2.y.lb ? x1.y.lb : x2.y.lb;
  double h = add_ru(top, -bot);
  double dx6 = add_rd(r6.lb, -mul_ru(0.25, mul_ru(h, h)));
  if (dx6 <= 0.0) return pos;
  dx6 = sqrt_rd(dx6);
  pos.p6 = Point{left + IV::pt(dx6), IV(mul_rd(0.5, add_rd(bot, top)), mul_ru(0.5, add_ru(bot, top)))};
  double dyb = (pos.p6.y - lower).ub;
  double dyt = (upper - pos.p6.y).ub;
  dyb = mul_ru(dyb, dyb);
  dyt = mul_ru(dyt, dyt);
  double m = dyb < dyt ? dyt : dyb;
  double sw = add_rd(r6.lb, -m);
  if (sw <= 0.0) return pos;
  sw = sqrt_rd(sw);
  pos.left = add_rd(pos.p6.x.lb, sw);
  return pos;
}
// seven_disk_strategies.cu:264-316
static void r7_t_shaped_r5_large(const Cell& c, IV& r7, const Circle& c1, const Circle& c3, const Circle& c5,
                                 double R6, double w12, double wrem, double hrem, double strip_w5, const IV& lower,
                                 const IV& upper) {
  Intersection x15 = intersection(c1, c5), x35 = intersection(c3, c5);
  if (!x15.definitely_intersecting || !x35.definitely_intersecting) {
    g_trap = 1;
    return;
  }
  double left = x15.p[1].x.lb;
  if (left > x35.p[1].x.lb) left = x35.p[1].x.lb;
  double hz = add_ru(w12, -left);
  if (hz <= 0.0) {
    if (can_recurse(R6, c.r[5].ub, wrem, 1.0)) r7.tighten_lb(kInf);
    return;
  }
  split_recursion(c, r7, R6, 5, hz, hrem, wrem, 1.0);
  split_recursion(c, r7, R6, 5, wrem, 1.0, hz, hrem);
  if (r7.empty()) return;
  double R7 = add_rd(R6, -c.r[5].ub);
  if (R7 <= 0.0) return;
  Point x1 = x15.p[1];
  Point x2 = x35.p[1];
  if (possibly(x1.y >= upper) || possibly(x1.y <= lower) || possibly(x2.y >= upper) || possibly(x2.y <= lower)) {
    x1 = Point{IV::pt(strip_w5), upper};
    x2 = Point{IV::pt(strip_w5), lower};
  }
  R6Pos pos = compute_r6(c.r[5], x1, x2, lower, upper);
  if (pos.left <= 0.0) return;
  if (pos.left >= w12) {
    r7.tighten_lb(recursion_bound_size(R7, wrem, 1.0));
  } else {
    hz = add_ru(w12, -left);  // the reference reuses the r5 intersection here
    split_recursion(c, r7, R7, 6, hz, hrem, wrem, 1.0);
    split_recursion(c, r7, R7, 6, wrem, 1.0, hz, hrem);
  }
}
// seven_disk_strategies.cu:318-518
static void r7_t_shaped(const Cell& c, const Vals& v, IV& r7) {
  (void)v;
  double w2 = sqrt_rd(mul_rd(2.0, c.r[1].lb));
  double h12 = w2;
  double h12sq = mul_ru(h12, h12);
  double w1 = sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lb), -h12sq));
  double w12 = add_rd(w1, w2);
  double h34 = two_disks_maximize_height(c.r[2], c.r[3], w12);
  if (h34 <= 0.0) return;
  double wtot = mul_rd(c.la.lb, kRsbRatio);
  double R5 = add_rd(wtot, -add_ru(c.r[0].ub, add_ru(c.r[1].ub, add_ru(c.r[2].ub, c.r[3].ub))));
  double hrem = add_ru(1.0, -add_rd(h12, h34));
  double wrem = add_ru(c.la.ub, -w12);
  if (hrem <= 0.0) {
    if (wrem <= 0.0 || can_recurse(R5, c.r[4].ub, wrem, 1.0)) r7.tighten_lb(kInf);
    return;
  }
  if (wrem <= 0.0) {
    if (can_recurse(R5, c.r[4].ub, c.la.ub, hrem)) r7.tighten_lb(kInf);
    return;
  }
  split_recursion(c, r7, R5, 4, w12, hrem, wrem, 1.0);
  split_recursion(c, r7, R5, 4, wrem, 1.0, w12, hrem);
  if (r7.empty()) return;
  double R6 = add_rd(R5, -c.r[4].ub);
  if (R6 <= 0.0) return;
  double hrem_sq = mul_ru(hrem, hrem);
  double x5 = add_rd(c.r[4].lb, -mul_ru(0.25, hrem_sq));
  if (x5 <= 0.0) return;
  x5 = sqrt_rd(x5);
  double strip_w5 = mul_rd(2.0, x5);
  IV y5(add_rd(h12, mul_rd(0.5, hrem)), add_ru(h12, mul_ru(0.5, hrem)));
  Point p5{IV::pt(x5), y5};
  IV p2x(mul_rd(0.5, w2), mul_ru(0.5, w2));
  IV p2y(mul_rd(0.5, h12), mul_ru(0.5, h12));
  Point p2{p2x, p2y};
  Point p1{w2 + IV(0.5, 0.5) * w1, p2y};
  IV x4 = iv_sqrt(c.r[3] - 0.25 * iv_square(IV::pt(h34)));
  IV y4 = 1.0 - 0.5 * IV::pt(h34);
  Point p4{x4, y4};
  IV w3 = w12 - 2.0 * p4.x;
  Point p3{w12 - 0.5 * w3, y4};
  Circle c1{p1, IV::pt(c.r[0].lb)};
  Circle c2{p2, IV::pt(c.r[1].lb)};
  Circle c3{p3, IV::pt(c.r[2].lb)};
  Circle c4{p4, IV::pt(c.r[3].lb)};
  Circle c5{p5, IV::pt(c.r[4].lb)};
  Point p43{2.0 * p4.x, IV(1.0, 1.0) - h34};
  Point p21{IV::pt(w2), IV::pt(h12)};
  IV lower = IV::pt(h12);
  IV upper = IV(1.0, 1.0) - h34;
  if (definitely(squared_distance(p5, p43) <= c.r[4]) && definitely(squared_distance(p5, p21) <= c.r[4])) {
    r7_t_shaped_r5_large(c, r7, c1, c3, c5, R6, w12, wrem, hrem, strip_w5, lower, upper);
    return;
  }
  Intersection x24 = intersection(c2, c4);
  (void)x24;
  Intersection x25 = intersection(c2, c5);
  Intersection x45 = intersection(c4, c5);
  if (!x25.definitely_intersecting || !x45.definitely_intersecting) {
    g_trap = 1;
    return;
  }
  if (x25.p[0].x.ub >= x25.p[1].x.lb || x45.p[0].x.ub >= x45.p[1].x.lb) return;
  Point x1 = x25.p[1];
  Point x2 = x45.p[1];
  if (possibly(x1.y >= upper) || possibly(x1.y <= lower) || possibly(x2.y >= upper) || possibly(x2.y <= lower)) {
    x1 = Point{IV::pt(strip_w5), upper};
    x2 = Point{IV::pt(strip_w5), lower};
  }
  double left = x1.x.lb;
  if (left > x2.x.lb) left = x2.x.lb;
  double hz = add_ru(w12, -left);
  split_recursion(c, r7, R6, 5, hz, hrem, wrem, 1.0);
  split_recursion(c, r7, R6, 5, wrem, 1.0, hz, hrem);
  if (r7.empty()) return;
  double R7 = add_rd(R6, -c.r[5].ub);
  if (R7 <= 0.0) return;
  R6Pos pos = compute_r6(c.r[5], x1, x2, lower, upper);
  if (pos.left <= 0.0) return;
  Circle c6{pos.p6, IV::pt(c.r[5].lb)};
  Point d71, d72;
  Intersection x16 = intersection(c1, c6), x36 = intersection(c3, c6);
  if (!x16.definitely_intersecting || !x36.definitely_intersecting) {
    g_trap = 1;
    return;
  }
  bool r6_right_x = false;
  if (definitely(squared_distance(p43, pos.p6) <= c.r[5]) && definitely(squared_distance(p21, pos.p6) <= c.r[5])) {
    left = x16.p[1].x.lb < x36.p[1].x.lb ? x16.p[1].x.lb : x36.p[1].x.lb;
    d71 = x36.p[1];
    d72 = x16.p[1];
    r6_right_x = true;
  } else {
    left = pos.left;
    // the reference builds these points with the strip borders as x
    d71 = Point{upper, IV::pt(left)};
    d72 = Point{lower, IV::pt(left)};
  }
  hz = add_ru(w12, -left);
  if (hz <= 0.0) {
    r7.tighten_lb(recursion_bound_size(R7, wrem, 1.0));
  } else {
    split_recursion(c, r7, R7, 6, hz, hrem, wrem, 1.0);
    split_recursion(c, r7, R7, 6, wrem, 1.0, hz, hrem);
  }
  if (r7.empty()) return;
  double R8 = add_rd(R7, -r7.ub);
  if (R8 <= 0.0) return;
  Intersection x13 = intersection(c1, c3);
  if (r6_right_x && x13.definitely_intersecting && definitely(squared_distance(x13.p[0], pos.p6) <= c.r[5]) &&
      definitely(x13.p[1].y >= lower) && definitely(x13.p[1].y <= upper)) {
    d71 = x13.p[1];
    d72 = x13.p[1];
  }
  R6Pos p7 = compute_r6(r7, d71, d72, lower, upper);
  if (p7.left <= 0.0) return;
  if (p7.left >= c.la.ub) {
    split_largest_r7(r7, R8, wrem, h12, wrem, h34);
    split_largest_r7(r7, R8, wrem, h34, wrem, h12);
  }
  if (p7.left >= w12) {
    r7.tighten_lb(recursion_bound_size(R8, wrem, 1.0));
  } else {
    hz = add_ru(w12, -p7.left);
    split_largest_r7(r7, R8, hz, hrem, wrem, 1.0);
    split_largest_r7(r7, R8, wrem, 1.0, hz, hrem);
  }
}
// seven_disk_strategies.cu:520-554
static bool three_disk_top_strip(const Cell& c, const IV& R, const IV& r7) {
  const double R_no7 = add_rd(R.lb, -r7.ub);
  const double d[7] = {c.r[0].lb, c.r[1].lb, c.r[2].lb, c.r[3].lb, c.r[4].lb, c.r[5].lb, r7.lb};
  for (int i = 0; i < 5; ++i)
    for (int j = i + 1; j < 6; ++j)
      for (int k = j + 1; k < 7; ++k) {
        double rem[4];
        int n = 0;
        for (int g = 0; g < 7; ++g)
          if (g != i && g != j && g != k) rem[n++] = d[g];
        if (three_disk_top_strip_can_cover(c.la.ub, d[i], d[j], d[k], rem, r7.ub, R_no7)) return true;
      }
  return false;
}
// seven_disk_strategies.cu:556-718 (rem_width adds: reference quirk)
static bool r435_strip(const Cell& c, const IV& r7, const Vals& v) {
  double cov = three_disks_maximize_height(c.r[2], c.r[3], c.r[4], 1.0);
  if (cov <= 0.0) return false;
  IV wsq435(mul_rd(cov, cov), mul_ru(cov, cov));
  IV r3 = IV::pt(c.r[2].lb), r4 = IV::pt(c.r[3].lb);
  IV h3 = 4.0 * r3 - wsq435;
  IV h4 = 4.0 * r4 - wsq435;
  if (possibly(h3 <= 0.0) || possibly(h4 <= 0.0)) return false;
  h3 = iv_sqrt(h3);
  h4 = iv_sqrt(h4);
  Point pos3{IV(mul_rd(0.5, cov), mul_ru(0.5, cov)), 0.5 * h3 + h4};
  IV r1 = IV::pt(c.r[0].lb), r2 = IV::pt(c.r[1].lb);
  IV r1s = iv_sqrt(r1), r2s = iv_sqrt(r2);
  if (possibly(r1s + r2s <= 0.5)) return false;
  double lo = 0.0;
  double hi = mul_rd(2.0, r2s.lb);
  double wrem = add_ru(c.la.ub, -cov);
  IV wrem_sq(mul_rd(wrem, wrem), mul_ru(wrem, wrem));
  if (wrem < hi) {
    IV y1 = r1 - 0.25 * wrem_sq;
    IV y2 = r2 - 0.25 * wrem_sq;
    if (definitely(y1 > 0.0) && definitely(y2 > 0.0)) {
      y1 = 1.0 - iv_sqrt(y1);
      y2 = iv_sqrt(y2);
      Point pos1{0.5 * IV::pt(wrem) + cov, y1};
      Point pos2{pos1.x, y2};
      Intersection x12 = intersection(Circle{pos1, r1}, Circle{pos2, r2});
      if (x12.definitely_intersecting && x12.p[0].x.ub < x12.p[1].x.lb &&
          definitely(squared_distance(pos3, x12.p[0]) <= r3)) {
        double h1 = mul_rd(2.0, sqrt_rd(add_rd(r1.lb, -mul_ru(0.25, wrem_sq.ub))));
        double h2 = mul_rd(2.0, sqrt_rd(add_rd(r2.lb, -mul_ru(0.25, wrem_sq.ub))));
        double rem_h = add_ru(1.0, -add_rd(h1, h2));
        double rem_w = add_ru(c.la.ub, x12.p[1].x.lb);
        if (can_recurse(add_rd(v.R.lb, c.r[5].lb), c.r[5].ub, rem_w, rem_h)) return true;
      }
    }
    hi = wrem;
  }
  double w12 = 0.0;
  Intersection best;
  best.definitely_intersecting = false;
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) {
      w12 = lo;
      break;
    }
    IV wsq(mul_rd(mid, mid), mul_ru(mid, mid));
    IV x = 0.5 * IV::pt(mid) + cov;
    IV hh2 = r2 - 0.25 * wsq;
    if (hh2.lb <= 0) {
      hi = mid;
      continue;
    }
    hh2 = iv_sqrt(hh2);
    IV hh1 = iv_sqrt(r1 - 0.25 * wsq);
    Intersection x12 = intersection(Circle{Point{x, 1.0 - hh1}, r1}, Circle{Point{x, hh2}, r2});
    if (x12.definitely_intersecting && x12.p[0].x.ub < x12.p[1].x.lb &&
        definitely(squared_distance(pos3, x12.p[0]) <= r3)) {
      lo = mid;
      best = x12;
    } else {
      hi = mid;
    }
  }
  if (!best.definitely_intersecting) return false;
  double w67 = add_ru(c.la.ub, -add_rd(cov, w12));
  double w67sq = mul_ru(w67, w67);
  {
    double h7 = add_rd(mul_rd(4.0, r7.lb), -w67sq);
    double h6 = add_rd(mul_rd(4.0, c.r[5].lb), -w67sq);
    if (h7 > 0 && h6 > 0) {
      h7 = sqrt_rd(h7);
      h6 = sqrt_rd(h6);
      Point pos6{c.la - IV(mul_rd(0.5, w67), mul_ru(0.5, w67)), 1.0 - IV::pt(h7) - 0.5 * IV::pt(h6)};
      double hr = add_ru(1.0, -add_rd(h7, h6));
      if (definitely(squared_distance(best.p[1], pos6) <= c.r[5].lb)) {
        if (can_recurse(add_rd(v.R.lb, -r7.ub), r7.ub, w67, hr)) return true;
      }
    }
  }
  {
    double wide = add_ru(c.la.ub, -best.p[1].x.lb);
    double widesq = mul_ru(wide, wide);
    double h7 = add_rd(mul_rd(4.0, r7.lb), -widesq);
    double h6 = add_rd(mul_rd(4.0, c.r[5].lb), -widesq);
    if (h7 > 0 && h6 > 0) {
      h7 = sqrt_rd(h7);
      h6 = sqrt_rd(h6);
      double hr = add_ru(1.0, -add_rd(h6, h7));
      double wr = hr <= best.p[1].y.lb ? w67 : wide;
      if (can_recurse(add_rd(v.R.lb, -r7.ub), r7.ub, wr, hr)) return true;
    }
  }
  return false;
}
// seven_disk_strategies.cu:961-1061 (`return false` from a double: 0.0)
static double five_disks_maximize_height_23(double ub_w, double t1, double t2, double m1, double m2, double m3) {
  Strip2 bot = two_disks_maximal_height_strip(IV::pt(ub_w), IV::pt(t1), IV::pt(t2));
  if (bot.lb_height <= 0.0) return 0.0;
  Intersection xt = intersection(bot.c1, bot.c2);
  if (!xt.definitely_intersecting) {
    g_trap = 1;
    return 0.0;
  }
  UB low = xt.p[0].y <= xt.p[1].y;
  if (!is_certain(low)) return 0.0;
  Point xtup = xt.p[low.lo ? 1 : 0];
  double lo = 0.0, hi = mul_rd(2.0, sqrt_rd(m3));
  for (int it = 0; it < kBisectCap; ++it) {
    double h = rn_mid(lo, hi);
    if (h <= lo || h >= hi) return lo == 0.0 ? 0.0 : add_rd(bot.lb_height, lo);
    double hsq = mul_ru(h, h);
    double w1 = add_rd(mul_rd(4.0, m1), -hsq);
    double w2 = add_rd(mul_rd(4.0, m2), -hsq);
    double w3 = add_rd(mul_rd(4.0, m3), -hsq);
    if (w1 <= 0.0 || w2 <= 0.0 || w3 <= 0.0) {
      hi = h;
      continue;
    }
    w1 = sqrt_rd(w1);
    w2 = sqrt_rd(w2);
    w3 = sqrt_rd(w3);
    if (add_rd(w1, add_rd(w2, w3)) >= ub_w) {
      lo = h;
      continue;
    }
    double wr = add_ru(ub_w, -add_rd(w2, w3));
    double q = mul_ru(0.25, mul_ru(wr, wr));
    double h1 = add_rd(m1, -q);
    if (h1 <= 0.0) {
      hi = h;
      continue;
    }
    h1 = sqrt_rd(h1);
    IV y23(add_rd(bot.lb_height, mul_rd(0.5, h)), add_ru(bot.lb_height, mul_ru(0.5, h)));
    Point pc2{IV::pt(w2) * 0.5, y23}, pc3{ub_w - IV::pt(w3) * 0.5, y23};
    IV x1 = 0.5 * (IV::pt(w2) + ub_w - IV::pt(w3));
    IV y1(add_rd(bot.lb_height, add_rd(h, -h1)), add_ru(bot.lb_height, add_ru(h, -h1)));
    Point pc1{x1, y1};
    if (squared_distance(pc1, xtup).ub > m1) {
      hi = h;
      continue;
    }
    Intersection x12 = intersection(Circle{pc1, IV::pt(m1)}, Circle{pc2, IV::pt(m2)});
    Intersection x13 = intersection(Circle{pc1, IV::pt(m1)}, Circle{pc3, IV::pt(m3)});
    if (!x12.definitely_intersecting || !x13.definitely_intersecting) {
      hi = h;
      continue;
    }
    UB l12 = x12.p[0].y <= x12.p[1].y;
    UB l13 = x13.p[0].y <= x13.p[1].y;
    if (!is_certain(l12) || !is_certain(l13)) {
      hi = h;
      continue;
    }
    const Point& a = x12.p[l12.lo ? 0 : 1];
    const Point& b = x13.p[l13.lo ? 0 : 1];
    if (squared_distance(a, bot.c1.center).ub <= t1 && squared_distance(b, bot.c2.center).ub <= t2) lo = h;
    else hi = h;
  }
  return lo == 0.0 ? 0.0 : add_rd(bot.lb_height, lo);
}
// seven_disk_strategies.cu:1197-1239
static IV compute_r7_range(const Cell& c, const Vals& v, double ub_w_remaining) {
  IV range(0.0, c.r[5].ub);
  range.tighten_ub(v.R.ub);
  double eff = div_rd(v.R.lb, ub_w_remaining);
  if (eff >= kRsbRatio) {
    double strip = div_rd(v.R.lb, kRsbRatio);
    if (strip >= 1.0) return IV(1.0, 0.0);
    range.tighten_lb(bound_allowed_weight(strip));
  }
  if (eff >= kWorstRatio && !range.empty()) {
    double lo_skew = rcp_rd(ub_w_remaining);
    double hi_skew = rcp_ru(ub_w_remaining);
    double crs = tr::critical_ratio(IV(lo_skew, hi_skew)).ub;
    if (crs <= eff) return IV(1.0, 0.0);
    double s = sqrt_rd(add_rd(mul_rd(4.0, mul_rd(eff, eff)), -2.0));
    double lam = add_rd(mul_rd(2.0, eff), s);
    double sigma = mul_rd(eff, add_rd(lam, -div_ru(2.0, lam)));
    range.tighten_lb(mul_rd(sigma, mul_rd(ub_w_remaining, ub_w_remaining)));
  }
  return range;
}
// seven_disk_strategies.cu:1241-1297 (h1 used without a sign check: quirk)
static bool six_disks_222(const Cell& c, const Vals& v, const IV& r7) {
  double c12 = two_disks_maximize_height(c.r[0], c.r[1], 1.0);
  double c34 = two_disks_maximize_height(c.r[2], c.r[3], 1.0);
  if (c34 <= 0.0) return false;
  double r5 = c.r[4].lb, r6 = c.r[5].lb;
  double wrem = add_ru(c.la.ub, -add_rd(c12, c34));
  double wrem_sq = mul_ru(wrem, wrem);
  double h1 = add_rd(mul_rd(4.0, r5), -wrem_sq);
  double h2 = add_rd(mul_rd(4.0, r6), -wrem_sq);
  if (h2 <= 0.0) return false;
  h1 = sqrt_rd(h1);
  h2 = sqrt_rd(h2);
  double hrem = add_ru(1.0, -add_rd(h1, h2));
  if (hrem <= 0) return true;
  double dy = add_ru(1.0, -add_rd(h1, mul_rd(0.5, h2)));
  dy = mul_ru(dy, dy);
  double wtop = add_rd(r6, -dy);
  if (wtop <= 0) return false;
  wtop = mul_rd(2.0, sqrt_rd(wtop));
  double wrect = mul_ru(0.5, add_ru(wrem, -wtop));
  double need7 = mul_ru(0.25, add_ru(mul_ru(wrect, wrect), mul_ru(hrem, hrem)));
  if (need7 <= r7.lb) return can_recurse(add_rd(v.R.lb, -r7.ub), r7.ub, wrect, hrem);
  double smaller = mul_rd(0.5, v.R.lb);
  smaller = add_rd(smaller, -mul_ru(0.5, r7.ub));
  return can_recurse(smaller, r7.ub, wrect, hrem);
}
// seven_disk_strategies.cu:1299-1349 (h1 uses w12 unsquared: quirk)
static bool different_height_r12_2x2(const Cell& c, const Vals& v, const IV& r7) {
  double w12 = sqrt_rd(mul_rd(2.0, c.r[1].lb));
  double h2 = w12, h1 = sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lb), -w12));
  double h12 = add_rd(h1, h2);
  double wrem = add_ru(c.la.ub, -w12);
  double h34 = two_disks_maximize_height(c.r[2], c.r[3], wrem);
  double h56 = two_disks_maximize_height(c.r[4], c.r[5], wrem);
  double h3456 = add_rd(h34, h56);
  double hr12 = add_ru(1.0, -h12);
  double hr3456 = add_ru(1.0, -h3456);
  double hhigh = hr12 < hr3456 ? hr3456 : hr12;
  double hlow = hr12 < hr3456 ? hr12 : hr3456;
  double whigh = hr12 < hr3456 ? wrem : w12;
  double wlow = hr12 < hr3456 ? w12 : wrem;
  double wr7 = add_rd(mul_rd(4.0, r7.lb), -mul_ru(hhigh, hhigh));
  if (wr7 <= 0.0) return false;
  wr7 = sqrt_rd(wr7);
  whigh = add_rd(whigh, -wr7);
  double R8 = add_rd(v.R.lb, -r7.ub);
  if (R8 <= 0.0) return false;
  if (whigh <= 0.0) return can_recurse(R8, r7.ub, add_ru(whigh, wlow), hlow);
  double shortest = hlow < wlow ? (hlow < whigh ? hlow : whigh) : (wlow < whigh ? wlow : whigh);
  double sig = div_ru(r7.ub, mul_rd(shortest, shortest));
  sig = mul_ru(sig, sig);
  double cr = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(sig, 1.0)), 1.0)));
  if (cr < kWorstRatio) cr = kWorstRatio;
  double extra = div_ru(r7.ub, mul_rd(cr, hhigh));
  extra = mul_ru(extra, add_ru(hhigh, -hlow));
  double area = extra;
  area = add_ru(area, mul_ru(whigh, hhigh));
  area = add_ru(area, mul_ru(wlow, hlow));
  return R8 >= mul_ru(area, cr);
}
// seven_disk_strategies.cu:1351-1394
static void r7_four_disk_strip_r1_square(const Cell& c, const Vals& v, IV& r7) {
  double h2345 = nd_maximize_covered(c.la.ub, c.r + 1, 4);
  if (h2345 <= 0.0) return;
  double wh1 = sqrt_rd(mul_rd(2.0, c.r[0].lb));
  double hrem = add_ru(1.0, -h2345);
  double hrem_bl = add_ru(1.0, -add_rd(h2345, wh1));
  if (hrem_bl <= 0.0) {
    double w1 = sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lb), -mul_ru(hrem, hrem)));
    double wrem = add_ru(c.la.ub, -w1);
    double R6 = mul_rd(kRsbRatio, c.la.lb);
    for (int i = 0; i < 5; ++i) R6 = add_rd(R6, -c.r[i].ub);
    if (can_recurse(R6, c.r[5].ub, wrem, hrem)) r7.lb = kInf;
    return;
  }
  double need6 = mul_ru(0.25, add_ru(mul_ru(wh1, wh1), mul_ru(hrem_bl, hrem_bl)));
  double wrem = add_ru(c.la.ub, -wh1);
  if (need6 <= c.r[5].lb) {
    r7.tighten_lb(recursion_bound_size(v.R.lb, wrem, hrem));
    return;
  }
  double w6 = add_rd(mul_rd(4.0, c.r[5].lb), -mul_ru(hrem_bl, hrem_bl));
  if (w6 <= 0.0) return;
  w6 = sqrt_rd(w6);
  split_largest_r7(r7, v.R.lb, wrem, hrem, add_ru(wh1, -w6), hrem_bl);
}

// ------------------------------------------------------------------ C14
// seven_disk_strategies.cu:1396-1496.  The reference prints a diagnostic line
// for every failing cell; the oracle only reports the verdict.
bool seven_disk_strategies(const Cell& c, const Vals& v) {
  double covered = six_disks_maximize_covered_width(c);
  if (covered >= c.la.ub) return true;
  IV w_rem = c.la - covered;
  w_rem.tighten_lb(0.0);
  IV range = compute_r7_range(c, v, w_rem.ub);
  if (range.empty()) return true;
  r7_t_shaped(c, v, range);
  if (range.empty()) return true;
  r7_four_disk_strip_r1_square(c, v, range);
  if (range.empty()) return true;
  double h34 = two_disks_maximize_height(c.r[2], c.r[3], c.la.ub);
  for (int k = 0; k < 16; ++k) {
    IV r7 = get_subinterval(range, k, 16);
    if (r7_in_corner_and_recurse(v.R, r7, w_rem.ub)) continue;
    if (three_disk_top_strip(c, v.R, r7)) continue;
    if (six_disks_222(c, v, r7)) continue;
    if (r435_strip(c, r7, v)) continue;
    if (different_height_r12_2x2(c, v, r7)) continue;
    if (h34 > 0.0) {
      double h5 = five_disks_maximize_height_23(c.la.ub, c.r[0].lb, c.r[1].lb, c.r[4].lb, c.r[5].lb, r7.lb);
      if (add_rd(h34, h5) >= 1.0) continue;
    }
    return false;
  }
  return true;
}

// C14's stages on one cell, for the tests (no reference counterpart):
// out[0] the width r1..r6 cover, out[1..6] the r7 range (lb, ub) after
// compute_r7_range, the T-shaped restriction and the four-disk strip, and
// for each of the np r7 pieces (lb, ub) in pieces, bits[k] = which of the
// six piece strategies of seven_disk_strategies.cu:1430-1475 cover it, each
// evaluated on its own: 1 r7 in the corner, 2 three-disk top strip, 4 the
// 2-2-2 placement, 8 the r4-r3-r5 strip, 16 the different-height cover, 32
// the five-disk height (with h34 > 0).
void seven_disk_probe(const Cell& in, const double* pieces, int np, double* out, uint32_t* bits) {
  Cell c = tighten(in);
  Vals v = make_vals(c);
  double covered = six_disks_maximize_covered_width(c);
  out[0] = covered;
  IV w_rem = c.la - covered;
  w_rem.tighten_lb(0.0);
  IV range = compute_r7_range(c, v, w_rem.ub);
  out[1] = range.lb;
  out[2] = range.ub;
  if (!range.empty()) r7_t_shaped(c, v, range);
  out[3] = range.lb;
  out[4] = range.ub;
  if (!range.empty()) r7_four_disk_strip_r1_square(c, v, range);
  out[5] = range.lb;
  out[6] = range.ub;
  double h34 = two_disks_maximize_height(c.r[2], c.r[3], c.la.ub);
  for (int k = 0; k < np; ++k) {
    IV r7(pieces[2 * k], pieces[2 * k + 1]);
    uint32_t b = 0;
    if (r7_in_corner_and_recurse(v.R, r7, w_rem.ub)) b |= 1;
    if (three_disk_top_strip(c, v.R, r7)) b |= 2;
    if (six_disks_222(c, v, r7)) b |= 4;
    if (r435_strip(c, r7, v)) b |= 8;
    if (different_height_r12_2x2(c, v, r7)) b |= 16;
    if (h34 > 0.0) {
      double h5 = five_disks_maximize_height_23(c.la.ub, c.r[0].lb, c.r[1].lb, c.r[4].lb, c.r[5].lb, r7.lb);
      if (add_rd(h34, h5) >= 1.0) b |= 32;
    }
    bits[k] = b;
  }
}

// ------------------------------------------------------------------ cell verdict
// rectangle_size_bound.cu:162-229 for one 7-D cell.
int cell_verdict(const Cell& in) {
  const IV& la = in.la;
  if (shortcut_even_split_recursion(la, in.r[0], in.r[1])) return 0;
  if (shortcut_uneven_split_recursion(la, in.r[0], in.r[1])) return 1;
  if (shortcut_uneven_split_recursion(la, in.r[0], in.r[1], in.r[2])) return 2;
  if (shortcut_r1_r2_opposite_corners_recursion(la, in.r[0], in.r[1], in.r[2])) return 3;
  if (shortcut_r1_r2_vertical_strip(la, in.r[0], in.r[1], in.r[2])) return 4;
  return chain_verdict(in);
}

// The leaf part of cell_verdict (rectangle_size_bound.cu:186-229): tighten,
// Intermediate_values, emptiness, C1..C14.
int chain_verdict(const Cell& in) {
  Cell c = tighten(in);
  Vals v = make_vals(c);
  if (is_empty(c) || v.R.empty()) return 5;
  if (even_split_recursion(c, v)) return 6;
  if (uneven_split_recursion(c, v)) return 7;
  if (multi_disk_strip_vertical(c, v)) return 8;
  if (vertical_wall_building_recursion(c, v)) return 9;
  if (r1_in_corner_explicit_recursion(c, v)) return 10;
  if (r1_r2_opposite_corners_recursion(c, v)) return 11;
  if (r1_r2_vertical_strip(c, v)) return 12;
  v.cover_2x2 = compute_two_by_two_cover(c.r);
  if (v.cover_2x2.width > 0) {
    if (v.cover_2x2.width > c.la.ub || two_by_two_cover_with_strip_and_recursion(c, v)) return 13;
  }
  if (l_shaped_recursion(c, v)) return 14;
  if (r1_in_corner_wall_building_recursion(c, v)) return 15;
  if (advanced_multi_disk_strip_vertical(c, v)) return 16;
  if (five_disk_full_cover(c, v)) return 17;
  if (six_disk_full_cover(c, v)) return 18;
  if (seven_disk_strategies(c, v)) return 19;
  return -1;
}

// Every step of cell_verdict evaluated independently (no short-circuit):
// bit k set when step k succeeds (steps numbered as in cell_verdict); bit
// 20+k when step k reached a trap site.
uint64_t cell_mask(const Cell& in, uint64_t* step_ops) {
  uint64_t m = 0;
  uint64_t last = CCOR_OPS();
  auto put = [&](int k, bool ok) {
    if (ok) m |= 1ull << k;
    if (g_trap) m |= 1ull << (20 + k);
    g_trap = 0;
    if (step_ops) step_ops[k] = CCOR_OPS() - last;
    last = CCOR_OPS();
  };
  g_trap = 0;
  const IV& la = in.la;
  put(0, shortcut_even_split_recursion(la, in.r[0], in.r[1]));
  put(1, shortcut_uneven_split_recursion(la, in.r[0], in.r[1]));
  put(2, shortcut_uneven_split_recursion(la, in.r[0], in.r[1], in.r[2]));
  put(3, shortcut_r1_r2_opposite_corners_recursion(la, in.r[0], in.r[1], in.r[2]));
  put(4, shortcut_r1_r2_vertical_strip(la, in.r[0], in.r[1], in.r[2]));
  Cell c = tighten(in);
  Vals v = make_vals(c);
  put(5, is_empty(c) || v.R.empty());
  put(6, even_split_recursion(c, v));
  put(7, uneven_split_recursion(c, v));
  put(8, multi_disk_strip_vertical(c, v));
  put(9, vertical_wall_building_recursion(c, v));
  put(10, r1_in_corner_explicit_recursion(c, v));
  put(11, r1_r2_opposite_corners_recursion(c, v));
  put(12, r1_r2_vertical_strip(c, v));
  v.cover_2x2 = compute_two_by_two_cover(c.r);
  put(13, v.cover_2x2.width > 0 && (v.cover_2x2.width > c.la.ub || two_by_two_cover_with_strip_and_recursion(c, v)));
  put(14, l_shaped_recursion(c, v));
  put(15, r1_in_corner_wall_building_recursion(c, v));
  put(16, advanced_multi_disk_strip_vertical(c, v));
  put(17, five_disk_full_cover(c, v));
  put(18, six_disk_full_cover(c, v));
  put(19, seven_disk_strategies(c, v));
  return m;
}

}  // namespace rsb
}  // namespace ccor
