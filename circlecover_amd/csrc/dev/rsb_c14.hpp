// rectangle_size_bound C14 (seven_disk_strategies) for the gfx950 kernels:
// seven_disk_strategies.cu and six_disks_cover_width.cu, cited per function.
#pragma once

namespace ccd {
namespace rsbc {

// ------------------------------------------------------------------ six-disk width
// six_disks_cover_width.cu:30-93
static __device__ __noinline__ bool six_r56_center(double w, const cell7& c, Trap& tp) {
  const double r1 = c.r[0].lo, r2 = c.r[1].lo, r3 = c.r[2].lo, r4 = c.r[3].lo, r5 = c.r[4].lo, r6 = c.r[5].lo;
  double h14 = trc::two_disk_height(r1, r4, w);
  double h23 = trc::two_disk_height(r2, r3, w);
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
  iv y56{add_rd(h14, mul_rd(0.5, hrem)), add_ru(h14, mul_ru(0.5, hrem))};
  iv y14{mul_rd(0.5, h14), mul_ru(0.5, h14)};
  iv y23{add_rd(1.0, -mul_ru(0.5, h23)), add_ru(1.0, -mul_rd(0.5, h23))};
  iv h14sq = iv{mul_rd(h14, h14), mul_ru(h14, h14)} * 0.25;
  iv h23sq = iv{mul_rd(h23, h23), mul_ru(h23, h23)} * 0.25;
  iv w1 = sqrt(r1 - h14sq, tp);
  iv w2 = sqrt(r2 - h23sq, tp);
  iv x1 = w - w1;
  iv x2 = w - w2;
  w1 = w1 * 2.0;
  w2 = w2 * 2.0;
  point p5{pt(x5), y56};
  point p6{pt(x6), y56};
  point ps1{w - w1, pt(h14)};
  point ps2{w - w2, 1.0 - pt(h23)};
  crossing x12 = intersect_circles(circle{{x1, y14}, pt(r1)}, circle{{x2, y23}, pt(r2)}, tp);
  if (!x12.sure || !definitely(x12.p[0].x < x12.p[1].x)) return false;
  return dist2(p5, ps1).hi <= r5 && dist2(p5, ps2).hi <= r5 && dist2(p5, x12.p[0]).hi <= r5 &&
         dist2(p6, x12.p[1]).hi <= r6;
}
// Disk pairings and triples of six_disks_cover_width.cu:111-140 (index codes).
__constant__ static const unsigned char kTriples[10][6] = {
    {0, 1, 2, 3, 4, 5}, {0, 1, 3, 2, 4, 5}, {0, 1, 4, 2, 3, 5}, {0, 1, 5, 2, 3, 4}, {0, 2, 3, 1, 4, 5},
    {0, 2, 4, 1, 3, 5}, {0, 2, 5, 1, 3, 4}, {0, 3, 4, 1, 2, 5}, {0, 3, 5, 1, 2, 4}, {0, 4, 5, 1, 2, 3}};
__constant__ static const unsigned char kPairs[15][6] = {
    {0, 1, 2, 3, 4, 5}, {0, 1, 2, 4, 3, 5}, {0, 1, 2, 5, 3, 4}, {0, 2, 1, 3, 4, 5}, {0, 2, 1, 4, 3, 5},
    {0, 2, 1, 5, 3, 4}, {0, 3, 1, 2, 4, 5}, {0, 3, 1, 4, 2, 5}, {0, 3, 1, 5, 2, 4}, {0, 4, 1, 2, 3, 5},
    {0, 4, 1, 3, 2, 5}, {0, 4, 1, 5, 2, 3}, {0, 5, 1, 2, 3, 4}, {0, 5, 1, 3, 2, 4}, {0, 5, 1, 4, 2, 3}};
// six_disks_cover_width.cu:156-233
static __device__ __noinline__ double six_max_width(const cell7& c, Trap& tp) {
  double lbs[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) lbs[i] = c.r[i].lo;
  double w = 0.0;
  for (int k = 0; k < 15; ++k) {
    const unsigned char* p = kPairs[k];
    double a = trc::two_disk_height(lbs[p[0]], lbs[p[1]], 1.0);
    double b = trc::two_disk_height(lbs[p[2]], lbs[p[3]], 1.0);
    double d = trc::two_disk_height(lbs[p[4]], lbs[p[5]], 1.0);
    double t = add_rd(a, add_rd(b, d));
    if (t > w) w = t;
  }
  double w2 = 0.0;
  for (int k = 0; k < 10; ++k) {
    const unsigned char* p = kTriples[k];
    double a = three_height(lbs[p[0]], lbs[p[1]], lbs[p[2]], 1.0);
    double b = three_height(lbs[p[3]], lbs[p[4]], lbs[p[5]], 1.0);
    double t = add_rd(a, b);
    if (t > w2) w2 = t;
  }
  if (w2 > w) w = w2;
  if (six_r56_center(w, c, tp)) {
    double lo = w, hi = next_up(c.la.hi);  // nextafter(la.ub, DBL_MAX) for a finite positive bound
    for (int it = 0; it < kBisectCap; ++it) {
      double mid = rn_mid(lo, hi);
      if (mid <= lo || mid >= hi) {
        w = lo;
        break;
      }
      if (six_r56_center(mid, c, tp)) lo = mid;
      else hi = mid;
    }
  }
  return w;
}

// ------------------------------------------------------------------ C14 helpers
// seven_disk_strategies.cu:38-47
CCD bool r7_in_corner(const Konst& K, iv R, iv r7, double ub_w) {
  double R_no7 = add_rd(R.lo, -r7.hi);
  double h7 = add_rd(mul_rd(4.0, r7.lo), -mul_ru(ub_w, ub_w));
  if (h7 < 0) return false;
  return can_recurse(K, R_no7, r7.hi, ub_w, add_ru(1.0, -sqrt_rd(h7)));
}
// seven_disk_strategies.cu:49-115 (ub_hrem rounded down, as the reference)
static __device__ __noinline__ bool top_strip_can_cover(double la, double t1, double t2, double t3, double a0, double a1,
                                                 double a2, double a3, double ub_max_disk, double R_no7) {
  double htop = three_height(t1, t2, t3, la);
  if (htop <= 0) return false;
  double hrem = add_rd(1.0, -htop);
  if (hrem <= 0) return true;
  double hrem_sq = mul_ru(hrem, hrem);
  double wrem;
  double width_sb = div_rd(div_rd(R_no7, kRsbRatio), hrem);
  double min_sb = width_sb < hrem ? width_sb : hrem;
  if (disk_fits(ub_max_disk, min_sb)) {
    wrem = add_ru(la, -width_sb);
  } else {
    double width_wc = div_rd(div_rd(R_no7, kWorstRatio), hrem);
    double long_wc = width_wc < hrem ? hrem : width_wc;
    double short_wc = width_wc < hrem ? width_wc : hrem;
    if (div_ru(long_wc, short_wc) > kRsbToleratedLambda) {
      if (width_wc > hrem) {
        width_wc = mul_rd(hrem, kRsbToleratedLambda);
      } else {
        width_wc = add_rd(mul_rd(2.0, R_no7), -mul_ru(0.5, hrem_sq));
        width_wc = width_wc <= 0 ? 0.0 : sqrt_rd(width_wc);
      }
    }
    wrem = add_ru(la, -width_wc);
  }
  double hbot = trc::two_disk_height(a1, a2, wrem);
  if (hbot <= 0) return false;
  if (hbot >= hrem) return true;
  double hrem2 = add_ru(hrem, -hbot);
  double hl2 = trc::two_disk_height(a0, a3, wrem);
  if (hl2 <= 0.0) return false;
  return hl2 >= hrem2;
}
// seven_disk_strategies.cu:117-129
CCD bool split_largest_works(const Konst& K, double ub_r7, double wrem, double wf, double hf, double ws, double hs) {
  double cr = bound_crit_ratio(K, wrem, ub_r7, wf, hf);
  if (cr >= __builtin_inf()) return false;
  double rem2 = add_rd(wrem, -add_ru(mul_ru(cr, mul_ru(wf, hf)), ub_r7));
  return can_recurse(K, rem2, ub_r7, ws, hs);
}
// seven_disk_strategies.cu:131-154
static __device__ __noinline__ void split_largest_r7(const Konst& K, iv& r7, double wrem, double wf, double hf, double ws,
                                              double hs) {
  if (split_largest_works(K, r7.hi, wrem, wf, hf, ws, hs)) {
    r7.lo = __builtin_inf();
    return;
  }
  double lo = r7.lo, hi = r7.hi;
#if CC_EXACT_SHORTCUTS
  // split_largest_works passes, then fails, as its r7 bound grows (for
  // finite operands): bound_crit_ratio is non-decreasing in the largest
  // disk's bound (every candidate ratio grows with it and feasibility only
  // gets lost), so the remaining weight shrinks, and the size-bound
  // can_recurse is non-decreasing in the weight and non-increasing in the
  // largest disk.
  if (lo >= 0.0 && lo < hi && hi < __builtin_inf() && __builtin_isfinite(wrem) && __builtin_isfinite(wf) &&
      __builtin_isfinite(hf) && __builtin_isfinite(ws) && __builtin_isfinite(hs)) {
    tighten_lo(r7, monotone_last_pass(lo, hi, [&](double m) { return split_largest_works(K, m, wrem, wf, hf, ws, hs); }));
    return;
  }
#endif
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) {
      tighten_lo(r7, lo);
      return;
    }
    if (split_largest_works(K, mid, wrem, wf, hf, ws, hs)) lo = mid;
    else hi = mid;
  }
}
// seven_disk_strategies.cu:156-222
static __device__ __noinline__ void split_recursion(const Konst& K, const cell7& c, iv& r7, double wrem, int largest,
                                             double wf, double hf, double ws, double hs) {
  if (largest > 5) {
    split_largest_r7(K, r7, wrem, wf, hf, ws, hs);
    return;
  }
  double cr = bound_crit_ratio(K, wrem, c.r[largest].hi, wf, hf);
  if (!(cr < __builtin_inf())) return;
  double need = mul_ru(cr, mul_ru(wf, hf));
  double in_disks = 0.0;
  for (int i = largest; i <= 5; ++i) in_disks = add_ru(in_disks, c.r[i].hi);
  if (in_disks > need) {
    double lb_in = 0.0;
    for (int i = largest; i <= 5; ++i) lb_in = add_rd(lb_in, c.r[i].lo);
    if (lb_in >= need) tighten_lo(r7, recursion_bound(K, add_rd(wrem, -in_disks), ws, hs));
    return;
  }
  tighten_lo(r7, recursion_bound(K, add_rd(wrem, -add_ru(need, r7.hi)), ws, hs));
  if (empty(r7)) return;
  double hi = r7.hi, lo = r7.lo;
#if CC_EXACT_SHORTCUTS
  // recursion_bound is non-decreasing in the weight it gets, which shrinks
  // as mid grows, so "recursion_bound(...) >= mid" passes, then fails.
  if (lo >= 0.0 && lo < hi && hi < __builtin_inf() && __builtin_isfinite(wrem) && __builtin_isfinite(need) &&
      __builtin_isfinite(ws) && __builtin_isfinite(hs)) {
    r7.lo = monotone_last_pass(lo, hi, [&](double m) { return !(recursion_bound(K, add_rd(wrem, -add_ru(need, m)), ws, hs) < m); });
    return;
  }
#endif
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) {
      r7.lo = lo;
      return;
    }
    if (recursion_bound(K, add_rd(wrem, -add_ru(need, mid)), ws, hs) < mid) hi = mid;
    else lo = mid;
  }
}
struct r6pos {
  point p6;
  double left;
};
// seven_disk_strategies.cu:233-262
CCD r6pos place_r6(iv r6, const point& x1, const point& x2, iv lower, iv upper) {
  r6pos pos;
  pos.left = 0.0;
  pos.p6 = point{pt(0.0), pt(0.0)};
  double left = x1.x.lo < x2.x.lo ? x1.x.lo : x2.x.lo;
  double top = x1.y.hi < x2.y.hi ? x2.y.hi : x1.y.hi;
  double bot = x1.y.lo < x2.y.lo ? x1.y.lo : x2.y.lo;
  double h = add_ru(top, -bot);
  double dx6 = add_rd(r6.lo, -mul_ru(0.25, mul_ru(h, h)));
  if (dx6 <= 0.0) return pos;
  dx6 = sqrt_rd(dx6);
  pos.p6 = point{left + pt(dx6), iv{mul_rd(0.5, add_rd(bot, top)), mul_ru(0.5, add_ru(bot, top))}};
  double dyb = (pos.p6.y - lower).hi;
  double dyt = (upper - pos.p6.y).hi;
  dyb = mul_ru(dyb, dyb);
  dyt = mul_ru(dyt, dyt);
  double sw = add_rd(r6.lo, -(dyb < dyt ? dyt : dyb));
  if (sw <= 0.0) return pos;
  pos.left = add_rd(pos.p6.x.lo, sqrt_rd(sw));
  return pos;
}
// seven_disk_strategies.cu:264-316
static __device__ __noinline__ void r7_t_r5_large(const Konst& K, const cell7& c, iv& r7, const circle& c1,
                                           const circle& c3, const circle& c5, double R6, double w12, double wrem,
                                           double hrem, double strip_w5, iv lower, iv upper, Trap& tp) {
  crossing x15 = intersect_circles(c1, c5, tp), x35 = intersect_circles(c3, c5, tp);
  if (!x15.sure || !x35.sure) {
    tp.flag = 1;
    return;
  }
  double left = x15.p[1].x.lo;
  if (left > x35.p[1].x.lo) left = x35.p[1].x.lo;
  double hz = add_ru(w12, -left);
  if (hz <= 0.0) {
    if (can_recurse(K, R6, c.r[5].hi, wrem, 1.0)) tighten_lo(r7, __builtin_inf());
    return;
  }
  split_recursion(K, c, r7, R6, 5, hz, hrem, wrem, 1.0);
  split_recursion(K, c, r7, R6, 5, wrem, 1.0, hz, hrem);
  if (empty(r7)) return;
  double R7 = add_rd(R6, -c.r[5].hi);
  if (R7 <= 0.0) return;
  point x1 = x15.p[1], x2 = x35.p[1];
  if (possibly(x1.y >= upper) || possibly(x1.y <= lower) || possibly(x2.y >= upper) || possibly(x2.y <= lower)) {
    x1 = point{pt(strip_w5), upper};
    x2 = point{pt(strip_w5), lower};
  }
  r6pos pos = place_r6(c.r[5], x1, x2, lower, upper);
  if (pos.left <= 0.0) return;
  if (pos.left >= w12) {
    tighten_lo(r7, recursion_bound(K, R7, wrem, 1.0));
  } else {
    hz = add_ru(w12, -left);  // the reference reuses the r5 intersection here
    split_recursion(K, c, r7, R7, 6, hz, hrem, wrem, 1.0);
    split_recursion(K, c, r7, R7, 6, wrem, 1.0, hz, hrem);
  }
}
// seven_disk_strategies.cu:318-518
static __device__ __noinline__ void r7_t_shaped(const Konst& K, const cell7& c, iv& r7, Trap& tp) {
  double w2 = sqrt_rd(mul_rd(2.0, c.r[1].lo));
  double h12 = w2;
  double w1 = sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lo), -mul_ru(h12, h12)));
  double w12 = add_rd(w1, w2);
  double h34 = trc::two_disk_height(c.r[2].lo, c.r[3].lo, w12);
  if (h34 <= 0.0) return;
  double R5 = add_rd(mul_rd(c.la.lo, kRsbRatio), -add_ru(c.r[0].hi, add_ru(c.r[1].hi, add_ru(c.r[2].hi, c.r[3].hi))));
  double hrem = add_ru(1.0, -add_rd(h12, h34));
  double wrem = add_ru(c.la.hi, -w12);
  if (hrem <= 0.0) {
    if (wrem <= 0.0 || can_recurse(K, R5, c.r[4].hi, wrem, 1.0)) tighten_lo(r7, __builtin_inf());
    return;
  }
  if (wrem <= 0.0) {
    if (can_recurse(K, R5, c.r[4].hi, c.la.hi, hrem)) tighten_lo(r7, __builtin_inf());
    return;
  }
  split_recursion(K, c, r7, R5, 4, w12, hrem, wrem, 1.0);
  split_recursion(K, c, r7, R5, 4, wrem, 1.0, w12, hrem);
  if (empty(r7)) return;
  double R6 = add_rd(R5, -c.r[4].hi);
  if (R6 <= 0.0) return;
  double x5 = add_rd(c.r[4].lo, -mul_ru(0.25, mul_ru(hrem, hrem)));
  if (x5 <= 0.0) return;
  x5 = sqrt_rd(x5);
  double strip_w5 = mul_rd(2.0, x5);
  point p5{pt(x5), iv{add_rd(h12, mul_rd(0.5, hrem)), add_ru(h12, mul_ru(0.5, hrem))}};
  iv p2y{mul_rd(0.5, h12), mul_ru(0.5, h12)};
  point p2{iv{mul_rd(0.5, w2), mul_ru(0.5, w2)}, p2y};
  point p1{w2 + pt(0.5) * w1, p2y};
  iv x4 = sqrt(c.r[3] - 0.25 * square(pt(h34)), tp);
  iv y4 = 1.0 - 0.5 * pt(h34);
  point p4{x4, y4};
  iv w3 = w12 - 2.0 * x4;
  point p3{w12 - 0.5 * w3, y4};
  circle c1{p1, pt(c.r[0].lo)};
  circle c2{p2, pt(c.r[1].lo)};
  circle c3{p3, pt(c.r[2].lo)};
  circle c4{p4, pt(c.r[3].lo)};
  circle c5{p5, pt(c.r[4].lo)};
  point p43{2.0 * x4, pt(1.0) - h34};
  point p21{pt(w2), pt(h12)};
  const iv lower = pt(h12);
  const iv upper = pt(1.0) - h34;
  if (definitely(dist2(p5, p43) <= c.r[4]) && definitely(dist2(p5, p21) <= c.r[4])) {
    r7_t_r5_large(K, c, r7, c1, c3, c5, R6, w12, wrem, hrem, strip_w5, lower, upper, tp);
    return;
  }
  (void)intersect_circles(c2, c4, tp);  // computed and unused by the reference (x24)
  crossing x25 = intersect_circles(c2, c5, tp);
  crossing x45 = intersect_circles(c4, c5, tp);
  if (!x25.sure || !x45.sure) {
    tp.flag = 1;
    return;
  }
  if (x25.p[0].x.hi >= x25.p[1].x.lo || x45.p[0].x.hi >= x45.p[1].x.lo) return;
  point x1 = x25.p[1], x2 = x45.p[1];
  if (possibly(x1.y >= upper) || possibly(x1.y <= lower) || possibly(x2.y >= upper) || possibly(x2.y <= lower)) {
    x1 = point{pt(strip_w5), upper};
    x2 = point{pt(strip_w5), lower};
  }
  double left = x1.x.lo > x2.x.lo ? x2.x.lo : x1.x.lo;
  double hz = add_ru(w12, -left);
  split_recursion(K, c, r7, R6, 5, hz, hrem, wrem, 1.0);
  split_recursion(K, c, r7, R6, 5, wrem, 1.0, hz, hrem);
  if (empty(r7)) return;
  double R7 = add_rd(R6, -c.r[5].hi);
  if (R7 <= 0.0) return;
  r6pos pos = place_r6(c.r[5], x1, x2, lower, upper);
  if (pos.left <= 0.0) return;
  circle c6{pos.p6, pt(c.r[5].lo)};
  crossing x16 = intersect_circles(c1, c6, tp), x36 = intersect_circles(c3, c6, tp);
  if (!x16.sure || !x36.sure) {
    tp.flag = 1;
    return;
  }
  point d71, d72;
  bool r6_right_x = false;
  if (definitely(dist2(p43, pos.p6) <= c.r[5]) && definitely(dist2(p21, pos.p6) <= c.r[5])) {
    left = x16.p[1].x.lo < x36.p[1].x.lo ? x16.p[1].x.lo : x36.p[1].x.lo;
    d71 = x36.p[1];
    d72 = x16.p[1];
    r6_right_x = true;
  } else {
    left = pos.left;
    // the reference builds these points with the strip borders as x
    d71 = point{upper, pt(left)};
    d72 = point{lower, pt(left)};
  }
  hz = add_ru(w12, -left);
  if (hz <= 0.0) {
    tighten_lo(r7, recursion_bound(K, R7, wrem, 1.0));
  } else {
    split_recursion(K, c, r7, R7, 6, hz, hrem, wrem, 1.0);
    split_recursion(K, c, r7, R7, 6, wrem, 1.0, hz, hrem);
  }
  if (empty(r7)) return;
  double R8 = add_rd(R7, -r7.hi);
  if (R8 <= 0.0) return;
  crossing x13 = intersect_circles(c1, c3, tp);
  if (r6_right_x && x13.sure && definitely(dist2(x13.p[0], pos.p6) <= c.r[5]) && definitely(x13.p[1].y >= lower) &&
      definitely(x13.p[1].y <= upper)) {
    d71 = x13.p[1];
    d72 = x13.p[1];
  }
  r6pos p7 = place_r6(r7, d71, d72, lower, upper);
  if (p7.left <= 0.0) return;
  if (p7.left >= c.la.hi) {
    split_largest_r7(K, r7, R8, wrem, h12, wrem, h34);
    split_largest_r7(K, r7, R8, wrem, h34, wrem, h12);
  }
  if (p7.left >= w12) {
    tighten_lo(r7, recursion_bound(K, R8, wrem, 1.0));
  } else {
    hz = add_ru(w12, -p7.left);
    split_largest_r7(K, r7, R8, hz, hrem, wrem, 1.0);
    split_largest_r7(K, r7, R8, wrem, 1.0, hz, hrem);
  }
}
// seven_disk_strategies.cu:520-554
static __device__ __noinline__ bool three_top_strip(const cell7& c, iv R, iv r7) {
  const double R_no7 = add_rd(R.lo, -r7.hi);
  double d[7];
#pragma unroll
  for (int i = 0; i < 6; ++i) d[i] = c.r[i].lo;
  d[6] = r7.lo;
  for (int i = 0; i < 5; ++i)
    for (int j = i + 1; j < 6; ++j)
      for (int k = j + 1; k < 7; ++k) {
        double rem[4];
        int n = 0;
        for (int g = 0; g < 7; ++g)
          if (g != i && g != j && g != k) rem[n++] = d[g];
        if (top_strip_can_cover(c.la.hi, d[i], d[j], d[k], rem[0], rem[1], rem[2], rem[3], r7.hi, R_no7)) return true;
      }
  return false;
}
// seven_disk_strategies.cu:556-718 (rem_width adds, as the reference)
static __device__ __noinline__ bool r435_strip(const Konst& K, const cell7& c, iv r7, iv R, Trap& tp) {
  double cov = three_height(c.r[2].lo, c.r[3].lo, c.r[4].lo, 1.0);
  if (cov <= 0.0) return false;
  iv wsq435{mul_rd(cov, cov), mul_ru(cov, cov)};
  const iv r3 = pt(c.r[2].lo), r4 = pt(c.r[3].lo);
  iv h3 = 4.0 * r3 - wsq435;
  iv h4 = 4.0 * r4 - wsq435;
  if (possibly(h3 <= 0.0) || possibly(h4 <= 0.0)) return false;
  h3 = sqrt(h3, tp);
  h4 = sqrt(h4, tp);
  point pos3{iv{mul_rd(0.5, cov), mul_ru(0.5, cov)}, 0.5 * h3 + h4};
  const iv r1 = pt(c.r[0].lo), r2 = pt(c.r[1].lo);
  iv r1s = sqrt(r1, tp), r2s = sqrt(r2, tp);
  if (possibly(r1s + r2s <= 0.5)) return false;
  double lo = 0.0;
  double hi = mul_rd(2.0, r2s.lo);
  double wrem = add_ru(c.la.hi, -cov);
  iv wrem_sq{mul_rd(wrem, wrem), mul_ru(wrem, wrem)};
  if (wrem < hi) {
    iv y1 = r1 - 0.25 * wrem_sq;
    iv y2 = r2 - 0.25 * wrem_sq;
    if (definitely(y1 > 0.0) && definitely(y2 > 0.0)) {
      y1 = 1.0 - sqrt(y1, tp);
      y2 = sqrt(y2, tp);
      iv x = 0.5 * pt(wrem) + cov;
      crossing x12 = intersect_circles(circle{{x, y1}, r1}, circle{{x, y2}, r2}, tp);
      if (x12.sure && x12.p[0].x.hi < x12.p[1].x.lo && definitely(dist2(pos3, x12.p[0]) <= r3)) {
        double h1 = mul_rd(2.0, sqrt_rd(add_rd(r1.lo, -mul_ru(0.25, wrem_sq.hi))));
        double h2 = mul_rd(2.0, sqrt_rd(add_rd(r2.lo, -mul_ru(0.25, wrem_sq.hi))));
        double rem_h = add_ru(1.0, -add_rd(h1, h2));
        double rem_w = add_ru(c.la.hi, x12.p[1].x.lo);
        if (can_recurse(K, add_rd(R.lo, c.r[5].lo), c.r[5].hi, rem_w, rem_h)) return true;
      }
    }
    hi = wrem;
  }
  double w12 = 0.0;
  crossing best;
  best.sure = false;
  for (int it = 0; it < kBisectCap; ++it) {
    double mid = rn_mid(lo, hi);
    if (mid <= lo || mid >= hi) {
      w12 = lo;
      break;
    }
    iv wsq{mul_rd(mid, mid), mul_ru(mid, mid)};
    iv x = 0.5 * pt(mid) + cov;
    iv hh2 = r2 - 0.25 * wsq;
    if (hh2.lo <= 0) {
      hi = mid;
      continue;
    }
    hh2 = sqrt(hh2, tp);
    iv hh1 = sqrt(r1 - 0.25 * wsq, tp);
    crossing x12 = intersect_circles(circle{{x, 1.0 - hh1}, r1}, circle{{x, hh2}, r2}, tp);
    if (x12.sure && x12.p[0].x.hi < x12.p[1].x.lo && definitely(dist2(pos3, x12.p[0]) <= r3)) {
      lo = mid;
      best = x12;
    } else {
      hi = mid;
    }
  }
  if (!best.sure) return false;
  double w67 = add_ru(c.la.hi, -add_rd(cov, w12));
  double w67sq = mul_ru(w67, w67);
  {
    double h7 = add_rd(mul_rd(4.0, r7.lo), -w67sq);
    double h6 = add_rd(mul_rd(4.0, c.r[5].lo), -w67sq);
    if (h7 > 0 && h6 > 0) {
      h7 = sqrt_rd(h7);
      h6 = sqrt_rd(h6);
      point pos6{c.la - iv{mul_rd(0.5, w67), mul_ru(0.5, w67)}, 1.0 - pt(h7) - 0.5 * pt(h6)};
      double hr = add_ru(1.0, -add_rd(h7, h6));
      if (dist2(best.p[1], pos6).hi <= c.r[5].lo && can_recurse(K, add_rd(R.lo, -r7.hi), r7.hi, w67, hr))
        return true;
    }
  }
  double wide = add_ru(c.la.hi, -best.p[1].x.lo);
  double widesq = mul_ru(wide, wide);
  double h7 = add_rd(mul_rd(4.0, r7.lo), -widesq);
  double h6 = add_rd(mul_rd(4.0, c.r[5].lo), -widesq);
  if (!(h7 > 0 && h6 > 0)) return false;
  h7 = sqrt_rd(h7);
  h6 = sqrt_rd(h6);
  double hr = add_ru(1.0, -add_rd(h6, h7));
  return can_recurse(K, add_rd(R.lo, -r7.hi), r7.hi, hr <= best.p[1].y.lo ? w67 : wide, hr);
}
// seven_disk_strategies.cu:961-1061 (`return false` from a double: 0.0)
static __device__ __noinline__ double five_height_23(double ub_w, double t1, double t2, double m1, double m2, double m3,
                                              Trap& tp) {
  strip2 bot = max_height_strip(ub_w, t1, t2, tp);
  if (bot.h <= 0.0) return 0.0;
  crossing xt = intersect_circles(bot.c1, bot.c2, tp);
  if (!xt.sure) {
    tp.flag = 1;
    return 0.0;
  }
  UB low = xt.p[0].y <= xt.p[1].y;
  if (!is_certain(low)) return 0.0;
  const point xtup = xt.p[low.lo ? 1 : 0];
  const double q1 = mul_rd(4.0, m1), q2 = mul_rd(4.0, m2), q3 = mul_rd(4.0, m3);
  double lo = 0.0, hi = mul_rd(2.0, sqrt_rd(m3));
  for (int it = 0; it < kBisectCap; ++it) {
    double h = rn_mid(lo, hi);
    if (h <= lo || h >= hi) break;
    double hsq = mul_ru(h, h);
    double w1 = add_rd(q1, -hsq), w2 = add_rd(q2, -hsq), w3 = add_rd(q3, -hsq);
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
    double h1 = add_rd(m1, -mul_ru(0.25, mul_ru(wr, wr)));
    if (h1 <= 0.0) {
      hi = h;
      continue;
    }
    h1 = sqrt_rd(h1);
    iv y23{add_rd(bot.h, mul_rd(0.5, h)), add_ru(bot.h, mul_ru(0.5, h))};
    point pc1{0.5 * (pt(w2) + ub_w - pt(w3)), iv{add_rd(bot.h, add_rd(h, -h1)), add_ru(bot.h, add_ru(h, -h1))}};
    if (dist2(pc1, xtup).hi > m1) {
      hi = h;
      continue;
    }
    circle d1{pc1, pt(m1)};
    crossing x12 = intersect_circles(d1, circle{{pt(w2) * 0.5, y23}, pt(m2)}, tp);
    crossing x13 = intersect_circles(d1, circle{{ub_w - pt(w3) * 0.5, y23}, pt(m3)}, tp);
    if (!x12.sure || !x13.sure) {
      hi = h;
      continue;
    }
    UB l12 = x12.p[0].y <= x12.p[1].y;
    UB l13 = x13.p[0].y <= x13.p[1].y;
    if (!is_certain(l12) || !is_certain(l13)) {
      hi = h;
      continue;
    }
    if (dist2(x12.p[l12.lo ? 0 : 1], bot.c1.c).hi <= t1 && dist2(x13.p[l13.lo ? 0 : 1], bot.c2.c).hi <= t2) lo = h;
    else hi = h;
  }
  return lo == 0.0 ? 0.0 : add_rd(bot.h, lo);
}
// seven_disk_strategies.cu:1197-1239
CCD iv r7_range_for(const Konst& K, const cell7& c, iv R, double ub_w_remaining) {
  iv range{0.0, c.r[5].hi};
  tighten_hi(range, R.hi);
  double eff = div_rd(R.lo, ub_w_remaining);
  if (eff >= kRsbRatio) {
    double strip = div_rd(R.lo, kRsbRatio);
    if (strip >= 1.0) return iv{1.0, 0.0};
    tighten_lo(range, allowed_weight(strip));
  }
  if (eff >= kWorstRatio && !empty(range)) {
    double crs = trc::crit_ratio(K, iv{rcp_rd(ub_w_remaining), rcp_ru(ub_w_remaining)}).hi;
    if (crs <= eff) return iv{1.0, 0.0};
    double lam = add_rd(mul_rd(2.0, eff), sqrt_rd(add_rd(mul_rd(4.0, mul_rd(eff, eff)), -2.0)));
    double sigma = mul_rd(eff, add_rd(lam, -div_ru(2.0, lam)));
    tighten_lo(range, mul_rd(sigma, mul_rd(ub_w_remaining, ub_w_remaining)));
  }
  return range;
}
// seven_disk_strategies.cu:1241-1297 (h1 not sign-checked, as the reference)
static __device__ __noinline__ bool six_222(const Konst& K, const cell7& c, iv R, iv r7) {
  double c12 = trc::two_disk_height(c.r[0].lo, c.r[1].lo, 1.0);
  double c34 = trc::two_disk_height(c.r[2].lo, c.r[3].lo, 1.0);
  if (c34 <= 0.0) return false;
  const double r6 = c.r[5].lo;
  double wrem = add_ru(c.la.hi, -add_rd(c12, c34));
  double wrem_sq = mul_ru(wrem, wrem);
  double h1 = add_rd(mul_rd(4.0, c.r[4].lo), -wrem_sq);
  double h2 = add_rd(mul_rd(4.0, r6), -wrem_sq);
  if (h2 <= 0.0) return false;
  h1 = sqrt_rd(h1);
  h2 = sqrt_rd(h2);
  double hrem = add_ru(1.0, -add_rd(h1, h2));
  if (hrem <= 0) return true;
  double dy = add_ru(1.0, -add_rd(h1, mul_rd(0.5, h2)));
  double wtop = add_rd(r6, -mul_ru(dy, dy));
  if (wtop <= 0) return false;
  wtop = mul_rd(2.0, sqrt_rd(wtop));
  double wrect = mul_ru(0.5, add_ru(wrem, -wtop));
  double need7 = mul_ru(0.25, add_ru(mul_ru(wrect, wrect), mul_ru(hrem, hrem)));
  if (need7 <= r7.lo) return can_recurse(K, add_rd(R.lo, -r7.hi), r7.hi, wrect, hrem);
  double smaller = add_rd(mul_rd(0.5, R.lo), -mul_ru(0.5, r7.hi));
  return can_recurse(K, smaller, r7.hi, wrect, hrem);
}
// seven_disk_strategies.cu:1299-1349 (h1 uses w12 unsquared, as the reference)
static __device__ __noinline__ bool diff_height_r12_22(const Konst& K, const cell7& c, iv R, iv r7) {
  double w12 = sqrt_rd(mul_rd(2.0, c.r[1].lo));
  double h12 = add_rd(sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lo), -w12)), w12);
  double wrem = add_ru(c.la.hi, -w12);
  double h3456 = add_rd(trc::two_disk_height(c.r[2].lo, c.r[3].lo, wrem), trc::two_disk_height(c.r[4].lo, c.r[5].lo, wrem));
  double hr12 = add_ru(1.0, -h12);
  double hr3456 = add_ru(1.0, -h3456);
  const bool a = hr12 < hr3456;
  double hhigh = a ? hr3456 : hr12, hlow = a ? hr12 : hr3456;
  double whigh = a ? wrem : w12, wlow = a ? w12 : wrem;
  double wr7 = add_rd(mul_rd(4.0, r7.lo), -mul_ru(hhigh, hhigh));
  if (wr7 <= 0.0) return false;
  whigh = add_rd(whigh, -sqrt_rd(wr7));
  double R8 = add_rd(R.lo, -r7.hi);
  if (R8 <= 0.0) return false;
  if (whigh <= 0.0) return can_recurse(K, R8, r7.hi, add_ru(whigh, wlow), hlow);
  double shortest = hlow < wlow ? (hlow < whigh ? hlow : whigh) : (wlow < whigh ? wlow : whigh);
  double sig = div_ru(r7.hi, mul_rd(shortest, shortest));
  sig = mul_ru(sig, sig);
  double cr = mul_ru(0.5, sqrt_ru(add_ru(sqrt_ru(add_ru(sig, 1.0)), 1.0)));
  if (cr < kWorstRatio) cr = kWorstRatio;
  double area = mul_ru(div_ru(r7.hi, mul_rd(cr, hhigh)), add_ru(hhigh, -hlow));
  area = add_ru(area, mul_ru(whigh, hhigh));
  area = add_ru(area, mul_ru(wlow, hlow));
  return R8 >= mul_ru(area, cr);
}
// seven_disk_strategies.cu:1351-1394
static __device__ __noinline__ void r7_four_disk_r1_square(const Konst& K, const cell7& c, iv R, iv& r7) {
  double lbs[4] = {c.r[1].lo, c.r[2].lo, c.r[3].lo, c.r[4].lo};
  double h2345 = nd_maximize(c.la.hi, lbs, 4);
  if (h2345 <= 0.0) return;
  double wh1 = sqrt_rd(mul_rd(2.0, c.r[0].lo));
  double hrem = add_ru(1.0, -h2345);
  double hrem_bl = add_ru(1.0, -add_rd(h2345, wh1));
  if (hrem_bl <= 0.0) {
    double wrem = add_ru(c.la.hi, -sqrt_rd(add_rd(mul_rd(4.0, c.r[0].lo), -mul_ru(hrem, hrem))));
    double R6 = mul_rd(kRsbRatio, c.la.lo);
    for (int i = 0; i < 5; ++i) R6 = add_rd(R6, -c.r[i].hi);
    if (can_recurse(K, R6, c.r[5].hi, wrem, hrem)) r7.lo = __builtin_inf();
    return;
  }
  double need6 = mul_ru(0.25, add_ru(mul_ru(wh1, wh1), mul_ru(hrem_bl, hrem_bl)));
  double wrem = add_ru(c.la.hi, -wh1);
  if (need6 <= c.r[5].lo) {
    tighten_lo(r7, recursion_bound(K, R.lo, wrem, hrem));
    return;
  }
  double w6 = add_rd(mul_rd(4.0, c.r[5].lo), -mul_ru(hrem_bl, hrem_bl));
  if (w6 <= 0.0) return;
  split_largest_r7(K, r7, R.lo, wrem, hrem, add_ru(wh1, -sqrt_rd(w6)), hrem_bl);
}

// ------------------------------------------------------------------ C14
// seven_disk_strategies.cu:1396-1496, in two parts: the prefix (the widest
// strip of r1..r6 and the restrictions of the r7 range) and one r7 piece of
// the loop over its 16 pieces.  The reference prints a diagnostic line for
// each failing cell; here the failing cell is simply left unproven.

// What the loop needs from the prefix.
struct c14_state {
  iv range;        // the restricted r7 range
  double w_rem_hi; // upper bound of the width r1..r6 leave uncovered
  double h34;      // two_disks_maximize_height(r3, r4, la.ub)
};

// seven_disk_strategies.cu:1398-1432: true if the prefix alone proves the
// cell, else false with `s` set for the loop.
CCD bool c14_prefix_body(const Konst& K, const cell7& c, const vals7& v, Trap& tp, c14_state& s) {
  double covered = six_max_width(c, tp);
  if (covered >= c.la.hi) return true;
  iv w_rem = c.la - covered;
  tighten_lo(w_rem, 0.0);
  iv range = r7_range_for(K, c, v.R, w_rem.hi);
  if (empty(range)) return true;
  r7_t_shaped(K, c, range, tp);
  if (empty(range)) return true;
  r7_four_disk_r1_square(K, c, v.R, range);
  if (empty(range)) return true;
  s.range = range;
  s.w_rem_hi = w_rem.hi;
  s.h34 = trc::two_disk_height(c.r[2].lo, c.r[3].lo, c.la.hi);
  return false;
}

// seven_disk_strategies.cu:1434-1471: one of the strategies covers the
// cell with r7 in piece k of the 16 pieces of s.range.
CCD bool c14_piece_body(const Konst& K, const cell7& c, const vals7& v, const c14_state& s, int k, Trap& tp) {
  iv r7 = subinterval(s.range, k, 16);
  if (r7_in_corner(K, v.R, r7, s.w_rem_hi)) return true;
  if (three_top_strip(c, v.R, r7)) return true;
  if (six_222(K, c, v.R, r7)) return true;
  if (r435_strip(K, c, r7, v.R, tp)) return true;
  if (diff_height_r12_22(K, c, v.R, r7)) return true;
  if (s.h34 > 0.0) {
    double h5 = five_height_23(c.la.hi, c.r[0].lo, c.r[1].lo, c.r[4].lo, c.r[5].lo, r7.lo, tp);
    if (add_rd(s.h34, h5) >= 1.0) return true;
  }
  return false;
}

// The whole of C14 in one call: the prefix, then the 16 pieces in order
// until one fails.
static __device__ __noinline__ bool c14_seven_disk(const Konst& K, const cell7& c, const vals7& v, Trap& tp) {
  c14_state s;
  if (c14_prefix_body(K, c, v, tp, s)) return true;
  for (int k = 0; k < 16; ++k)
    if (!c14_piece_body(K, c, v, s, k, tp)) return false;
  return true;
}

// The two parts as calls of their own, for drivers that run the pieces of
// one cell on several lanes.
static __device__ __noinline__ bool c14_prefix(const Konst& K, const cell7& c, const vals7& v, Trap& tp,
                                                c14_state& s) {
  return c14_prefix_body(K, c, v, tp, s);
}
static __device__ __noinline__ bool c14_piece(const Konst& K, const cell7& c, const vals7& v, const c14_state& s,
                                               int k, Trap& tp) {
  return c14_piece_body(K, c, v, s, k, tp);
}

}  // namespace rsbc
}  // namespace ccd
