// rectangle_size_bound C8..C14 for the gfx950 kernels (see rsb_chain.hpp).
#pragma once

namespace ccd {
namespace rsbc {

// ------------------------------------------------------------------ 2x2 cover
// two_by_two_cover.cu:29-76 for one pairing; keeps width and disk 4's circle.
struct cover22 {
  double width;
  circle c4;
};
CCD cover22 two_by_two_concrete(double r1, double r2, double r3, double r4, Trap& tp) {
  cover22 res;
  res.width = 0.0;
  const iv z = pt(0.0);
  iv w1;
  {
    iv d = pt(r1) - r2;
    iv dsq = square(d);
    iv h1 = 0.5 + 2.0 * d;
    iv h2 = 0.5 - 2.0 * d;
    w1 = 2.0 * (pt(r1) + r2) - 0.25 - 4.0 * dsq;
    if (possibly(r1 - 0.25 * square(h1) <= 0.0) || possibly(r2 - 0.25 * square(h2) <= 0.0) || possibly(w1 <= 0.0)) {
      w1 = z;
    } else {
      w1 = sqrt(w1, tp);
      res.width = w1.lo;
    }
  }
  {
    iv d = pt(r3) - r4;
    iv dsq = square(d);
    iv h1 = 0.5 + 2.0 * d;
    iv h2 = 0.5 - 2.0 * d;
    iv w2 = 2.0 * (pt(r3) + r4) - 0.25 - 4.0 * dsq;
    if (possibly(r3 - 0.25 * square(h1) <= 0.0) || possibly(r4 - 0.25 * square(h2) <= 0.0) || possibly(w2 <= 0.0)) {
      res.c4 = circle{{z, z}, pt(r3)};
    } else {
      w2 = sqrt(w2, tp);
      res.c4 = circle{{w1 + 0.5 * w2, 1.0 - 0.5 * h1}, pt(r3)};
      res.width = add_rd(res.width, w2.lo);
    }
  }
  return res;
}
// two_by_two_cover.cu:78-88
CC_CHAIN_FN cover22 two_by_two(const cell7& c, Trap& tp) {
  cover22 a = two_by_two_concrete(c.r[0].lo, c.r[1].lo, c.r[2].lo, c.r[3].lo, tp);
  cover22 b = two_by_two_concrete(c.r[0].lo, c.r[2].lo, c.r[1].lo, c.r[3].lo, tp);
  cover22 d = two_by_two_concrete(c.r[0].lo, c.r[3].lo, c.r[1].lo, c.r[2].lo, tp);
  if (a.width > b.width) return a.width > d.width ? a : d;
  return b.width > d.width ? b : d;
}

// ------------------------------------------------------------------ C8
// two_by_two_cover_with_strip_and_recursion.cu:26-105
CC_CHAIN_FN bool c8_two_by_two_strip(const Konst& K, const cell7& c, const vals7& v, Trap& tp) {
  const double r5 = c.r[4].lo, r6 = c.r[5].lo, R = v.R.lo;
  double ub_w = add_ru(c.la.hi, -v.cover_width);
  double R5 = mul_rd(kRsbRatio, c.la.lo);
  double R6 = add_rd(R5, -c.r[4].hi);
  for (int i = 0; i < 4; ++i) {
    R5 = add_rd(R5, -c.r[i].hi);
    R6 = add_rd(R6, -c.r[i].hi);
  }
  if (can_recurse(K, R5, c.r[4].hi, ub_w, 1.0)) return true;
  double wsq = mul_ru(ub_w, ub_w);
  double h5 = add_rd(mul_rd(4.0, r5), -wsq);
  if (h5 <= 0.0) return false;
  h5 = sqrt_rd(h5);
  double hrem = add_ru(1.0, -h5);
  if (h5 >= 1.0 || can_recurse(K, R6, c.r[5].hi, ub_w, hrem)) return true;
  double h6 = add_rd(mul_rd(4.0, r6), -wsq);
  if (h6 > 0.0) {
    h6 = sqrt_rd(h6);
    if (h6 >= hrem) return true;
    if (can_recurse(K, R, c.r[5].hi, ub_w, add_ru(hrem, -h6))) return true;
  }
  double w6 = add_rd(mul_rd(4.0, r6), -mul_ru(hrem, hrem));
  if (!(w6 > 0.0)) return false;
  w6 = sqrt_rd(w6);
  if (w6 >= ub_w) return true;
  circle c5{{c.la - pt(0.5) * ub_w, pt(0.5) * h5}, pt(r5)};
  circle c6{{c.la - pt(0.5) * w6, 1.0 - pt(0.5) * hrem}, pt(r6)};
  crossing x56 = intersect_circles(c5, c6, tp);
  if (!(x56.sure && x56.p[0].x.hi < x56.p[1].x.lo)) return false;
  if (!definitely(dist2(x56.p[0], v.cover_c4.c) <= v.cover_c4.r2)) return false;
  crossing x46 = intersect_circles(c6, v.cover_c4, tp);
  if (!x46.sure) return false;
  UB higher = x46.p[1].y < x46.p[0].y;
  if (!is_certain(higher)) return false;
  double hrect = add_ru(1.0, -x46.p[higher.lo ? 0 : 1].y.lo);
  return can_recurse(K, v.R.lo, c.r[5].hi, add_ru(ub_w, -w6), hrect);
}

// ------------------------------------------------------------------ C9
// l_shaped_recursion.cu:28-120
CC_CHAIN_FN bool l_one_horizontal(const Konst& K, const cell7& c, double ub_w, double lb_weight) {
  double R4 = add_rd(lb_weight, -c.r[2].hi);
  double side3 = sqrt_rd(mul_rd(2.0, c.r[2].lo));
  iv h_top = pt(1.0) - side3;
  double min_top = ub_w < h_top.lo ? ub_w : h_top.lo;
  if (disk_fits(c.r[3].hi, min_top)) {
    double rem_br = add_rd(R4, -add_ru(mul_ru(kRsbRatio, mul_ru(h_top.hi, ub_w)), c.r[3].hi));
    if (rem_br > 0) {
      double w_br = add_ru(ub_w, -side3);
      double m = side3 < w_br ? side3 : w_br;
      if (disk_fits(c.r[4].hi, m)) {
        if (mul_ru(kRsbRatio, mul_ru(side3, w_br)) <= rem_br) return true;
      } else if (can_recurse(K, rem_br, c.r[4].hi, w_br, side3)) {
        return true;
      }
    }
  }
  iv w_right = c.la - side3;
  double min_right = w_right.lo > 1.0 ? 1.0 : w_right.lo;
  if (min_right >= 1.0 || disk_fits(c.r[3].hi, min_right)) {
    double rem_tl = add_rd(R4, -add_ru(mul_ru(kRsbRatio, w_right.hi), c.r[3].hi));
    if (rem_tl > 0) {
      double h_tl = add_ru(1.0, -side3);
      double m = side3 < h_tl ? side3 : h_tl;
      if (disk_fits(c.r[4].hi, m)) {
        if (mul_ru(kRsbRatio, mul_ru(side3, h_tl)) <= rem_tl) return true;
      } else if (can_recurse(K, rem_tl, c.r[4].hi, side3, h_tl)) {
        return true;
      }
    }
  }
  double h3 = add_rd(mul_rd(4.0, c.r[2].lo), -mul_ru(ub_w, ub_w));
  if (h3 > 0.0) {
    h3 = sqrt_rd(h3);
    double hrem = add_ru(1.0, -h3);
    double area = mul_ru(ub_w, hrem);
    double m = ub_w < hrem ? ub_w : hrem;
    if (disk_fits(c.r[3].hi, m)) {
      if (mul_ru(area, kRsbRatio) <= R4) return true;
    } else if (can_recurse(K, R4, c.r[3].hi, ub_w, hrem)) {
      return true;
    }
  }
  return false;
}
// l_shaped_recursion.cu:122-154
CCD bool c9_l_shaped(const Konst& K, const cell7& c) {
  {
    double rem = add_rd(mul_rd(c.la.lo, kRsbRatio), -add_ru(c.r[0].hi, c.r[1].hi));
    double wrem = add_ru(c.la.hi, -trc::two_disk_height(c.r[0].lo, c.r[1].lo, 1.0));
    if (!(mul_rd(wrem, kRsbRatio) > rem) && l_one_horizontal(K, c, wrem, rem)) return true;
  }
  double w12 = trc::two_disk_height(c.r[0].lo, c.r[1].lo, 1.0);
  if (w12 <= 0.0) return false;
  double wrem = add_ru(c.la.hi, -w12);
  double h34 = trc::two_disk_height(c.r[2].lo, c.r[3].lo, wrem);
  if (h34 <= 0.0) return false;
  double R5 = add_rd(mul_rd(c.la.lo, kRsbRatio), -add_ru(c.r[0].hi, add_ru(c.r[1].hi, add_ru(c.r[2].hi, c.r[3].hi))));
  return can_recurse(K, R5, c.r[5].hi, wrem, add_ru(1.0, -h34));
}

// ------------------------------------------------------------------ C10
// r1_in_corner_wall_building_recursion.cu:28-149
CC_CHAIN_FN bool c10_corner_wall(const Konst& K, const cell7& c, Trap& tp) {
  iv w1 = sqrt(2.0 * c.r[0], tp);
  double extra_r1 = mul_rd(add_rd(mul_rd(2.0, kRsbRatio), -1.0), c.r[0].lo);
  iv wB = c.la - w1;
  iv hA = 1.0 - w1;
  iv minA{hA.lo < w1.lo ? hA.lo : w1.lo, hA.hi < w1.hi ? hA.hi : w1.hi};
  iv maxA{hA.lo < w1.lo ? w1.lo : hA.lo, hA.hi < w1.hi ? w1.hi : hA.hi};
  iv minA_sq = square(minA);
  iv maxA_sq = square(maxA);
  iv boundA = kDiskWeightMax * minA_sq;
  double costA = boundA.hi >= c.r[5].hi ? c.r[5].hi : boundA.hi;
  if (boundA.lo <= 0.0 || costA > extra_r1) return false;
  if (possibly(wB < 1.0) && !disk_fits(c.r[1].hi, wB.lo)) {
    const double ninf = -__builtin_inf();
    double e23 = ninf;
    {
      double h23 = trc::two_disk_height(c.r[1].lo, c.r[2].lo, wB.hi);
      double hremB = add_rd(1.0, -h23);
      if (disk_fits(c.r[3].hi, hremB < wB.lo ? hremB : wB.lo))
        e23 = add_rd(mul_rd(kRsbRatio, mul_rd(h23, wB.lo)), -add_ru(c.r[1].hi, c.r[2].hi));
    }
    double e234 = ninf;
    {
      double h234 = three_height(c.r[1].lo, c.r[2].lo, c.r[3].lo, wB.hi);
      double hremB = add_rd(1.0, -h234);
      if (disk_fits(c.r[4].hi, hremB < wB.lo ? hremB : wB.lo))
        e234 = add_rd(mul_rd(kRsbRatio, mul_rd(h234, wB.lo)), -add_ru(c.r[1].hi, add_ru(c.r[2].hi, c.r[3].hi)));
    }
    extra_r1 = add_rd(extra_r1, e23 > e234 ? e23 : e234);
    if (extra_r1 <= 0.0) return false;
  }
  double wb_short = mul_rd(K.weight_bound_wall_building, minA_sq.lo);
  double wb_long = mul_rd(K.weight_bound_wall_building, maxA_sq.lo);
  if (wb_long < c.r[5].hi) return false;
  const bool use_short = wb_short >= c.r[5].hi;
  const double sideA = use_short ? minA.hi : maxA.hi;
  double costA_wb = mul_ru(mul_ru(kRsbRatio, sideA), sqrt_ru(mul_ru(2.0, c.r[5].hi)));
  if (costA_wb > costA) costA = costA_wb;
  if (costA > extra_r1) return false;
  if (c.r[5].hi <= boundA.lo) return true;
  double R6 = mul_rd(kRsbRatio, c.la.lo);
  for (int i = 0; i < 5; ++i) R6 = add_rd(R6, -c.r[i].hi);
  double max_incomplete = mul_ru(sideA, K.width_fraction_incomplete);
  double num = mul_ru(kRsbRatio, mul_ru(max_incomplete, sqrt_ru(mul_ru(2.0, c.r[5].hi))));
  double kept_inf = add_rd(R6, -div_ru(num, K.lb_w_infty_denom));
  double needA = mul_ru(kRsbRatio, mul_ru(w1.hi, hA.hi));
  if (kept_inf >= needA) return true;
  double dropped = needA;
  double cur = c.r[5].hi;
  for (int it = 0; it < kBisectCap && cur > boundA.lo; ++it) {
    dropped = add_ru(dropped, mul_ru(kRsbRatio, mul_ru(max_incomplete, sqrt_ru(mul_ru(2.0, cur)))));
    cur = mul_ru(cur, K.phase_factor_wall_building.hi);
  }
  return add_rd(R6, -dropped) >= needA;
}

// ------------------------------------------------------------------ C11
// multi_disk_strip.cu:264-317: all 62 nonempty proper disk subsets as one strip.
CC_CHAIN_FN bool c11_advanced_strip(const cell7& c) {
  for (unsigned s = 1; s < 63; ++s) {
    double wgt = 0.0;
    for (int j = 0; j < 6; ++j)
      if (s & (1u << j)) wgt = add_ru(wgt, c.r[j].hi);
    double w = div_ru(wgt, kRsbRatio);
    double wsq = mul_ru(w, w);
    double h = 0.0;
    int largest = -1;
    bool ok = true;
    for (int j = 0; j < 6; ++j) {
      if (!(s & (1u << j))) {
        if (largest < 0) largest = j;
        continue;
      }
      double hc = add_rd(mul_rd(4.0, c.r[j].lo), -wsq);
      if (hc <= 0.0) {
        ok = false;
        break;
      }
      h = add_rd(h, sqrt_rd(hc));
    }
    if (!ok || h < 1.0) continue;
    double rem = add_rd(c.la.lo, -w);
    if (rem >= 1.0 || disk_fits(c.r[largest].hi, rem)) return true;
  }
  return false;
}

// ------------------------------------------------------------------ C12
// five_disk_full_cover.cu:91-177
CC_CHAIN_FN bool five_3v2_center(const Konst& K, const cell7& c, const vals7& v, Trap& tp) {
  const double la = c.la.hi;
  iv R6 = c.r[5] + v.R;
  strip2 bot = max_height_strip(la, c.r[0].lo, c.r[1].lo, tp);
  if (bot.h <= 0.0) return false;
  const double r3 = c.r[2].lo, r4 = c.r[3].lo, r5 = c.r[4].lo;
  double hrem = add_ru(1.0, -bot.h);
  double hrem_sq = mul_ru(hrem, hrem);
  double w4 = add_rd(mul_rd(4.0, r4), -hrem_sq);
  double w5 = add_rd(mul_rd(4.0, r5), -hrem_sq);
  if (w4 <= 0.0 || w5 <= 0.0) return false;
  w4 = sqrt_rd(w4);
  w5 = sqrt_rd(w5);
  iv y45 = 1.0 - pt(0.5) * hrem;
  circle c4{{la - pt(0.5) * w4, y45}, pt(r4)};
  circle c5{{pt(0.5) * w5, y45}, pt(r5)};
  double wrem = add_ru(la, -add_rd(w4, w5));
  double h3 = add_rd(mul_rd(4.0, r3), -mul_ru(wrem, wrem));
  if (h3 <= 0.0) return false;
  h3 = sqrt_rd(h3);
  if (h3 >= hrem) return true;
  circle c3{{w5 + pt(0.5) * wrem, 1.0 - pt(0.5) * h3}, pt(r3)};
  crossing x15 = intersect_circles(bot.c1, c5, tp);
  crossing x12 = intersect_circles(bot.c1, bot.c2, tp);
  crossing x24 = intersect_circles(bot.c2, c4, tp);
  if (!x12.sure || !x15.sure || !x24.sure) {
    tp.flag = 1;
    return false;
  }
  if (x15.p[0].x.hi >= x15.p[1].x.lo || x24.p[0].x.hi >= x24.p[1].x.lo) return false;
  UB higher = x12.p[1].y < x12.p[0].y;
  if (!is_certain(higher)) return false;
  const point top12 = x12.p[higher.lo ? 0 : 1];
  bool in12 = definitely(dist2(top12, c3.c) <= c3.r2);
  bool in15 = definitely(dist2(x15.p[1], c3.c) <= c3.r2);
  bool in24 = definitely(dist2(x24.p[0], c3.c) <= c3.r2);
  if (in12 && in15 && in24) return true;
  if (!(in15 && in24)) return false;
  crossing x31 = intersect_circles(bot.c1, c3, tp);
  crossing x32 = intersect_circles(bot.c2, c3, tp);
  if (!(x31.sure && x32.sure && definitely(x31.p[0].x < x31.p[1].x) && definitely(x32.p[0].x < x32.p[1].x)))
    return false;
  double pw = (x32.p[0].x - x31.p[1].x).hi;
  double y31 = x31.p[1].y.hi, y32 = x32.p[0].y.hi;
  double ph = add_ru(y31 > y32 ? y31 : y32, -top12.y.lo);
  return can_recurse(K, R6.lo, c.r[5].hi, pw, ph);
}
// five_disk_full_cover.cu:179-203
CCD bool five_border32(const Konst& K, const cell7& c, int t1, int t2, int t3, int a, int b) {
  double htop = three_height(c.r[t1].lo, c.r[t2].lo, c.r[t3].lo, c.la.hi);
  if (htop <= 0.0) return false;
  double hrem = add_ru(1.0, -htop);
  double wright = trc::two_disk_height(c.r[a].lo, c.r[b].lo, hrem);
  if (wright <= 0.0) return false;
  double wrem = add_ru(c.la.hi, -wright);
  double R6 = mul_rd(c.la.lo, kRsbRatio);
  for (int i = 0; i < 5; ++i) R6 = add_rd(R6, -c.r[i].hi);
  return can_recurse(K, R6, c.r[5].hi, wrem, hrem);
}
CCD bool c12_five_disk(const Konst& K, const cell7& c, const vals7& v, Trap& tp) {
  return five_3v2_center(K, c, v, tp) || five_border32(K, c, 0, 1, 2, 3, 4) || five_border32(K, c, 0, 1, 4, 2, 3);
}

// ------------------------------------------------------------------ C13
// six_disk_full_cover.cu:29-108
CC_CHAIN_FN bool six_4_plus_2(const cell7& c, Trap& tp) {
  const double la = c.la.hi;
  strip2 bot = max_height_strip(la, c.r[0].lo, c.r[1].lo, tp);
  if (bot.h <= 0.0) return false;
  double hrem = add_ru(1.0, -bot.h);
  double q = mul_ru(hrem, hrem);
  iv w4 = pt(c.r[3].lo) - pt(0.25) * q;
  iv w3 = pt(c.r[2].lo) - pt(0.25) * q;
  if (possibly(w4 <= 0.0) || possibly(w3 <= 0.0)) return false;
  w4 = sqrt(w4, tp);
  w3 = sqrt(w3, tp);
  iv y = 1.0 - pt(0.5) * hrem;
  circle c4{{w4, y}, pt(c.r[3].lo)};
  circle c3{{la - w3, y}, pt(c.r[2].lo)};
  crossing x34 = intersect_circles(c3, c4, tp);
  if (!x34.sure) return false;
  const bool p1_up = x34.p[0].y.lo < x34.p[1].y.lo;
  const point up_top = p1_up ? x34.p[1] : x34.p[0];
  const point low_top = p1_up ? x34.p[0] : x34.p[1];
  crossing x12 = intersect_circles(bot.c1, bot.c2, tp);
  if (!x12.sure) {
    tp.flag = 1;
    return false;
  }
  const point up_bot = x12.p[0].y.lo < x12.p[1].y.lo ? x12.p[1] : x12.p[0];
  crossing x14 = intersect_circles(bot.c1, c4, tp);
  crossing x23 = intersect_circles(bot.c2, c3, tp);
  if (!x14.sure || !x23.sure) return false;
  point top_c{up_top.x, 0.5 * (1.0 + up_top.y)};
  iv t1 = dist2(top_c, point{2.0 * w4, pt(1.0)});
  iv t2 = dist2(top_c, point{la - 2.0 * w3, pt(1.0)});
  iv t3 = dist2(top_c, up_top);
  point bc = midpoint(low_top, up_bot);
  iv b1 = dist2(bc, x14.p[1]);
  iv b2 = dist2(bc, x23.p[0]);
  iv b3 = dist2(bc, low_top);
  iv b4 = dist2(bc, up_bot);
  const double r5 = c.r[4].lo, r6 = c.r[5].lo;
  if (t1.hi <= r6 && t2.hi <= r6 && t3.hi <= r6) return b1.hi <= r5 && b2.hi <= r5 && b3.hi <= r5 && b4.hi <= r5;
  if (t1.hi <= r5 && t2.hi <= r5 && t3.hi <= r5) return b1.hi <= r6 && b2.hi <= r6 && b3.hi <= r6 && b4.hi <= r6;
  return false;
}
// six_disk_full_cover.cu:110-167
CC_CHAIN_FN bool six_2v3_r6_center(const cell7& c, const vals7& v, Trap& tp) {
  strip2 bot = max_height_strip_wc(c.r[0].lo, c.r[1].lo, v.R.lo, c.la.hi);
  if (bot.h <= 0.0) return false;
  double hrem = add_ru(1.0, -bot.h);
  double q = mul_ru(hrem, hrem);
  double w3 = add_rd(mul_rd(4.0, c.r[2].lo), -q);
  double w4 = add_rd(mul_rd(4.0, c.r[3].lo), -q);
  if (w3 <= 0 || w4 <= 0) return false;
  w3 = sqrt_rd(w3);
  w4 = sqrt_rd(w4);
  double wrem = add_ru(c.la.hi, -add_rd(w3, w4));
  if (wrem <= 0.0) return true;
  const double r5 = c.r[4].lo;
  double dy5 = add_rd(r5, -mul_ru(0.25, mul_ru(wrem, wrem)));
  if (dy5 <= 0.0) return false;
  dy5 = sqrt_rd(dy5);
  if (add_rd(mul_rd(2.0, dy5), bot.h) >= 1.0) return true;
  circle c5{{w4 + pt(0.5) * wrem, 1.0 - pt(dy5)}, pt(r5)};
  crossing x15 = intersect_circles(bot.c1, c5, tp);
  crossing x25 = intersect_circles(bot.c2, c5, tp);
  if (!x15.sure || !x25.sure) return false;
  const point p15 = x15.p[1], p25 = x25.p[0];
  if (p15.y.hi <= bot.h && p25.y.hi <= bot.h) return true;
  double w6 = add_ru(p25.x.hi, -p15.x.lo);
  double ymax = p15.y.hi < p25.y.hi ? p25.y.hi : p15.y.hi;
  double h6 = add_ru(ymax, -bot.h);
  return add_ru(mul_ru(w6, w6), mul_ru(h6, h6)) <= mul_rd(4.0, c.r[5].lo);
}
// six_disk_full_cover.cu:169-193
CCD bool six_r12_r456_r3(const Konst& K, const cell7& c, const vals7& v) {
  double w12 = trc::two_disk_height(c.r[0].lo, c.r[1].lo, 1.0);
  if (w12 <= 0.0) return false;
  double w456 = three_height(c.r[3].lo, c.r[4].lo, c.r[5].lo, 1.0);
  if (w456 <= 0.0) return false;
  double wrem = add_ru(c.la.hi, -add_rd(w12, w456));
  if (wrem <= 0.0) return true;
  double h3 = add_rd(mul_rd(4.0, c.r[2].lo), -mul_ru(wrem, wrem));
  if (h3 <= 0.0) return false;
  return can_recurse(K, v.R.lo, c.r[5].hi, wrem, add_ru(1.0, -sqrt_rd(h3)));
}
CCD bool c13_six_disk(const Konst& K, const cell7& c, const vals7& v, Trap& tp) {
  return six_r12_r456_r3(K, c, v) || six_4_plus_2(c, tp) || six_2v3_r6_center(c, v, tp);
}

}  // namespace rsbc
}  // namespace ccd

#include "rsb_c14.hpp"
