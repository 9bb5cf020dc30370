// FP32 pre-filter of the ten-criterion box set: an FP32 "shadow" of the FP64
// computation.  Every double the FP64 criterion computes is enclosed by an
// FP32 interval [lo, hi] (struct sf), and every FP64 directed operation maps
// to the outward FP32 interval operation on the enclosures: RD64(x + y) and
// RU64(x + y) both lie in [RD32(x.lo + y.lo), RU32(x.hi + y.hi)], since
// directed roundings are monotone and every FP32 number is an FP64 number.
// A branch of the FP64 code is followed only when the enclosures decide it;
// otherwise the shadow gives up.  So when the shadow ends with a certainly
// true verdict, the FP64 computation takes the same branches and proves the
// box too: the pre-filter accepts nothing the FP64 path rejects.
//
// Kernels run with MODE.FP_ROUND = toward -infinity for FP32 as well, so an
// add/mul is one VALU instruction in that direction and the upward result
// comes from RU(x op y) = -RD((-x) op (-y)).  Division and sqrt are computed
// in FP64 with the directed FP64 primitives and converted with one more
// directed rounding (RD_f32(RD_f64(q)) = RD_f32(q)).
#pragma once
#include "tr.hpp"

namespace ccd {
namespace f32 {

CCD float add_rd(float a, float b) {
  float r;
  asm("v_add_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
CCD float add_ru(float a, float b) {
  float r;
  asm("v_add_f32 %0, -%1, -%2" : "=v"(r) : "v"(a), "v"(b));
  return -r;
}
CCD float mul_rd(float a, float b) {
  float r;
  asm("v_mul_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
CCD float mul_ru(float a, float b) {
  float r;
  asm("v_mul_f32 %0, -%1, %2" : "=v"(r) : "v"(a), "v"(b));
  return -r;
}
CCD float from_rd(double x) {
  float r;
  asm("v_cvt_f32_f64 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
CCD float from_ru(double x) {
  float r;
  asm("v_cvt_f32_f64 %0, -%1" : "=v"(r) : "v"(x));
  return -r;
}
CCD float div_rd(float a, float b) { return from_rd(ccd::div_rd((double)a, (double)b)); }
CCD float div_ru(float a, float b) { return from_ru(ccd::div_ru((double)a, (double)b)); }
CCD float rcp_ru(float a) { return from_ru(ccd::rcp_ru((double)a)); }
CCD float rcp_rd(float a) { return from_rd(ccd::rcp_rd((double)a)); }
CCD float sqrt_rd(float a) { return from_rd(ccd::sqrt_rd((double)a)); }
CCD float sqrt_ru(float a) { return from_ru(ccd::sqrt_ru((double)a)); }

// ------------------------------------------------------------- shadows
struct sf {  // an FP64 value known to lie in [lo, hi]
  float lo, hi;
};
// State of one shadow evaluation: ok turns false when a branch of the FP64
// code cannot be decided from the enclosures.
struct Sh {
  bool ok;
};

CCD sf shd(double x) { return sf{from_rd(x), from_ru(x)}; }
CCD sf sneg(sf a) { return sf{-a.hi, -a.lo}; }
CCD sf sadd(sf a, sf b) { return sf{add_rd(a.lo, b.lo), add_ru(a.hi, b.hi)}; }
CCD sf ssub(sf a, sf b) { return sadd(a, sneg(b)); }
CCD sf smul(sf a, sf b) {
  float l1 = mul_rd(a.lo, b.lo), l2 = mul_rd(a.lo, b.hi), l3 = mul_rd(a.hi, b.lo), l4 = mul_rd(a.hi, b.hi);
  float u1 = mul_ru(a.lo, b.lo), u2 = mul_ru(a.lo, b.hi), u3 = mul_ru(a.hi, b.lo), u4 = mul_ru(a.hi, b.hi);
  float m1 = l1 < l2 ? l1 : l2, m2 = l3 < l4 ? l3 : l4, x1 = u1 > u2 ? u1 : u2, x2 = u3 > u4 ? u3 : u4;
  return sf{m1 < m2 ? m1 : m2, x1 > x2 ? x1 : x2};
}
static __device__ __noinline__ sf sdiv(sf a, sf b, Sh& st) {
  if (!(b.lo > 0.f || b.hi < 0.f)) {  // the FP64 divisor's sign is not known
    st.ok = false;
    return sf{0.f, 0.f};
  }
  float q1 = div_rd(a.lo, b.lo), q2 = div_rd(a.lo, b.hi), q3 = div_rd(a.hi, b.lo), q4 = div_rd(a.hi, b.hi);
  float p1 = div_ru(a.lo, b.lo), p2 = div_ru(a.lo, b.hi), p3 = div_ru(a.hi, b.lo), p4 = div_ru(a.hi, b.hi);
  float m1 = q1 < q2 ? q1 : q2, m2 = q3 < q4 ? q3 : q4, x1 = p1 > p2 ? p1 : p2, x2 = p3 > p4 ? p3 : p4;
  return sf{m1 < m2 ? m1 : m2, x1 > x2 ? x1 : x2};
}
CCD sf ssqrt(sf a, Sh& st) {
  if (!(a.lo >= 0.f)) {  // the FP64 argument may be negative (a NaN there)
    st.ok = false;
    return sf{0.f, 0.f};
  }
  return sf{sqrt_rd(a.lo), sqrt_ru(a.hi)};
}
CCD sf smin(sf a, sf b) { return sf{a.lo < b.lo ? a.lo : b.lo, a.hi < b.hi ? a.hi : b.hi}; }
CCD sf smax(sf a, sf b) { return sf{a.lo > b.lo ? a.lo : b.lo, a.hi > b.hi ? a.hi : b.hi}; }
// Comparisons of FP64 values: decided from the enclosures or give up.
CCD bool slt(sf a, sf b, Sh& st) {
  if (a.hi < b.lo) return true;
  if (a.lo >= b.hi) return false;
  st.ok = false;
  return false;
}
CCD bool sle(sf a, sf b, Sh& st) {
  if (a.hi <= b.lo) return true;
  if (a.lo > b.hi) return false;
  st.ok = false;
  return false;
}
CCD bool sgt(sf a, sf b, Sh& st) { return slt(b, a, st); }
CCD bool sge(sf a, sf b, Sh& st) { return sle(b, a, st); }

struct siv {  // an FP64 interval: shadows of its two bounds
  sf lo, hi;
};
CCD siv sv(iv x) { return siv{shd(x.lo), shd(x.hi)}; }
CCD siv spt(double v) { return siv{shd(v), shd(v)}; }
CCD siv operator+(siv a, siv b) { return siv{sadd(a.lo, b.lo), sadd(a.hi, b.hi)}; }
CCD siv operator-(siv a, siv b) { return siv{ssub(a.lo, b.hi), ssub(a.hi, b.lo)}; }
CCD siv operator+(siv a, double b) { return a + spt(b); }
CCD siv operator*(siv a, siv b) {  // the non-branching product: min/max of corners
  sf c1 = smul(a.lo, b.lo), c2 = smul(a.lo, b.hi), c3 = smul(a.hi, b.lo), c4 = smul(a.hi, b.hi);
  return siv{smin(smin(c1, c2), smin(c3, c4)), smax(smax(c1, c2), smax(c3, c4))};
}
CCD siv scale(double n, siv a) {  // n >= 0 constant: both bounds rounded down in FP64
  const sf k = shd(n);
  return siv{smul(a.lo, k), smul(a.hi, k)};
}
CCD siv div(siv a, siv b, Sh& st) {  // sign cases of the FP64 interval division
  if (sgt(b.lo, shd(0.0), st)) {
    if (sge(a.lo, shd(0.0), st)) return siv{sdiv(a.lo, b.hi, st), sdiv(a.hi, b.lo, st)};
    if (!st.ok) return a;
    if (sle(a.hi, shd(0.0), st)) return siv{sdiv(a.lo, b.lo, st), sdiv(a.hi, b.hi, st)};
    return siv{sdiv(a.lo, b.lo, st), sdiv(a.hi, b.lo, st)};
  }
  st.ok = false;  // negative or zero-straddling divisors do not occur on this path
  return a;
}
CCD siv recip(siv x, Sh& st) {
  if (sgt(x.lo, shd(0.0), st)) return siv{sdiv(shd(1.0), x.hi, st), sdiv(shd(1.0), x.lo, st)};
  st.ok = false;
  return x;
}
CCD siv square_pos(siv x, Sh& st) {  // square() of an interval with x.lo >= 0
  if (!sge(x.lo, shd(0.0), st)) st.ok = false;
  return siv{smul(x.lo, x.lo), smul(x.hi, x.hi)};
}

struct KonstF {
  sf be_lo, be_hi;
};
CCD KonstF konst(const Konst& K) { return KonstF{shd(K.be_lo), shd(K.be_hi)}; }

// critical_ratio(IV la) of tight_rectangle/critical_ratio.cu:58-78, shadowed
// (la > 0 on this path).
static __device__ __noinline__ siv crit_ratio(const KonstF& K, siv la, Sh& st) {
  siv a{}, b{};
  const bool lo_side = sle(la.hi, K.be_lo, st);
  const bool hi_side = !lo_side && sge(la.lo, K.be_hi, st);
  if (!st.ok) return la;
  if (!hi_side) {
    const siv sq = square_pos(la, st);
    const siv inner = scale(16.0, sq) + 40.0 + div(spt(9.0), sq, st);
    a = div(spt(3.0 / 256.0), la, st) * inner;
    if (lo_side) return a;
  }
  b = scale(0.25, la) + scale(0.5, recip(la, st));
  if (hi_side) return b;
  return siv{smin(a.lo, b.lo), smax(a.hi, b.hi)};
}
// critical_ratio(IV w, IV h): critical_ratio.cu:32-52
static __device__ __noinline__ siv crit_ratio(const KonstF& K, siv w, siv h, Sh& st) {
  if (sle(w.hi, h.lo, st)) return crit_ratio(K, div(h, w, st), st);
  if (!st.ok) return w;
  if (sle(h.hi, w.lo, st)) return crit_ratio(K, div(w, h, st), st);
  if (!st.ok) return w;
  siv hw = div(h, w, st), wh = div(w, h, st);
  hw.lo = smax(hw.lo, shd(1.0));  // tighten_lb(1.0)
  wh.lo = smax(wh.lo, shd(1.0));
  const siv c1 = crit_ratio(K, hw, st), c2 = crit_ratio(K, wh, st);
  return siv{smin(c1.lo, c2.lo), smax(c1.hi, c2.hi)};
}

CCD bool disk_fits(sf r_sq_ub, sf h_lb, Sh& st) {  // disk_satisfies_size_bound.cu:25-28
  const sf s = sdiv(shd(1.0), h_lb, st);
  return sle(smul(r_sq_ub, smul(s, s)), shd(kDiskWeightMax), st);
}

// can_recurse(ub_w, ub_h, ub_largest, lb_rem): tight_rectangle/can_recurse.cu:32-62
static __device__ __noinline__ bool can_recurse(const KonstF& K, sf ub_w, sf ub_h, sf ub_largest, sf lb_rem,
                                               Sh& st) {
  sf cr = crit_ratio(K, siv{ub_w, ub_w}, siv{ub_h, ub_h}, st).hi;
  const sf area = smul(ub_w, ub_h);
  const sf side = smin(ub_w, ub_h);
  const sf scl = sdiv(shd(1.0), side, st);
  const sf norm = smul(smul(scl, scl), ub_largest);
  sf cor = smul(shd(0.5), ssqrt(sadd(ssqrt(sadd(smul(norm, norm), shd(1.0)), st), shd(1.0)), st));
  cor = smax(cor, shd(kWorstRatio));
  cr = smin(cr, cor);
  if (sle(norm, shd(kDiskWeightMax), st)) cr = shd(kRsbRatio);
  return sge(lb_rem, smul(cr, area), st);
}

// r1_strategies (tight_rectangle/r1_strategies.cu:27-87) shadowed.  Returns
// true only when the FP64 computation certainly returns true.
CCD bool r1_strategies(const KonstF& K, siv la, siv W, siv r1, sf ub_r2, Sh& st) {
  const sf sq_w = ssqrt(smul(shd(2.0), r1.lo), st);
  sf strip_w = ssub(smul(shd(4.0), r1.lo), shd(1.0));
  const sf rem = (W - r1).lo;
  if (!st.ok) return false;
  if (sgt(strip_w, shd(0.0), st)) {
    strip_w = ssqrt(strip_w, st);
    if (can_recurse(K, ssub(la.hi, strip_w), shd(1.0), ub_r2, rem, st) && st.ok) return true;
  }
  if (!st.ok) return false;
  if (slt(sq_w, shd(1.0), st)) {
    const sf wright = ssub(la.hi, sq_w);
    const sf minright = smin(wright, shd(1.0));
    sf cost;
    const bool fits = disk_fits(ub_r2, minright, st);
    if (!st.ok) return false;
    if (fits) {
      cost = smul(wright, shd(kRsbRatio));
    } else {
      const siv wr{wright, wright};
      cost = (crit_ratio(K, wr, spt(1.0), st) * wr * spt(1.0)).hi;
    }
    cost = sadd(cost, ub_r2);
    const sf rem_top = ssub(rem, cost);
    if (sgt(rem_top, shd(0.0), st) && can_recurse(K, sq_w, ssub(shd(1.0), sq_w), ub_r2, rem_top, st) && st.ok)
      return true;
  }
  if (!st.ok) return false;
  sf s1 = ssub(smul(shd(4.0), r1.lo), shd(1.0));
  if (sgt(s1, shd(0.0), st)) {
    s1 = ssqrt(s1, st);
    const sf t = ssub(la.hi, smul(shd(0.5), s1));
    sf h1 = ssub(r1.lo, smul(t, t));
    if (sgt(h1, shd(0.0), st)) {
      h1 = smul(shd(2.0), ssqrt(h1, st));
      const sf ph = smul(shd(0.5), ssub(shd(1.0), h1));
      const sf pw = ssub(la.hi, s1);
      const sf ru = smul(shd(0.5), ssub(rem, ub_r2));
      if (sgt(ru, shd(0.0), st) && can_recurse(K, pw, ph, ub_r2, ru, st) && st.ok) return true;
    }
  }
  return false;
}

// two_disks_maximize_height(r1, r2, 1) (maximize_covered.cu:106-130)
// shadowed; the FP64 value is 0 on its early exits (returned exactly).
CCD sf two_disk_height_unit(sf r1, sf r2, Sh& st) {
  const sf d = ssub(r1, r2);  // encloses both RD(r1 - r2) and RU(r1 - r2)
  const sf dsq = smul(d, d);
  sf w1 = sadd(shd(0.5), smul(shd(2.0), d));
  sf w2 = ssub(shd(0.5), smul(shd(2.0), d));
  w1 = smul(w1, w1);
  w2 = smul(w2, w2);
  if (sle(r1, smul(shd(0.25), w1), st) || !st.ok) return shd(0.0);
  if (sle(r2, smul(shd(0.25), w2), st) || !st.ok) return shd(0.0);
  const sf h = ssub(smul(shd(2.0), sadd(r1, r2)), sadd(shd(0.25), smul(shd(4.0), dsq)));
  if (sle(h, shd(0.0), st) || !st.ok) return shd(0.0);
  return ssqrt(h, st);
}

// required_weight_for(IV(w, w), IV(h, h)).ub (critical_ratio.cu:80-82)
CCD sf weight_for_pt_hi(const KonstF& K, sf w, sf h, Sh& st) {
  const siv sw{w, w}, sh{h, h};
  return (crit_ratio(K, sw, sh, st) * sw * sh).hi;
}

// manually_proved_two_pocket (r1_r2_strategies.cu:29-44).  Its sqrt argument
// r1 - 1/4 is positive once r1.lb >= 0.49, so the FP64 form never traps here.
CCD bool two_pocket(siv la, siv r1, siv r2, Sh& st) {
  const bool low = slt(r2.lo, shd(0.2400000000000000188), st) || slt(r1.lo, shd(0.4900000000000000466), st);
  if (low || !st.ok) return false;
  if (!sgt(la.lo, shd(0.0), st)) {  // square(la) = [la.lb^2, la.ub^2] needs la > 0
    st.ok = false;
    return false;
  }
  const sf r1req = ssub(smul(shd(0.25), sadd(smul(la.hi, la.hi), shd(1.0))), shd(0.00999999999999999847));
  if (slt(r1.lo, r1req, st) || !st.ok) return false;
  const sf s1_lo = smul(shd(2.0), ssqrt(ssub(r1.lo, shd(0.25)), st));
  const sf pw_hi = ssub(la.hi, s1_lo);  // tighten_lb(0) leaves the upper bound
  if (!sgt(pw_hi, shd(0.0), st)) {      // square(pw).ub = pw.ub^2 needs pw.ub > 0
    st.ok = false;
    return false;
  }
  const sf r2req = ssub(sadd(smul(pw_hi, pw_hi), shd(0.25)), shd(0.00999999999999999847));
  return sge(r2.lo, r2req, st) && st.ok;
}

// r1_r2_strategies_simple (r1_r2_strategies.cu:46-72)
CCD bool r1_r2_simple(const KonstF& K, siv la, siv W, siv r1, siv r2, sf ub_r3, Sh& st) {
  const bool pocket = two_pocket(la, r1, r2, st);
  if (!st.ok) return false;
  if (pocket) return true;
  sf width = two_disk_height_unit(r1.lo, r2.lo, st);
  if (!st.ok || sle(width, shd(0.0), st) || !st.ok) return false;
  sf w2 = ssub(smul(shd(4.0), r2.lo), shd(1.0));
  if (sgt(w2, shd(0.0), st)) {
    w2 = ssqrt(w2, st);
    const sf w = sadd(ssqrt(ssub(smul(shd(4.0), r1.lo), shd(1.0)), st), w2);
    if (sgt(w, width, st)) {
      width = w;
      if (sgt(w, la.hi, st) && st.ok) return true;
    }
  }
  if (!st.ok) return false;
  // the reference adds the covered width here (r1_r2_strategies.cu:69)
  return can_recurse(K, sadd(la.hi, width), shd(1.0), ub_r3, (W - r1 - r2).lo, st) && st.ok;
}

// r1_r2_strategies_stacked (r1_r2_strategies.cu:74-130)
CCD bool r1_r2_stacked(const KonstF& K, siv la, siv W, siv r1, siv r2, sf ub_r3, Sh& st) {
  const sf side = ssqrt(smul(shd(2.0), r1.lo), st);
  if (!st.ok || sgt(side, shd(1.0), st) || !st.ok) return false;
  const sf top_h = ssub(shd(1.0), side);
  sf w2 = ssub(smul(shd(4.0), r2.lo), smul(top_h, top_h));
  if (slt(w2, shd(0.0), st) || !st.ok) return false;
  w2 = ssqrt(w2, st);
  const sf w_top = ssub(la.hi, w2), w_bot = ssub(la.hi, side);
  const sf top_min = smin(w_top, top_h), bot_min = smin(w_bot, side);
  const sf R3 = (W - r1 - r2).lo;
  sf need, cw = w_bot, ch = side;
  if (disk_fits(ub_r3, top_min, st)) {
    need = sadd(smul(shd(kRsbRatio), smul(w_top, top_h)), ub_r3);
  } else if (st.ok && disk_fits(ub_r3, bot_min, st)) {
    need = sadd(smul(shd(kRsbRatio), smul(w_bot, side)), ub_r3);
    cw = w_top;
    ch = top_h;
  } else {
    need = sadd(weight_for_pt_hi(K, w_top, top_h, st), ub_r3);
  }
  if (!st.ok) return false;
  const sf left = ssub(R3, need);
  return sgt(left, shd(0.0), st) && can_recurse(K, cw, ch, ub_r3, left, st) && st.ok;
}

// r1_r2_r3_strategy_horizontal (r1_r2_r3_strategies.cu:28-46)
CCD bool r123_horizontal(const KonstF& K, siv la, siv W, siv r1, siv r2, siv r3, sf ub_r4, Sh& st) {
  sf w3 = ssub(smul(shd(4.0), r3.lo), shd(1.0));
  if (sle(w3, shd(0.0), st) || !st.ok) return false;
  w3 = ssqrt(w3, st);
  const sf w2 = ssqrt(ssub(smul(shd(4.0), r2.lo), shd(1.0)), st);
  const sf w1 = ssqrt(ssub(smul(shd(4.0), r1.lo), shd(1.0)), st);
  const sf cov = sadd(w1, sadd(w2, w3));
  const sf wrem = ssub(la.hi, cov);
  if (!st.ok) return false;
  if (sle(wrem, shd(0.0), st) && st.ok) return true;
  if (!st.ok) return false;
  return can_recurse(K, wrem, shd(1.0), ub_r4, (W - r1 - r2 - r3).lo, st) && st.ok;
}

// ------------------------------------------- rectangle_size_bound shortcuts

CCD sf required_height(sf r_sq_ub, Sh& st) {  // disk_satisfies_size_bound.cu:30-32
  return sdiv(ssqrt(r_sq_ub, st), shd(kDiskRadiusMax), st);
}

// can_recurse (rectangle_size_bound/can_recurse.cu:26-88)
static __device__ __noinline__ bool rsb_can_recurse(const KonstF& K, sf rem, sf largest, sf w, sf h, Sh& st) {
  if (sle(rem, shd(0.0), st)) {
    if (slt(rem, shd(0.0), st)) return false;
    const bool a = sle(w, shd(0.0), st), b = sle(h, shd(0.0), st);
    return (a || b) && st.ok;
  }
  if (!st.ok) return false;
  const bool w0 = sle(w, shd(0.0), st), h0 = sle(h, shd(0.0), st);
  if (!st.ok) return false;
  if (w0 || h0) return true;
  const sf smn = smin(w, h), smx = smax(w, h);  // enclose the FP64 min and max of the sides
  const sf area = sdiv(rem, shd(kRsbRatio), st);
  sf side = sdiv(area, smx, st);
  if (slt(side, smn, st) || !st.ok) return false;
  if (sgt(side, smx, st)) side = ssqrt(area, st);
  if (!st.ok) return false;
  if (sge(smul(shd(kDiskWeightMax), smul(side, side)), largest, st) && st.ok) return true;
  if (!st.ok) return false;
  const sf skew = sdiv(smx, smn, st);
  if (sle(skew, shd(kRsbToleratedLambda), st)) {
    const sf cr = crit_ratio(K, siv{skew, skew}, st).hi;
    return sle(smul(cr, smul(w, h)), rem, st) && st.ok;
  }
  if (!st.ok) return false;
  sf sig = sdiv(largest, smul(smn, smn), st);
  sig = smul(sig, sig);
  const sf cr = smax(smul(shd(0.5), ssqrt(sadd(ssqrt(sadd(sig, shd(1.0)), st), shd(1.0)), st)), shd(kWorstRatio));
  if (sle(smul(cr, smul(w, h)), rem, st) && st.ok) return true;
  if (!st.ok) return false;
  const sf ww = sadd(smul(shd(0.25), smul(smx, smx)), smul(shd(0.5), smul(smn, smn)));
  return sle(ww, rem, st) && st.ok;
}

// S1 shortcut_even_split_recursion (even_split_recursion.cu:42-57)
CCD bool s1_even_split(siv la, siv r1, siv r2, Sh& st) {
  const sf big = smul(shd(0.5), la.lo);
  if (slt(big, shd(1.0), st)) {
    if (!disk_fits(r1.hi, big, st) || !st.ok) return false;
  }
  if (!st.ok) return false;
  const sf small = ssub(big, sdiv(r2.hi, shd(kRsbRatio), st));
  if (sge(small, shd(1.0), st) && st.ok) return true;
  if (!st.ok) return false;
  return disk_fits(r2.hi, small, st) && st.ok;
}

// S2/S3 shortcut_uneven_split_recursion (uneven_split_recursion.cu:78-125)
CCD bool s23_uneven_split(siv la, siv r1, siv rk, Sh& st) {
  if (sle(r1.lo, shd(0.0), st) || !st.ok) return false;
  const sf need = required_height(r1.hi, st);
  if (sge(need, la.lo, st) || !st.ok) return false;
  const sf g1 = sadd(rk.hi, sdiv(need, shd(kRsbRatio), st));
  const sf rem = ssub(smul(la.lo, shd(kRsbRatio)), g1);
  if (sle(rem, shd(0.0), st) || !st.ok) return false;
  const sf rem_w = sdiv(rem, shd(kRsbRatio), st);
  if (sge(rem_w, shd(1.0), st) && st.ok) return true;
  if (!st.ok) return false;
  return disk_fits(rk.hi, rem_w, st) && st.ok;
}

// bound_worst_case_ratio (can_recurse.cu:221-235) for positive w, h
CCD sf worst_case_ratio(siv w, siv h, Sh& st) {
  const sf a = sdiv(w.hi, h.lo, st), b = sdiv(h.hi, w.lo, st);
  const sf skew = smax(a, b);
  const sf r = sadd(smul(shd(0.25), skew), smul(shd(0.5), sdiv(shd(1.0), skew, st)));
  return smax(r, shd(kWorstRatio));
}

// S4 shortcut_r1_r2_opposite_corners_recursion (r1_r2_opposite_corners_recursion.cu:28-70)
CCD bool s4_opposite_corners(siv la, siv r1, siv r2, siv r3, Sh& st) {
  const siv w1{ssqrt(smul(shd(2.0), r1.lo), st), ssqrt(smul(shd(2.0), r1.hi), st)};
  const siv w2{ssqrt(smul(shd(2.0), r2.lo), st), ssqrt(smul(shd(2.0), r2.hi), st)};
  const siv both = w1 + w2;
  if (!st.ok || !slt(both.hi, shd(1.0), st) || !st.ok) return false;
  const siv wA = la - w2, wB = la - w1, hA = spt(1.0) - w1;
  const sf minAB = smin(smin(hA.lo, wA.lo), smin(w1.lo, wB.lo));
  if (!disk_fits(r3.hi, minAB, st) || !st.ok) return false;
  const siv hC = spt(1.0) - both;
  const sf minC = smin(hC.lo, w2.lo);
  const sf extra12 = smul(ssub(smul(shd(2.0), shd(kRsbRatio)), shd(1.0)), sadd(r1.lo, r2.lo));
  const sf two_r3 = smul(shd(2.0), r3.hi);
  if (disk_fits(r3.hi, minC, st)) return !sgt(two_r3, extra12, st) && st.ok;
  if (!st.ok) return false;
  if (sle(w2.lo, shd(0.0), st) || !st.ok || sle(hC.lo, shd(0.0), st) || !st.ok) return false;  // ratio infinite
  const sf before = ssub(extra12, two_r3);
  const sf extra = smul(ssub(worst_case_ratio(w2, hC, st), shd(kRsbRatio)), smul(w2.hi, hC.hi));
  return !slt(before, extra, st) && st.ok;
}

// S5 shortcut_r1_r2_vertical_strip (r1_r2_vertical_strip.cu:28-41)
CCD bool s5_vertical_strip(const KonstF& K, siv la, siv r1, siv r2, siv r3, Sh& st) {
  const sf w = two_disk_height_unit(r1.lo, r2.lo, st);
  if (!st.ok || sle(w, shd(0.0), st) || !st.ok) return false;
  const sf R3 = ssub(smul(la.lo, shd(kRsbRatio)), sadd(r1.hi, r2.hi));
  const siv wrem{ssub(la.lo, w), ssub(la.hi, w)};
  bool strip = sge(wrem.lo, shd(1.0), st);
  if (!st.ok) return false;
  if (!strip) {
    strip = disk_fits(r3.hi, wrem.lo, st);
    if (!st.ok) return false;
  }
  if (strip) return sge(R3, smul(shd(kRsbRatio), wrem.hi), st) && st.ok;
  return rsb_can_recurse(K, R3, r3.hi, wrem.hi, shd(1.0), st);
}

// ------------------------------------------------- the ten-criterion set

// Inputs of the ten criteria in shadow form: the box, required_weight_for(la)
// and ub_r4 = min(r3.ub, (W - r1 - r2 - r3).ub), all computed in FP32.
struct C2In {
  KonstF K;
  siv la, r1, r2, r3, W;
  sf ub_r4;
  bool ok;  // false when required_weight_for(la) could not be shadowed
};
CCD C2In c2_inputs(const Konst& K, iv la, iv r1, iv r2, iv r3) {
  C2In c;
  c.K = konst(K);
  c.la = sv(la);
  c.r1 = sv(r1);
  c.r2 = sv(r2);
  c.r3 = sv(r3);
  Sh st{true};
  c.W = crit_ratio(c.K, c.la, st) * c.la;
  c.ub_r4 = smin(shd(r3.hi), (c.W - c.r1 - c.r2 - c.r3).hi);
  c.ok = st.ok;
  return c;
}

// Criterion k (0..9 except 4, whose width bisection stays in FP64) in shadow
// form: true only when the FP64 criterion certainly proves the box.
CCD bool c2_shadow(int k, const C2In& c) {
  if (!c.ok) return false;
  Sh st{true};
  bool r = false;
  switch (k) {
    case 0: r = r1_strategies(c.K, c.la, c.W, c.r1, c.r2.hi, st); break;
    case 1: r = r1_r2_simple(c.K, c.la, c.W, c.r1, c.r2, c.r3.hi, st); break;
    case 2: r = r1_r2_stacked(c.K, c.la, c.W, c.r1, c.r2, c.r3.hi, st); break;
    case 3: r = r123_horizontal(c.K, c.la, c.W, c.r1, c.r2, c.r3, c.ub_r4, st); break;
    case 5: r = s1_even_split(c.la, c.r1, c.r2, st); break;
    case 6: r = s23_uneven_split(c.la, c.r1, c.r2, st); break;
    case 7: r = s23_uneven_split(c.la, c.r1, c.r3, st); break;
    case 8: r = s4_opposite_corners(c.la, c.r1, c.r2, c.r3, st); break;
    case 9: r = s5_vertical_strip(c.K, c.la, c.r1, c.r2, c.r3, st); break;
    default: break;
  }
  return r && st.ok;
}

// Bit k set: the shadow of criterion k proves the box (bit 4 never set).
CCD uint32_t c2_shadow_mask(const Konst& K, iv la, iv r1, iv r2, iv r3) {
  const C2In c = c2_inputs(K, la, r1, r2, r3);
  uint32_t m = 0;
#pragma unroll 1
  for (int k = 0; k < 10; ++k)
    if (k != 4 && c2_shadow(k, c)) m |= 1u << k;
  return m;
}

}  // namespace f32
}  // namespace ccd
