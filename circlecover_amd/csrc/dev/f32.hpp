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

}  // namespace f32
}  // namespace ccd
