// FP32 outward-rounded intervals for the pre-filter of the ten-criterion box
// set.  Kernels run with MODE.FP_ROUND = toward -infinity for FP32 as well,
// so an add/mul/fma is one VALU instruction in that direction and the
// upward result comes from RU(x op y) = -RD((-x) op (-y)).  Division and
// sqrt are computed in FP64 with the directed FP64 primitives and converted
// with one more directed rounding: RD_f32(RD_f64(q)) = RD_f32(q) because every
// FP32 number is an FP64 number.  Unlike the FP64 path (which reproduces the
// reference's operation sequence, quirks included) every operation here is
// outward: n*[a, b] with n >= 0 is [RD(n*a), RU(n*b)], and constants that are
// not FP32 numbers are rounded in the safe direction for each use.
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

// 0.61 rounded up / down to FP32, 9/64 and 195/256 are FP32 numbers.
constexpr float kRatioUp = 0.610000014305114746f;   // RU_f32(0.61) > 0.61
constexpr float kRatioDown = 0.609999954700469971f; // RD_f32(0.61) < 0.61
constexpr float kWeightMax = 0.140625f;
constexpr float kWorst = 0.76171875f;

struct ivf {
  float lo, hi;
};
CCD ivf out(iv x) { return ivf{from_rd(x.lo), from_ru(x.hi)}; }
CCD ivf ptf(float v) { return ivf{v, v}; }
CCD ivf operator+(ivf a, ivf b) { return ivf{add_rd(a.lo, b.lo), add_ru(a.hi, b.hi)}; }
CCD ivf operator-(ivf a, ivf b) { return ivf{add_rd(a.lo, -b.hi), add_ru(a.hi, -b.lo)}; }
CCD ivf operator+(ivf a, float b) { return ivf{add_rd(a.lo, b), add_ru(a.hi, b)}; }
CCD ivf operator*(ivf a, ivf b) {
  float l1 = mul_rd(a.lo, b.lo), l2 = mul_rd(a.lo, b.hi), l3 = mul_rd(a.hi, b.lo), l4 = mul_rd(a.hi, b.hi);
  float u1 = mul_ru(a.lo, b.lo), u2 = mul_ru(a.lo, b.hi), u3 = mul_ru(a.hi, b.lo), u4 = mul_ru(a.hi, b.hi);
  float m1 = l1 < l2 ? l1 : l2, m2 = l3 < l4 ? l3 : l4, x1 = u1 > u2 ? u1 : u2, x2 = u3 > u4 ? u3 : u4;
  return ivf{m1 < m2 ? m1 : m2, x1 > x2 ? x1 : x2};
}
CCD ivf scale(float n, ivf a) {  // n >= 0, outward
  return ivf{mul_rd(a.lo, n), mul_ru(a.hi, n)};
}
CCD ivf operator/(ivf a, ivf b) {  // b > 0 or b < 0, else unbounded
  if (b.lo > 0) {
    if (a.lo >= 0) return ivf{div_rd(a.lo, b.hi), div_ru(a.hi, b.lo)};
    if (a.hi <= 0) return ivf{div_rd(a.lo, b.lo), div_ru(a.hi, b.hi)};
    return ivf{div_rd(a.lo, b.lo), div_ru(a.hi, b.lo)};
  }
  return ivf{-__builtin_inff(), __builtin_inff()};
}
CCD ivf recip_pos(ivf x) {  // x > 0
  return x.lo > 0 ? ivf{rcp_rd(x.hi), rcp_ru(x.lo)} : ivf{-__builtin_inff(), __builtin_inff()};
}
CCD ivf sq(ivf x) {
  float l = 0.f, u;
  float al = x.lo < 0 ? -x.lo : x.lo, ah = x.hi < 0 ? -x.hi : x.hi;
  u = al > ah ? al : ah;
  if (!(x.lo <= 0 && x.hi >= 0)) l = al < ah ? al : ah;
  return ivf{mul_rd(l, l), mul_ru(u, u)};
}

// Break-even bounds of critical_ratio as FP32: the lower one rounded down,
// the upper one rounded up, so the one-sided branches stay one-sided.
struct KonstF {
  float be_lo, be_hi;
};
CCD KonstF konst(const Konst& K) { return KonstF{from_rd(K.be_lo), from_ru(K.be_hi)}; }

// critical_ratio (tight_rectangle/critical_ratio.cu:58-78) as an enclosure.
CCD ivf crit_ratio(const KonstF& K, ivf la) {
  ivf a, b;
  const bool lo_side = la.hi <= K.be_lo;
  const bool hi_side = !lo_side && la.lo >= K.be_hi;
  if (!hi_side) {
    ivf s = sq(la);
    ivf inner = scale(16.f, s) + 40.f + ptf(9.f) / s;
    a = (ptf(3.f / 256.f) / la) * inner;
    if (lo_side) return a;
  }
  b = scale(0.25f, la) + scale(0.5f, recip_pos(la));
  if (hi_side) return b;
  return ivf{a.lo < b.lo ? a.lo : b.lo, a.hi < b.hi ? b.hi : a.hi};
}
CCD ivf crit_ratio(const KonstF& K, ivf w, ivf h) {  // critical_ratio.cu:32-52
  if (w.hi <= h.lo) return crit_ratio(K, h / w);
  if (h.hi <= w.lo) return crit_ratio(K, w / h);
  ivf hw = h / w, wh = w / h;
  if (hw.lo < 1.f) hw.lo = 1.f;
  if (wh.lo < 1.f) wh.lo = 1.f;
  ivf c1 = crit_ratio(K, hw), c2 = crit_ratio(K, wh);
  return ivf{c1.lo < c2.lo ? c1.lo : c2.lo, c1.hi < c2.hi ? c2.hi : c1.hi};
}
CCD ivf weight_for(const KonstF& K, ivf la) { return crit_ratio(K, la) * la; }

CCD bool disk_fits(float r_sq_ub, float h_lb) {  // disk_satisfies_size_bound.cu:25-28
  const float s = rcp_ru(h_lb);
  return mul_ru(r_sq_ub, mul_ru(s, s)) <= kWeightMax;
}

// can_recurse(ub_w, ub_h, ub_largest, lb_rem): tight_rectangle/can_recurse.cu:32-62
CCD bool can_recurse(const KonstF& K, float ub_w, float ub_h, float ub_largest, float lb_rem) {
  float cr = crit_ratio(K, ptf(ub_w), ptf(ub_h)).hi;
  const float area = mul_ru(ub_w, ub_h);
  const float side = ub_w < ub_h ? ub_w : ub_h;
  const float scl = rcp_ru(side);
  const float norm = mul_ru(mul_ru(scl, scl), ub_largest);
  float cor = mul_ru(0.5f, sqrt_ru(add_ru(sqrt_ru(add_ru(mul_ru(norm, norm), 1.f)), 1.f)));
  if (cor < kWorst) cor = kWorst;
  if (cr > cor) cr = cor;
  if (norm <= kWeightMax) cr = kRatioUp;
  return lb_rem >= mul_ru(cr, area);
}

// r1_strategies (tight_rectangle/r1_strategies.cu:27-87) as an enclosure.
CCD bool r1_strategies(const KonstF& K, ivf la, ivf W, ivf r1, float ub_r2) {
  float sq_w = sqrt_rd(mul_rd(2.f, r1.lo));
  float strip_w = add_rd(mul_rd(4.f, r1.lo), -1.f);
  const float rem = (W - r1).lo;
  if (strip_w > 0.f) {
    strip_w = sqrt_rd(strip_w);
    if (can_recurse(K, add_ru(la.hi, -strip_w), 1.f, ub_r2, rem)) return true;
  }
  if (sq_w < 1.f) {
    const float wright = add_ru(la.hi, -sq_w);
    const float minright = wright < 1.f ? wright : 1.f;
    float cost = disk_fits(ub_r2, minright) ? mul_ru(wright, kRatioUp) : (crit_ratio(K, ptf(wright), ptf(1.f)) * ptf(wright)).hi;
    cost = add_ru(cost, ub_r2);
    const float rem_top = add_rd(rem, -cost);
    if (rem_top > 0 && can_recurse(K, sq_w, add_ru(1.f, -sq_w), ub_r2, rem_top)) return true;
  }
  float s1 = add_rd(mul_rd(4.f, r1.lo), -1.f);
  if (s1 > 0.f) {
    s1 = sqrt_rd(s1);
    const float t = add_ru(la.hi, -mul_rd(0.5f, s1));
    float h1 = add_rd(r1.lo, -mul_ru(t, t));
    if (h1 > 0.f) {
      h1 = mul_rd(2.f, sqrt_rd(h1));
      const float ph = mul_ru(0.5f, add_ru(1.f, -h1));
      const float pw = add_ru(la.hi, -s1);
      const float ru = mul_rd(0.5f, add_rd(rem, -ub_r2));
      if (ru > 0 && can_recurse(K, pw, ph, ub_r2, ru)) return true;
    }
  }
  return false;
}

}  // namespace f32
}  // namespace ccd
