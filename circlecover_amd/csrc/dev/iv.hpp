// FP64 interval arithmetic and circle geometry for the gfx950 prover kernels.
//
// Semantics follow the reference interval type operation by operation
// (Prover/algcuda/cuda-utils/cpp/include/algcuda/interval.hpp:32-357 and
// .../cuh/include/algcuda/interval.cuh:30-355), including its quirks:
// `IV * n` with n >= 0 rounds BOTH bounds down (interval.cuh:186-188), and a
// NaN scalar divisor leaves the bounds undefined (here: NaN).  Every bound is
// produced by the directed primitives of round.hpp.  A sqrt of a possibly
// negative interval, which traps the reference's whole grid, sets a per-lane
// trap flag instead (`Trap`), so the cell can be reported.
#pragma once
#include "round.hpp"

namespace ccd {

struct UB {  // tri-state boolean {definitely, possibly}
  bool lo, hi;
};
CCD bool definitely(UB b) { return b.lo; }
CCD bool possibly(UB b) { return b.hi; }
CCD UB operator!(UB b) { return UB{!b.hi, !b.lo}; }
CCD bool is_certain(UB b) { return b.lo == b.hi; }

struct iv {
  double lo, hi;
};
CCD iv pt(double v) { return iv{v, v}; }
CCD bool empty(iv x) { return x.lo > x.hi; }
CCD iv intersect(iv a, iv b) { return iv{a.lo < b.lo ? b.lo : a.lo, a.hi > b.hi ? b.hi : a.hi}; }
CCD void join_into(iv& a, iv b) {
  if (b.lo < a.lo) a.lo = b.lo;
  if (b.hi > a.hi) a.hi = b.hi;
}
CCD void tighten_hi(iv& a, double v) {
  if (v < a.hi) a.hi = v;
}
CCD void tighten_lo(iv& a, double v) {
  if (a.lo < v) a.lo = v;
}
CCD iv operator-(iv a) { return iv{-a.hi, -a.lo}; }

CCD iv abs(iv x) {
  double l = 0.0, u = 0.0;
  if (x.hi > 0) u = x.hi;
  else l = -x.hi;
  if (x.lo < 0) u = u < -x.lo ? -x.lo : u;
  else l = x.lo;
  return iv{l, u};
}

CCD iv operator+(iv a, iv b) { return iv{add_rd(a.lo, b.lo), add_ru(a.hi, b.hi)}; }
CCD iv operator+(iv a, double b) { return iv{add_rd(a.lo, b), add_ru(a.hi, b)}; }
CCD iv operator+(double a, iv b) { return iv{add_rd(b.lo, a), add_ru(b.hi, a)}; }
CCD iv operator-(iv a, iv b) { return iv{add_rd(a.lo, -b.hi), add_ru(a.hi, -b.lo)}; }
CCD iv operator-(iv a, double b) { return iv{add_rd(a.lo, -b), add_ru(a.hi, -b)}; }
CCD iv operator-(double a, iv b) { return iv{add_rd(a, -b.hi), add_ru(a, -b.lo)}; }
CCD iv& operator+=(iv& a, iv b) { return a = a + b; }
CCD iv& operator-=(iv& a, iv b) { return a = a - b; }

// Non-branching interval product, min/max formed with '<' / '>' selects.
CCD iv operator*(iv a, iv b) {
  double l1 = mul_rd(a.lo, b.lo), l2 = mul_rd(a.lo, b.hi), l3 = mul_rd(a.hi, b.lo), l4 = mul_rd(a.hi, b.hi);
  double u1 = mul_ru(a.lo, b.lo), u2 = mul_ru(a.lo, b.hi), u3 = mul_ru(a.hi, b.lo), u4 = mul_ru(a.hi, b.hi);
  double m1 = l1 < l2 ? l1 : l2, m2 = l3 < l4 ? l3 : l4;
  double x1 = u1 > u2 ? u1 : u2, x2 = u3 > u4 ? u3 : u4;
  return iv{m1 < m2 ? m1 : m2, x1 > x2 ? x1 : x2};
}
CCD iv operator*(iv a, double n) {
  if (n >= 0) return iv{mul_rd(a.lo, n), mul_rd(a.hi, n)};
  return iv{mul_rd(a.hi, n), mul_ru(a.lo, n)};
}
CCD iv operator*(double n, iv a) { return a * n; }

CCD iv operator/(iv a, double n) {
  if (n > 0) return iv{div_rd(a.lo, n), div_ru(a.hi, n)};
  if (n <= 0) return iv{div_rd(a.hi, n), div_ru(a.lo, n)};
  return iv{__builtin_nan(""), __builtin_nan("")};
}
CCD iv operator/(iv a, iv b) {
  if (b.lo > 0) {
    if (a.lo >= 0) return iv{div_rd(a.lo, b.hi), div_ru(a.hi, b.lo)};
    if (a.hi <= 0) return iv{div_rd(a.lo, b.lo), div_ru(a.hi, b.hi)};
    return iv{div_rd(a.lo, b.lo), div_ru(a.hi, b.lo)};
  }
  if (b.hi < 0) {
    if (a.lo >= 0) return iv{div_rd(a.hi, b.hi), div_ru(a.lo, b.lo)};
    if (a.hi <= 0) return iv{div_rd(a.hi, b.lo), div_ru(a.lo, b.hi)};
    return iv{div_rd(a.hi, b.hi), div_ru(a.lo, b.hi)};
  }
  return iv{-__builtin_inf(), __builtin_inf()};
}
CCD iv operator/(double a, iv b) { return pt(a) / b; }

CCD iv reciprocal(iv x) {
  if (x.lo > 0) return iv{rcp_rd(x.hi), rcp_ru(x.lo)};
  if (x.hi < 0) return iv{rcp_rd(x.lo), rcp_ru(x.hi)};
  return iv{-__builtin_inf(), __builtin_inf()};
}
CCD iv square(iv x) {
  iv a = abs(x);
  return iv{mul_rd(a.lo, a.lo), mul_ru(a.hi, a.hi)};
}
// Trap bookkeeping: set when a sqrt argument may be negative.
struct Trap {
  int flag;
};
CCD iv sqrt(iv x, Trap& t) {
  if (x.lo < 0.0) t.flag = 1;
  return iv{sqrt_rd(x.lo), sqrt_ru(x.hi)};
}

CCD UB operator<(iv x, double n) { return UB{x.hi < n, x.lo < n}; }
CCD UB operator>(iv x, double n) { return UB{x.lo > n, x.hi > n}; }
CCD UB operator<=(iv x, double n) { return UB{x.hi <= n, x.lo <= n}; }
CCD UB operator>=(iv x, double n) { return UB{x.lo >= n, x.hi >= n}; }
CCD UB operator<(iv x, iv o) { return UB{x.hi < o.lo, x.lo < o.hi}; }
CCD UB operator<=(iv x, iv o) { return UB{x.hi <= o.lo, x.lo <= o.hi}; }
CCD UB operator>(iv x, iv o) { return o < x; }
CCD UB operator>=(iv x, iv o) { return o <= x; }

// ----------------------------------------------------------------- geometry
struct point {
  iv x, y;
};
struct circle {
  point c;
  iv r2;  // squared radius
};
struct crossing {
  point p[2];
  bool sure;  // definitely intersecting
};

CCD iv dist2(const point& a, const point& b) { return square(a.x - b.x) + square(a.y - b.y); }

// Two-circle intersection (reference Prover/operations.cu:30-83).
CCD crossing intersect_circles(const circle& c1, const circle& c2, Trap& t) {
  crossing out;
  out.sure = false;
  out.p[0] = out.p[1] = point{pt(0.0), pt(0.0)};
  iv d2 = dist2(c1.c, c2.c);
  if (d2.lo == 0) return out;
  iv rd = c1.r2 - c2.r2;
  iv f1 = (2.0 * (c1.r2 + c2.r2)) / d2 - 1.0 - square(rd) / square(d2);
  if (f1.lo < 0) return out;
  f1 = 0.5 * sqrt(f1, t);
  out.sure = true;
  iv f2 = (0.5 * rd) / d2;
  iv dx = c2.c.x - c1.c.x;
  iv dy = c2.c.y - c1.c.y;
  iv cx = 0.5 * (c1.c.x + c2.c.x);
  iv cy = 0.5 * (c1.c.y + c2.c.y);
  cx = cx + f2 * dx;
  cy = cy + f2 * dy;
  point a{cx + f1 * dy, cy - f1 * dx};
  point b{cx - f1 * dy, cy + f1 * dx};
  bool a_first = dy.hi < 0 || (dy.hi == 0 && dx.lo >= 0);
  out.p[0] = a_first ? a : b;
  out.p[1] = a_first ? b : a;
  return out;
}
CCD point midpoint(const point& a, const point& b) { return point{a.x + 0.5 * (b.x - a.x), a.y + 0.5 * (b.y - a.y)}; }

// i-th of n near-equal pieces of x, the last snapped to x.hi
// (reference Prover/subintervals.hpp:32-43, lb + diff*i fused).
CCD iv subinterval(iv x, int i, int n) {
  double l, u;
  rn_piece(x.hi, x.lo, (double)n, (double)i, (double)(i + 1), l, u);
  return iv{l, (i + 1 == n) ? x.hi : u};
}

}  // namespace ccd
