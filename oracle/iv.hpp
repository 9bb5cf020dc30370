// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// This directory is the correctness oracle for the HIP prover: a CPU
// restatement of the reference's FP64 interval arithmetic and covering
// criteria.  Only the test-suite, the smoke check and the CPU-baseline leg of
// the benchmark may load it; the product library never links or calls it.
//
// Directed rounding.  Every lower/upper bound is produced by one IEEE-754
// operation executed with the x86 MXCSR rounding field set to round-down or
// round-up (SSE2 addsd/mulsd/divsd/sqrtsd are correctly rounded in every mode),
// which is by definition what the reference's __dadd_rd/__dadd_ru,
// __dmul_rd/_ru, __ddiv_rd/_ru, __dsqrt_rd/_ru and __drcp_rd/_ru compute
// (reference: Prover/algcuda/cuda-utils/cuh/include/algcuda/interval.cuh:30-355).
// The mode switch, the operation and the switch back sit in ONE asm volatile
// block so the compiler can neither constant-fold nor move the operation out
// of its rounding mode.
//
// Op counting.  Built with -DCC_ORACLE_COUNT every primitive bumps a
// thread-local counter; that build provides the algorithmic FP64-op counts the
// benchmark's roofline uses.
#pragma once
#include <cstdint>
#include <cmath>
#include <cfloat>
#include <limits>

namespace ccor {

#ifdef CC_ORACLE_COUNT
extern thread_local uint64_t g_op_count;
#define CCOR_COUNT() (++::ccor::g_op_count)
#define CCOR_OPS() (::ccor::g_op_count)
#else
#define CCOR_COUNT() ((void)0)
#define CCOR_OPS() (uint64_t{0})
#endif

// MXCSR images: all exceptions masked, FTZ/DAZ off; bits 13-14 select rounding.
static const uint32_t kCsrNearest = 0x1F80u;
static const uint32_t kCsrDown = 0x3F80u;
static const uint32_t kCsrUp = 0x5F80u;

#define CCOR_BINOP(name, insn, csr)                                            \
  static inline double name(double a, double b) {                             \
    CCOR_COUNT();                                                              \
    asm volatile("ldmxcsr %[m]\n\t" insn " %[b], %[a]\n\tldmxcsr %[n]"        \
                 : [a] "+x"(a)                                                 \
                 : [b] "x"(b), [m] "m"(csr), [n] "m"(kCsrNearest));           \
    return a;                                                                  \
  }

CCOR_BINOP(add_rd, "addsd", kCsrDown)
CCOR_BINOP(add_ru, "addsd", kCsrUp)
CCOR_BINOP(mul_rd, "mulsd", kCsrDown)
CCOR_BINOP(mul_ru, "mulsd", kCsrUp)
CCOR_BINOP(div_rd, "divsd", kCsrDown)
CCOR_BINOP(div_ru, "divsd", kCsrUp)
#undef CCOR_BINOP

static inline double sqrt_rd(double a) {
  CCOR_COUNT();
  double r;
  asm volatile("ldmxcsr %[m]\n\tsqrtsd %[a], %[r]\n\tldmxcsr %[n]"
               : [r] "=x"(r)
               : [a] "x"(a), [m] "m"(kCsrDown), [n] "m"(kCsrNearest));
  return r;
}
static inline double sqrt_ru(double a) {
  CCOR_COUNT();
  double r;
  asm volatile("ldmxcsr %[m]\n\tsqrtsd %[a], %[r]\n\tldmxcsr %[n]"
               : [r] "=x"(r)
               : [a] "x"(a), [m] "m"(kCsrUp), [n] "m"(kCsrNearest));
  return r;
}
// __drcp_rd / __drcp_ru: the correctly rounded reciprocal.
static inline double rcp_rd(double a) { return div_rd(1.0, a); }
static inline double rcp_ru(double a) { return div_ru(1.0, a); }

// Plain round-to-nearest operations that the reference writes as ordinary C
// (bisection midpoints, get_subinterval).  Wrapped so that they are counted
// and cannot be contracted into an FMA behind our back.
static inline double rn_add(double a, double b) {
  CCOR_COUNT();
  asm volatile("addsd %[b], %[a]" : [a] "+x"(a) : [b] "x"(b));
  return a;
}
static inline double rn_sub(double a, double b) {
  CCOR_COUNT();
  asm volatile("subsd %[b], %[a]" : [a] "+x"(a) : [b] "x"(b));
  return a;
}
static inline double rn_mul(double a, double b) {
  CCOR_COUNT();
  asm volatile("mulsd %[b], %[a]" : [a] "+x"(a) : [b] "x"(b));
  return a;
}
static inline double rn_div(double a, double b) {
  CCOR_COUNT();
  asm volatile("divsd %[b], %[a]" : [a] "+x"(a) : [b] "x"(b));
  return a;
}
// Fused multiply-add, rounded once to nearest (std::fma is exact-then-round).
static inline double rn_fma(double a, double b, double c) {
  CCOR_COUNT();
  return std::fma(a, b, c);
}
// 0.5*(a+b) as the reference's bisection loops write it.
static inline double rn_mid(double a, double b) { return rn_mul(0.5, rn_add(a, b)); }

static const double kInf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------
// Tri-state boolean (reference interval.hpp:32-82): {definitely, possibly}.
struct UB {
  bool lo, hi;
};
static inline bool definitely(UB b) { return b.lo; }
static inline bool possibly(UB b) { return b.hi; }
static inline bool definitely_not(UB b) { return !b.hi; }
static inline bool possibly_not(UB b) { return !b.lo; }
static inline UB operator!(UB b) { return UB{!b.hi, !b.lo}; }
static inline bool is_certain(UB b) { return b.lo == b.hi; }

// Device traps of the reference (sqrt of a possibly negative interval and the
// geometric "must intersect" checks) abort the whole CUDA grid.  The oracle
// records them in a thread-local flag instead; the caller reports the cell.
extern thread_local int g_trap;

// ---------------------------------------------------------------------------
// FP64 interval with the reference's exact rounding recipe per operation.
struct IV {
  double lb, ub;
  IV() = default;
  constexpr IV(double l, double u) : lb(l), ub(u) {}
  static constexpr IV pt(double v) { return IV(v, v); }

  bool empty() const { return lb > ub; }
  IV intersect(const IV& o) const { return IV(lb < o.lb ? o.lb : lb, ub > o.ub ? o.ub : ub); }
  IV join(const IV& o) const { return IV(lb > o.lb ? o.lb : lb, ub < o.ub ? o.ub : ub); }
  void do_join(const IV& o) {
    if (o.lb < lb) lb = o.lb;
    if (o.ub > ub) ub = o.ub;
  }
  void do_intersect(const IV& o) {
    if (o.lb > lb) lb = o.lb;
    if (o.ub < ub) ub = o.ub;
  }
  void tighten_ub(double v) {
    if (v < ub) ub = v;
  }
  void tighten_lb(double v) {
    if (lb < v) lb = v;
  }
  IV operator-() const { return IV(-ub, -lb); }
};

// abs (interval.hpp:214-232)
static inline IV iv_abs(const IV& x) {
  double l = 0.0, u = 0.0;
  if (x.ub > 0) u = x.ub;
  else l = -x.ub;
  if (x.lb < 0) u = u < -x.lb ? -x.lb : u;
  else l = x.lb;
  return IV(l, u);
}

// + and - (interval.cuh:38-104)
static inline IV operator+(const IV& a, const IV& b) { return IV(add_rd(a.lb, b.lb), add_ru(a.ub, b.ub)); }
static inline IV operator+(const IV& a, double b) { return a + IV::pt(b); }
static inline IV operator+(double a, const IV& b) { return b + IV::pt(a); }
static inline IV operator-(const IV& a, const IV& b) { return IV(add_rd(a.lb, -b.ub), add_ru(a.ub, -b.lb)); }
static inline IV operator-(const IV& a, double b) { return a - IV::pt(b); }
static inline IV operator-(double a, const IV& b) { return IV::pt(a) - b; }
static inline IV& operator+=(IV& a, const IV& b) { return a = a + b; }
static inline IV& operator-=(IV& a, const IV& b) { return a = a - b; }
static inline IV& operator+=(IV& a, double b) { return a = a + b; }
static inline IV& operator-=(IV& a, double b) { return a = a - b; }

// IV*IV: the non-branching form (interval.cuh:171-182), min/max by '<'/'>'.
static inline IV operator*(const IV& a, const IV& b) {
  double l1 = mul_rd(a.lb, b.lb), l2 = mul_rd(a.lb, b.ub), l3 = mul_rd(a.ub, b.lb), l4 = mul_rd(a.ub, b.ub);
  double u1 = mul_ru(a.lb, b.lb), u2 = mul_ru(a.lb, b.ub), u3 = mul_ru(a.ub, b.lb), u4 = mul_ru(a.ub, b.ub);
  double m1 = l1 < l2 ? l1 : l2;
  double m2 = l3 < l4 ? l3 : l4;
  double x1 = u1 > u2 ? u1 : u2;
  double x2 = u3 > u4 ? u3 : u4;
  return IV(m1 < m2 ? m1 : m2, x1 > x2 ? x1 : x2);
}
// IV*scalar (interval.cuh:185-197).  For n >= 0 BOTH bounds round down: the
// reference's upper bound uses the round-down product (kept for parity).
static inline IV operator*(const IV& a, double n) {
  if (n >= 0) return IV(mul_rd(a.lb, n), mul_rd(a.ub, n));
  return IV(mul_rd(a.ub, n), mul_ru(a.lb, n));
}
static inline IV operator*(double n, const IV& a) { return a * n; }
static inline IV& operator*=(IV& a, const IV& b) { return a = a * b; }
static inline IV& operator*=(IV& a, double b) { return a = a * b; }

// IV/scalar (interval.cuh:261-275).  A NaN divisor leaves the reference's
// bounds uninitialised; the oracle yields NaN bounds there.
static inline IV operator/(const IV& a, double n) {
  if (n > 0) return IV(div_rd(a.lb, n), div_ru(a.ub, n));
  if (n <= 0) return IV(div_rd(a.ub, n), div_ru(a.lb, n));
  return IV(std::nan(""), std::nan(""));
}
// IV/IV (interval.cuh:277-310)
static inline IV operator/(const IV& a, const IV& b) {
  if (b.lb > 0) {
    if (a.lb >= 0) return IV(div_rd(a.lb, b.ub), div_ru(a.ub, b.lb));
    if (a.ub <= 0) return IV(div_rd(a.lb, b.lb), div_ru(a.ub, b.ub));
    return IV(div_rd(a.lb, b.lb), div_ru(a.ub, b.lb));
  }
  if (b.ub < 0) {
    if (a.lb >= 0) return IV(div_rd(a.ub, b.ub), div_ru(a.lb, b.lb));
    if (a.ub <= 0) return IV(div_rd(a.ub, b.lb), div_ru(a.lb, b.ub));
    return IV(div_rd(a.ub, b.ub), div_ru(a.lb, b.ub));
  }
  return IV(-kInf, kInf);
}
static inline IV operator/(double a, const IV& b) { return IV::pt(a) / b; }

// reciprocal (interval.cuh:227-235)
static inline IV iv_reciprocal(const IV& x) {
  if (x.lb > 0) return IV(rcp_rd(x.ub), rcp_ru(x.lb));
  if (x.ub < 0) return IV(rcp_rd(x.lb), rcp_ru(x.ub));
  return IV(-kInf, kInf);
}
// sqrt (interval.cuh:339-350): a possibly negative argument traps.
static inline IV iv_sqrt(const IV& x) {
  if (x.lb < 0.0) g_trap = 1;
  return IV(sqrt_rd(x.lb), sqrt_ru(x.ub));
}
// square (interval.cuh:352-355)
static inline IV iv_square(const IV& x) {
  IV a = iv_abs(x);
  return IV(mul_rd(a.lb, a.lb), mul_ru(a.ub, a.ub));
}

// Comparisons (interval.hpp:134-170, 251-289)
static inline UB operator<(const IV& x, double n) { return UB{x.ub < n, x.lb < n}; }
static inline UB operator>(const IV& x, double n) { return UB{x.lb > n, x.ub > n}; }
static inline UB operator<=(const IV& x, double n) { return UB{x.ub <= n, x.lb <= n}; }
static inline UB operator>=(const IV& x, double n) { return UB{x.lb >= n, x.ub >= n}; }
static inline UB operator<(const IV& x, const IV& o) { return UB{x.ub < o.lb, x.lb < o.ub}; }
static inline UB operator<=(const IV& x, const IV& o) { return UB{x.ub <= o.lb, x.lb <= o.ub}; }
static inline UB operator>(const IV& x, const IV& o) { return o < x; }
static inline UB operator>=(const IV& x, const IV& o) { return o <= x; }
static inline UB operator<(double n, const IV& x) { return x > n; }
static inline UB operator>(double n, const IV& x) { return x < n; }
static inline UB operator<=(double n, const IV& x) { return x >= n; }
static inline UB operator>=(double n, const IV& x) { return x <= n; }
static inline UB operator==(const IV& x, const IV& o) {
  return UB{x.lb == x.ub && o.lb == o.ub && x.lb == o.lb, !(x.lb > o.ub) && !(x.ub < o.lb)};
}
static inline UB operator==(const IV& x, double n) { return UB{x.lb == x.ub && x.lb == n, !(x.lb > n) && !(x.ub < n)}; }

// Lexicographic (lb, ub) order used by the host sorts (interval.hpp:353-357).
static inline bool iv_less(const IV& a, const IV& b) { return a.lb < b.lb || (a.lb == b.lb && a.ub < b.ub); }

// ---------------------------------------------------------------------------
// Geometry (reference Prover/operations.hpp:29-44, operations.cu:26-87).
struct Point {
  IV x, y;
};
struct Circle {
  Point center;
  IV squared_radius;
};
struct Intersection {
  Point p[2];
  bool definitely_intersecting;
};

static inline IV squared_distance(const Point& a, const Point& b) {
  return iv_square(a.x - b.x) + iv_square(a.y - b.y);
}

Intersection intersection(const Circle& c1, const Circle& c2);

static inline Point center(const Point& a, const Point& b) {
  return Point{a.x + 0.5 * (b.x - a.x), a.y + 0.5 * (b.y - a.y)};
}

// i-th of n near-equal pieces of [lb,ub], the last snapped to ub
// (reference Prover/subintervals.hpp:32-43).  The reference's `lb + diff*i`
// is FMA-contracted by nvcc's default device code generation; the fused form
// is used here and in the GPU kernels alike.
static inline IV get_subinterval(const IV& x, int i, int n) {
  double diff = rn_div(rn_sub(x.ub, x.lb), (double)n);
  double l = rn_fma(diff, (double)i, x.lb);
  double u = rn_fma(diff, (double)(i + 1), x.lb);
  if (i + 1 == n) u = x.ub;
  return IV(l, u);
}

}  // namespace ccor
