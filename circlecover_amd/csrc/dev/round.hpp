// Directed-rounding FP64/FP32 primitives for gfx950 prover kernels.
//
// Execution model.  Every prover kernel ("ccrd_*" symbols) starts its waves
// with MODE.FP_ROUND = round-toward-negative-infinity for f32 and f64: the
// build sets `.amdhsa_float_round_mode_32/_16_64 2` in the kernel descriptor
// (see build.py), so no instruction of the kernel can ever run in another
// mode by accident.  In that mode:
//   * a round-down add/mul/fma is ONE VALU instruction;
//   * a round-up op is the exact IEEE identity RU(x op y) = -RD((-x) op y)
//     (for add: -RD((-x)+(-y))), realised with VALU negation modifiers;
//   * div and sqrt start from a Newton approximation (rcp/rsq seeds, all in
//     round-down) that is within two ulps, and are then stepped to the exact
//     directed result by the sign of the residual a - q*b (resp. x - s*s),
//     which one FMA yields with the correct sign and zero-ness whenever the
//     residual's granularity does not underflow (guaranteed on the fast path's
//     exponent range; the slow path rescales by powers of two first);
//   * the few round-to-nearest operations of the reference (bisection
//     midpoints and get_subinterval) run inside a single asm block that
//     switches MODE to nearest-even, computes, and switches back.
// All rounding arithmetic is inline asm so the compiler, which assumes
// round-to-nearest, can neither constant-fold nor algebraically rewrite it.
// Denormals stay enabled (descriptor default), as IEEE directed rounding
// requires.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CCD __device__ __forceinline__

namespace ccd {

CCD uint64_t bits(double x) { return (uint64_t)__double_as_longlong(x); }
CCD double from_bits(uint64_t b) { return __longlong_as_double((long long)b); }

// ---------------------------------------------------------------- add / mul
// Operands the compiler knows as constants are not copied into VGPRs: the
// inline constants 0.5, 1, 2 and 4 are written into the instruction, any
// other constant goes in as an SGPR pair (one per instruction: the VALU reads
// at most one scalar operand).  Same instruction, same rounding; a VGPR
// operand would cost two moves per use.
// A translation unit sets CC_CONST_OPERANDS 0 before including this header
// to keep every operand in VGPRs: in the RSB enumeration's inlined leaf
// chain, which already uses every SGPR, the constant operands made its
// stage A 8% slower (SGPR spills through VGPR lanes, a different schedule).
#ifndef CC_CONST_OPERANDS
#define CC_CONST_OPERANDS 1
#endif
#define CC_K(x, v) (CC_CONST_OPERANDS && __builtin_constant_p(x) && (x) == (v))
#define CC_KS(x) (CC_CONST_OPERANDS && __builtin_constant_p(x))
// op %0, M1 a, M2 b with an inline constant, an SGPR or a VGPR for b, or an
// SGPR for a (both operand orders are listed by the callers of commutative
// operations).
#define CC_BIN(OP, M1, M2, a, b, r)                                                            \
  do {                                                                                         \
    if (CC_K(b, 0.5)) asm(OP " %0, " M1 "%1, " M2 "0.5" : "=v"(r) : "v"(a));                   \
    else if (CC_K(b, 1.0)) asm(OP " %0, " M1 "%1, " M2 "1.0" : "=v"(r) : "v"(a));              \
    else if (CC_K(b, 2.0)) asm(OP " %0, " M1 "%1, " M2 "2.0" : "=v"(r) : "v"(a));              \
    else if (CC_K(b, 4.0)) asm(OP " %0, " M1 "%1, " M2 "4.0" : "=v"(r) : "v"(a));              \
    else if (CC_KS(b)) asm(OP " %0, " M1 "%1, " M2 "%2" : "=v"(r) : "v"(a), "s"(b));            \
    else if (CC_KS(a)) asm(OP " %0, " M1 "%1, " M2 "%2" : "=v"(r) : "s"(a), "v"(b));            \
    else asm(OP " %0, " M1 "%1, " M2 "%2" : "=v"(r) : "v"(a), "v"(b));                          \
  } while (0)

CCD double add_rd(double a, double b) {
  double r;
  if (CC_KS(a) && !CC_KS(b)) CC_BIN("v_add_f64", "", "", b, a, r);
  else CC_BIN("v_add_f64", "", "", a, b, r);
  return r;
}
CCD double add_ru(double a, double b) {
  double r;
  if (CC_KS(a) && !CC_KS(b)) CC_BIN("v_add_f64", "-", "-", b, a, r);
  else CC_BIN("v_add_f64", "-", "-", a, b, r);
  return -r;
}
CCD double mul_rd(double a, double b) {
  double r;
  if (CC_KS(a) && !CC_KS(b)) CC_BIN("v_mul_f64", "", "", b, a, r);
  else CC_BIN("v_mul_f64", "", "", a, b, r);
  return r;
}
CCD double mul_ru(double a, double b) {
  double r;
  // -RD((-a) * b) = -RD(a * (-b)): the sign goes on whichever operand stays a VGPR
  if (CC_KS(a) && !CC_KS(b)) CC_BIN("v_mul_f64", "-", "", b, a, r);
  else CC_BIN("v_mul_f64", "-", "", a, b, r);
  return -r;
}
// a*b + c rounded once, downward; c (the accumulator) may be a constant.
#define CC_FMA(M1, a, b, c, r)                                                                     \
  do {                                                                                             \
    if (CC_K(c, 0.5)) asm("v_fma_f64 %0, " M1 "%1, %2, 0.5" : "=v"(r) : "v"(a), "v"(b));          \
    else if (CC_K(c, 1.0)) asm("v_fma_f64 %0, " M1 "%1, %2, 1.0" : "=v"(r) : "v"(a), "v"(b));     \
    else if (CC_KS(c)) asm("v_fma_f64 %0, " M1 "%1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));   \
    else if (CC_KS(b)) asm("v_fma_f64 %0, " M1 "%1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));   \
    else asm("v_fma_f64 %0, " M1 "%1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));                 \
  } while (0)
CCD double fma_rd(double a, double b, double c) {
  double r;
  CC_FMA("", a, b, c, r);
  return r;
}
// c - a*b rounded once, downward.
CCD double fnma_rd(double a, double b, double c) {
  double r;
  CC_FMA("-", a, b, c, r);
  return r;
}

// ------------------------------------------------------------ round-nearest
// 0.5*(a+b) with both operations rounded to nearest-even.
CCD double rn_mid(double a, double b) {
  double r;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 0\n\t"
      "s_nop 1\n\t"
      "v_add_f64 %0, %1, %2\n\t"
      "v_mul_f64 %0, 0.5, %0\n\t"
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10\n\t"
      "s_nop 1"
      : "=&v"(r)
      : "v"(a), "v"(b));
  return r;
}
// fma(d, k, lo) rounded to nearest-even: the reference's `lb + diff*i`
// as contracted by nvcc's default device code generation.
CCD double rn_fma(double d, double k, double lo) {
  double r;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 0\n\t"
      "s_nop 1\n\t"
      "v_fma_f64 %0, %1, %2, %3\n\t"
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10\n\t"
      "s_nop 1"
      : "=&v"(r)
      : "v"(d), "v"(k), "v"(lo));
  return r;
}
// (hi - lo) / n rounded to nearest-even (the diff of get_subinterval): the
// standard scaled Newton division sequence executed in nearest-even mode.
CCD double rn_span_div(double hi, double lo, double n) {
  double r;
  double t0, t1, t2, t3, t4;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 0\n\t"
      "s_nop 1\n\t"
      "v_add_f64 %[a], %[hi], -%[lo]\n\t"
      "v_div_scale_f64 %[d], vcc, %[n], %[n], %[a]\n\t"
      "v_rcp_f64 %[y], %[d]\n\t"
      "v_div_scale_f64 %[s], vcc, %[a], %[n], %[a]\n\t"
      "v_fma_f64 %[t], -%[d], %[y], 1.0\n\t"
      "v_fma_f64 %[y], %[y], %[t], %[y]\n\t"
      "v_fma_f64 %[t], -%[d], %[y], 1.0\n\t"
      "v_fma_f64 %[y], %[y], %[t], %[y]\n\t"
      "v_mul_f64 %[t], %[s], %[y]\n\t"
      "v_fma_f64 %[d], -%[d], %[t], %[s]\n\t"
      "s_nop 4\n\t"
      "v_div_fmas_f64 %[d], %[d], %[y], %[t]\n\t"
      "v_div_fixup_f64 %[r], %[d], %[n], %[a]\n\t"
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10\n\t"
      "s_nop 1"
      : [r] "=&v"(r), [a] "=&v"(t0), [d] "=&v"(t1), [y] "=&v"(t2), [s] "=&v"(t3), [t] "=&v"(t4)
      : [hi] "v"(hi), [lo] "v"(lo), [n] "v"(n)
      : "vcc");
  return r;
}

// fma(d, i, lo) and fma(d, j, lo) rounded to nearest-even under one mode
// switch: the two bounds of the reference's `lb + diff*i`, `lb + diff*(i+1)`.
CCD void rn_fma2(double d, double i, double j, double lo, double& l, double& u) {
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 0\n\t"
      "s_nop 1\n\t"
      "v_fma_f64 %0, %2, %3, %5\n\t"
      "v_fma_f64 %1, %2, %4, %5\n\t"
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10\n\t"
      "s_nop 1"
      : "=&v"(l), "=&v"(u)
      : "v"(d), "v"(i), "v"(j), "v"(lo));
}
// get_subinterval's three nearest-even operations under one mode switch:
// diff = (hi - lo) / n (the scaled Newton division of rn_span_div), then
// l = fma(diff, i, lo) and u = fma(diff, i + 1, lo).
CCD void rn_piece(double hi, double lo, double n, double i, double j, double& l, double& u) {
  double t0, t1, t2, t3, t4, t5;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 0\n\t"
      "s_nop 1\n\t"
      "v_add_f64 %[a], %[hi], -%[lo]\n\t"
      "v_div_scale_f64 %[d], vcc, %[n], %[n], %[a]\n\t"
      "v_rcp_f64 %[y], %[d]\n\t"
      "v_div_scale_f64 %[s], vcc, %[a], %[n], %[a]\n\t"
      "v_fma_f64 %[t], -%[d], %[y], 1.0\n\t"
      "v_fma_f64 %[y], %[y], %[t], %[y]\n\t"
      "v_fma_f64 %[t], -%[d], %[y], 1.0\n\t"
      "v_fma_f64 %[y], %[y], %[t], %[y]\n\t"
      "v_mul_f64 %[t], %[s], %[y]\n\t"
      "v_fma_f64 %[d], -%[d], %[t], %[s]\n\t"
      "s_nop 4\n\t"
      "v_div_fmas_f64 %[d], %[d], %[y], %[t]\n\t"
      "v_div_fixup_f64 %[q], %[d], %[n], %[a]\n\t"
      "v_fma_f64 %[l], %[q], %[i], %[lo]\n\t"
      "v_fma_f64 %[u], %[q], %[j], %[lo]\n\t"
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10\n\t"
      "s_nop 1"
      : [l] "=&v"(l), [u] "=&v"(u), [a] "=&v"(t0), [d] "=&v"(t1), [y] "=&v"(t2), [s] "=&v"(t3), [t] "=&v"(t4),
        [q] "=&v"(t5)
      : [hi] "v"(hi), [lo] "v"(lo), [n] "v"(n), [i] "v"(i), [j] "v"(j)
      : "vcc");
}

// ------------------------------------------------------------ neighbours
// Next representable double toward +inf / -inf for finite nonzero x.
CCD double next_up(double x) {
  uint64_t b = bits(x);
  return from_bits((int64_t)b >= 0 ? b + 1 : b - 1);
}
CCD double next_down(double x) {
  uint64_t b = bits(x);
  return from_bits((int64_t)b >= 0 ? b - 1 : b + 1);
}
CCD double pow2(int k) {  // 2^k for -1022 <= k <= 1023
  return from_bits((uint64_t)(k + 1023) << 52);
}
CCD int exponent_of(double x) {  // unbiased exponent of a normal double
  return (int)((bits(x) >> 52) & 0x7ff) - 1023;
}
// Reciprocal / reciprocal-square-root seeds.  They are transcendental ops:
// on gfx950 a dependent VALU instruction must not issue in the next cycle
// (the hardware does not interlock it), and the compiler's hazard tracking
// does not look into the inline asm that consumes these seeds, so the wait
// states are part of the asm block itself.
CCD double rcp_seed(double x) {
  double r;
  asm volatile("v_rcp_f64 %0, %1\n\ts_nop 1" : "=v"(r) : "v"(x));
  return r;
}
CCD double rsq_seed(double x) {
  double r;
  asm volatile("v_rsq_f64 %0, %1\n\ts_nop 1" : "=v"(r) : "v"(x));
  return r;
}

// ------------------------------------------------------------------ division
// Round-down quotient for b > 0, a and b finite with |a|,|b| in
// [2^-500, 2^500]: Newton approximation then residual stepping.  The seed is
// good to about 2^-25, one Newton step makes y good to about 2^-50, and the
// corrected quotient q + r*y then misses a/b by at most 2^-48 ulp, so q is
// RD(a/b) or a neighbour of it: one residual test each way settles it
// (sign(a - q*b) == sign(a/b - q), exact from one FMA here).
// The dependent chain of the quotient, q and its residual r = a - q*b, in
// one asm block: the same instructions, operands and modifiers as the
// one-operation primitives above (seed, fnma_rd(b, y, 1), fma_rd(y, e, y),
// mul_rd(a, y), fnma_rd(b, q, a), fma_rd(r0, y, q), fnma_rd(q, b, a)).
// Between separate asm blocks the compiler inserts a wait state for a
// result-forwarding hazard of sub-dword destinations, which these
// instructions do not have; on a dependent chain those waits add up.
CCD void div_core(double a, double b, double& q, double& r) {
  double y, e, r0;
  asm volatile(
      "v_rcp_f64 %[y], %[b]\n\t"
      "s_nop 1\n\t"
      "v_fma_f64 %[e], -%[b], %[y], 1.0\n\t"
      "v_fma_f64 %[y], %[y], %[e], %[y]\n\t"
      "v_mul_f64 %[q], %[a], %[y]\n\t"
      "v_fma_f64 %[r0], -%[b], %[q], %[a]\n\t"
      "v_fma_f64 %[q], %[r0], %[y], %[q]\n\t"
      "v_fma_f64 %[r], -%[q], %[b], %[a]"
      : [y] "=&v"(y), [e] "=&v"(e), [q] "=&v"(q), [r0] "=&v"(r0), [r] "=&v"(r)
      : [a] "v"(a), [b] "v"(b));
}
CCD double div_rd_fast_pos(double a, double b) {
  double q, r;
  div_core(a, b, q, r);
  const double qu = next_up(q);
  const double ru = fnma_rd(qu, b, a);
  return r < 0.0 ? next_down(q) : (ru >= 0.0 ? qu : q);
}

static __device__ __noinline__ double div_rd_slow(double a, double b);

CCD bool div_fast_ok(double a, double b) {
  // |a| and |b| in [2^-500, 2^500) (false for zeros, subnormals, infinities
  // and NaN): four compares with absolute-value operands
  const double lo = 0x1p-500, hi = 0x1p500;
  const double fa = __builtin_fabs(a), fb = __builtin_fabs(b);
  return (fa >= lo) & (fa < hi) & (fb >= lo) & (fb < hi);
}

CCD double div_rd(double a, double b) {
  if (div_fast_ok(a, b)) return b > 0.0 ? div_rd_fast_pos(a, b) : div_rd_fast_pos(-a, -b);
  return div_rd_slow(a, b);
}
CCD double div_ru(double a, double b) { return -div_rd(-a, b); }
CCD double rcp_rd(double x) { return div_rd(1.0, x); }
CCD double rcp_ru(double x) { return div_ru(1.0, x); }

// Exact-sign slow path: special operands by IEEE rules, extreme exponents by
// power-of-two rescaling (round-down compositions on nested grids are exact).
static __device__ __noinline__ double div_rd_slow(double a, double b) {
  const double inf = __builtin_inf();
  const bool neg = ((bits(a) ^ bits(b)) >> 63) != 0;
  if (a != a || b != b) return __builtin_nan("");
  if (b == 0.0) return a == 0.0 ? __builtin_nan("") : (neg ? -inf : inf);
  if (__builtin_isinf(a)) return __builtin_isinf(b) ? __builtin_nan("") : (neg ? -inf : inf);
  if (__builtin_isinf(b) || a == 0.0) return neg ? -0.0 : 0.0;
  // finite nonzero: a = ma*2^ka, b = mb*2^kb with |ma|,|mb| in [1/2,1)
  int ka, kb;
  double ma = __builtin_frexp(a, &ka);  // frexp is exact (power-of-two scaling)
  double mb = __builtin_frexp(b, &kb);
  double q = mb > 0.0 ? div_rd_fast_pos(ma, mb) : div_rd_fast_pos(-ma, -mb);
  int k = ka - kb;
  while (k > 1023) {
    q = mul_rd(q, pow2(1023));
    k -= 1023;
  }
  while (k < -1022) {
    q = mul_rd(q, pow2(-1022));
    k += 1022;
  }
  return mul_rd(q, pow2(k));
}

// --------------------------------------------------------------------- sqrt
// Round-down square root for x in [2^-960, DBL_MAX].  One Goldschmidt step
// from the ~2^-25 rsq seed gives g and h = 1/(2 sqrt x) to about 2^-49, one
// Newton step then leaves g + d*h within 2^-48 ulp of sqrt(x), so g is
// RD(sqrt x) or a neighbour of it (a non-square's root lies at least 2^-54
// ulp from every double, so both neighbours can occur, rarely): one residual
// test each way settles it (sign(x - g*g) == sign(sqrt(x) - g); g > 0).
// The dependent chain of the root, g and its residual d = x - g*g, in one
// asm block (as div_core: the instructions of rsq_seed, mul_rd(x, y),
// mul_rd(y, 0.5), fnma_rd(g, h, 0.5), fma_rd(g, r, g), fma_rd(h, r, h),
// fnma_rd(g, g, x), fma_rd(d0, h, g), fnma_rd(g, g, x)).
CCD void sqrt_core(double x, double& g, double& d) {
  double y, h, r, d0;
  asm volatile(
      "v_rsq_f64 %[y], %[x]\n\t"
      "s_nop 1\n\t"
      "v_mul_f64 %[g], %[x], %[y]\n\t"
      "v_mul_f64 %[h], %[y], 0.5\n\t"
      "v_fma_f64 %[r], -%[g], %[h], 0.5\n\t"
      "v_fma_f64 %[g], %[g], %[r], %[g]\n\t"
      "v_fma_f64 %[h], %[h], %[r], %[h]\n\t"
      "v_fma_f64 %[d0], -%[g], %[g], %[x]\n\t"
      "v_fma_f64 %[g], %[d0], %[h], %[g]\n\t"
      "v_fma_f64 %[d], -%[g], %[g], %[x]"
      : [y] "=&v"(y), [h] "=&v"(h), [r] "=&v"(r), [d0] "=&v"(d0), [g] "=&v"(g), [d] "=&v"(d)
      : [x] "v"(x));
}
CCD double sqrt_rd_fast(double x) {
  double g, d;
  sqrt_core(x, g, d);
  const double gu = from_bits(bits(g) + 1u);
  const double du = fnma_rd(gu, gu, x);
  return d < 0.0 ? from_bits(bits(g) - 1u) : (du >= 0.0 ? gu : g);
}

static __device__ __noinline__ double sqrt_rd_slow(double x);

CCD bool sqrt_fast_ok(double x) {
  // x in [2^-960, DBL_MAX] (false for NaN): biased exponent in [63, 2046],
  // sign clear
  return (x >= 0x1p-960) & (x <= __DBL_MAX__);
}
CCD double sqrt_rd(double x) { return sqrt_fast_ok(x) ? sqrt_rd_fast(x) : sqrt_rd_slow(x); }
CCD double sqrt_ru(double x) {
  double s = sqrt_rd(x);
  if (!sqrt_fast_ok(x)) {
    // exact for 0, -0, +inf, NaN and negative inputs; tiny x rescaled
    if (!(x > 0.0) || __builtin_isinf(x)) return s;
    double xs = mul_rd(x, pow2(1000));
    double ss = mul_rd(s, pow2(500));
    return fnma_rd(ss, ss, xs) == 0.0 ? s : next_up(s);
  }
  return fnma_rd(s, s, x) == 0.0 ? s : next_up(s);
}
static __device__ __noinline__ double sqrt_rd_slow(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0 || __builtin_isinf(x)) return x;
  // 0 < x < 2^-960: sqrt(x*2^1000) = sqrt(x)*2^500, both scalings exact
  return mul_rd(sqrt_rd_fast(mul_rd(x, pow2(1000))), pow2(-500));
}

}  // namespace ccd
