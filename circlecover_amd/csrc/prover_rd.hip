// Round-down-mode prover kernels (gfx950).  Every kernel in this translation
// unit is named ccrd_* and is built through the patched-descriptor path of
// build.py, so its waves start with MODE.FP_ROUND = toward -infinity.
#include "dev/rsb_verdict.hpp"

using namespace ccd;

static_assert(sizeof(Konst) <= 256, "device constant block exceeds its host allocation");

// Cell-independent constants (reference critical_ratio.cu:60-61 and
// rectangle_size_bound/variables.cuh:31-51), evaluated once per context.
extern "C" __global__ void ccrd_init_konst(Konst* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Trap t{0};
  Konst K;
  K.be_lo = sqrt_rd(add_rd(mul_rd(0.5, sqrt_rd(7.0)), -0.25));
  K.be_hi = sqrt_ru(add_ru(mul_ru(0.5, sqrt_ru(7.0)), -0.25));
  const double cr = kRsbRatio;
  K.critical_ratio_sq = iv{mul_rd(cr, cr), mul_ru(cr, cr)};
  iv additive = reciprocal(pt(cr)) * sqrt(4.0 * K.critical_ratio_sq - 1.0, t);
  K.min_width_factor = sqrt(2.0 - additive, t);
  K.max_width_factor = sqrt(2.0 + additive, t);
  const double den = sqrt_rd(add_rd(1.0, sqrt_rd(add_rd(1.0, -div_ru(0.25, K.critical_ratio_sq.lo)))));
  const double fac = add_rd(1.0, -rcp_ru(den));
  K.weight_bound_wall_building = mul_rd(0.5, mul_rd(fac, fac));
  K.phase_factor_wall_building = 4.0 * K.critical_ratio_sq - cr * sqrt(16.0 * K.critical_ratio_sq - 4.0, t);
  K.lb_w_infty_denom = add_rd(1.0, -sqrt_ru(K.phase_factor_wall_building.hi));
  K.width_fraction_incomplete =
      rcp_ru(sqrt_rd(add_rd(1.0, sqrt_rd(add_rd(1.0, -div_ru(0.25, K.critical_ratio_sq.lo))))));
  *out = K;
}

// Primitive self-test: out[8*i + k] for op k in {add, mul, div, sqrt} x
// {rd, ru}, inputs a[i], b[i] (sqrt uses a[i]).
extern "C" __global__ void ccrd_prim(const double* __restrict__ a, const double* __restrict__ b, uint64_t n,
                                     double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  out[8 * i + 0] = add_rd(x, y);
  out[8 * i + 1] = add_ru(x, y);
  out[8 * i + 2] = mul_rd(x, y);
  out[8 * i + 3] = mul_ru(x, y);
  out[8 * i + 4] = div_rd(x, y);
  out[8 * i + 5] = div_ru(x, y);
  out[8 * i + 6] = sqrt_rd(x);
  out[8 * i + 7] = sqrt_ru(x);
}

// Round-to-nearest helpers: out[3*i + k] = rn_mid(a,b), rn_fma(a,b,c),
// rn_span_div(a, b, c) = (a-b)/c.
extern "C" __global__ void ccrd_prim_rn(const double* __restrict__ a, const double* __restrict__ b,
                                        const double* __restrict__ c, uint64_t n, double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[3 * i + 0] = rn_mid(a[i], b[i]);
  out[3 * i + 1] = rn_fma(a[i], b[i], c[i]);
  out[3 * i + 2] = rn_span_div(a[i], b[i], c[i]);
}

extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_verdict(const double* __restrict__ cells, uint64_t n,
                                                                    const Konst* __restrict__ kp,
                                                                    int* __restrict__ out,
                                                                    uint8_t* __restrict__ trap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Konst K = *kp;
  const double* p = cells + 14 * i;
  cell7 c;
  c.la = iv{p[0], p[1]};
#pragma unroll
  for (int k = 0; k < 6; ++k) c.r[k] = iv{p[2 + 2 * k], p[3 + 2 * k]};
  Trap tp{0};
  out[i] = rsb_cell_verdict(K, c, tp);
  trap[i] = (uint8_t)tp.flag;
}

// Every rectangle_size_bound step evaluated independently on each cell (no
// short-circuit): bit k = step k succeeds (numbering of rsb_cell_verdict),
// bit 20+k = step k reached a reference trap site.
extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_masks(const double* __restrict__ cells, uint64_t n,
                                                                  const Konst* __restrict__ kp,
                                                                  uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Konst K = *kp;
  const double* p = cells + 14 * i;
  cell7 in;
  in.la = iv{p[0], p[1]};
#pragma unroll
  for (int k = 0; k < 6; ++k) in.r[k] = iv{p[2 + 2 * k], p[3 + 2 * k]};
  uint64_t m = 0;
  Trap tp{0};
  auto put = [&](int k, bool ok) {
    if (ok) m |= 1ull << k;
    if (tp.flag) m |= 1ull << (20 + k);
    tp.flag = 0;
  };
  put(0, rsbc::s1_even_split(in.la, in.r[0], in.r[1]));
  put(1, rsbc::s23_uneven_split(in.la, in.r[0], in.r[1]));
  put(2, rsbc::s23_uneven_split(in.la, in.r[0], in.r[2]));
  put(3, rsbc::s4_opposite_corners(in.la, in.r[0], in.r[1], in.r[2]));
  put(4, rsbc::s5_vertical_strip(K, in.la, in.r[0], in.r[1], in.r[2]));
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  put(5, rsbc::is_empty(c) || empty(v.R));
  put(6, rsbc::c1_even_split(c, v));
  put(7, rsbc::c2_uneven_split(c));
  put(8, rsbc::c3_multi_disk_strip(c, v, tp));
  put(9, rsbc::c4_wall_building(K, c, v));
  put(10, rsbc::c5_corner_explicit(c, v, tp));
  put(11, rsbc::c6_opposite_corners(c, v, tp));
  put(12, rsbc::c7_vertical_strip(K, c, v));
  const rsbc::cover22 cov = rsbc::two_by_two(c, tp);
  v.cover_width = cov.width;
  v.cover_c4 = cov.c4;
  put(13, cov.width > 0 && (cov.width > c.la.hi || rsbc::c8_two_by_two_strip(K, c, v, tp)));
  put(14, rsbc::c9_l_shaped(K, c));
  put(15, rsbc::c10_corner_wall(K, c, tp));
  put(16, rsbc::c11_advanced_strip(c));
  put(17, rsbc::c12_five_disk(K, c, v, tp));
  put(18, rsbc::c13_six_disk(K, c, v, tp));
  put(19, rsbc::c14_seven_disk(K, c, v, tp));
  out[i] = m;
}
