// FP32 pre-filter kernel of the ten-criterion box set (gfx950, round-down
// mode for FP32 and FP64).
#define CC_EXACT_SHORTCUTS 1
#include "dev/c2_criteria.hpp"
#include "dev/f32.hpp"

using namespace ccd;

// FP32 pre-filter (plain order): criterion 0 (r1_strategies, which proves
// most boxes first) is run as an FP32 shadow of its FP64 computation (see
// dev/f32.hpp); a box the shadow proves is one the FP64 criterion proves, so
// it is decided there (first = 0, f32_decided[i] = 1).
// Every other box is evaluated in FP64 in plain order as ccrd_c2_first
// does.  widths (may be null) receives, per box, the widths of
// critical_ratio(lambda) and of required_weight_for(lambda) - r1 in FP32 and
// in FP64 (4 floats: cr32, cr64, rem32, rem64) for the inflation report; the
// FP32 width is that of the hull of the shadows of both FP64 bounds.
extern "C" __global__ void __launch_bounds__(256) ccrd_c2_first_f32(const double* __restrict__ soa, uint64_t ld,
                                                                     uint64_t n, const Konst* __restrict__ kp,
                                                                     uint8_t* __restrict__ first,
                                                                     uint8_t* __restrict__ evals,
                                                                     uint8_t* __restrict__ trap,
                                                                     uint8_t* __restrict__ f32_decided,
                                                                     float* __restrict__ widths) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Konst K = *kp;
  const iv la{soa[0 * ld + i], soa[1 * ld + i]};
  const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
  const iv r2{soa[4 * ld + i], soa[5 * ld + i]};
  const iv r3{soa[6 * ld + i], soa[7 * ld + i]};
  const f32::KonstF KF = f32::konst(K);
  const f32::siv las = f32::sv(la), r1s = f32::sv(r1);
  f32::Sh st{true};
  const f32::siv Ws = f32::crit_ratio(KF, las, st) * las;  // required_weight_for(la)
  if (widths) {
    f32::Sh st2{true};
    const f32::siv crs = f32::crit_ratio(KF, las, st2);
    const iv cr = trc::crit_ratio(K, la);
    const iv W = trc::weight_for(K, la);
    const f32::siv rems = Ws - r1s;
    const iv rem = W - r1;
    widths[4 * i + 0] = f32::add_ru(crs.hi.hi, -crs.lo.lo);
    widths[4 * i + 1] = (float)(cr.hi - cr.lo);
    widths[4 * i + 2] = f32::add_ru(rems.hi.hi, -rems.lo.lo);
    widths[4 * i + 3] = (float)(rem.hi - rem.lo);
  }
  if (st.ok && f32::r1_strategies(KF, las, Ws, r1s, f32::shd(r2.hi), st) && st.ok) {
    first[i] = 0;
    evals[i] = 1;
    trap[i] = 0;
    f32_decided[i] = 1;
    return;
  }
  f32_decided[i] = 0;
  const iv W = trc::weight_for(K, la);
  const double rem4 = (W - r1 - r2 - r3).hi;
  const double ub_r4 = r3.hi < rem4 ? r3.hi : rem4;
  Trap tp{0};
  int found = 255, count = 1;  // the FP32 attempt counts as one evaluation
  for (int k = 0; k < 10; ++k) {
    ++count;
    if (c2_criterion(k, K, la, W, r1, r2, r3, ub_r4, tp)) {
      found = k;
      break;
    }
  }
  first[i] = (uint8_t)found;
  evals[i] = (uint8_t)(count > 255 ? 255 : count);
  trap[i] = (uint8_t)tp.flag;
}

// Per-criterion report of the FP32 shadows (no short-circuit): bit k of
// f32_mask[i] = the shadow of criterion k proves box i (bit 4 never: the
// cover-strip bisection has no shadow).  infl[10*i+k] (may be null) = the
// width of criterion k's weight budget in the FP32 shadow over its width in
// FP64, or -1 where the FP64 width is 0 or criterion k has no shadow.  The
// budget is the weight a criterion starts from minus the disks it places:
//   0: W - r1   1, 2: W - r1 - r2   3: W - r1 - r2 - r3   (W = required_weight_for(la))
//   5, 6, 9: R - r1 - r2   7: R - r1 - r3   8: R - r1 - r2 - r3   (R = 0.61 la)
// The FP32 width is that of the hull of the shadows of both FP64 bounds.
CCD float budget_ratio(const f32::siv& s, iv d) {
  const double w64 = d.hi - d.lo;
  if (!(w64 > 0.0)) return -1.f;
  const float w32 = f32::add_ru(s.hi.hi, -s.lo.lo);
  return (float)((double)w32 / w64);
}

extern "C" __global__ void __launch_bounds__(256) ccrd_c2_f32_report(const double* __restrict__ soa, uint64_t ld,
                                                                      uint64_t n, const Konst* __restrict__ kp,
                                                                      uint32_t* __restrict__ f32_mask,
                                                                      float* __restrict__ infl) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Konst K = *kp;
  const iv la{soa[0 * ld + i], soa[1 * ld + i]};
  const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
  const iv r2{soa[4 * ld + i], soa[5 * ld + i]};
  const iv r3{soa[6 * ld + i], soa[7 * ld + i]};
  const f32::C2In c = f32::c2_inputs(K, la, r1, r2, r3);
  uint32_t m = 0;
#pragma unroll 1
  for (int k = 0; k < 10; ++k)
    if (k != 4 && f32::c2_shadow(k, c)) m |= 1u << k;
  f32_mask[i] = m;
  if (!infl) return;
  float* o = infl + 10 * i;
  const iv W = trc::weight_for(K, la);
  const iv R = kRsbRatio * la;
  const f32::siv Rs = f32::scale(kRsbRatio, c.la);
  if (c.ok) {
    o[0] = budget_ratio(c.W - c.r1, W - r1);
    o[1] = o[2] = budget_ratio(c.W - c.r1 - c.r2, W - r1 - r2);
    o[3] = budget_ratio(c.W - c.r1 - c.r2 - c.r3, W - r1 - r2 - r3);
  } else {
    o[0] = o[1] = o[2] = o[3] = -1.f;
  }
  o[4] = -1.f;
  o[5] = o[6] = o[9] = budget_ratio(Rs - c.r1 - c.r2, R - r1 - r2);
  o[7] = budget_ratio(Rs - c.r1 - c.r3, R - r1 - r3);
  o[8] = budget_ratio(Rs - c.r1 - c.r2 - c.r3, R - r1 - r2 - r3);
}
