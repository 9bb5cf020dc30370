// First-success evaluation of the ten-criterion 4-D box set (gfx950,
// round-down mode).  Boxes are visited through an index list grouped into
// buckets; bucket b tries criterion b first and then the others in plain
// order (0..9), stopping at the first criterion that proves the box.  Since
// a wave's boxes come from one bucket (except at the few bucket borders),
// the wave runs one criterion at a time instead of a different one per lane.
// With every box in bucket 0 this is the plain order.
#include "dev/f32.hpp"
#include "dev/rsb.hpp"

using namespace ccd;

CCD bool c2_criterion(int k, const Konst& K, iv la, iv W, iv r1, iv r2, iv r3, double ub_r4, Trap& tp) {
  switch (k) {
    case 0: return trc::r1_strategies(K, la, W, r1, r2.hi);
    case 1: return trc::r1_r2_simple(K, la, W, r1, r2, r3.hi, tp);
    case 2: return trc::r1_r2_stacked(K, la, W, r1, r2, r3.hi);
    case 3: return trc::r123_horizontal(K, la, W, r1, r2, r3, ub_r4);
    case 4: return trc::r123_cover_strip(K, la, W, r1, r2, r3, ub_r4);
    case 5: return rsbc::s1_even_split(la, r1, r2);
    case 6: return rsbc::s23_uneven_split(la, r1, r2);
    case 7: return rsbc::s23_uneven_split(la, r1, r3);
    case 8: return rsbc::s4_opposite_corners(la, r1, r2, r3);
    default: return rsbc::s5_vertical_strip(K, la, r1, r2, r3);
  }
}

// idx[0..n) = box indices; bucket_start[b] = first position of bucket b
// (bucket_start[10] = n).  first[box] = proving criterion or 255,
// evals[box] = criteria evaluated, trap[box] = 1 if an evaluated criterion
// reached a trap site.
extern "C" __global__ void __launch_bounds__(256) ccrd_c2_first(const double* __restrict__ soa, uint64_t ld,
                                                                 const uint32_t* __restrict__ idx, uint64_t n,
                                                                 const uint64_t* __restrict__ bucket_start,
                                                                 const Konst* __restrict__ kp,
                                                                 uint8_t* __restrict__ first,
                                                                 uint8_t* __restrict__ evals,
                                                                 uint8_t* __restrict__ trap) {
  const uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  int b = 0;
  while (b < 9 && pos >= bucket_start[b + 1]) ++b;
  const uint64_t i = idx[pos];
  const Konst K = *kp;
  const iv la{soa[0 * ld + i], soa[1 * ld + i]};
  const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
  const iv r2{soa[4 * ld + i], soa[5 * ld + i]};
  const iv r3{soa[6 * ld + i], soa[7 * ld + i]};
  const iv W = trc::weight_for(K, la);
  const double rem4 = (W - r1 - r2 - r3).hi;
  const double ub_r4 = r3.hi < rem4 ? r3.hi : rem4;
  Trap tp{0};
  int found = 255, count = 0;
  for (int s = 0; s < 10; ++s) {
    // step 0 tries the bucket's criterion, then 0..9 without it
    const int k = s == 0 ? b : (s - 1 < b ? s - 1 : s);
    ++count;
    if (c2_criterion(k, K, la, W, r1, r2, r3, ub_r4, tp)) {
      found = k;
      break;
    }
  }
  first[i] = (uint8_t)found;
  evals[i] = (uint8_t)count;
  trap[i] = (uint8_t)tp.flag;
}

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
