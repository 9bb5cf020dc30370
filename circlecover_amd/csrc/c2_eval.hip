// The ten-criterion batch kernel in its own translation unit, built with the
// can_recurse helpers inlined (no call stack: this kernel runs at full
// occupancy with 81 VGPRs and no scratch).
#define CC_INLINE_RECURSE 1
#include "dev/rsb.hpp"

using namespace ccd;

// Synthetic batched-criterion workload: one lane per 4-D box
// (lambda, r1, r2, r3) stored as 8 SoA FP64 arrays of stride ld; all ten
// criteria evaluated without short-circuit.  Bit k of mask[i] = criterion k
// proved the box; bit 16+k = criterion k hit a sqrt-domain trap.  first[i]
// = lowest proving criterion or 255.  Criterion order:
//   0 TR r1_strategies(la, r1, r2.ub)            3 TR r1_r2_r3 horizontal
//   1 TR r1_r2_strategies_simple(.., r3.ub)      4 TR r1_r2_r3 cover_strip
//   2 TR r1_r2_strategies_stacked(.., r3.ub)     5..9 RSB shortcuts S1..S5
extern "C" __global__ void __launch_bounds__(256) ccrd_c2_eval(const double* __restrict__ soa, uint64_t n,
                                                                uint64_t ld, const Konst* __restrict__ kp,
                                                                uint32_t* __restrict__ mask,
                                                                uint8_t* __restrict__ first) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Konst K = *kp;
  const iv la{soa[0 * ld + i], soa[1 * ld + i]};
  const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
  const iv r2{soa[4 * ld + i], soa[5 * ld + i]};
  const iv r3{soa[6 * ld + i], soa[7 * ld + i]};

  const iv W = trc::weight_for(K, la);
  const double rem4 = (W - r1 - r2 - r3).hi;
  const double ub_r4 = r3.hi < rem4 ? r3.hi : rem4;

  uint32_t m = 0;
  Trap tp{0};
  if (trc::r1_strategies(K, la, W, r1, r2.hi)) m |= 1u << 0;
  if (trc::r1_r2_simple(K, la, W, r1, r2, r3.hi, tp)) m |= 1u << 1;
  if (tp.flag) m |= 1u << 17;
  if (trc::r1_r2_stacked(K, la, W, r1, r2, r3.hi)) m |= 1u << 2;
  if (trc::r123_horizontal(K, la, W, r1, r2, r3, ub_r4)) m |= 1u << 3;
  if (trc::r123_cover_strip(K, la, W, r1, r2, r3, ub_r4)) m |= 1u << 4;
  if (rsbc::s1_even_split(la, r1, r2)) m |= 1u << 5;
  if (rsbc::s23_uneven_split(la, r1, r2)) m |= 1u << 6;
  if (rsbc::s23_uneven_split(la, r1, r3)) m |= 1u << 7;
  if (rsbc::s4_opposite_corners(la, r1, r2, r3)) m |= 1u << 8;
  if (rsbc::s5_vertical_strip(K, la, r1, r2, r3)) m |= 1u << 9;
  mask[i] = m;
  const uint32_t v = m & 0x3ffu;
  first[i] = v ? (uint8_t)__builtin_ctz(v) : (uint8_t)255;
}

