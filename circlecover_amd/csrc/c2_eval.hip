// The ten-criterion batch kernel in its own translation unit, built with the
// can_recurse helpers inlined (no call stack: this kernel runs at full
// occupancy with 81 VGPRs and no scratch).
#define CC_INLINE_RECURSE 1
#define CC_EXACT_SHORTCUTS 1
#include "dev/queue.hpp"
#include "dev/rsb.hpp"
#include "dev/f32.hpp"

using namespace ccd;

// Bisection steps a box may take before the cover-strip shortcut applies;
// boxes needing more go to the second pass.
static constexpr int kStripCap = 4;

// Synthetic batched-criterion workload: one lane per 4-D box
// (lambda, r1, r2, r3) stored as 8 SoA FP64 arrays of stride ld; all ten
// criteria evaluated without short-circuit.  Bit k of mask[i] = criterion k
// proved the box; bit 16+k = criterion k hit a sqrt-domain trap.  first[i]
// = lowest proving criterion or 255.  Criterion order:
//   0 TR r1_strategies(la, r1, r2.ub)            3 TR r1_r2_r3 horizontal
//   1 TR r1_r2_strategies_simple(.., r3.ub)      4 TR r1_r2_r3 cover_strip
//   2 TR r1_r2_strategies_stacked(.., r3.ub)     5..9 RSB shortcuts S1..S5
//
// The cover-strip criterion (4) runs its width search with the exact
// shortcut of r123_cover_strip_capped; the few boxes that would need a long
// plain bisection first are appended to `deferred` (count in
// *n_deferred, which must be 0 at launch), their data to the same slot of
// d_stage (8 rows of stride n), with bit 4 left clear, and
// ccrd_c2_strip_deferred completes them.  Criterion 0 runs only its first
// strategy here, which proves most boxes; ccrd_c2_r1_pass tries the others
// on the boxes whose bit 0 is clear, so that a wave does not run them for
// the one or two of its lanes that need them.  Each wave stages those boxes
// (lambda, r1, r2.hi and the box index) at the front of its own 64 slots of
// r1_stage (5 rows of stride ns = the launch's lane count) and r1_id, and
// their number in r1_cnt[wave]: the second pass reads them densely, with no
// atomics on either side.
//
// kF32Pre: every criterion but 4 is first tried as the FP32 shadow of its
// FP64 computation (dev/f32.hpp), which proves a box only when the FP64
// criterion proves it too; the FP64 criterion runs on the boxes the shadow
// leaves undecided, so the outputs are those of the FP64 form.
template <bool kF32Pre>
__device__ __forceinline__ void c2_eval_body(const double* __restrict__ soa, uint64_t n, uint64_t ld,
                                             const Konst* __restrict__ kp, uint32_t* __restrict__ mask,
                                             uint8_t* __restrict__ first, uint32_t* __restrict__ deferred,
                                             unsigned long long* __restrict__ n_deferred,
                                             double* __restrict__ d_stage, double* __restrict__ r1_stage,
                                             uint32_t* __restrict__ r1_id, uint32_t* __restrict__ r1_cnt,
                                             uint64_t ns) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // Lanes past n stay to the wave-wide append below and store nothing.
  const bool live = i < n;
  const uint64_t j = live ? i : 0;
  const Konst K = *kp;
  const iv la{soa[0 * ld + j], soa[1 * ld + j]};
  const iv r1{soa[2 * ld + j], soa[3 * ld + j]};
  const iv r2{soa[4 * ld + j], soa[5 * ld + j]};
  const iv r3{soa[6 * ld + j], soa[7 * ld + j]};

  const iv W = trc::weight_for(K, la);
  const double rem4 = (W - r1 - r2 - r3).hi;
  const double ub_r4 = r3.hi < rem4 ? r3.hi : rem4;

  // pre bit k: the FP32 shadow proved criterion k, whose FP64 form is then
  // skipped (criterion 1's FP64 form reaches its sqrt only with r1.lb >= 0.49,
  // where the argument r1 - 1/4 is positive: skipping it skips no trap).
  uint32_t pre = 0;
  if constexpr (kF32Pre) pre = f32::c2_shadow_mask(K, la, r1, r2, r3);
  uint32_t m = 0;
  Trap tp{0};
  if ((pre & 1u) || trc::r1_strip(K, la, W, r1, r2.hi)) m |= 1u << 0;
  if ((pre & 2u) || trc::r1_r2_simple(K, la, W, r1, r2, r3.hi, tp)) m |= 1u << 1;
  if (tp.flag) m |= 1u << 17;
  if ((pre & 4u) || trc::r1_r2_stacked(K, la, W, r1, r2, r3.hi)) m |= 1u << 2;
  if ((pre & 8u) || trc::r123_horizontal(K, la, W, r1, r2, r3, ub_r4)) m |= 1u << 3;
  bool defer = false;
  if (trc::r123_cover_strip_capped(K, la, W, r1, r2, r3, ub_r4, kStripCap, defer)) m |= 1u << 4;
  if ((pre & 32u) || rsbc::s1_even_split(la, r1, r2)) m |= 1u << 5;
  if ((pre & 64u) || rsbc::s23_uneven_split(la, r1, r2)) m |= 1u << 6;
  if ((pre & 128u) || rsbc::s23_uneven_split(la, r1, r3)) m |= 1u << 7;
  if ((pre & 256u) || rsbc::s4_opposite_corners(la, r1, r2, r3)) m |= 1u << 8;
  if ((pre & 512u) || rsbc::s5_vertical_strip(K, la, r1, r2, r3)) m |= 1u << 9;
  const uint64_t slot = wave_append(live && defer, n_deferred);
  const bool r1_open = live && !(m & 1u);
  const uint64_t ob = __ballot(r1_open);
  const uint64_t wave = i >> 6;  // the block is a whole number of waves
  if (r1_open) {
    const uint64_t d = wave * 64 + (uint64_t)__popcll(ob & ((1ull << __lane_id()) - 1ull));  // < ns
    r1_stage[0 * ns + d] = la.lo;
    r1_stage[1 * ns + d] = la.hi;
    r1_stage[2 * ns + d] = r1.lo;
    r1_stage[3 * ns + d] = r1.hi;
    r1_stage[4 * ns + d] = r2.hi;
    r1_id[d] = (uint32_t)i;
  }
  if (__lane_id() == 0) r1_cnt[wave] = (uint32_t)__popcll(ob);
  if (!live) return;
  if (defer) {
    deferred[slot] = (uint32_t)i;  // slot < n: one slot per deferring box
    d_stage[0 * n + slot] = la.lo;
    d_stage[1 * n + slot] = la.hi;
    d_stage[2 * n + slot] = r1.lo;
    d_stage[3 * n + slot] = r1.hi;
    d_stage[4 * n + slot] = r2.lo;
    d_stage[5 * n + slot] = r2.hi;
    d_stage[6 * n + slot] = r3.lo;
    d_stage[7 * n + slot] = r3.hi;
  }
  mask[i] = m;
  const uint32_t v = m & 0x3ffu;
  first[i] = v ? (uint8_t)__builtin_ctz(v) : (uint8_t)255;
}

extern "C" __global__ void __launch_bounds__(256) ccrd_c2_eval(const double* __restrict__ soa, uint64_t n,
                                                                uint64_t ld, const Konst* __restrict__ kp,
                                                                uint32_t* __restrict__ mask,
                                                                uint8_t* __restrict__ first,
                                                                uint32_t* __restrict__ deferred,
                                                                unsigned long long* __restrict__ n_deferred,
                                                                double* __restrict__ d_stage,
                                                                double* __restrict__ r1_stage,
                                                                uint32_t* __restrict__ r1_id,
                                                                uint32_t* __restrict__ r1_cnt, uint64_t ns) {
  c2_eval_body<false>(soa, n, ld, kp, mask, first, deferred, n_deferred, d_stage, r1_stage, r1_id, r1_cnt, ns);
}

extern "C" __global__ void __launch_bounds__(256) ccrd_c2_eval_f32pre(const double* __restrict__ soa, uint64_t n,
                                                                       uint64_t ld, const Konst* __restrict__ kp,
                                                                       uint32_t* __restrict__ mask,
                                                                       uint8_t* __restrict__ first,
                                                                       uint32_t* __restrict__ deferred,
                                                                       unsigned long long* __restrict__ n_deferred,
                                                                       double* __restrict__ d_stage,
                                                                       double* __restrict__ r1_stage,
                                                                       uint32_t* __restrict__ r1_id,
                                                                       uint32_t* __restrict__ r1_cnt, uint64_t ns) {
  c2_eval_body<true>(soa, n, ld, kp, mask, first, deferred, n_deferred, d_stage, r1_stage, r1_id, r1_cnt, ns);
}

// Second pass over the boxes ccrd_c2_eval deferred (their data read densely
// from d_stage, stride ld = the batch size): the plain cover-strip
// bisection, then bit 4 of mask and first updated.  The bisection is
// sequential, and the deferred boxes are few, so one lane per box would
// leave the GPU mostly idle; instead a group of 16 lanes takes a box and
// advances its bisection four levels per round: lane j (1..15) tests the
// midpoint of node j of the next four levels' binary tree (heap order; the
// bits of j below its leading one are the path, 1 = test accepted, lo moves
// up), then every lane walks the tree with the shuffled outcomes.  Each
// midpoint is computed exactly as the sequential loop computes it, so the
// result is the sequential one, test by test.
static constexpr int kGroup = 16, kLevels = 4;

extern "C" __global__ void __launch_bounds__(256) ccrd_c2_strip_deferred(const double* __restrict__ d_stage,
                                                                          uint64_t ld,
                                                                          const Konst* __restrict__ kp,
                                                                          const uint32_t* __restrict__ deferred,
                                                                          const unsigned long long* __restrict__ n_deferred,
                                                                          uint32_t* __restrict__ mask,
                                                                          uint8_t* __restrict__ first) {
  const uint64_t cnt = *n_deferred;
  const Konst K = *kp;
  const int lane = threadIdx.x % kGroup;
  const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / kGroup;
  for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kGroup; k < cnt; k += groups) {
    const uint64_t i = deferred[k];
    const iv la{d_stage[0 * ld + k], d_stage[1 * ld + k]};
    const iv r1{d_stage[2 * ld + k], d_stage[3 * ld + k]};
    const iv r2{d_stage[4 * ld + k], d_stage[5 * ld + k]};
    const iv r3{d_stage[6 * ld + k], d_stage[7 * ld + k]};
    const double q1 = mul_rd(4.0, r1.lo), q2 = mul_rd(4.0, r2.lo), q3 = mul_rd(4.0, r3.lo);
    double lo = 0.3, hi = la.hi;
    if (!trc::strip_width_ok(q1, q2, q3, lo)) continue;  // (deferred boxes pass this test)
    for (int round = 0; round <= kBisectCap / kLevels; ++round) {
      bool valid = false;
      double pt = 0.0;
      if (lane != 0) {
        const int depth = 31 - __builtin_clz((unsigned)lane);
        double l = lo, h = hi;
        valid = true;
        for (int b = depth - 1; b >= 0; --b) {
          const double m = rn_mid(l, h);
          if (m <= l || m >= h) {
            valid = false;
            break;
          }
          if ((lane >> b) & 1) l = m;
          else h = m;
        }
        if (valid) {
          pt = rn_mid(l, h);
          valid = !(pt <= l || pt >= h);
        }
      }
      const int r = valid && trc::strip_width_ok(q1, q2, q3, pt);
      int node = 1;
      bool done = false;
      for (int lev = 0; lev < kLevels; ++lev) {
        const double m = rn_mid(lo, hi);
        if (m <= lo || m >= hi) {
          done = true;
          break;
        }
        if (__shfl(r, node, kGroup)) {
          lo = m;
          node = 2 * node + 1;
        } else {
          hi = m;
          node = 2 * node;
        }
      }
      if (done) break;
    }
    const iv W = trc::weight_for(K, la);
    const double rem4 = (W - r1 - r2 - r3).hi;
    const double ub_r4 = r3.hi < rem4 ? r3.hi : rem4;
    if (lane == 0 && trc::can_recurse(K, add_ru(la.hi, -lo), 1.0, ub_r4, (W - r1 - r2 - r3).lo)) {
      const uint32_t m = mask[i] | (1u << 4);
      mask[i] = m;
      first[i] = (uint8_t)__builtin_ctz(m & 0x3ffu);
    }
  }
}

// Criterion 0 second pass: r1_strategies' remaining strategies for the boxes
// ccrd_c2_eval staged (those whose bit 0 is clear), then bit 0 of mask and
// first.  Block b takes the staging slots of waves [32 b, 32 b + 32): a
// prefix over their 32 counts in LDS, then one lane per staged box.
static constexpr uint32_t kR1Waves = 32;

extern "C" __global__ void __launch_bounds__(256) ccrd_c2_r1_pass(const double* __restrict__ r1_stage,
                                                                   const uint32_t* __restrict__ r1_id,
                                                                   const uint32_t* __restrict__ r1_cnt,
                                                                   uint64_t nwaves, uint64_t ns,
                                                                   const Konst* __restrict__ kp,
                                                                   uint32_t* __restrict__ mask,
                                                                   uint8_t* __restrict__ first) {
  __shared__ uint32_t pre[kR1Waves + 1];
  const uint64_t w0 = (uint64_t)blockIdx.x * kR1Waves;
  if (threadIdx.x < 64) {  // wave 0: inclusive scan of the 32 counts
    const uint32_t t = threadIdx.x;
    uint32_t x = t < kR1Waves && w0 + t < nwaves ? r1_cnt[w0 + t] : 0u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d);
      if ((int)t >= d) x += y;
    }
    if (t == 0) pre[0] = 0;
    if (t < kR1Waves) pre[t + 1] = x;
  }
  __syncthreads();
  const uint32_t total = pre[kR1Waves];
  const Konst K = *kp;
  for (uint32_t k = threadIdx.x; k < total; k += blockDim.x) {
    uint32_t j = 0;  // the wave whose slots hold entry k: pre[j] <= k < pre[j + 1]
#pragma unroll
    for (uint32_t step = kR1Waves / 2; step > 0; step >>= 1)
      if (pre[j + step] <= k) j += step;
    const uint64_t d = (w0 + j) * 64 + (k - pre[j]);
    const iv la{r1_stage[0 * ns + d], r1_stage[1 * ns + d]};
    const iv r1{r1_stage[2 * ns + d], r1_stage[3 * ns + d]};
    if (trc::r1_rest(K, la, trc::weight_for(K, la), r1, r1_stage[4 * ns + d])) {
      const uint32_t i = r1_id[d];
      mask[i] |= 1u;
      first[i] = 0;
    }
  }
}
