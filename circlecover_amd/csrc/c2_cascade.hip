// Cascade form of first-success evaluation, first stage (gfx950,
// round-down mode): ccrd_c2_first0 tries criterion 0 on every box,
// ccrd_c2_first0_list packs the boxes it leaves open into a dense SoA, and
// ccrd_c2_first (c2_first.hip, done0 > 0) runs the other criteria on those
// only, in plain or bucket order.  The first proving criterion in plain
// order is criterion 0 for every box criterion 0 proves, so the plain
// cascade gives plain order's first/evals/trap exactly.  Built with the
// can_recurse helpers inlined, as the batch kernel is (no call stack).
#define CC_INLINE_RECURSE 1
#define CC_EXACT_SHORTCUTS 1
#include "dev/c2_criteria.hpp"
#include "dev/f32.hpp"

using namespace ccd;

// Cascade, first stage: criterion 0 (it proves most boxes first) for the
// 2048 boxes [2048 * blockIdx.x, ...) of each block.  Its first strategy
// runs on every box; the boxes that strategy leaves open are gathered in LDS
// (a ballot and one LDS atomic per wave) and the remaining strategies run on
// them densely.  first/evals/trap as ccrd_c2_first writes them, first = 255
// for the boxes left to the second stage.  Those boxes are also copied, with
// their indices, to the front of the block's range of `stage` / `stage_id`
// (SoA of stride n), so that the second stage reads them densely; their
// number goes to pending[blockIdx.x].  The order of the copies within a
// block depends on wave timing; no result depends on it.
static constexpr uint32_t kFirstChunk = 2048;

extern "C" __global__ void __launch_bounds__(256) ccrd_c2_first0(const double* __restrict__ soa, uint64_t ld,
                                                                  uint64_t n, const Konst* __restrict__ kp,
                                                                  uint8_t* __restrict__ first,
                                                                  uint8_t* __restrict__ evals,
                                                                  uint8_t* __restrict__ trap,
                                                                  uint32_t* __restrict__ pending,
                                                                  double* __restrict__ stage,
                                                                  uint32_t* __restrict__ stage_id) {
  __shared__ uint32_t list[kFirstChunk];
  __shared__ uint32_t count, open;
  if (threadIdx.x == 0) count = open = 0;
  __syncthreads();
  const Konst K = *kp;
  const uint64_t base = (uint64_t)blockIdx.x * kFirstChunk;
  const int lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t j = threadIdx.x; j < kFirstChunk; j += blockDim.x) {  // kFirstChunk is a multiple of the block
    const uint64_t i = base + j;
    bool keep = false;
    if (i < n) {
      const iv la{soa[0 * ld + i], soa[1 * ld + i]};
      const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
      keep = !trc::r1_strip(K, la, trc::weight_for(K, la), r1, soa[5 * ld + i]);
      first[i] = keep ? 255 : 0;
      evals[i] = 1;
      trap[i] = 0;  // criterion 0 has no trap site
    }
    const uint64_t b = __ballot(keep);
    if (b) {
      const int leader = __ffsll((unsigned long long)b) - 1;
      uint32_t at = 0;
      if (lane == leader) at = atomicAdd(&count, (uint32_t)__popcll(b));
      at = __shfl(at, leader);
      if (keep) list[at + (uint32_t)__popcll(b & below)] = j;
    }
  }
  __syncthreads();
  const uint32_t total = count;
  for (uint32_t k0 = 0; k0 < total; k0 += blockDim.x) {  // whole waves stay in the loop for the ballot
    const uint32_t k = k0 + threadIdx.x;
    bool still = false;
    uint64_t i = 0;
    if (k < total) {
      i = base + list[k];
      const iv la{soa[0 * ld + i], soa[1 * ld + i]};
      const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
      if (trc::r1_rest(K, la, trc::weight_for(K, la), r1, soa[5 * ld + i])) first[i] = 0;
      else still = true;
    }
    const uint64_t b = __ballot(still);
    if (b) {
      const int leader = __ffsll((unsigned long long)b) - 1;
      uint32_t at = 0;
      if (lane == leader) at = atomicAdd(&open, (uint32_t)__popcll(b));
      at = __shfl(at, leader);
      if (still) {
        // base + at + rank < base + (boxes of this chunk) <= n
        const uint64_t d = base + at + (uint32_t)__popcll(b & below);
#pragma unroll
        for (int r = 0; r < 8; ++r) stage[r * n + d] = soa[r * ld + i];
        stage_id[d] = (uint32_t)i;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) pending[blockIdx.x] = open;
}

// FP32 form of ccrd_c2_first0 (the FP32 pre-filter in cascade form): the
// FP32 shadow of criterion 0 (dev/f32.hpp) on every box; a box it proves is
// one the FP64 criterion proves (first = 0, evals = 1, f32_decided = 1).
// The others are gathered in LDS and run FP64 criterion 0 densely (evals =
// 2: the FP32 attempt counts as one); those it leaves open are staged as in
// ccrd_c2_first0.  widths (may be null) as ccrd_c2_first_f32 writes them.
extern "C" __global__ void __launch_bounds__(256) ccrd_c2_first0_f32(const double* __restrict__ soa, uint64_t ld,
                                                                      uint64_t n, const Konst* __restrict__ kp,
                                                                      uint8_t* __restrict__ first,
                                                                      uint8_t* __restrict__ evals,
                                                                      uint8_t* __restrict__ trap,
                                                                      uint8_t* __restrict__ f32_decided,
                                                                      float* __restrict__ widths,
                                                                      uint32_t* __restrict__ pending,
                                                                      double* __restrict__ stage,
                                                                      uint32_t* __restrict__ stage_id) {
  __shared__ uint32_t list[kFirstChunk];
  __shared__ uint32_t count, open;
  if (threadIdx.x == 0) count = open = 0;
  __syncthreads();
  const Konst K = *kp;
  const f32::KonstF KF = f32::konst(K);
  const uint64_t base = (uint64_t)blockIdx.x * kFirstChunk;
  const int lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t j = threadIdx.x; j < kFirstChunk; j += blockDim.x) {  // kFirstChunk is a multiple of the block
    const uint64_t i = base + j;
    bool keep = false;
    if (i < n) {
      const iv la{soa[0 * ld + i], soa[1 * ld + i]};
      const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
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
      const bool dec = st.ok && f32::r1_strategies(KF, las, Ws, r1s, f32::shd(soa[5 * ld + i]), st) && st.ok;
      keep = !dec;
      first[i] = dec ? 0 : 255;
      evals[i] = dec ? 1 : 2;
      trap[i] = 0;  // criterion 0 has no trap site
      f32_decided[i] = dec ? 1 : 0;
    }
    const uint64_t b = __ballot(keep);
    if (b) {
      const int leader = __ffsll((unsigned long long)b) - 1;
      uint32_t at = 0;
      if (lane == leader) at = atomicAdd(&count, (uint32_t)__popcll(b));
      at = __shfl(at, leader);
      if (keep) list[at + (uint32_t)__popcll(b & below)] = j;
    }
  }
  __syncthreads();
  const uint32_t total = count;
  for (uint32_t k0 = 0; k0 < total; k0 += blockDim.x) {  // whole waves stay in the loop for the ballot
    const uint32_t k = k0 + threadIdx.x;
    bool still = false;
    uint64_t i = 0;
    if (k < total) {
      i = base + list[k];
      const iv la{soa[0 * ld + i], soa[1 * ld + i]};
      const iv r1{soa[2 * ld + i], soa[3 * ld + i]};
      const iv W = trc::weight_for(K, la);
      const double r2hi = soa[5 * ld + i];
      if (trc::r1_strip(K, la, W, r1, r2hi) || trc::r1_rest(K, la, W, r1, r2hi)) first[i] = 0;
      else still = true;
    }
    const uint64_t b = __ballot(still);
    if (b) {
      const int leader = __ffsll((unsigned long long)b) - 1;
      uint32_t at = 0;
      if (lane == leader) at = atomicAdd(&open, (uint32_t)__popcll(b));
      at = __shfl(at, leader);
      if (still) {
        const uint64_t d = base + at + (uint32_t)__popcll(b & below);
#pragma unroll
        for (int r = 0; r < 8; ++r) stage[r * n + d] = soa[r * ld + i];
        stage_id[d] = (uint32_t)i;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) pending[blockIdx.x] = open;
}

// Cascade: the boxes ccrd_c2_first0 left open, moved from the front of each
// block's range of stage / stage_id to positions offset[b] .. of the dense
// SoA `packed` (stride n) and of `list` (their box indices).
extern "C" __global__ void __launch_bounds__(256) ccrd_c2_first0_list(const double* __restrict__ stage,
                                                                       const uint32_t* __restrict__ stage_id,
                                                                       uint64_t n,
                                                                       const uint32_t* __restrict__ pending,
                                                                       const uint32_t* __restrict__ offset,
                                                                       double* __restrict__ packed,
                                                                       uint32_t* __restrict__ list) {
  const uint32_t cnt = pending[blockIdx.x];
  const uint64_t src0 = (uint64_t)blockIdx.x * kFirstChunk, dst0 = offset[blockIdx.x];
  for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
    const uint64_t s = src0 + t, d = dst0 + t;
#pragma unroll
    for (int r = 0; r < 8; ++r) packed[r * n + d] = stage[r * n + s];
    list[d] = stage_id[s];
  }
}
