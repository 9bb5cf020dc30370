// Cascade form of first-success evaluation, first stage (gfx950,
// round-down mode): ccrd_c2_first0 tries criterion 0 on every box,
// ccrd_c2_first0_list packs the boxes it leaves open into a dense SoA, and
// ccrd_c2_first (c2_first.hip, skip0 = 1) runs the other criteria on those
// only, in plain or bucket order.  The first proving criterion in plain
// order is criterion 0 for every box criterion 0 proves, so the plain
// cascade gives plain order's first/evals/trap exactly.  Built with the
// can_recurse helpers inlined, as the batch kernel is (no call stack).
#define CC_INLINE_RECURSE 1
#define CC_EXACT_SHORTCUTS 1
#include "dev/c2_criteria.hpp"

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
