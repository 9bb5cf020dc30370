// First-success evaluation of the ten-criterion 4-D box set (gfx950,
// round-down mode).  Boxes are visited through an index list grouped into
// buckets; bucket b tries criterion b first and then the others in plain
// order (0..9), stopping at the first criterion that proves the box.  Since
// a wave's boxes come from one bucket (except at the few bucket borders),
// the wave runs one criterion at a time instead of a different one per lane.
// With every box in bucket 0 this is the plain order.
#define CC_EXACT_SHORTCUTS 1
#include "dev/c2_criteria.hpp"

using namespace ccd;

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

