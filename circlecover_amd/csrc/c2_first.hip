// First-success evaluation of the ten-criterion 4-D box set (gfx950,
// round-down mode).  Boxes are visited through an index list grouped into
// buckets; bucket b < 10 tries criterion b first and then the others in
// plain order (0..9), stopping at the first criterion that proves the box;
// bucket 10 (boxes predicted to be unprovable) uses plain order.  Since
// a wave's boxes come from one bucket (except at the few bucket borders),
// the wave runs one criterion at a time instead of a different one per lane.
// With every box in bucket 0 this is the plain order.
//
// Cascade form (c2_cascade.hip): criterion 0 first on every box, then this
// kernel with done0 > 0 on the packed boxes it leaves open.
#define CC_EXACT_SHORTCUTS 1
#include "dev/c2_criteria.hpp"

using namespace ccd;

// Positions [0, n) (n = *n_dev when n_dev is not null, a count another
// kernel wrote), grid-stride.  idx[pos] = box index (idx null: pos); the
// box's data is read at the box index, or at pos when by_pos = 1 (a packed
// SoA).  bucket_start[b] = first position of bucket b (0..10),
// bucket_start[11] = n (bucket_start null: every position in bucket 0).
// first[box] = proving criterion or 255, evals[box] = criteria evaluated,
// trap[box] = 1 if an evaluated criterion reached a trap site.  done0 > 0:
// criterion 0, which has no trap site, was already tried and failed in
// ccrd_c2_first0 (done0 = 1) or ccrd_c2_first0_f32 (done0 = 2, the FP32
// attempt included), so it is skipped here and evals starts at done0.
extern "C" __global__ void __launch_bounds__(256) ccrd_c2_first(const double* __restrict__ soa, uint64_t ld,
                                                                 const uint32_t* __restrict__ idx, uint64_t n,
                                                                 const uint32_t* __restrict__ n_dev,
                                                                 const uint64_t* __restrict__ bucket_start,
                                                                 const Konst* __restrict__ kp,
                                                                 uint8_t* __restrict__ first,
                                                                 uint8_t* __restrict__ evals,
                                                                 uint8_t* __restrict__ trap, int done0, int by_pos) {
  const uint64_t cnt = n_dev ? *n_dev : n;
  const Konst K = *kp;
  for (uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < cnt;
       pos += (uint64_t)gridDim.x * blockDim.x) {
    int b = 0;
    if (bucket_start)
      while (b < 10 && pos >= bucket_start[b + 1]) ++b;
    const uint64_t i = idx ? idx[pos] : pos;
    const uint64_t d = by_pos ? pos : i;
    const iv la{soa[0 * ld + d], soa[1 * ld + d]};
    const iv r1{soa[2 * ld + d], soa[3 * ld + d]};
    const iv r2{soa[4 * ld + d], soa[5 * ld + d]};
    const iv r3{soa[6 * ld + d], soa[7 * ld + d]};
    const iv W = trc::weight_for(K, la);
    const double rem4 = (W - r1 - r2 - r3).hi;
    const double ub_r4 = r3.hi < rem4 ? r3.hi : rem4;
    Trap tp{0};
    int found = 255, count = done0;
    for (int s = 0; s < 10; ++s) {
      // step 0 tries the bucket's criterion, then 0..9 without it
      const int k = b == 10 ? s : s == 0 ? b : (s - 1 < b ? s - 1 : s);
      if (done0 && k == 0) continue;
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
}
