// rectangle_size_bound level kernel (gfx950, round-down mode): the
// widest-dimension branch-and-bound over 7-D cells, one launch per depth.
// This is the plain baseline of the batched form (rsb_fast.hip): it runs
// the reference's operation sequence without the exact shortcuts
// (rsb_bnb_x.hip builds the same kernel with them).
#include "dev/rsb_level.hpp"

using namespace ccd;

// One level of the branch-and-bound: one lane per cell of `in` (the body
// and the meaning of every argument: bnb_level_body in dev/rsb_level.hpp).
extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_bnb_level(const rsb_rec* __restrict__ in, uint64_t n,
                                                                      int at_limit, const Konst* __restrict__ kp,
                                                                      rsb_rec* __restrict__ children,
                                                                      rsb_rec* __restrict__ unproven,
                                                                      uint64_t cap_unproven,
                                                                      unsigned long long* __restrict__ counts,
                                                                      unsigned long long* __restrict__ hist,
                                                                      bnb_log_rec* __restrict__ log, uint64_t log_cap,
                                                                      int depth) {
  bnb_level_body(in, n, at_limit, kp, children, unproven, cap_unproven, counts, hist, log, log_cap, depth);
}
