// The level kernel of rsb_bnb.hip built with the exact shortcuts of the
// batched form (bisections whose test fails at their smallest argument
// exit at once and return what the full loop returns), so that the
// level-synchronous and batched forms can be compared with the same
// arithmetic per cell.
#define CC_EXACT_SHORTCUTS 1
#include "dev/rsb_level.hpp"

using namespace ccd;

extern "C" __global__ void __launch_bounds__(128) ccrd_rsb_bnb_level_x(const rsb_rec* __restrict__ in, uint64_t n,
                                                                        int at_limit, const Konst* __restrict__ kp,
                                                                        rsb_rec* __restrict__ children,
                                                                        rsb_rec* __restrict__ unproven,
                                                                        uint64_t cap_unproven,
                                                                        unsigned long long* __restrict__ counts,
                                                                        unsigned long long* __restrict__ hist,
                                                                        bnb_log_rec* __restrict__ log,
                                                                        uint64_t log_cap, int depth) {
  bnb_level_body(in, n, at_limit, kp, children, unproven, cap_unproven, counts, hist, log, log_cap, depth);
}
