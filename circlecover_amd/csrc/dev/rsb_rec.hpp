// The 112-byte 7-D cell record of the device queues (the reference's
// rectangle_size_bound::Variables layout, variables.hpp:33-110).
#pragma once
#include "rsb_verdict.hpp"

namespace ccd {

struct rsb_rec {
  double v[14];
};

CCD void store_cell7(rsb_rec* __restrict__ out, const cell7& c) {
  rsb_rec r;
  r.v[0] = c.la.lo;
  r.v[1] = c.la.hi;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    r.v[2 + 2 * k] = c.r[k].lo;
    r.v[3 + 2 * k] = c.r[k].hi;
  }
  *out = r;
}

}  // namespace ccd
