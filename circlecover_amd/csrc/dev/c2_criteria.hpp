// Criterion k of the ten-criterion 4-D box set (the definitions of the batch
// kernel, see c2_eval.hip), selected by index for first-success evaluation.
#pragma once
#include "rsb.hpp"

namespace ccd {

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

}  // namespace ccd
