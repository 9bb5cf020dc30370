// rectangle_size_bound verdict of one 7-D cell in the reference's order
// (Prover/rectangle_size_bound/rectangle_size_bound.cu:162-229).  Return
// codes: index of the first succeeding step (S1..S5 = 0..4, empty/vacuous
// cell = 5, C1..C14 = 6..19) or -1 when the cell is not proven; `tp` is set
// when a reference trap site was reached.
#pragma once
#include "rsb_chain.hpp"

namespace ccd {

// The leaf part (rectangle_size_bound.cu:186-229): tighten, the remaining
// weight R, emptiness, then C1..C14.
CCD int rsb_chain_verdict(const Konst& K, const cell7& in, Trap& tp) {
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  if (rsbc::is_empty(c) || empty(v.R)) return 5;
  if (rsbc::c1_even_split(c, v)) return 6;
  if (rsbc::c2_uneven_split(c)) return 7;
  if (rsbc::c3_multi_disk_strip(c, v, tp)) return 8;
  if (rsbc::c4_wall_building(K, c, v)) return 9;
  if (rsbc::c5_corner_explicit(c, v, tp)) return 10;
  if (rsbc::c6_opposite_corners(c, v, tp)) return 11;
  if (rsbc::c7_vertical_strip(K, c, v)) return 12;
  const rsbc::cover22 cov = rsbc::two_by_two(c, tp);
  v.cover_width = cov.width;
  v.cover_c4 = cov.c4;
  if (cov.width > 0 && (cov.width > c.la.hi || rsbc::c8_two_by_two_strip(K, c, v, tp))) return 13;
  if (rsbc::c9_l_shaped(K, c)) return 14;
  if (rsbc::c10_corner_wall(K, c, tp)) return 15;
  if (rsbc::c11_advanced_strip(c)) return 16;
  if (rsbc::c12_five_disk(K, c, v, tp)) return 17;
  if (rsbc::c13_six_disk(K, c, v, tp)) return 18;
  if (rsbc::c14_seven_disk(K, c, v, tp)) return 19;
  return -1;
}

// The leaf chain of rsb_chain_verdict up to C13: its verdict if one of
// those steps proves the cell, -2 if only C14 is left to try.
CCD int rsb_chain_before_c14(const Konst& K, const cell7& in, Trap& tp) {
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  if (rsbc::is_empty(c) || empty(v.R)) return 5;
  if (rsbc::c1_even_split(c, v)) return 6;
  if (rsbc::c2_uneven_split(c)) return 7;
  if (rsbc::c3_multi_disk_strip(c, v, tp)) return 8;
  if (rsbc::c4_wall_building(K, c, v)) return 9;
  if (rsbc::c5_corner_explicit(c, v, tp)) return 10;
  if (rsbc::c6_opposite_corners(c, v, tp)) return 11;
  if (rsbc::c7_vertical_strip(K, c, v)) return 12;
  const rsbc::cover22 cov = rsbc::two_by_two(c, tp);
  v.cover_width = cov.width;
  v.cover_c4 = cov.c4;
  if (cov.width > 0 && (cov.width > c.la.hi || rsbc::c8_two_by_two_strip(K, c, v, tp))) return 13;
  if (rsbc::c9_l_shaped(K, c)) return 14;
  if (rsbc::c10_corner_wall(K, c, tp)) return 15;
  if (rsbc::c11_advanced_strip(c)) return 16;
  if (rsbc::c12_five_disk(K, c, v, tp)) return 17;
  if (rsbc::c13_six_disk(K, c, v, tp)) return 18;
  return -2;
}

// Steps 0..18 of rsb_cell_verdict (everything before C14): the verdict if
// one of them proves the cell, -2 if only C14 is left to try.
CCD int rsb_verdict_before_c14(const Konst& K, const cell7& in, Trap& tp) {
  if (rsbc::s1_even_split(in.la, in.r[0], in.r[1])) return 0;
  if (rsbc::s23_uneven_split(in.la, in.r[0], in.r[1])) return 1;
  if (rsbc::s23_uneven_split(in.la, in.r[0], in.r[2])) return 2;
  if (rsbc::s4_opposite_corners(in.la, in.r[0], in.r[1], in.r[2])) return 3;
  if (rsbc::s5_vertical_strip(K, in.la, in.r[0], in.r[1], in.r[2])) return 4;
  return rsb_chain_before_c14(K, in, tp);
}

// C14 alone on a cell that reached it: 19 if it proves the cell, else -1.
CCD int rsb_verdict_c14(const Konst& K, const cell7& in, Trap& tp) {
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  v.cover_width = 0.0;
  return rsbc::c14_seven_disk(K, c, v, tp) ? 19 : -1;
}

// Step s (0..18) of rsb_verdict_before_c14 alone: S1..S5 on the untightened
// cell `in`, the emptiness test and C1..C13 on its tightened form `c` with
// remaining weight v.R.  C8's 2x2 cover is computed when first needed.
CCD bool rsb_step_before_c14(const Konst& K, const cell7& in, const cell7& c, vals7& v, bool& have_cover, int s,
                             Trap& tp) {
  switch (s) {
    case 0: return rsbc::s1_even_split(in.la, in.r[0], in.r[1]);
    case 1: return rsbc::s23_uneven_split(in.la, in.r[0], in.r[1]);
    case 2: return rsbc::s23_uneven_split(in.la, in.r[0], in.r[2]);
    case 3: return rsbc::s4_opposite_corners(in.la, in.r[0], in.r[1], in.r[2]);
    case 4: return rsbc::s5_vertical_strip(K, in.la, in.r[0], in.r[1], in.r[2]);
    case 5: return rsbc::is_empty(c) || empty(v.R);
    case 6: return rsbc::c1_even_split(c, v);
    case 7: return rsbc::c2_uneven_split(c);
    case 8: return rsbc::c3_multi_disk_strip(c, v, tp);
    case 9: return rsbc::c4_wall_building(K, c, v);
    case 10: return rsbc::c5_corner_explicit(c, v, tp);
    case 11: return rsbc::c6_opposite_corners(c, v, tp);
    case 12: return rsbc::c7_vertical_strip(K, c, v);
    case 13: {
      if (!have_cover) {
        const rsbc::cover22 cov = rsbc::two_by_two(c, tp);
        v.cover_width = cov.width;
        v.cover_c4 = cov.c4;
        have_cover = true;
      }
      return v.cover_width > 0 && (v.cover_width > c.la.hi || rsbc::c8_two_by_two_strip(K, c, v, tp));
    }
    case 14: return rsbc::c9_l_shaped(K, c);
    case 15: return rsbc::c10_corner_wall(K, c, tp);
    case 16: return rsbc::c11_advanced_strip(c);
    case 17: return rsbc::c12_five_disk(K, c, v, tp);
    case 18: return rsbc::c13_six_disk(K, c, v, tp);
    default: return false;
  }
}

// rsb_verdict_before_c14 with step `first` (0..18; -1: none) tried before
// the others, which follow in the reference's order.  A cell is proven here
// exactly when rsb_verdict_before_c14 proves it (a step proves only what
// the cell's verdict proves: an empty cell, vacuous in the reference's
// order, is proven either way); the step reported is the first that
// succeeds in this order, and trap sites of steps tried out of order can be
// reached.
CCD int rsb_verdict_before_c14_ranked(const Konst& K, const cell7& in, int first, Trap& tp) {
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  v.cover_width = 0.0;
  bool have_cover = false;
  for (int k = -1; k < 19; ++k) {
    const int s = k < 0 ? first : k;
    if (s < 0 || (k >= 0 && s == first)) continue;
    if (rsb_step_before_c14(K, in, c, v, have_cover, s, tp)) return s;
  }
  return -2;
}

// C14 in two parts, for the batched driver, which runs the prefix and piece
// 0 on one lane per cell and each further piece on a lane of its own:
// rsb_c14_head returns 19 if the prefix proves the cell, -1 if piece 0 is
// not covered (the cell is not proven), and -3 if pieces 1..15 decide
// (`s` set).  The cell is proven exactly when every piece is covered; the
// trap sites the reference reaches are those of the prefix and of the
// pieces up to the first uncovered one.
CCD int rsb_c14_head(const Konst& K, const cell7& in, Trap& tp, rsbc::c14_state& s) {
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  v.cover_width = 0.0;
  if (rsbc::c14_prefix(K, c, v, tp, s)) return 19;
  return rsbc::c14_piece(K, c, v, s, 0, tp) ? -3 : -1;
}
CCD bool rsb_c14_piece(const Konst& K, const cell7& in, const rsbc::c14_state& s, int k, Trap& tp) {
  const cell7 c = rsbc::tighten(in);
  vals7 v;
  v.R = rsbc::remaining_weight(c);
  v.cover_width = 0.0;
  return rsbc::c14_piece(K, c, v, s, k, tp);
}

// The cell-level shortcuts S1..S5 (rectangle_size_bound.cu:162 and :174-176)
// on the untightened cell, then the leaf part.
CCD int rsb_cell_verdict(const Konst& K, const cell7& in, Trap& tp) {
  if (rsbc::s1_even_split(in.la, in.r[0], in.r[1])) return 0;
  if (rsbc::s23_uneven_split(in.la, in.r[0], in.r[1])) return 1;
  if (rsbc::s23_uneven_split(in.la, in.r[0], in.r[2])) return 2;
  if (rsbc::s4_opposite_corners(in.la, in.r[0], in.r[1], in.r[2])) return 3;
  if (rsbc::s5_vertical_strip(K, in.la, in.r[0], in.r[1], in.r[2])) return 4;
  return rsb_chain_verdict(K, in, tp);
}

CCD cell7 load_cell7(const double* __restrict__ p) {
  cell7 c;
  c.la = iv{p[0], p[1]};
#pragma unroll
  for (int k = 0; k < 6; ++k) c.r[k] = iv{p[2 + 2 * k], p[3 + 2 * k]};
  return c;
}

}  // namespace ccd
