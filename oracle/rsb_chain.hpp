// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// rectangle_size_bound leaf-criterion chain (reference strategies.cuh:33-147).
#pragma once
#include "criteria.hpp"

namespace ccor {
namespace rsb {

struct TwoByTwo {  // variables.hpp:120-123
  Circle circles[4];
  double width;
};

struct Vals {  // Intermediate_values, variables.hpp:125-138
  IV R;
  IV min_width_factor, max_width_factor, critical_ratio_sq;
  double weight_bound_wall_building;
  double width_fraction_incomplete;
  IV phase_factor_wall_building;
  double lb_w_infty_denom;
  TwoByTwo cover_2x2;
};

struct Strip2 {  // Max_height_strip_2 / Max_height_strip_wc_recursion_2
  Circle c1, c2;
  double lb_height;
};

Cell tighten(const Cell& c);
bool is_empty(const Cell& c);
Vals make_vals(const Cell& c);
Strip2 two_disks_maximal_height_strip(const IV& la, const IV& r1, const IV& r2);
Strip2 two_disks_maximal_height_strip_wc_recursion(const IV& r1, const IV& r2, const IV& R, double ub_w);
double six_disks_maximize_covered_width(const Cell& c);
TwoByTwo compute_two_by_two_cover(const IV* disks);

// C1..C14 in the reference's leaf order (rectangle_size_bound.cu:201-229)
bool even_split_recursion(const Cell& c, const Vals& v);                        // C1
bool uneven_split_recursion(const Cell& c, const Vals& v);                      // C2
bool multi_disk_strip_vertical(const Cell& c, const Vals& v);                   // C3
bool vertical_wall_building_recursion(const Cell& c, const Vals& v);            // C4
bool r1_in_corner_explicit_recursion(const Cell& c, const Vals& v);             // C5
bool r1_r2_opposite_corners_recursion(const Cell& c, const Vals& v);            // C6
bool r1_r2_vertical_strip(const Cell& c, const Vals& v);                        // C7
bool two_by_two_cover_with_strip_and_recursion(const Cell& c, const Vals& v);   // C8
bool l_shaped_recursion(const Cell& c, const Vals& v);                          // C9
bool r1_in_corner_wall_building_recursion(const Cell& c, const Vals& v);        // C10
bool advanced_multi_disk_strip_vertical(const Cell& c, const Vals& v);          // C11
bool five_disk_full_cover(const Cell& c, const Vals& v);                        // C12
bool six_disk_full_cover(const Cell& c, const Vals& v);                         // C13
bool seven_disk_strategies(const Cell& c, const Vals& v);                       // C14
void seven_disk_probe(const Cell& in, const double* pieces, int np, double* out, uint32_t* bits);

// Verdict of one 7-D cell exactly as the reference's enumeration kernel
// treats a leaf (rectangle_size_bound.cu:162-229): cell-level shortcuts S1,S2,
// r3-level shortcuts S3..S5, then tighten, Intermediate_values and C1..C14.
// Returns the index of the first succeeding step (0..19: S1..S5 = 0..4,
// empty/vacuous = 5, C1..C14 = 6..19) or -1 when the cell stays unproven.
int cell_verdict(const Cell& c);
// cell_verdict without the shortcut steps S1..S5 (same return codes).
int chain_verdict(const Cell& c);
// All steps evaluated independently: bit k = step k succeeds, bit 20+k = trap.
// step_ops (20 entries, may be null) receives the directed-op count of each
// step in the counting build (the tighten/vals set-up is charged to step 5).
uint64_t cell_mask(const Cell& c, uint64_t* step_ops = nullptr);

}  // namespace rsb
}  // namespace ccor
