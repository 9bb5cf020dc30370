// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// Declarations of the restated covering criteria.
#pragma once
#include "iv.hpp"

namespace ccor {

// Proof constants (reference Prover/rectangle_size_bound/values.hpp:28-35).
static const double kRsbLambdaLo = 1.0;
static const double kRsbLambdaHi = 2.5;
static const double kRsbRatio = 0.61;        // the double is below 61/100
static const double kDiskRadiusMax = 3.0 / 8.0;
static const double kDiskWeightMax = 9.0 / 64.0;
static const double kWorstRatio = 195.0 / 256.0;
// Two literals for the tolerated skew, one ulp apart (reference quirk):
// tight_rectangle.cu:462 uses ...d7p+1, the size-bound code ...d6p+1.
static const double kTrLambdaMax = 2.0898841580413818342;
static const double kRsbToleratedLambda = 2.0898841580413813901;
// Bisections over doubles end within ~2300 halvings for finite bounds; the
// cap only matters for NaN bounds, on which the reference would spin forever.
static const int kBisectCap = 4096;

namespace tr {
struct Cell {  // reference Prover/tight_rectangle/strategies.cuh:33-50
  IV la;
  IV r[4];
};
IV critical_ratio(const IV& la);
IV critical_ratio(const IV& w, const IV& h);
IV required_weight_for(const IV& w, const IV& h);
IV required_weight_for(const IV& la);
double required_weight_for_bounded_size(double ub_w, double ub_h);
bool can_recurse(double ub_w, double ub_h, double ub_largest, double lb_rem);
bool r1_strategies(const IV& la, const IV& r1, double ub_r2);
bool r1_r2_strategies_simple(const IV& la, const IV& r1, const IV& r2, double ub_r3);
bool r1_r2_strategies_stacked(const IV& la, const IV& r1, const IV& r2, double ub_r3);
bool r1_r2_strategies(const IV& la, const IV& r1, const IV& r2, double ub_r3);
bool r1_r2_r3_strategy_horizontal(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4);
bool r1_r2_r3_strategy_cover_strip(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4);
bool r1_r2_r3_strategies(const IV& la, const IV& r1, const IV& r2, const IV& r3, double ub_r4);
bool r1_r2_r3_strategies(const Cell& c, const IV& R4);
bool two_by_two_strategies(const Cell& c, const IV& R);
bool r1_r2_large_r3_r4_gaps_strategy(const Cell& c, const IV& R);
uint32_t leaf_strategy_bits(const Cell& c, const IV& R4, const IV& R);
IV r1_range(const IV& la);
IV r2_range(const IV& la, const IV& r1);
IV r3_range(const IV& la, const IV& r1, const IV& r2);
IV r4_range(const IV& la, const IV& r1, const IV& r2, const IV& r3);
}  // namespace tr

namespace rsb {
struct Cell {  // reference Prover/rectangle_size_bound/variables.hpp:33-110
  IV la;
  IV r[6];
};
bool disk_satisfies_size_bound(double r_sq_ub, double h_lb);
double bound_required_height(double r_sq_ub);
double bound_allowed_weight(double h_lb);
double two_disks_maximize_height(const IV& r1, const IV& r2, double ub_w);
double three_disks_maximize_height(const IV& r1, const IV& r2, const IV& r3, double ub_w);
bool three_disks_can_cover(const IV& r1, const IV& r2, const IV& r3, double w, double h);
bool nd_can_cover(double ub_w, double ub_h, const IV* disks, int n);
double nd_maximize_covered(double ub_w, const IV* disks, int n);
bool can_recurse(double lb_rem, double ub_largest, double ub_w, double ub_h);
double bound_critical_ratio(double lb_rem, double ub_largest, double ub_w, double ub_h);
double recursion_bound_size(double lb_rem, double ub_w, double ub_h);
double bound_worst_case_ratio(const IV& w, const IV& h);
bool can_recurse_worst_case(double lb_rem, double ub_w, double ub_h);
IV compute_efficient_rectangle_cover_weight_range(double lb_w, const IV& h);
bool shortcut_even_split_recursion(const IV& la, const IV& r1, const IV& r2);
bool shortcut_uneven_split_recursion(const IV& la, const IV& r1, const IV& r2);
bool shortcut_uneven_split_recursion(const IV& la, const IV& r1, const IV& r2, const IV& r3);
bool shortcut_r1_r2_opposite_corners_recursion(const IV& la, const IV& r1, const IV& r2, const IV& r3);
bool shortcut_r1_r2_vertical_strip(const IV& la, const IV& r1, const IV& r2, const IV& r3);
}  // namespace rsb

}  // namespace ccor
