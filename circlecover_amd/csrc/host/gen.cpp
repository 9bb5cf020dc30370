// Synthetic box generator for the batched-criterion workload (host only).
// Built with -ffp-contract=off: the arithmetic below is plain
// round-to-nearest, and the upper bounds are rounded upward exactly via an
// error-free two-sum.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "circlecover_amd.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * 0x1.0p-53; }  // exact, in [0,1)
};

// lb + w rounded toward +infinity (w > 0).
double add_up(double lb, double w) {
  double s = lb + w;
  double bb = s - lb;
  double err = (lb - (s - bb)) + (w - bb);  // exact error of the rounded sum
  return err > 0.0 ? std::nextafter(s, INFINITY) : s;
}

}  // namespace

extern "C" void cc_gen_c2_boxes(uint64_t seed, size_t n, double* soa) {
  SplitMix64 g{seed};
  for (size_t i = 0; i < n; ++i) {
    double la = 1.0 + 1.5 * g.unit();
    double r[3];
    for (double& v : r) v = 0.02 + 0.68 * g.unit();
    std::sort(r, r + 3, [](double a, double b) { return a > b; });
    double lbs[4] = {la, r[0], r[1], r[2]};
    for (int d = 0; d < 4; ++d) {
      double w = std::exp2(-20.0 + 14.0 * g.unit());
      soa[(2 * d) * n + i] = lbs[d];
      soa[(2 * d + 1) * n + i] = add_up(lbs[d], w);
    }
  }
}

extern "C" {

// get_subinterval (Prover/subintervals.hpp:32-43) on the host: the n pieces of
// [lb, ub] as 2n doubles, lb + diff*i fused (one rounding) as on the device.
void cc_subintervals(double lb, double ub, int n, double* out) {
  const double diff = (ub - lb) / (double)n;
  for (int i = 0; i < n; ++i) {
    out[2 * i] = std::fma(diff, (double)i, lb);
    out[2 * i + 1] = (i + 1 == n) ? ub : std::fma(diff, (double)(i + 1), lb);
  }
}

}  // extern "C"
