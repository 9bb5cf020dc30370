// Plain (reference-shaped) enumeration kernels of both proofs, built with
// the reference's operation sequence.
#define CC_EXACT_SHORTCUTS 0
#include "dev/plain_enum.hpp"

using namespace ccd;

extern "C" __global__ void __launch_bounds__(256) ccrd_tr_plain(const double* __restrict__ cells, int n1, int n2,
                                                                  int n3, int n4, const Konst* __restrict__ kp,
                                                                  trc::cell5* __restrict__ leaves, uint64_t cap,
                                                                  uint32_t* __restrict__ leaf_cell,
                                                                  uint32_t* __restrict__ leaf_piece,
                                                                  unsigned long long* __restrict__ count,
                                                                  int* __restrict__ trap) {
  const Konst K = *kp;
  tr_plain_cell(cells, n1, n2, n3, n4, K, leaves, cap, leaf_cell, leaf_piece, count, trap);
}

// two waves per SIMD (at most 256 VGPRs): the loop nest needs about that
// many, and one register above it would halve the occupancy
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) ccrd_rsb_plain(const rsb_rec* __restrict__ cells, int n3, int n4,
                                                                   int n5, int n6, const Konst* __restrict__ kp,
                                                                   rsb_rec* __restrict__ leaves, uint64_t cap,
                                                                   unsigned long long* __restrict__ count,
                                                                   int* __restrict__ trap) {
  const Konst K = *kp;
  rsb_plain_cell(cells, n3, n4, n5, n6, K, leaves, cap, count, trap);
}
