// Plain streaming copy (16 bytes per lane and iteration): the calibration
// kernel of known traffic for the HBM counters (FETCH_SIZE / WRITE_SIZE).
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" __global__ void __launch_bounds__(256) cc_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                              uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Zero n 64-bit words (counters cleared on the stream by a kernel rather
// than a memset node, so the sequence can also be captured in a graph).
extern "C" __global__ void __launch_bounds__(64) cc_zero_u64(unsigned long long* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}
