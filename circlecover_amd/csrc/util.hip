// Plain streaming copy (16 bytes per lane and iteration): the calibration
// kernel of known traffic for the HBM counters (FETCH_SIZE / WRITE_SIZE).
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" __global__ void __launch_bounds__(256) cc_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                              uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
