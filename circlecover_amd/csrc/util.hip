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

// Exclusive prefix sum of in[0..m) into out[0..m), total in out[m]; one
// workgroup of 1024 lanes walking the input 1024 values at a time.
extern "C" __global__ void __launch_bounds__(1024) cc_scan_u32(const uint32_t* __restrict__ in, uint32_t m,
                                                                uint32_t* __restrict__ out) {
  __shared__ uint32_t sc[1024];
  const uint32_t t = threadIdx.x;
  uint32_t base = 0;
  for (uint32_t c = 0; c < m; c += 1024) {
    const uint32_t v = c + t < m ? in[c + t] : 0u;
    sc[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
      const uint32_t add = t >= d ? sc[t - d] : 0u;
      __syncthreads();
      sc[t] += add;
      __syncthreads();
    }
    if (c + t < m) out[c + t] = base + sc[t] - v;
    base += sc[1023];
    __syncthreads();
  }
  if (t == 0) out[m] = base;
}
