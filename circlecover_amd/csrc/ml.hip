// Learned criterion ranker: an 8 -> 64 -> 64 -> 11 MLP (ReLU) that predicts,
// from a 4-D box, which criterion of the ten-criterion set proves it first
// (classes 0..9), or that none does (class 10).
// Weights are kept as FP32 masters and used as BF16 (round to nearest even)
// in every forward pass; accumulation is FP32.  Kernels here run in the
// default round-to-nearest mode (no ccrd_ prefix).
//
//   ccml_fwd_bwd    one lane per training sample: forward, softmax
//                   cross-entropy, backward; stores what the weight
//                   gradients need
//   ccml_grad_partial / ccml_grad_adam  the batch gradient in a fixed
//                   summation order (deterministic), then an Adam step
//   ccml_rank       batched inference on MFMA (v_mfma_f32_32x32x8bf16_1k):
//                   each wave takes 32 boxes per tile; top-1 criterion out
//   ccml_bucket_*   groups boxes by predicted criterion (counting sort)
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kIn = 8, kH = 64, kOut = 11;
// packed parameter offsets (floats)
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kOut, kParams = oB3 + kOut;
// per-sample record of ccml_fwd_bwd: x, h1, h2, dz1, dz2, dz3
constexpr int rX = 0, rH1 = rX + kIn, rH2 = rH1 + kH, rD1 = rH2 + kH, rD2 = rD1 + kH, rD3 = rD2 + kH,
              kRec = rD3 + kOut;

__device__ __forceinline__ float bf16_round(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}
__device__ __forceinline__ uint16_t bf16_bits(float x) { return (uint16_t)(__float_as_uint(bf16_round(x)) >> 16); }

// Features of box i: centre and log2 width of lambda, r1, r2, r3, each
// normalised by (value - mean) / std.  A zero width counts as 2^-60.
// Training evaluates log2 in FP64; inference (fast = true) uses the FP32
// log2 of the width rounded to FP32, which differs by float rounding only.
template <bool fast>
__device__ __forceinline__ float feature(const double* __restrict__ soa, uint64_t ld, uint64_t i, int k,
                                         const float* __restrict__ norm) {
  const double lo = soa[(2 * (k >> 1)) * ld + i], hi = soa[(2 * (k >> 1) + 1) * ld + i];
  float v;
  if (k & 1) {
    const double w = hi - lo;
    if (fast) v = w > 0x1p-60 ? log2f((float)w) : -60.0f;
    else v = (float)(w > 0x1p-60 ? log2(w) : -60.0);
  } else {
    v = (float)(0.5 * (lo + hi));
  }
  return (v - norm[k]) / norm[kIn + k];
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256) ccml_fwd_bwd(const double* __restrict__ soa, uint64_t ld,
                                                                const uint32_t* __restrict__ sample,
                                                                const uint8_t* __restrict__ label, int B,
                                                                const float* __restrict__ norm,
                                                                const float* __restrict__ params,
                                                                float* __restrict__ rec, float* __restrict__ loss) {
  __shared__ float w[kParams];
  for (int p = threadIdx.x; p < kParams; p += blockDim.x) {
    const bool bias = (p >= oB1 && p < oW2) || (p >= oB2 && p < oW3) || p >= oB3;
    w[p] = bias ? params[p] : bf16_round(params[p]);
  }
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const uint64_t i = sample[s];
  float* r = rec + (uint64_t)s * kRec;
  float x[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) {
    x[k] = bf16_round(feature<false>(soa, ld, i, k, norm));
    r[rX + k] = x[k];
  }
  float h1[kH], h2[kH], z3[kOut];
  for (int j = 0; j < kH; ++j) {
    float a = w[oB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) a += x[k] * w[oW1 + k * kH + j];
    h1[j] = a > 0.f ? a : 0.f;
    r[rH1 + j] = h1[j];
  }
  for (int j = 0; j < kH; ++j) {
    float a = w[oB2 + j];
    for (int k = 0; k < kH; ++k) a += h1[k] * w[oW2 + k * kH + j];
    h2[j] = a > 0.f ? a : 0.f;
    r[rH2 + j] = h2[j];
  }
  float zmax = -3.0e38f;
  for (int c = 0; c < kOut; ++c) {
    float a = w[oB3 + c];
    for (int k = 0; k < kH; ++k) a += h2[k] * w[oW3 + k * kOut + c];
    z3[c] = a;
    zmax = a > zmax ? a : zmax;
  }
  float se = 0.f;
  for (int c = 0; c < kOut; ++c) se += expf(z3[c] - zmax);
  const int y = label[i];
  loss[s] = logf(se) - (z3[y] - zmax);
  const float inv_b = 1.0f / (float)B;
  float d3[kOut];
  for (int c = 0; c < kOut; ++c) {
    d3[c] = (expf(z3[c] - zmax) / se - (c == y ? 1.f : 0.f)) * inv_b;
    r[rD3 + c] = d3[c];
  }
  // dz2 = (W3 dz3) * [h2 > 0]
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
    for (int c = 0; c < kOut; ++c) a += d3[c] * w[oW3 + k * kOut + c];
    const float d = h2[k] > 0.f ? a : 0.f;
    h2[k] = d;  // h2 no longer needed: reuse as dz2
    r[rD2 + k] = d;
  }
  // dz1 = (W2 dz2) * [h1 > 0]
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
    for (int j = 0; j < kH; ++j) a += h2[j] * w[oW2 + k * kH + j];
    r[rD1 + k] = h1[k] > 0.f ? a : 0.f;
  }
}

// Gradient in two deterministic passes: ccml_grad_partial sums, for every
// parameter p and slice q of kSlice consecutive samples, a[s] * b[s] over
// the slice (one lane per (p, q)); ccml_grad_adam adds the slices in order,
// then takes an Adam step (lr, beta1, beta2, eps; step t >= 1) on the FP32
// master unless update == 0.  grad (may be null) receives the raw gradient.
constexpr int kSlice = 64;

__device__ __forceinline__ void grad_operands(int p, int& a_off, int& b_off) {
  // gradient of parameter p = sum_s a[s] * b[s] (a_off < 0: a = 1)
  if (p < oB1) {
    a_off = rX + p / kH;
    b_off = rD1 + p % kH;
  } else if (p < oW2) {
    a_off = -1;
    b_off = rD1 + (p - oB1);
  } else if (p < oB2) {
    a_off = rH1 + (p - oW2) / kH;
    b_off = rD2 + (p - oW2) % kH;
  } else if (p < oW3) {
    a_off = -1;
    b_off = rD2 + (p - oB2);
  } else if (p < oB3) {
    a_off = rH2 + (p - oW3) / kOut;
    b_off = rD3 + (p - oW3) % kOut;
  } else {
    a_off = -1;
    b_off = rD3 + (p - oB3);
  }
}

extern "C" __global__ void __launch_bounds__(256) ccml_grad_partial(const float* __restrict__ rec, int B,
                                                                     float* __restrict__ partial) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (p >= kParams) return;
  int a_off, b_off;
  grad_operands(p, a_off, b_off);
  const int s0 = q * kSlice, s1 = s0 + kSlice < B ? s0 + kSlice : B;
  float g = 0.f;
  for (int s = s0; s < s1; ++s) {
    const float* r = rec + (uint64_t)s * kRec;
    g += (a_off < 0 ? 1.f : r[a_off]) * r[b_off];
  }
  partial[(uint64_t)q * kParams + p] = g;
}

extern "C" __global__ void __launch_bounds__(256) ccml_grad_adam(const float* __restrict__ partial, int nslices,
                                                                  float* __restrict__ params, float* __restrict__ m,
                                                                  float* __restrict__ v, int t, float lr,
                                                                  float beta1, float beta2, float eps, int update,
                                                                  float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  float g = 0.f;
  for (int q = 0; q < nslices; ++q) g += partial[(uint64_t)q * kParams + p];
  if (grad) grad[p] = g;
  if (!update) return;
  const float mm = beta1 * m[p] + (1.f - beta1) * g;
  const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.f - powf(beta1, (float)t));
  const float vh = vv / (1.f - powf(beta2, (float)t));
  params[p] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- inference
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// Two floats to a packed BF16 pair, rounding to nearest even (one
// v_cvt_pk_bf16_f32 on gfx950); with relu, negative inputs become 0 first.
template <bool relu>
__device__ __forceinline__ int bf16_pair(float a, float b) {
  if (relu) {
    a = __builtin_fmaxf(a, 0.f);
    b = __builtin_fmaxf(b, 0.f);
  }
  const f32x2 v = {a, b};
  return __builtin_bit_cast(int, __builtin_convertvector(v, bf16x2));
}
template <bool relu>
__device__ __forceinline__ bf16x4 bf16_quad(float a, float b, float c, float d) {
  const i32x2 w = {bf16_pair<relu>(a, b), bf16_pair<relu>(c, d)};
  return __builtin_bit_cast(bf16x4, w);
}
// Accumulator elements r0..r0+7 of h as the BF16 octet of an MFMA operand.
template <bool relu>
__device__ __forceinline__ bf16x8 bf16_oct(const f32x16& h, int r0) {
  const i32x4 w = {bf16_pair<relu>(h[r0], h[r0 + 1]), bf16_pair<relu>(h[r0 + 2], h[r0 + 3]),
                   bf16_pair<relu>(h[r0 + 4], h[r0 + 5]), bf16_pair<relu>(h[r0 + 6], h[r0 + 7])};
  return __builtin_bit_cast(bf16x8, w);
}
// The hidden unit that K index k (0..63) of layers 2 and 3 reads.  Those
// layers use v_mfma_f32_32x32x16_bf16, whose B operand gives lane l the K
// indices 16 s + 8 (l / 32) + (0..7) at step s; a lane holds the previous
// layer's units 32 t + 8 q + 4 (l / 32) + (0..3) as accumulator elements
// 4 q + (0..3) of block t.  Step s takes elements 8 (s % 2) .. + 7 of block
// s / 2 (q = 2 (s % 2) + (0, 1)), so K index 16 s + 8 g + 4 b + e is unit
// 32 (s / 2) + 16 (s % 2) + 8 b + 4 g + e, and the weight rows are read in
// that order: no activation moves between lanes.
__device__ __forceinline__ int unit_of_k(int k) {
  const int s = k >> 4, g = (k >> 3) & 1, b = (k >> 2) & 1, e = k & 3;
  return 32 * (s >> 1) + 16 * (s & 1) + 8 * b + 4 * g + e;
}

// Batched forward on MFMA, with activations kept in registers.  Operand
// layouts of v_mfma_f32_32x32x8bf16_1k (D = A B + C) for lane l, g = l / 32,
// j = l % 32: A[i][k] with i = j, k = 4g + e (element e = 0..3); B[k][j'] with
// j' = j, k = 4g + e; D[i][j'] with j' = j, i = 8 (r / 4) + 4g + (r % 4) for
// accumulator r = 0..15.  Each layer is computed transposed, D = W^T X^T
// (rows: units, columns: the wave's 32 boxes), because then the rows a lane
// holds of one layer's output, units 8s + 4g + e for step s, are exactly the
// k-slice 8s + 4g + e that it supplies as B for the next layer's step s:
// bias, ReLU and the BF16 rounding are applied in registers and no
// activation goes through LDS.  The biases enter as the accumulators' start;
// float-to-BF16 conversions use the hardware pair conversion.
// wb: BF16 weights as [W1 (8x64) | W2 (64x64) | W3 (64x10)] row-major (k, n);
// bias: FP32 [b1 | b2 | b3].  top1[i] = argmax logit (lowest index on
// ties); logits (may be null) receives the 10 logits of each box.  list
// (may be null): position i ranks box list[i] (outputs at position i); the
// positions are [0, n), n = *n_dev when n_dev is not null.  Value r of box
// b is soa[r * ld + b * bs]: a SoA of stride ld (bs = 1) or the lambda, r1,
// r2, r3 bounds at the start of 14-double cell records (ld = 1, bs = 14).
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) ccml_rank(const double* __restrict__ soa, uint64_t n, uint64_t ld, uint64_t bs,
                                                             const uint32_t* __restrict__ list,
                                                             const uint32_t* __restrict__ n_dev,
                                                             const float* __restrict__ norm,
                                                             const uint16_t* __restrict__ wb,
                                                             const float* __restrict__ bias,
                                                             uint8_t* __restrict__ top1,
                                                             float* __restrict__ logits) {
  // biases in LDS (read four consecutive units at a time as accumulator
  // starts), so that the weights and two tiles of work fit in the registers
  // of two waves per SIMD
  __shared__ float sb[2 * kH + 16];
  for (int k = threadIdx.x; k < 2 * kH + 16; k += blockDim.x) sb[k] = k < 2 * kH + kOut ? bias[k] : 0.f;
  __syncthreads();
  const int l = threadIdx.x & 63, g = l >> 5, col = l & 31, wv = threadIdx.x >> 6;
  const uint16_t* W1 = wb;
  const uint16_t* W2 = wb + kIn * kH;
  const uint16_t* W3 = W2 + kH * kH;
  bf16x4 a1[2];
  bf16x8 a2[4][2], a3[4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) a1[t][e] = (short)W1[(4 * g + e) * kH + 32 * t + col];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    u16x8 w2[2], w3;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int u = unit_of_k(16 * s + 8 * g + e);
#pragma unroll
      for (int t = 0; t < 2; ++t) w2[t][e] = W2[u * kH + 32 * t + col];
      w3[e] = col < kOut ? W3[u * kOut + col] : (unsigned short)0;
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) a2[s][t] = __builtin_bit_cast(bf16x8, w2[t]);
    a3[s] = __builtin_bit_cast(bf16x8, w3);
  }
  // accumulator start r..r+3 of the units 8q + 4g + (0..3), q = r / 4
  auto start = [&](int off, f32x16& c) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = *reinterpret_cast<const float4*>(&sb[off + 8 * q + 4 * g]);
      c[4 * q] = v.x;
      c[4 * q + 1] = v.y;
      c[4 * q + 2] = v.z;
      c[4 * q + 3] = v.w;
    }
  };
  if (n_dev) n = *n_dev;
  const uint64_t tiles = (n + 31) / 32;
  const uint64_t stride = (uint64_t)gridDim.x * 4;
  uint64_t tile = (uint64_t)blockIdx.x * 4 + wv;
  // this lane's two dimensions (2g, 2g + 1) of its box, loaded two of the
  // wave's tiles ahead into two alternating buffers: a tile's work is far
  // shorter than the memory latency under load
  double v0[4] = {0.0, 0.0, 0.0, 0.0}, v1[4] = {0.0, 0.0, 0.0, 0.0};
  auto load = [&](double (&v)[4], uint64_t t) {
    const uint64_t i = t * 32 + col;
    if (t < tiles && i < n) {
      const uint64_t box = list ? list[i] : i;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = soa[(4 * g + e) * ld + box * bs];
    }
  };
  auto step = [&](double (&v)[4], uint64_t tile) {
    const uint64_t i = tile * 32 + col;
    const bool live = i < n;
    // layer 1: B = features 4g..4g+3 of box i (centre and log2 width of
    // dimensions 2g and 2g + 1)
    float f[4];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const double lo = v[2 * d], hi = v[2 * d + 1];
      const double w = hi - lo;
      const float c = (float)(0.5 * (lo + hi));
      const float lw = w > 0x1p-60 ? log2f((float)w) : -60.0f;
      const int k = 4 * g + 2 * d;
      f[2 * d] = live ? (c - norm[k]) / norm[kIn + k] : 0.f;
      f[2 * d + 1] = live ? (lw - norm[k + 1]) / norm[kIn + k + 1] : 0.f;
    }
    const bf16x4 x = bf16_quad<false>(f[0], f[1], f[2], f[3]);
    load(v, tile + 2 * stride);
    f32x16 h1[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      start(32 * t, h1[t]);
      h1[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1[t], x, h1[t], 0, 0, 0);
    }
    // layer 2: step s takes accumulator elements 8 (s % 2) .. + 7 of h1[s / 2]
    // (units unit_of_k(16 s + 8 g + 0..7))
    f32x16 h2[2];
    start(kH, h2[0]);
    start(kH + 32, h2[1]);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const bf16x8 b = bf16_oct<true>(h1[s >> 1], 8 * (s & 1));
#pragma unroll
      for (int t = 0; t < 2; ++t) h2[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2[s][t], b, h2[t], 0, 0, 0);
    }
    // layer 3: logits of classes 8 (r / 4) + 4g + (r % 4) of box i
    f32x16 z3;
    start(2 * kH, z3);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      z3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a3[s], bf16_oct<true>(h2[s >> 1], 8 * (s & 1)), z3, 0, 0, 0);
    // classes here: g = 0 holds 0..3 and 8..10; g = 1 holds 4..7.  Best of
    // each lane in increasing class order, then the pair of lanes l, l ^ 32
    // combined (larger logit; on equal logits the lower class).
    float bv = z3[0];
    int best = 4 * g;
#pragma unroll
    for (int r = 1; r < 7; ++r) {
      const int c = 8 * (r >> 2) + 4 * g + (r & 3);
      if (c < kOut && z3[r] > bv) {
        bv = z3[r];
        best = c;
      }
    }
    const float ov = __shfl_xor(bv, 32);
    const int ob = __shfl_xor(best, 32);
    if (ov > bv || (ov == bv && ob < best)) best = ob;
    if (live) {
      if (g == 0) top1[i] = (uint8_t)best;
      if (logits) {
#pragma unroll
        for (int r = 0; r < 7; ++r) {
          const int c = 8 * (r >> 2) + 4 * g + (r & 3);
          if (c < kOut) logits[i * kOut + c] = z3[r];
        }
      }
    }
  };
  load(v0, tile);
  load(v1, tile + stride);
  for (; tile < tiles; tile += 2 * stride) {  // tile and stride are the same on every lane of the wave
    step(v0, tile);
    if (tile + stride < tiles) step(v1, tile + stride);
  }
}

// Counting sort of box indices by predicted criterion, in three passes over
// fixed block ranges (no contended global atomics): per-block bucket counts,
// a prefix over (bucket, block), then a scatter from per-block offsets.
// Order inside a bucket is unspecified.  Positions [0, n), n = *n_dev when
// n_dev is not null; block b takes the b-th of gridDim.x equal ranges.
extern "C" __global__ void __launch_bounds__(256) ccml_bucket_count(const uint8_t* __restrict__ top1, uint64_t n,
                                                                     const uint32_t* __restrict__ n_dev,
                                                                     unsigned long long* __restrict__ counts) {
  __shared__ unsigned int c[kOut];
  if (threadIdx.x < kOut) c[threadIdx.x] = 0;
  __syncthreads();
  if (n_dev) n = *n_dev;
  const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&c[top1[i] < kOut ? top1[i] : 0], 1u);
  __syncthreads();
  if (threadIdx.x < kOut) counts[(uint64_t)blockIdx.x * kOut + threadIdx.x] = c[threadIdx.x];
}

// One workgroup of 1024 lanes: offsets[b*11 + k] = first position of block
// b's boxes of bucket k (bucket-major, then block order); bstart[k] = first
// position of bucket k, bstart[11] = total.  Wave k scans bucket k's counts
// 64 blocks at a time (shuffles, no workgroup barrier inside the scan), then
// adds the bucket's base once every bucket total is known.
extern "C" __global__ void __launch_bounds__(1024) ccml_bucket_prefix(const unsigned long long* __restrict__ counts,
                                                                       uint32_t nblocks,
                                                                       unsigned long long* __restrict__ offsets,
                                                                       uint64_t* __restrict__ bstart) {
  static_assert(kOut <= 16, "one wave per bucket");
  __shared__ unsigned long long tot[kOut];
  const int w = threadIdx.x >> 6, lane = __lane_id();
  if (w < kOut) {
    unsigned long long run = 0;
    for (uint32_t c = 0; c < nblocks; c += 64) {
      const uint32_t b = c + lane;
      const unsigned long long v = b < nblocks ? counts[(uint64_t)b * kOut + w] : 0ull;
      unsigned long long x = v;  // inclusive scan over the wave
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long y = __shfl_up(x, d);
        if (lane >= d) x += y;
      }
      if (b < nblocks) offsets[(uint64_t)b * kOut + w] = run + x - v;
      run += __shfl(x, 63);
    }
    if (lane == 0) tot[w] = run;
  }
  __syncthreads();
  if (w < kOut) {
    unsigned long long base = 0;
    for (int k = 0; k < w; ++k) base += tot[k];
    // each lane revisits the entries it wrote above
    for (uint32_t b = lane; b < nblocks; b += 64) offsets[(uint64_t)b * kOut + w] += base;
    if (lane == 0) {
      bstart[w] = base;
      if (w == kOut - 1) bstart[kOut] = base + tot[w];
    }
  }
}

// list (may be null): position i stands for box list[i].  src (may be
// null): a packed SoA of stride ld by position, copied into dst in bucket
// order along with the indices.
extern "C" __global__ void __launch_bounds__(256) ccml_bucket_scatter(const uint8_t* __restrict__ top1, uint64_t n,
                                                                       const uint32_t* __restrict__ n_dev,
                                                                       const unsigned long long* __restrict__ offsets,
                                                                       const uint32_t* __restrict__ list,
                                                                       uint32_t* __restrict__ idx,
                                                                       const double* __restrict__ src, uint64_t ld,
                                                                       double* __restrict__ dst) {
  __shared__ unsigned long long cur[kOut];
  if (threadIdx.x < kOut) cur[threadIdx.x] = offsets[(uint64_t)blockIdx.x * kOut + threadIdx.x];
  __syncthreads();
  if (n_dev) n = *n_dev;
  const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const int b = top1[i] < kOut ? top1[i] : 0;
    const unsigned long long at = atomicAdd(&cur[b], 1ull);
    idx[at] = list ? list[i] : (uint32_t)i;
    if (src)
#pragma unroll
      for (int r = 0; r < 8; ++r) dst[r * ld + at] = src[r * ld + i];
  }
}
