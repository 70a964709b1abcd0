// Workgroup-wide reductions and prefix sums built from wave64 shuffles.
// lds must hold blockDim.x/64 + 1 words; blockDim.x is a multiple of 64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace svd {

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ uint32_t block_reduce_sum(uint32_t v, uint32_t* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  uint32_t s = 0;
  for (int w = 0; w < nw; ++w) s += lds[w];
  __syncthreads();
  return s;
}

// Exclusive prefix of v over the workgroup; agg = workgroup total.
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* lds, uint32_t& agg) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t inc = wave_inclusive_sum(v);
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    uint32_t x = lds[w];
    if (w < wid) before += x;
    tot += x;
  }
  __syncthreads();
  agg = tot;
  return before + inc - v;
}

}  // namespace svd
