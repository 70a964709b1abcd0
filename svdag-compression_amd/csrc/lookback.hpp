// Single-pass prefix sums over tiles with decoupled look-back.
//
// A workgroup takes its tile number from a device counter (so every tile it
// waits for belongs to a workgroup that is already running), publishes the
// tile's aggregate, then walks back over the published words of the tiles
// before it, a workgroup's width at a time, summing aggregates until it meets an inclusive
// prefix, and publishes its own inclusive prefix.  A published word holds
// epoch (30 bits) | flag (2 bits) | value (32 bits): every pass over the
// status array uses a fresh epoch, so words of earlier passes read as not
// yet published and the array is never cleared.  The workgroup that takes
// the last tile number of a pass sets the counter back to zero for the next
// pass on the stream.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace svd {

constexpr unsigned long long kLbAggregate = 1ull << 32;
constexpr unsigned long long kLbInclusive = 2ull << 32;

struct LbPass {
  unsigned long long* status;   // one word per tile
  uint32_t* counter;            // zero between passes
  uint32_t epoch;               // 1 .. 2^30 - 1
  uint32_t tiles;
};

__device__ __forceinline__ void lb_publish(const LbPass& p, uint32_t tile, unsigned long long flag, uint32_t v) {
  const unsigned long long w = ((unsigned long long)p.epoch << 34) | flag | v;
  __hip_atomic_store(&p.status[tile], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile number of the calling workgroup (all threads call; one atomic).
__device__ __forceinline__ uint32_t lb_take_tile(const LbPass& p, uint32_t* ldsWord) {
  if (threadIdx.x == 0) {
    const uint32_t t = atomicAdd(p.counter, 1u);
    if (t == p.tiles - 1) atomicExch(p.counter, 0u);   // the last number: reset for the next pass
    *ldsWord = t;
  }
  __syncthreads();
  const uint32_t t = *ldsWord;
  __syncthreads();
  return t;
}

// Exclusive prefix of tile `tile` whose own sum is agg; all threads call,
// every one gets the prefix.  Publishes the tile's inclusive prefix.  The
// whole workgroup looks back, blockDim.x tiles per round trip: thread j reads
// the word of tile look - j.  lds: blockDim.x / 64 + 2 words.
// first: the first tile of the calling tile's scan (several independent scans
// can share one pass, one after another; the tiles before `first` read as
// an inclusive zero).
// The caller may still be reading lds (a workgroup total it staged there):
// the barrier on entry orders those reads before the look-back's writes, or
// a wave that reached the look-back first would overwrite a word a slower
// wave has not read yet.
__device__ __forceinline__ uint32_t lb_exclusive(const LbPass& p, uint32_t tile, uint32_t agg, uint32_t* lds,
                                                 uint32_t first = 0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  uint32_t excl = 0;
  __syncthreads();
  if (tile == first) {
    if (tid == 0) lb_publish(p, tile, kLbInclusive, agg);
  } else {
    if (tid == 0) lb_publish(p, tile, kLbAggregate, agg);
    int64_t look = (int64_t)tile - 1;
    while (true) {
      const int64_t q = look - tid;
      unsigned long long w = kLbInclusive;   // before the scan's first tile: an inclusive zero
      if (q >= (int64_t)first) {
        while (true) {
          w = __hip_atomic_load(&p.status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((uint32_t)(w >> 34) == p.epoch && (w & (3ull << 32)) != 0) break;
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // the nearest inclusive prefix: the lowest thread holding one
      const unsigned long long incMask = __ballot((w & (3ull << 32)) == kLbInclusive);
      if (lane == 0) lds[wave] = incMask ? (uint32_t)(wave * 64 + __ffsll((long long)incMask) - 1) : 0xFFFFFFFFu;
      __syncthreads();
      uint32_t first = 0xFFFFFFFFu;
      for (int v = 0; v < nw; ++v) first = min(first, lds[v]);
      uint32_t x = (uint32_t)tid <= first ? (uint32_t)w : 0u;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
      __syncthreads();
      if (lane == 0) lds[wave] = x;
      __syncthreads();
      for (int v = 0; v < nw; ++v) excl += lds[v];
      __syncthreads();
      if (first != 0xFFFFFFFFu) break;
      look -= blockDim.x;
    }
    if (tid == 0) lb_publish(p, tile, kLbInclusive, excl + agg);
  }
  return excl;
}

}  // namespace svd
