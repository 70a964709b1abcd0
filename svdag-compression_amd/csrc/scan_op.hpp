// A per-element value, its exclusive prefix sum and a per-element use of
// that prefix in one launch (single-pass scan with decoupled look-back).
// Op provides
//   __device__ uint32_t value(uint64_t i, uint32_t& flags) const;  // i < n
//   __device__ void apply(uint64_t i, uint32_t excl) const;        // i < n
//   __device__ void report(uint32_t flags) const;  // once per thread, flags != 0
// value() ORs error bits into flags instead of storing (a store between the
// elements' loads would keep them from overlapping).  The value of element i
// may not depend on another element's apply.
#pragma once
#include <algorithm>
#include "svd_internal.hpp"
#include "block_scan.hpp"
#include "lookback.hpp"

namespace svd {

// 1024 threads per tile: every tile takes its number from one device counter,
// and atomics on one address serialise (about 11 ns each on MI355X), so
// fewer, larger tiles start sooner.
constexpr int kScanOpT = 1024;

// Element k of thread (wave w, lane j): w * 64 * PER + 64 * k + j of the
// tile, so value() and apply() see 64 consecutive elements per row.  The
// values of a thread's PER elements are computed one after another (their
// loads sit behind data-dependent branches), so an op whose value() chains
// several memory round trips takes a small PER; the tile count, and with it
// the look-back, grows as PER shrinks.
template <int PER, typename Op>
__global__ void __launch_bounds__(kScanOpT) k_scan_op(uint64_t n, LbPass p, Op op, uint32_t* __restrict__ total) {
  constexpr int kScanOpPer = PER;
  constexpr uint32_t kScanOpTile = kScanOpT * PER;
  __shared__ uint32_t lds[kScanOpT / 64 + 2];
  __shared__ uint32_t word;
  const uint32_t tile = lb_take_tile(p, &word);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)tile * kScanOpTile + (uint64_t)wave * (64 * kScanOpPer) + lane;
  uint32_t ex[kScanOpPer];
  uint32_t flags = 0;
#pragma unroll
  for (int k = 0; k < kScanOpPer; ++k) {
    const uint64_t i = base + 64 * k;
    ex[k] = i < n ? op.value(i, flags) : 0u;
  }
  if (flags) op.report(flags);
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < kScanOpPer; ++k) {
    const uint32_t v = ex[k];
    const uint32_t inc = wave_inclusive_sum(v);
    ex[k] = run + inc - v;
    run += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
  }
  if (lane == 0) lds[wave] = run;
  __syncthreads();
  uint32_t off = 0, agg = 0;
#pragma unroll
  for (int w = 0; w < kScanOpT / 64; ++w) {
    const uint32_t x = lds[w];
    if (w < wave) off += x;
    agg += x;
  }
  off += lb_exclusive(p, tile, agg, lds);   // its entry barrier orders the reads above
#pragma unroll
  for (int k = 0; k < kScanOpPer; ++k) {
    const uint64_t i = base + 64 * k;
    if (i < n) op.apply(i, off + ex[k]);
  }
  if (total && tile + 1 == p.tiles && threadIdx.x == 0) *total = off + agg;   // wave 0's offset is the tile's prefix
}

// Two independent value-scan-apply passes (op0 over n0 elements, op1 over
// n1) in one launch: the first T0 tiles are op0's, the rest op1's, each with
// its own look-back chain in the shared pass.
template <int PER, typename Op>
__global__ void __launch_bounds__(kScanOpT) k_scan_op2(uint64_t n0, uint64_t n1, uint32_t T0, LbPass p, Op op0,
                                                       Op op1, uint32_t* __restrict__ total0,
                                                       uint32_t* __restrict__ total1) {
  constexpr uint32_t kTile = kScanOpT * PER;
  __shared__ uint32_t lds[kScanOpT / 64 + 2];
  __shared__ uint32_t word;
  const uint32_t tile = lb_take_tile(p, &word);
  const bool second = tile >= T0;
  const Op& op = second ? op1 : op0;
  const uint64_t n = second ? n1 : n0;
  const uint32_t first = second ? T0 : 0u;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)(tile - first) * kTile + (uint64_t)wave * (64 * PER) + lane;
  uint32_t ex[PER];
  uint32_t flags = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint64_t i = base + 64 * k;
    ex[k] = i < n ? op.value(i, flags) : 0u;
  }
  if (flags) op.report(flags);
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t v = ex[k];
    const uint32_t inc = wave_inclusive_sum(v);
    ex[k] = run + inc - v;
    run += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
  }
  if (lane == 0) lds[wave] = run;
  __syncthreads();
  uint32_t off = 0, agg = 0;
#pragma unroll
  for (int w = 0; w < kScanOpT / 64; ++w) {
    const uint32_t x = lds[w];
    if (w < wave) off += x;
    agg += x;
  }
  off += lb_exclusive(p, tile, agg, lds, first);   // its entry barrier orders the reads above
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint64_t i = base + 64 * k;
    if (i < n) op.apply(i, off + ex[k]);
  }
  uint32_t* total = second ? total1 : total0;
  const uint32_t last = second ? p.tiles - 1 : T0 - 1;
  if (total && tile == last && threadIdx.x == 0) *total = off + agg;
}

template <int PER, typename Op>
void scan_op2(svd_ctx* c, uint64_t n0, const Op& op0, uint32_t* total0, uint64_t n1, const Op& op1,
              uint32_t* total1) {
  const uint64_t T0 = div_up(std::max<uint64_t>(n0, 1), (uint64_t)kScanOpT * PER);
  const uint64_t T1 = div_up(std::max<uint64_t>(n1, 1), (uint64_t)kScanOpT * PER);
  k_scan_op2<PER, Op><<<(uint32_t)(T0 + T1), kScanOpT, 0, c->stream>>>(n0, n1, (uint32_t)T0, lb_pass(c, T0 + T1),
                                                                       op0, op1, total0, total1);
  check_launch("scan_op2");
}

// total (device, may be null) gets the sum of all values.
template <int PER, typename Op>
void scan_op(svd_ctx* c, uint64_t n, const Op& op, uint32_t* total) {
  if (n == 0) {
    if (total) fill_u32(c, total, 0u, 1);
    return;
  }
  const uint64_t tiles = div_up(n, (uint64_t)kScanOpT * PER);
  k_scan_op<PER, Op><<<(uint32_t)tiles, kScanOpT, 0, c->stream>>>(n, lb_pass(c, tiles), op, total);
  check_launch("scan_op");
}

}  // namespace svd
