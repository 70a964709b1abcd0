// Device-wide exclusive prefix sums: one 1024-thread workgroup up to 8192
// items, else one single-pass launch with decoupled look-back (4096 items per
// 256-thread workgroup).
#include "svd_internal.hpp"
#include "block_scan.hpp"
#include "lookback.hpp"

namespace svd {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;

// One tile of kLbItems items per thread; out may alias in (every item is
// read before any is written).  The last tile writes the total.
constexpr int kLbItems = 16;
constexpr uint32_t kLbTile = kThreads * kLbItems;
__global__ void __launch_bounds__(kThreads) k_scan_lb(const uint32_t* in, uint32_t* out, uint64_t n, LbPass p,
                                                      uint32_t* __restrict__ total) {
  __shared__ uint32_t lds[kThreads / 64 + 2];
  __shared__ uint32_t word;
  const uint32_t tile = lb_take_tile(p, &word);
  const uint64_t base = (uint64_t)tile * kLbTile + (uint64_t)threadIdx.x * kLbItems;
  uint32_t v[kLbItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kLbItems; ++k) {
    const uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t agg;
  uint32_t ex = block_exclusive_sum(s, lds, agg);
  ex += lb_exclusive(p, tile, agg, lds);
#pragma unroll
  for (int k = 0; k < kLbItems; ++k) {
    const uint64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
  if (total && tile + 1 == p.tiles && threadIdx.x == kThreads - 1) *total = ex;
}

// Small inputs: one 1024-thread workgroup walks the array in chunks of 8192
// with a running carry (one launch instead of three).  Workgroup k scans
// job k, so two arrays of the same length share a launch.
constexpr int kSmallThreads = 1024;
constexpr uint64_t kSmallMax = 8192;
struct ScanJob {
  const uint32_t* in;
  uint32_t* out;
  uint32_t* total;
};
struct ScanJobs {
  ScanJob j[2];
};
__global__ void __launch_bounds__(kSmallThreads) k_scan_small(ScanJobs jobs, uint64_t n) {
  const uint32_t* __restrict__ in = jobs.j[blockIdx.x].in;
  uint32_t* __restrict__ out = jobs.j[blockIdx.x].out;
  uint32_t* __restrict__ total = jobs.j[blockIdx.x].total;
  __shared__ uint32_t lds[kSmallThreads / 64 + 1];
  uint32_t carry = 0;
  for (uint64_t base = 0; base < n; base += (uint64_t)kSmallThreads * kItems) {
    const uint64_t b = base + (uint64_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const uint64_t i = b + k;
      v[k] = i < n ? in[i] : 0u;
      s += v[k];
    }
    uint32_t agg;
    uint32_t ex = block_exclusive_sum(s, lds, agg) + carry;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const uint64_t i = b + k;
      if (i < n) out[i] = ex;
      ex += v[k];
    }
    carry += agg;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_fill(uint32_t* d, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}
__global__ void k_fill8(uint8_t* d, uint8_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}
__global__ void k_iota(uint32_t* d, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (uint32_t)i;
}

}  // namespace

void scan_exclusive(svd_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total) {
  if (n == 0) {
    if (total) fill_u32(c, total, 0u, 1);
    return;
  }
  if (n <= kSmallMax) {
    ScanJobs jobs{};
    jobs.j[0] = ScanJob{in, out, total};
    k_scan_small<<<1, kSmallThreads, 0, c->stream>>>(jobs, n);
    check_launch("scan_exclusive");
    return;
  }
  const uint64_t tiles = div_up(n, kLbTile);
  const LbPass p = lb_pass(c, tiles);
  k_scan_lb<<<(uint32_t)tiles, kThreads, 0, c->stream>>>(in, out, n, p, total);
  check_launch("scan_exclusive");
}

void scan_exclusive2(svd_ctx* c, uint32_t* a, uint32_t* b, uint64_t n) {
  if (n == 0) return;
  if (n > kSmallMax) {
    scan_exclusive(c, a, a, n, nullptr);
    scan_exclusive(c, b, b, n, nullptr);
    return;
  }
  ScanJobs jobs{};
  jobs.j[0] = ScanJob{a, a, nullptr};
  jobs.j[1] = ScanJob{b, b, nullptr};
  k_scan_small<<<2, kSmallThreads, 0, c->stream>>>(jobs, n);
  check_launch("scan_exclusive2");
}

LbPass lb_pass(svd_ctx* c, uint64_t tiles) {
  if (tiles >= (1ull << 31)) throw Error{SVD_ERR_CAPACITY, "single-pass scan: too many tiles"};
  bool clear = false;
  if (tiles > c->lbCap) {
    dfree(c, c->lbStatus);
    c->lbCap = std::max<uint64_t>(tiles, 4096);
    c->lbStatus = dalloc<unsigned long long>(c, c->lbCap);
    clear = true;
  }
  if (!c->lbCounter) {
    c->lbCounter = dalloc<uint32_t>(c, 1);
    fill_u32(c, c->lbCounter, 0u, 1);
  }
  if (++c->lbEpoch >= (1u << 30)) {
    c->lbEpoch = 1;
    clear = true;
  }
  if (clear) fill_u32(c, (uint32_t*)c->lbStatus, 0u, 2 * c->lbCap);
  return LbPass{c->lbStatus, c->lbCounter, c->lbEpoch, (uint32_t)tiles};
}

void fill_u32(svd_ctx* c, uint32_t* d, uint32_t v, uint64_t n) {
  if (!n) return;
  k_fill<<<div_up(n, 256), 256, 0, c->stream>>>(d, v, n);
  check_launch("fill_u32");
}

void fill_u8(svd_ctx* c, uint8_t* d, uint8_t v, uint64_t n) {
  if (!n) return;
  k_fill8<<<div_up(n, 256), 256, 0, c->stream>>>(d, v, n);
  check_launch("fill_u8");
}

void iota_u32(svd_ctx* c, uint32_t* d, uint64_t n) {
  if (!n) return;
  k_iota<<<div_up(n, 256), 256, 0, c->stream>>>(d, n);
  check_launch("iota_u32");
}

}  // namespace svd
