// Device-wide exclusive prefix sums (tile reduce / scan of the tile sums / down-sweep),
// wave64 shuffles + LDS, 2048 items per 256-thread workgroup.
#include "svd_internal.hpp"
#include "block_scan.hpp"

namespace svd {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                          uint32_t* __restrict__ partial) {
  __shared__ uint32_t lds[kThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = block_reduce_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        uint64_t n, const uint32_t* __restrict__ partial) {
  __shared__ uint32_t lds[kThreads / 64 + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t agg;
  uint32_t ex = block_exclusive_sum(s, lds, agg) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    uint64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
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

// Down-sweep for up to kFuseMax tiles that also replaces the partials pass:
// each workgroup sums the partial sums of the tiles before it (they sit in
// L2), and the last one writes the total.
constexpr uint32_t kFuseMax = 1024;
__global__ void __launch_bounds__(kThreads) k_scan_down_fused(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                              uint64_t n, const uint32_t* __restrict__ partial,
                                                              uint32_t* __restrict__ total) {
  __shared__ uint32_t red[kThreads / 64];
  __shared__ uint32_t lds[kThreads / 64 + 1];
  uint32_t p = 0;
  for (uint32_t k = threadIdx.x; k < blockIdx.x; k += kThreads) p += partial[k];
  const uint32_t prefix = block_reduce_sum(p, red);
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t agg;
  uint32_t ex = block_exclusive_sum(s, lds, agg) + prefix;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    uint64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
  if (total && blockIdx.x + 1 == gridDim.x && threadIdx.x == 0) *total = prefix + agg;
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
  const uint32_t nb = div_up(n, kTile);
  uint32_t* partial = dalloc<uint32_t>(c, nb);
  k_scan_reduce<<<nb, kThreads, 0, c->stream>>>(in, n, partial);
  if (nb <= kFuseMax) {
    k_scan_down_fused<<<nb, kThreads, 0, c->stream>>>(in, out, n, partial, total);
  } else {
    // the tile sums are scanned by the same multi-workgroup scan (at most
    // three levels deep for 2^32 elements)
    scan_exclusive(c, partial, partial, nb, total);
    k_scan_down<<<nb, kThreads, 0, c->stream>>>(in, out, n, partial);
  }
  check_launch("scan_exclusive");
  dfree(c, partial);
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
