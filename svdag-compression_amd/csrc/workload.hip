// Synthetic scene generator on the device (benchmark and test inputs; no
// reference counterpart).  It evaluates exactly the float64 operations of
// workloads.py (value noise with a SplitMix64 lattice hash, fBm), so its
// output is bit-identical to the NumPy definition and large grids (4096^3 =
// 6.9e10 cells) are generated in seconds.
//
// caves: surface voxels (solid with an empty 6-neighbour, outside the grid is
// empty) of {fbm3(p) - 0.5 > threshold}, p = (i + 0.5) * base_cells / 2^L.
// The grid is walked in 8^3 Morton bricks: a 512-thread block evaluates the
// brick plus a one-voxel halo (10^3 cells) into LDS, each lane tests its
// voxel, and the wave ballots pack the brick's 512 surface bits (64 B); a
// popcount scan then emits the codes in ascending Morton order.
#include "svd_internal.hpp"

namespace svd {

namespace {

__device__ __forceinline__ double lattice(int64_t ix, int64_t iy, int64_t iz, uint64_t seedMix) {
  uint64_t h = (uint64_t)ix * 0x8DA6B343ull;
  h ^= (uint64_t)iy * 0xD8163841ull;
  h ^= (uint64_t)iz * 0xCB1AB31Full;
  h ^= seedMix;
  h = mix64(h);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

__device__ double value_noise3(double x, double y, double z, uint64_t seedMix) {
  const double fx0 = floor(x), fy0 = floor(y), fz0 = floor(z);
  const double fx = smooth(x - fx0), fy = smooth(y - fy0), fz = smooth(z - fz0);
  const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0, iz = (int64_t)fz0;
  double out = 0.0;
  for (int dz = 0; dz < 2; ++dz) {
    const double wz = dz ? fz : 1.0 - fz;
    for (int dy = 0; dy < 2; ++dy) {
      const double wy = dy ? fy : 1.0 - fy;
      for (int dx = 0; dx < 2; ++dx) {
        const double wx = dx ? fx : 1.0 - fx;
        out = out + wx * wy * wz * lattice(ix + dx, iy + dy, iz + dz, seedMix);
      }
    }
  }
  return out;
}

struct CaveArgs {
  uint32_t levels, octaves;
  uint64_t seed;
  double threshold, step;          // step = base_cells / 2^L
};

__device__ bool solid(const CaveArgs& a, int64_t i, int64_t j, int64_t k) {
  const int64_t n = 1ll << a.levels;
  if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
  const double x = ((double)i + 0.5) * a.step, y = ((double)j + 0.5) * a.step, z = ((double)k + 0.5) * a.step;
  double s = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
  for (uint32_t o = 0; o < a.octaves; ++o) {
    const uint64_t sd = a.seed * 1000ull + o;
    s = s + amp * value_noise3(x * freq, y * freq, z * freq, sd * 0x9E3779B97F4A7C15ull);
    norm += amp;
    amp *= 0.5;
    freq *= 2.0;
  }
  return s / norm - 0.5 > a.threshold;
}

__device__ __forceinline__ uint32_t compact3(uint32_t x) {   // every third bit -> packed
  x &= 0x49249249u;
  x = (x ^ (x >> 2)) & 0xC30C30C3u;
  x = (x ^ (x >> 4)) & 0x0F00F00Fu;
  x = (x ^ (x >> 8)) & 0xFF0000FFu;
  x = (x ^ (x >> 16)) & 0x000003FFu;
  return x;
}

// Brick b (Morton index of an 8^3 brick) -> 512 surface bits (8 x u64).
__global__ void __launch_bounds__(512) k_cave_bricks(CaveArgs a, uint64_t brick0, uint64_t nbricks,
                                                     uint64_t* __restrict__ masks) {
  __shared__ uint8_t cell[10 * 10 * 10];
  const uint64_t b = brick0 + blockIdx.x;
  if (blockIdx.x >= nbricks) return;
  // brick coordinates from its Morton index (x most significant)
  // (b < 2^30 for levels <= 13, so 10 bits per axis)
  const uint32_t bx = compact3((uint32_t)(b >> 2)), by = compact3((uint32_t)(b >> 1)), bz = compact3((uint32_t)b);
  const int64_t x0 = (int64_t)bx * 8 - 1, y0 = (int64_t)by * 8 - 1, z0 = (int64_t)bz * 8 - 1;
  for (int t = threadIdx.x; t < 1000; t += 512) {
    const int cx = t / 100, cy = (t / 10) % 10, cz = t % 10;
    cell[t] = solid(a, x0 + cx, y0 + cy, z0 + cz) ? 1 : 0;
  }
  __syncthreads();
  const uint32_t m = threadIdx.x;                 // local Morton code (3 bits per level)
  const int lx = (int)compact3(m >> 2), ly = (int)compact3(m >> 1), lz = (int)compact3(m);
  const int c = (lx + 1) * 100 + (ly + 1) * 10 + (lz + 1);
  const bool surf = cell[c] && !(cell[c - 100] && cell[c + 100] && cell[c - 10] && cell[c + 10] && cell[c - 1] && cell[c + 1]);
  const uint64_t bal = __ballot(surf);
  if ((threadIdx.x & 63) == 0) masks[(b - brick0) * 8 + (threadIdx.x >> 6)] = bal;
}

__global__ void k_brick_counts(const uint64_t* __restrict__ masks, uint64_t nb, uint32_t* __restrict__ cnt) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  uint32_t s = 0;
#pragma unroll
  for (int w = 0; w < 8; ++w) s += (uint32_t)__popcll(masks[i * 8 + w]);
  cnt[i] = s;
}

__global__ void k_brick_emit(const uint64_t* __restrict__ masks, uint64_t brick0, uint64_t nb,
                             const uint32_t* __restrict__ off, uint64_t base, uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  uint64_t o = base + off[i];
  const uint64_t hi = (brick0 + i) << 9;
  for (int w = 0; w < 8; ++w) {
    uint64_t m = masks[i * 8 + w];
    while (m) {
      const int t = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      out[o++] = hi | (uint64_t)(w * 64 + t);
    }
  }
}

}  // namespace

// Returns a device array (hipMalloc; free with hipFree) of the sorted codes.
uint64_t gen_caves(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double threshold, double baseCells,
                   uint64_t** out) {
  if (levels < 3 || levels > 13) throw Error{SVD_ERR_LEVELS, "caves: levels must be 3..13"};
  CaveArgs a;
  a.levels = levels; a.octaves = octaves; a.seed = seed; a.threshold = threshold;
  a.step = baseCells / (double)(1ull << levels);
  const uint64_t nBricks = 1ull << (3 * (levels - 3));
  // bricks are generated in chunks of 2^22 (2^31 voxels: the per-chunk
  // counts and offsets stay within u32, the masks within 256 MB)
  const uint64_t chunk = nBricks < (1ull << 22) ? nBricks : (1ull << 22);
  uint64_t* masks = dalloc<uint64_t>(c, chunk * 8);
  uint32_t* cnt = dalloc<uint32_t>(c, chunk + 1);
  uint32_t* total = dalloc<uint32_t>(c, 1);
  std::vector<uint64_t*> parts;
  std::vector<uint64_t> partN;
  uint64_t n = 0;
  for (uint64_t b0 = 0; b0 < nBricks; b0 += chunk) {
    const uint64_t nb = nBricks - b0 < chunk ? nBricks - b0 : chunk;
    k_cave_bricks<<<(uint32_t)nb, 512, 0, c->stream>>>(a, b0, nb, masks);
    k_brick_counts<<<div_up(nb, 256), 256, 0, c->stream>>>(masks, nb, cnt);
    check_launch("k_cave_bricks");
    scan_exclusive(c, cnt, cnt, nb, total);
    const uint64_t m = read_u32(c, total);
    uint64_t* part = nullptr;
    SVD_HIP(hipMalloc(&part, (m ? m : 1) * 8));
    k_brick_emit<<<div_up(nb, 256), 256, 0, c->stream>>>(masks, b0, nb, cnt, 0, part);
    check_launch("k_brick_emit");
    parts.push_back(part);
    partN.push_back(m);
    n += m;
  }
  uint64_t* res = nullptr;
  SVD_HIP(hipMalloc(&res, (n ? n : 1) * 8));
  uint64_t o = 0;
  for (size_t i = 0; i < parts.size(); ++i) {
    if (partN[i]) SVD_HIP(hipMemcpyAsync(res + o, parts[i], partN[i] * 8, hipMemcpyDeviceToDevice, c->stream));
    o += partN[i];
  }
  SVD_HIP(hipStreamSynchronize(c->stream));
  for (auto* p : parts) SVD_HIP(hipFree(p));
  dfree(c, masks); dfree(c, cnt); dfree(c, total);
  *out = res;
  return n;
}

}  // namespace svd
