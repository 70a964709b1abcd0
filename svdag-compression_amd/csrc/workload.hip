// Synthetic scene generator on the device (benchmark and test inputs; no
// reference counterpart).  It evaluates exactly the float64 operations of
// workloads.py (value noise with a SplitMix64 lattice hash, fBm), so its
// output is bit-identical to the NumPy definition and large grids (4096^3 =
// 6.9e10 cells) are generated in seconds.
//
// Scenes (surface voxels = solid with an empty 6-neighbour; outside the grid
// is empty):
//   caves   {fbm3(p) - 0.5 > threshold}, p = (i + 0.5) * base_cells / 2^L;
//   balls   union of balls |p - c| <= r, p = (i + 0.5) / 2^L (the spheres of
//           the icosphere scenes, centres and radii as workloads.spheres_mesh);
//   menger  Menger sponge of a given order filling the grid;
//   terrain each column filled from the lowest of its 4-neighbour heights up
//           to its own height (fBm2 heightfield stretched to [hmin, hmax]).
// The grid is walked in 8^3 Morton bricks: a 512-thread block evaluates the
// brick plus a one-voxel halo (10^3 cells) into LDS, each lane tests its
// voxel, and the wave ballots pack the brick's 512 surface bits (64 B); a
// popcount scan then emits the codes in ascending Morton order.
#include "svd_internal.hpp"
#include <rocprim/device/device_merge.hpp>
#include <rocprim/device/device_select.hpp>

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

__device__ double value_noise2(double x, double y, uint64_t seedMix) {
  const double fx0 = floor(x), fy0 = floor(y);
  const double fx = smooth(x - fx0), fy = smooth(y - fy0);
  const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0;
  auto lat = [&](int64_t a, int64_t b) {
    uint64_t h = (uint64_t)a * 0x8DA6B343ull;
    h ^= (uint64_t)b * 0xD8163841ull;
    h ^= seedMix;
    h = mix64(h);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
  };
  const double v00 = lat(ix, iy), v10 = lat(ix + 1, iy), v01 = lat(ix, iy + 1), v11 = lat(ix + 1, iy + 1);
  const double a = v00 + (v10 - v00) * fx;
  const double b = v01 + (v11 - v01) * fx;
  return a + (b - a) * fy;
}

__device__ __forceinline__ uint32_t compact3(uint32_t x) {   // every third bit -> packed
  x &= 0x49249249u;
  x = (x ^ (x >> 2)) & 0xC30C30C3u;
  x = (x ^ (x >> 4)) & 0x0F00F00Fu;
  x = (x ^ (x >> 8)) & 0xFF0000FFu;
  x = (x ^ (x >> 16)) & 0x000003FFu;
  return x;
}

struct BallArgs {
  static constexpr bool kSurface = true;
  const double* balls;     // (cx, cy, cz, r) per ball
  uint32_t nBalls;
};

struct MengerArgs {
  static constexpr bool kSurface = true;
  uint32_t order;
};

struct TerrainArgs {
  static constexpr bool kSurface = false;   // the columns are the voxel set itself
  const int64_t* h;        // n x n heights, [x][z]
  const int64_t* lo;       // n x n column bottoms
};

__device__ bool solid_at(const CaveArgs& a, int64_t i, int64_t j, int64_t k) { return solid(a, i, j, k); }

__device__ bool solid_at(const BallArgs& a, uint32_t levels, int64_t i, int64_t j, int64_t k) {
  const int64_t n = 1ll << levels;
  if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
  const double dn = (double)n;
  const double x = ((double)i + 0.5) / dn, y = ((double)j + 0.5) / dn, z = ((double)k + 0.5) / dn;
  for (uint32_t b = 0; b < a.nBalls; ++b) {
    const double* q = a.balls + 4 * b;
    const double dx = x - q[0], dy = y - q[1], dz = z - q[2];
    if (dx * dx + dy * dy + dz * dz <= q[3] * q[3]) return true;
  }
  return false;
}

__device__ bool solid_at(const MengerArgs& a, uint32_t levels, int64_t i, int64_t j, int64_t k) {
  const int64_t n = 1ll << levels;
  if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
  int64_t p = 3;
  for (uint32_t o = 0; o < a.order; ++o, p *= 3) {
    const int mx = ((i * p) / n) % 3 == 1, my = ((j * p) / n) % 3 == 1, mz = ((k * p) / n) % 3 == 1;
    if (mx + my + mz >= 2) return false;
  }
  return true;
}

__device__ bool solid_at(const TerrainArgs& a, uint32_t levels, int64_t i, int64_t j, int64_t k) {
  const int64_t n = 1ll << levels;
  if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
  const int64_t c = i * n + k;
  return j >= a.lo[c] && j <= a.h[c];
}

__device__ __forceinline__ void brick_origin(uint64_t b, int64_t& x0, int64_t& y0, int64_t& z0) {
  // brick coordinates from its Morton index (x most significant; b < 2^30 for levels <= 13)
  x0 = (int64_t)compact3((uint32_t)(b >> 2)) * 8 - 1;
  y0 = (int64_t)compact3((uint32_t)(b >> 1)) * 8 - 1;
  z0 = (int64_t)compact3((uint32_t)b) * 8 - 1;
}

// The 512 bits of brick b (surface voxels, or all solid voxels when surface is
// false) from the LDS cells of the brick + halo.
__device__ __forceinline__ void brick_emit_bits(const uint8_t* cell, uint64_t b, uint64_t brick0, uint64_t* masks,
                                                bool surface = true) {
  const uint32_t m = threadIdx.x;                 // local Morton code (3 bits per level)
  const int lx = (int)compact3(m >> 2), ly = (int)compact3(m >> 1), lz = (int)compact3(m);
  const int c = (lx + 1) * 100 + (ly + 1) * 10 + (lz + 1);
  const bool surf = cell[c] && (!surface || !(cell[c - 100] && cell[c + 100] && cell[c - 10] && cell[c + 10] &&
                                              cell[c - 1] && cell[c + 1]));
  const uint64_t bal = __ballot(surf);
  if ((threadIdx.x & 63) == 0) masks[(b - brick0) * 8 + (threadIdx.x >> 6)] = bal;
}

// Brick b (Morton index of an 8^3 brick) -> 512 surface bits (8 x u64).
__global__ void __launch_bounds__(512) k_cave_bricks(CaveArgs a, uint64_t brick0, uint64_t nbricks,
                                                     uint64_t* __restrict__ masks) {
  __shared__ uint8_t cell[10 * 10 * 10];
  const uint64_t b = brick0 + blockIdx.x;
  if (blockIdx.x >= nbricks) return;
  int64_t x0, y0, z0;
  brick_origin(b, x0, y0, z0);
  for (int t = threadIdx.x; t < 1000; t += 512) {
    const int cx = t / 100, cy = (t / 10) % 10, cz = t % 10;
    cell[t] = solid(a, x0 + cx, y0 + cy, z0 + cz) ? 1 : 0;
  }
  __syncthreads();
  brick_emit_bits(cell, b, brick0, masks);
}

template <typename A>
__global__ void __launch_bounds__(512) k_scene_bricks(A a, uint32_t levels, uint64_t brick0, uint64_t nbricks,
                                                      uint64_t* __restrict__ masks) {
  __shared__ uint8_t cell[10 * 10 * 10];
  const uint64_t b = brick0 + blockIdx.x;
  if (blockIdx.x >= nbricks) return;
  int64_t x0, y0, z0;
  brick_origin(b, x0, y0, z0);
  for (int t = threadIdx.x; t < 1000; t += 512) {
    const int cx = t / 100, cy = (t / 10) % 10, cz = t % 10;
    cell[t] = solid_at(a, levels, x0 + cx, y0 + cy, z0 + cz) ? 1 : 0;
  }
  __syncthreads();
  brick_emit_bits(cell, b, brick0, masks, A::kSurface);
}

// fBm2 of every column (terrain_heights before the stretch), [x][z].
__global__ void k_terrain_fbm(uint32_t levels, uint64_t seed, uint32_t octaves, double step, double* __restrict__ f) {
  const uint64_t n = 1ull << levels;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const uint64_t i = t / n, k = t % n;
  const double x = ((double)i + 0.5) * step, z = ((double)k + 0.5) * step;
  double s = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
  for (uint32_t o = 0; o < octaves; ++o) {
    const uint64_t sd = seed * 1000ull + o;
    s = s + amp * value_noise2(x * freq, z * freq, sd * 0x9E3779B97F4A7C15ull);
    norm += amp;
    amp *= 0.5;
    freq *= 2.0;
  }
  f[t] = s / norm;
}

__global__ void k_minmax_partial(const double* __restrict__ f, uint64_t n, double* __restrict__ part) {
  __shared__ double lo[256], hi[256];
  double a = 1e300, b = -1e300;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    a = fmin(a, f[i]);
    b = fmax(b, f[i]);
  }
  lo[threadIdx.x] = a; hi[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + w]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = lo[0]; part[2 * blockIdx.x + 1] = hi[0]; }
}

__global__ void k_terrain_heights(const double* __restrict__ f, uint32_t levels, double fmin, double range, double hmin,
                                  double hmax, int64_t* __restrict__ h) {
  const uint64_t n = 1ull << levels;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const double g = (f[t] - fmin) / range;
  int64_t v = (int64_t)floor((hmin + (hmax - hmin) * g) * (double)n);
  v = v < 0 ? 0 : (v > (int64_t)n - 1 ? (int64_t)n - 1 : v);
  h[t] = v;
}

__global__ void k_terrain_lo(const int64_t* __restrict__ h, uint32_t levels, int64_t* __restrict__ lo) {
  const int64_t n = 1ll << levels;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t i = t / n, k = t % n;
  auto H = [&](int64_t a, int64_t b) {        // edge padding
    a = a < 0 ? 0 : (a >= n ? n - 1 : a);
    b = b < 0 ? 0 : (b >= n ? n - 1 : b);
    return h[a * n + b];
  };
  int64_t m = h[t];
  m = min(m, H(i - 1, k)); m = min(m, H(i + 1, k)); m = min(m, H(i, k - 1)); m = min(m, H(i, k + 1));
  lo[t] = m;
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

// workloads.material_attrs: quantized 3D value noise at the voxel centre.
__global__ void k_materials(const uint64_t* __restrict__ codes, uint64_t n, double step, uint64_t seedMix,
                            uint32_t nValues, uint8_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = codes[i];
  uint32_t x = 0, y = 0, z = 0;
  for (int b = 0; b < 21; ++b) {
    x |= (uint32_t)((c >> (3 * b + 2)) & 1u) << b;
    y |= (uint32_t)((c >> (3 * b + 1)) & 1u) << b;
    z |= (uint32_t)((c >> (3 * b)) & 1u) << b;
  }
  const double v = value_noise3(((double)x + 0.5) * step, ((double)y + 0.5) * step, ((double)z + 0.5) * step, seedMix);
  double q = floor(v * (double)nValues);
  if (q > (double)(nValues - 1)) q = (double)(nValues - 1);
  out[i] = (uint8_t)q;
}

}  // namespace

void gen_materials(svd_ctx* c, const uint64_t* codes, uint64_t n, uint32_t levels, uint64_t seed, uint32_t nValues,
                   double baseCells, uint8_t* out) {
  if (!n) return;
  const double step = baseCells / (double)(1ull << levels);
  k_materials<<<div_up(n, 256), 256, 0, c->stream>>>(codes, n, step, seed * 0x9E3779B97F4A7C15ull, nValues, out);
  check_launch("k_materials");
  SVD_HIP(hipStreamSynchronize(c->stream));
}

// Sorted union of two sorted unique code lists (rocPRIM merge + unique).
uint64_t codes_union(svd_ctx* c, const uint64_t* a, uint64_t na, const uint64_t* b, uint64_t nb, uint64_t** out) {
  uint64_t* merged = dalloc<uint64_t>(c, na + nb + 1);
  size_t bytes = 0;
  SVD_HIP(rocprim::merge(nullptr, bytes, a, b, merged, (size_t)na, (size_t)nb, rocprim::less<uint64_t>(), c->stream));
  void* tmp = c->arena->get(bytes + 16);
  SVD_HIP(rocprim::merge(tmp, bytes, a, b, merged, (size_t)na, (size_t)nb, rocprim::less<uint64_t>(), c->stream));
  c->arena->put(tmp);
  uint64_t* res = nullptr;
  SVD_HIP(hipMalloc(&res, (na + nb + 1) * 8));
  uint64_t* cnt = dalloc<uint64_t>(c, 1);
  bytes = 0;
  SVD_HIP(rocprim::unique(nullptr, bytes, merged, res, cnt, (size_t)(na + nb), rocprim::equal_to<uint64_t>(), c->stream));
  tmp = c->arena->get(bytes + 16);
  SVD_HIP(rocprim::unique(tmp, bytes, merged, res, cnt, (size_t)(na + nb), rocprim::equal_to<uint64_t>(), c->stream));
  c->arena->put(tmp);
  uint64_t n = 0;
  SVD_HIP(hipMemcpyAsync(&n, cnt, 8, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
  dfree(c, merged); dfree(c, cnt);
  *out = res;
  return n;
}

// Walk all bricks of the 2^levels grid with `launch(b0, nb, masks)` writing
// their surface bits; returns a device array (hipMalloc; free with hipFree)
// of the sorted codes.
template <typename F>
uint64_t emit_bricks(svd_ctx* c, uint32_t levels, F launch, uint64_t** out) {
  if (levels < 3 || levels > 13) throw Error{SVD_ERR_LEVELS, "generated scenes need 3..13 levels"};
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
    launch(b0, nb, masks);
    k_brick_counts<<<div_up(nb, 256), 256, 0, c->stream>>>(masks, nb, cnt);
    check_launch("scene bricks");
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

uint64_t gen_caves(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double threshold, double baseCells,
                   uint64_t** out) {
  CaveArgs a;
  a.levels = levels; a.octaves = octaves; a.seed = seed; a.threshold = threshold;
  a.step = baseCells / (double)(1ull << levels);
  return emit_bricks(c, levels, [&](uint64_t b0, uint64_t nb, uint64_t* masks) {
    k_cave_bricks<<<(uint32_t)nb, 512, 0, c->stream>>>(a, b0, nb, masks);
  }, out);
}

uint64_t gen_balls(svd_ctx* c, uint32_t levels, const double* balls, uint32_t nBalls, uint64_t** out) {
  double* d = dalloc<double>(c, 4ull * nBalls + 4);
  SVD_HIP(hipMemcpyAsync(d, balls, 32ull * nBalls, hipMemcpyHostToDevice, c->stream));
  BallArgs a{d, nBalls};
  const uint64_t n = emit_bricks(c, levels, [&](uint64_t b0, uint64_t nb, uint64_t* masks) {
    k_scene_bricks<BallArgs><<<(uint32_t)nb, 512, 0, c->stream>>>(a, levels, b0, nb, masks);
  }, out);
  dfree(c, d);
  return n;
}

uint64_t gen_menger(svd_ctx* c, uint32_t levels, uint32_t order, uint64_t** out) {
  if (order > 20) throw Error{SVD_ERR_ARG, "menger order must be <= 20"};
  MengerArgs a{order};
  return emit_bricks(c, levels, [&](uint64_t b0, uint64_t nb, uint64_t* masks) {
    k_scene_bricks<MengerArgs><<<(uint32_t)nb, 512, 0, c->stream>>>(a, levels, b0, nb, masks);
  }, out);
}

uint64_t gen_terrain(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double hmin, double hmax,
                     double baseCells, uint64_t** out) {
  if (levels < 3 || levels > 13) throw Error{SVD_ERR_LEVELS, "generated scenes need 3..13 levels"};
  const uint64_t n = 1ull << levels, nn = n * n;
  double* f = dalloc<double>(c, nn);
  int64_t* h = dalloc<int64_t>(c, nn);
  int64_t* lo = dalloc<int64_t>(c, nn);
  const uint32_t nPart = 256;
  double* part = dalloc<double>(c, 2 * nPart);
  k_terrain_fbm<<<div_up(nn, 256), 256, 0, c->stream>>>(levels, seed, octaves, baseCells / (double)n, f);
  k_minmax_partial<<<nPart, 256, 0, c->stream>>>(f, nn, part);
  check_launch("terrain fbm");
  std::vector<double> hp(2 * nPart);
  SVD_HIP(hipMemcpyAsync(hp.data(), part, 16 * nPart, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
  double fmn = hp[0], fmx = hp[1];
  for (uint32_t i = 1; i < nPart; ++i) { fmn = std::min(fmn, hp[2 * i]); fmx = std::max(fmx, hp[2 * i + 1]); }
  // workloads.terrain_heights: f = (f - min) / max(max - min, 1e-12)
  const double range = std::max(fmx - fmn, 1e-12);
  k_terrain_heights<<<div_up(nn, 256), 256, 0, c->stream>>>(f, levels, fmn, range, hmin, hmax, h);
  k_terrain_lo<<<div_up(nn, 256), 256, 0, c->stream>>>(h, levels, lo);
  check_launch("terrain heights");
  TerrainArgs a{h, lo};
  const uint64_t m = emit_bricks(c, levels, [&](uint64_t b0, uint64_t nb, uint64_t* masks) {
    k_scene_bricks<TerrainArgs><<<(uint32_t)nb, 512, 0, c->stream>>>(a, levels, b0, nb, masks);
  }, out);
  dfree(c, f); dfree(c, h); dfree(c, lo); dfree(c, part);
  return m;
}

}  // namespace svd
