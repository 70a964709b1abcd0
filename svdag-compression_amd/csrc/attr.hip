// Gray-code attribute bit-tree (the reference names the feature in readme.md:8
// and has no code for it; the encoding is this library's own definition).
//
// Attributes are given in voxel (Morton / DFS) order.  Code = Gray b^(b>>1)
// or binary.  Per bit plane, a pruned binary tree over the sequence padded to
// P = 2^ceil(log2 V) with its last value is emitted in pre-order:
// single element -> its bit; uniform node -> 0, bit; mixed node -> 1, left,
// right.  Bits are packed LSB-first in u32 words; each plane starts on a word.
//
// Device form.  The sequence is cut into chunks of S = min(P, 64) elements:
// one wavefront reads a chunk (one byte per lane) and eight ballots give the
// chunk's eight plane masks, so the bit transpose costs no memory traffic.
//  1. chunk states: per chunk a u16 (bit k = plane k uniform, bit 8+k = its
//     value); a pyramid of the same u16 over chunk ranges up to the root.
//  2. count: lane (q, k) of a wave owns chunk q of the wave's eight and plane
//     k, and walks the emitted nodes that start inside that chunk in pre-order
//     (the chain of larger nodes that begin at the chunk's first element, read
//     from the pyramid, then the chunk's own subtree from its 64-bit mask).
//     Per workgroup and plane the bit counts are summed.
//  3. exclusive scan of the workgroup sums per plane.
//  4. emit: the same walk again, now writing bits at its pre-order offset.
// The attribute bytes are read three times; nothing of size P is stored.
#include "svd_internal.hpp"

namespace svd {

namespace {

constexpr int kBT = 256;                 // four wavefronts
constexpr int kWaves = kBT / 64;
constexpr int kChunksPerWave = 8;        // lane = 8 * q + k: chunk q of the wave, plane k
constexpr int kChunksPerBlock = kWaves * kChunksPerWave;

struct Geo {
  uint64_t V;        // attributes
  uint64_t C;        // chunks (P >> hc)
  uint32_t hc;       // element height of a chunk node (S = 2^hc)
  uint32_t G;        // chunk-level height of the root (C = 2^G)
  int gray;
  uint32_t nb;       // workgroups (ceil(C / kChunksPerBlock))
  uint64_t lvl[40];  // offset of pyramid level g (C >> g entries)
};

__device__ __forceinline__ uint32_t code_at(const uint8_t* __restrict__ a, uint64_t V, uint64_t i, int gray) {
  const uint32_t b = a[i < V ? i : V - 1];
  return gray ? (b ^ (b >> 1)) : b;
}

__device__ __forceinline__ uint64_t chunk_full(uint32_t hc) {
  return hc >= 6 ? ~0ull : ((1ull << (1u << hc)) - 1);
}

// The eight plane masks of one chunk (every lane of the wave must be here).
// Lane l holds element chunk*S + l when l < S.
struct Masks { uint64_t m[8]; };
__device__ __forceinline__ Masks chunk_masks(const uint8_t* __restrict__ a, const Geo& g, uint64_t chunk) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t S = 1u << g.hc;
  uint32_t code = 0;
  if (lane < S) code = code_at(a, g.V, chunk * S + lane, g.gray);
  const uint64_t full = chunk_full(g.hc);
  Masks r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.m[k] = (uint64_t)__ballot((code >> k) & 1u) & full;
  return r;
}

__device__ __forceinline__ uint16_t state_of(const Masks& ms, uint64_t full) {
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint64_t m = ms.m[k];
    if (m == 0 || m == full) s |= (1u << k) | ((uint32_t)(m & 1u) << (8 + k));
  }
  return (uint16_t)s;
}

__device__ __forceinline__ uint16_t combine(uint16_t a, uint16_t b) {
  const uint32_t uni = (uint32_t)a & (uint32_t)b & ~(((uint32_t)a ^ (uint32_t)b) >> 8) & 0xFFu;
  return (uint16_t)(uni | (((uint32_t)a >> 8) & uni) << 8);
}

// Pre-order walk of the emitted nodes of plane k that start in chunk j.
template <typename Put>
__device__ __forceinline__ void walk(uint64_t mask, uint32_t k, uint64_t j, const uint16_t* __restrict__ pyr,
                                     const Geo& g, Put& put) {
  const uint32_t top = (j == 0) ? g.G : min((uint32_t)__builtin_ctzll(j), g.G);
  if (j != 0) {   // not the root: emitted only under a mixed parent
    const uint16_t ps = pyr[g.lvl[top + 1] + (j >> (top + 1))];
    if ((ps >> k) & 1u) return;
  }
  for (uint32_t h = top; h >= 1; --h) {
    const uint16_t s = pyr[g.lvl[h] + (j >> h)];
    if (!((s >> k) & 1u)) { put(1u); continue; }
    put(0u);
    put((uint32_t)(s >> (8 + k)) & 1u);
    return;
  }
  // the chunk node (element height hc) and its subtree
  const uint32_t hc = g.hc;
  uint32_t h = hc, i = 0;
  while (true) {
    if (h == 0) {
      put((uint32_t)(mask >> i) & 1u);
    } else {
      const uint32_t w = 1u << h;
      const uint64_t full = (w >= 64) ? ~0ull : ((1ull << w) - 1);
      const uint64_t sub = (w >= 64) ? mask : (mask >> (i * w)) & full;
      if (sub != 0 && sub != full) { put(1u); --h; i <<= 1; continue; }
      put(0u);
      put((uint32_t)(sub & 1u));
    }
    while (h < hc && (i & 1u)) { ++h; i >>= 1; }
    if (h == hc) break;
    ++i;
  }
}

// 1a. chunk states (pyramid level 0)
__global__ void __launch_bounds__(kBT) k_bt_states(const uint8_t* __restrict__ a, Geo g,
                                                   uint16_t* __restrict__ pyr) {
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t full = chunk_full(g.hc);
  for (int q = 0; q < kChunksPerWave; ++q) {
    const uint64_t chunk = ((uint64_t)blockIdx.x * kWaves + w) * kChunksPerWave + q;
    if (chunk >= g.C) break;   // uniform across the wave
    const Masks ms = chunk_masks(a, g, chunk);
    if ((threadIdx.x & 63u) == 0) pyr[chunk] = state_of(ms, full);
  }
}

// 1b. up to 8 pyramid levels above level `g0` (n0 = its entry count, a power of 2)
__global__ void __launch_bounds__(kBT) k_bt_pyramid(uint16_t* __restrict__ pyr, Geo g, uint32_t g0,
                                                    uint64_t n0, uint32_t nlev) {
  __shared__ uint16_t s[kBT];
  const uint64_t base = (uint64_t)blockIdx.x * kBT;
  uint32_t cnt = (uint32_t)min<uint64_t>(kBT, n0 - base);
  if (threadIdx.x < cnt) s[threadIdx.x] = pyr[g.lvl[g0] + base + threadIdx.x];
  __syncthreads();
  uint64_t outBase = base;
  for (uint32_t l = 1; l <= nlev; ++l) {
    cnt >>= 1;
    outBase >>= 1;
    uint16_t v = 0;
    if (threadIdx.x < cnt) v = combine(s[2 * threadIdx.x], s[2 * threadIdx.x + 1]);
    __syncthreads();
    if (threadIdx.x < cnt) {
      s[threadIdx.x] = v;
      pyr[g.lvl[g0 + l] + outBase + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

struct Counter {
  uint32_t n = 0;
  __device__ __forceinline__ void operator()(uint32_t) { ++n; }
};

struct Writer {
  uint32_t* words;
  uint32_t pos, cur, acc;
  __device__ __forceinline__ void operator()(uint32_t b) {
    if ((pos >> 5) != cur) {
      if (acc) atomicOr(&words[cur], acc);
      cur = pos >> 5;
      acc = 0;
    }
    acc |= (b & 1u) << (pos & 31u);
    ++pos;
  }
  __device__ __forceinline__ void flush() { if (acc) atomicOr(&words[cur], acc); }
};

// Lane (q, k): its chunk's mask for plane k and the bits it emits.
__device__ __forceinline__ uint32_t lane_chunk_mask(const uint8_t* __restrict__ a, const Geo& g, uint64_t chunk0,
                                                    uint64_t& mine) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t myq = lane >> 3, myk = lane & 7u;
  mine = 0;
  for (uint32_t q = 0; q < (uint32_t)kChunksPerWave; ++q) {
    if (chunk0 + q >= g.C) break;   // uniform across the wave
    const Masks ms = chunk_masks(a, g, chunk0 + q);
    if (q == myq) {
      uint64_t m = ms.m[0];
#pragma unroll
      for (int k = 1; k < 8; ++k) m = (myk == (uint32_t)k) ? ms.m[k] : m;
      mine = m;
    }
  }
  return myk;
}

// 2. per-workgroup, per-plane bit counts: sums[k * nb + block]
__global__ void __launch_bounds__(kBT) k_bt_count(const uint8_t* __restrict__ a, Geo g,
                                                  const uint16_t* __restrict__ pyr, uint32_t* __restrict__ sums) {
  __shared__ uint32_t s[kWaves][8];
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint64_t chunk0 = ((uint64_t)blockIdx.x * kWaves + w) * kChunksPerWave;
  uint64_t mask;
  const uint32_t k = lane_chunk_mask(a, g, chunk0, mask);
  const uint64_t j = chunk0 + (lane >> 3);
  Counter cn;
  if (j < g.C) walk(mask, k, j, pyr, g, cn);
  uint32_t x = cn.n;
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  if (lane < 8) s[w][lane] = x;
  __syncthreads();
  if (threadIdx.x < 8) {
    uint32_t t = 0;
    for (int v = 0; v < kWaves; ++v) t += s[v][threadIdx.x];
    sums[(uint64_t)threadIdx.x * g.nb + blockIdx.x] = t;
  }
}

struct PlaneWords { uint64_t w[8]; };

// 4. emit at pre-order offsets
__global__ void __launch_bounds__(kBT) k_bt_emit(const uint8_t* __restrict__ a, Geo g,
                                                 const uint16_t* __restrict__ pyr, const uint32_t* __restrict__ offs,
                                                 PlaneWords pw, uint32_t* __restrict__ words) {
  __shared__ uint32_t s[kWaves][8];
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint64_t chunk0 = ((uint64_t)blockIdx.x * kWaves + w) * kChunksPerWave;
  uint64_t mask;
  const uint32_t k = lane_chunk_mask(a, g, chunk0, mask);
  const uint64_t j = chunk0 + (lane >> 3);
  Counter cn;
  if (j < g.C) walk(mask, k, j, pyr, g, cn);
  // exclusive prefix over the wave's chunks of the same plane (lanes k, k+8, ...)
  uint32_t x = cn.n;
#pragma unroll
  for (uint32_t d = 8; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane >= 56) s[w][lane - 56] = x;   // wave totals per plane
  __syncthreads();
  uint32_t pos = offs[(uint64_t)k * g.nb + blockIdx.x] + x - cn.n;
  for (uint32_t v = 0; v < w; ++v) pos += s[v][k];
  if (j < g.C && cn.n) {
    Writer wr{words + pw.w[k], pos, pos >> 5, 0u};
    walk(mask, k, j, pyr, g, wr);
    wr.flush();
  }
}

}  // namespace

// Returns the blob size; fills dst (host) when cap suffices.
uint64_t attr_bittree(svd_ctx* c, const uint8_t* dattrs, uint64_t V, int gray, uint8_t* dst, uint64_t cap) {
  if (V > (1ull << 30)) throw Error{SVD_ERR_CAPACITY, "bit-tree: more than 2^30 attributes"};
  uint64_t P = 1;
  uint32_t H = 0;
  while (P < V) { P <<= 1; ++H; }
  Geo g{};
  g.V = V;
  g.hc = H < 6 ? H : 6;
  g.G = H - g.hc;
  g.C = P >> g.hc;
  g.gray = gray;
  g.nb = (uint32_t)((g.C + kChunksPerBlock - 1) / kChunksPerBlock);
  uint64_t at = 0;
  for (uint32_t l = 0; l <= g.G; ++l) { g.lvl[l] = at; at += g.C >> l; }
  uint32_t bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint16_t* pyr = nullptr;
  uint32_t* sums = nullptr;
  uint32_t* total = dalloc<uint32_t>(c, 8);
  if (V) {
    pyr = dalloc<uint16_t>(c, at + 1);
    sums = dalloc<uint32_t>(c, 8ull * g.nb + 1);
    k_bt_states<<<g.nb, kBT, 0, c->stream>>>(dattrs, g, pyr);
    check_launch("k_bt_states");
    for (uint32_t l = 0; l < g.G;) {
      const uint32_t nlev = std::min<uint32_t>(8u, g.G - l);
      const uint64_t n0 = g.C >> l;
      k_bt_pyramid<<<div_up(n0, kBT), kBT, 0, c->stream>>>(pyr, g, l, n0, nlev);
      check_launch("k_bt_pyramid");
      l += nlev;
    }
    k_bt_count<<<g.nb, kBT, 0, c->stream>>>(dattrs, g, pyr, sums);
    check_launch("k_bt_count");
    for (int k = 0; k < 8; ++k) scan_exclusive(c, sums + (uint64_t)k * g.nb, sums + (uint64_t)k * g.nb, g.nb, total + k);
    read_u32s(c, total, bits, 8);
  }
  uint64_t nwords = 0;
  PlaneWords pw{};
  for (int k = 0; k < 8; ++k) { pw.w[k] = nwords; nwords += (bits[k] + 31) / 32; }
  const uint64_t bytes = 48 + 4 * nwords;
  if (dst && cap >= bytes) {
    uint32_t* words = dalloc<uint32_t>(c, nwords + 1);
    fill_u32(c, words, 0u, nwords + 1);
    if (V) k_bt_emit<<<g.nb, kBT, 0, c->stream>>>(dattrs, g, pyr, sums, pw, words);
    check_launch("k_bt_emit");
    const uint32_t hdr[4] = {0x42415653u, (uint32_t)V, (uint32_t)gray, 8u};
    memcpy(dst, hdr, 16);
    memcpy(dst + 16, bits, 32);
    if (nwords) {
      SVD_HIP(hipMemcpyAsync(dst + 48, words, 4 * nwords, hipMemcpyDeviceToHost, c->stream));
    }
    SVD_HIP(hipStreamSynchronize(c->stream));
    dfree(c, words);
  }
  dfree(c, pyr);
  dfree(c, sums);
  dfree(c, total);
  return bytes;
}

}  // namespace svd
