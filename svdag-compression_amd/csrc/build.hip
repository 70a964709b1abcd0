// Bottom-up level reduction on the device:
//   * SVO levels from a sorted Morton voxel list (point-major DFS order),
//   * toDAG  : exact-key hash dedup, first occurrence wins (geom_octree.cpp:462-548),
//   * toSDAG : 8 mirror variants per node resolved in the reference's order
//              (geom_octree.cpp:551-697).
// Node records are 9 u32 words: w0 = mask | mX<<8 | mY<<16 | mZ<<24,
// w1..w8 = children[0..7]; the finest level stores only the mask word.
#include "svd_internal.hpp"
#include "block_scan.hpp"
#include "lookback.hpp"
#include "scan_op.hpp"
#include <algorithm>

namespace svd {

namespace {

constexpr int kB = 256;

// ---------------------------------------------------------------- SVO build
__device__ __forceinline__ bool head64(const uint64_t* code, uint64_t i) {
  return i == 0 || (code[i] >> 3) != (code[i - 1] >> 3);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Sources of the codes the one-pass level builds read: plain codes, or a
// batch's per-scene codes tagged on the fly with their scene id above bit
// kSceneTagShift (scene-major, so the tagged list ascends when each scene's
// does).  A code at or above 2^(3 * its scene's levels) comes back with bit
// 63 set, which the leaf pass's input check catches.
// at_tile(tileBase) lets a source cache what the tile's codes share (run_tile
// calls it once the tile's code loads are issued); code i is tag(i, raw(i)),
// split so that a tile's loads all issue before any of the (branching)
// tagging.
struct PlainCodes {
  const uint64_t* __restrict__ code;
  __device__ __forceinline__ void at_tile(uint64_t) {}
  __device__ __forceinline__ uint64_t raw(uint64_t i) const { return code[i]; }
  __device__ __forceinline__ uint64_t tag(uint64_t, uint64_t c) const { return c; }
  __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return code[i]; }
};
struct TaggedCodes {
  const uint64_t* __restrict__ code;
  SceneTags t;
  uint32_t tileSize;
  // the scene s0 of the tile's first code and the next one: their code
  // ranges [b0, e0), [e0, e1) and level shifts, one 32-byte record per tile
  // (k_tile_scenes).  A tile whose codes span more scenes, or a code before
  // the tile, bisects: a slow tile holds up its successors' look-back.
  uint32_t s0 = 0, sh0 = 0, sh1 = 0;
  uint64_t b0 = 0, e0 = 0, e1 = 0;
  __device__ __forceinline__ void at_tile(uint64_t tileBase) {
    const uint32_t* r = t.tileScene + 8 * (tileBase / tileSize);
    b0 = ((uint64_t)r[1] << 32) | r[0];
    e0 = ((uint64_t)r[3] << 32) | r[2];
    e1 = ((uint64_t)r[5] << 32) | r[4];
    s0 = r[6];
    sh0 = r[7] & 0xFFu;
    sh1 = r[7] >> 8;
  }
  __device__ __forceinline__ uint64_t tagged(uint64_t c, uint32_t s, uint32_t sh) const {
    const uint64_t out = (c >> sh) != 0 ? (1ull << 63) : 0ull;
    return out | ((uint64_t)s << kSceneTagShift) | (c & ((1ull << kSceneTagShift) - 1));
  }
  __device__ __forceinline__ uint64_t raw(uint64_t i) const { return code[i]; }
  __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return tag(i, code[i]); }
  __device__ __forceinline__ uint64_t tag(uint64_t i, uint64_t c) const {
    if (i >= b0 && i < e0) return tagged(c, s0, sh0);
    if (i >= e0 && i < e1) return tagged(c, s0 + 1, sh1);
    // another scene: the last s with off[s] <= i, by bisection on the side
    // of the tile's scenes that holds it
    uint32_t lo = i < b0 ? 0u : s0 + 2, hi = i < b0 ? s0 : t.S;
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) >> 1;
      if (t.off[m] <= i) lo = m; else hi = m;
    }
    return tagged(c, lo, 3 * t.levels[lo]);
  }
};

// A tile of sorted codes for the one-pass level builds: thread (wave w, lane
// j) holds codes w * 64 * PER + 64 * k + j (k < PER) of the tile, so every
// load is coalesced; a run of codes sharing code >> 3 is one parent node.
// After run_tile: head bit k = code k starts a run, rank[k] = the run's
// number within the tile counted from the tile's first head (-1: the code
// continues the previous tile's last run), agg = the tile's runs.
// 512 threads per run tile: tile numbers come from one device counter whose
// same-address atomics serialise, so fewer, larger tiles start sooner.
constexpr int kRunT = 512;
template <int PER>
struct RunTile {
  uint64_t v[PER];
  int32_t rank[PER];
  uint32_t heads = 0;
  uint32_t agg = 0;
};
template <int PER, typename Src>
__device__ __forceinline__ void run_tile(Src& code, uint64_t n, uint64_t tileBase, RunTile<PER>& t,
                                         uint32_t* wtot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t wbase = tileBase + (uint64_t)wave * (64 * PER) + lane;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint64_t i = wbase + 64 * k;
    t.v[k] = i < n ? code.raw(i) : 0ull;
  }
  code.at_tile(tileBase);   // its loads queue behind the codes'
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint64_t i = wbase + 64 * k;
    if (i < n) t.v[k] = code.tag(i, t.v[k]);
  }
  const uint64_t first = tileBase + (uint64_t)wave * (64 * PER);
  const uint64_t before = (lane == 0 && first > 0 && first - 1 < n) ? code(first - 1) : 0ull;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint64_t i = wbase + 64 * k;
    const uint64_t up = __shfl_up(t.v[k], 1, 64);
    const uint64_t wrap = k ? __shfl(t.v[k - 1], 63, 64) : before;
    const uint64_t u = lane ? up : wrap;
    const bool h = i < n && (i == 0 || (t.v[k] >> 3) != (u >> 3));
    const uint64_t b = __ballot(h);
    t.heads |= (h ? 1u : 0u) << k;
    // inclusive head count up to this code, minus one
    t.rank[k] = (int32_t)(c + lanes_below(b) + (h ? 1u : 0u)) - 1;
    c += (uint32_t)__popcll(b);
  }
  if (lane == 0) wtot[wave] = c;
  __syncthreads();
  uint32_t off = 0, agg = 0;
#pragma unroll
  for (int w = 0; w < kRunT / 64; ++w) {
    if (w < wave) off += wtot[w];
    agg += wtot[w];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) t.rank[k] += (int32_t)off;
  t.agg = agg;
}

// Mask bits of the codes from position `end` on that continue the run of
// parent code pc (at most 7), loaded together.
template <typename Src>
__device__ __forceinline__ uint32_t run_tail_bits(const Src& code, uint64_t n, uint64_t end, uint64_t pc) {
  uint64_t y[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) y[j] = end + j < n ? code(end + j) : ~0ull;
  uint32_t m = 0;
  bool same = true;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    same = same && (y[j] >> 3) == pc;
    if (same) m |= 1u << (uint32_t)(y[j] & 7u);
  }
  return m;
}

// Input contract: strictly ascending codes with no bit of hiMask set (plain
// codes: every bit from 3 * levels up; scene-tagged codes: bit 63, which
// TaggedCodes sets on a code beyond its scene's depth).
// One pass over the codes for the finest level: the input check, every
// level's node count, the run heads, their numbering by a single-pass scan,
// and the leaf masks and parent codes.  Code i starts a node of the level k
// steps above the voxels (prefix code >> 3k) exactly when k <= K_i =
// floor(msb(code[i] ^ code[i-1]) / 3) (K_0 = L), so the level k steps up has
// #{i : K_i >= k} nodes; part[tile * (L + 1) + k] = that count over the tile
// (k >= 1), part[tile * (L + 1)] = the tile's bad codes.  Counts go per tile
// rather than to device atomics, whose same-address updates from thousands
// of workgroups would queue.  Each code ORs its bit into its run's LDS mask;
// a run that crosses the tile's end reads the codes past it.  With bad input
// the writes stay below n and the caller throws on the bad count.
constexpr int kLeafPer = 16;
constexpr uint32_t kLeafTile = kRunT * kLeafPer;
template <typename Src>
__global__ void __launch_bounds__(kRunT) k_leaf_build(Src src, uint64_t n, uint32_t L, uint64_t hiMask, LbPass p,
                                                      uint32_t* __restrict__ part, uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ mask, uint64_t* __restrict__ pcode) {
  __shared__ uint32_t lds[kRunT / 64 + 2];
  __shared__ uint32_t word;
  __shared__ uint32_t cpart[kMaxLevels + 1][kRunT / 64];
  __shared__ uint32_t lmask[kLeafTile];
  const uint32_t tile = lb_take_tile(p, &word);
  const uint64_t tileBase = (uint64_t)tile * kLeafTile;
  if (tile == 0 && threadIdx.x <= L) hist[threadIdx.x] = 0;   // k_level_counts adds to it
  for (uint32_t j = threadIdx.x; j < kLeafTile; j += kRunT) lmask[j] = 0;
  Src code = src;
  RunTile<kLeafPer> t;
  run_tile(code, n, tileBase, t, lds);   // its barrier also orders the clearing
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t wbase = tileBase + (uint64_t)wave * (64 * kLeafPer) + lane;
  // per-thread histogram of K over the good codes, 6-bit bins packed in three
  // 64-bit words (bins 0-9, 10-19, 20-21; at most kLeafPer < 64 per bin):
  // a few operations per code instead of one counter per level
  static_assert(kLeafPer < 64, "6-bit histogram bins");
  uint64_t hA = 0, hB = 0, hC = 0;
  uint32_t nBad = 0;
#pragma unroll
  for (int k = 0; k < kLeafPer; ++k) {
    const uint64_t i = wbase + 64 * k;
    const uint64_t x = t.v[k];
    const uint64_t up = __shfl_up(x, 1, 64);
    const uint64_t wrap = k ? __shfl(t.v[k - 1], 63, 64) : (lane == 0 && i > 0 && i - 1 < n ? code(i - 1) : 0ull);
    const uint64_t u = lane ? up : wrap;
    if (i >= n) continue;
    uint32_t K = L;
    bool bad = (x & hiMask) != 0;
    if (i > 0) {
      if (x <= u) bad = true;
      else K = min(L, (uint32_t)(63 - __clzll((long long)(x ^ u))) / 3u);
    }
    nBad += bad ? 1u : 0u;
    const uint32_t kb = K < 10 ? K : (K < 20 ? K - 10 : K - 20);
    const uint64_t one = bad ? 0ull : 1ull << (6 * kb);
    hA += K < 10 ? one : 0ull;
    hB += (K >= 10 && K < 20) ? one : 0ull;
    hC += K >= 20 ? one : 0ull;
    if (t.rank[k] >= 0) atomicOr(&lmask[t.rank[k]], 1u << (uint32_t)(x & 7u));
  }
  // level q (q >= 1) counts the good codes with K >= q: suffix sums of the bins
  uint32_t cnt[kMaxLevels + 1];
  cnt[0] = nBad;
  {
    uint32_t run = 0;
#pragma unroll
    for (int q = (int)kMaxLevels; q >= 1; --q) {
      const uint64_t w = q < 10 ? hA : (q < 20 ? hB : hC);
      const int sh = 6 * (q < 10 ? q : (q < 20 ? q - 10 : q - 20));
      run += (uint32_t)(w >> sh) & 63u;
      cnt[q] = run;
    }
  }
#pragma unroll
  for (int q = 0; q <= (int)kMaxLevels; ++q) {
    uint32_t x = cnt[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) cpart[q][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x <= L) {
    uint32_t x = 0;
#pragma unroll
    for (int w = 0; w < kRunT / 64; ++w) x += cpart[threadIdx.x][w];
    part[(uint64_t)tile * (L + 1) + threadIdx.x] = x;
  }
  const uint32_t prefix = lb_exclusive(p, tile, t.agg, lds);
  const uint64_t tileEnd = tileBase + kLeafTile;
#pragma unroll
  for (int k = 0; k < kLeafPer; ++k) {
    if (!((t.heads >> k) & 1u)) continue;
    const uint32_t r = (uint32_t)t.rank[k];
    uint32_t m = lmask[r];
    const uint64_t pc = t.v[k] >> 3;
    if (r + 1 == t.agg && tileEnd < n) m |= run_tail_bits(code, n, tileEnd, pc);
    mask[prefix + r] = m;
    pcode[prefix + r] = pc;
  }
}

// An inner level in one pass: run heads of the child codes, their numbering
// by a single-pass scan, the records (child pointer = the child's position
// in its level) and the parent codes.  Every code stores its own child
// pointer into its parent's record and ORs its slot into the run's LDS mask;
// the head writes the mask and kNull into the other slots.
constexpr int kInnerPer = 16;
constexpr uint32_t kInnerTile = kRunT * kInnerPer;
__global__ void __launch_bounds__(kRunT) k_inner_build(const uint64_t* __restrict__ ccode, uint64_t n, LbPass p,
                                                       uint32_t* __restrict__ rec, uint64_t* __restrict__ pcode) {
  __shared__ uint32_t lds[kRunT / 64 + 2];
  __shared__ uint32_t word;
  __shared__ uint32_t lmask[kInnerTile];
  const uint32_t tile = lb_take_tile(p, &word);
  const uint64_t tileBase = (uint64_t)tile * kInnerTile;
  for (uint32_t j = threadIdx.x; j < kInnerTile; j += kRunT) lmask[j] = 0;
  RunTile<kInnerPer> t;
  PlainCodes src{ccode};
  run_tile(src, n, tileBase, t, lds);
  const uint32_t prefix = lb_exclusive(p, tile, t.agg, lds);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t wbase = tileBase + (uint64_t)wave * (64 * kInnerPer) + lane;
#pragma unroll
  for (int k = 0; k < kInnerPer; ++k) {
    const uint64_t i = wbase + 64 * k;
    if (i >= n) continue;
    const uint32_t sl = (uint32_t)(t.v[k] & 7u);
    // rank -1: the run began in the previous tile, whose last node is prefix - 1
    rec[(uint64_t)(prefix + t.rank[k]) * kRecWords + 1 + sl] = (uint32_t)i;
    if (t.rank[k] >= 0) atomicOr(&lmask[t.rank[k]], 1u << sl);
  }
  __syncthreads();
  const uint64_t tileEnd = tileBase + kInnerTile;
#pragma unroll
  for (int k = 0; k < kInnerPer; ++k) {
    if (!((t.heads >> k) & 1u)) continue;
    const uint32_t r = (uint32_t)t.rank[k];
    uint32_t m = lmask[r];
    const uint64_t pc = t.v[k] >> 3;
    if (r + 1 == t.agg && tileEnd < n) m |= run_tail_bits(src, n, tileEnd, pc);
    uint32_t* o = rec + (uint64_t)(prefix + r) * kRecWords;
    o[0] = m;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (!((m >> c) & 1u)) o[1 + c] = kNull;
    if (pcode) pcode[prefix + r] = pc;
  }
}

// Level counts of k_leaf_build: hist[k] = sum over tiles of part[t][k], all
// counters in one pass over the tiles; each workgroup adds its tiles' sums
// to hist, which k_leaf_build's first tile cleared.
constexpr uint32_t kCountTilesPerWg = 8192;
__global__ void __launch_bounds__(1024) k_level_counts(const uint32_t* __restrict__ part, uint32_t tiles, uint32_t L,
                                                       uint32_t* __restrict__ hist) {
  __shared__ uint32_t red[kMaxLevels + 1][1024 / 64];
  uint32_t s[kMaxLevels + 1];
#pragma unroll
  for (int q = 0; q <= (int)kMaxLevels; ++q) s[q] = 0;
  for (uint32_t t = blockIdx.x * 1024 + threadIdx.x; t < tiles; t += gridDim.x * 1024) {
    const uint32_t* pt = part + (uint64_t)t * (L + 1);
#pragma unroll
    for (int q = 0; q <= (int)kMaxLevels; ++q)
      if ((uint32_t)q <= L) s[q] += pt[q];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q <= (int)kMaxLevels; ++q) {
    uint32_t x = s[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[q][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x <= L) {
    uint32_t x = 0;
    for (int v = 0; v < 1024 / 64; ++v) x += red[threadIdx.x][v];
    atomicAdd(&hist[threadIdx.x], x);
  }
}

// An inner SVO level from at most kSvoSmall child codes in one workgroup: the
// run heads (a code whose parent code differs from its predecessor's), their
// scan, and one record per run (child pointer = the child's position in its
// level), with workgroup barriers between the steps.
constexpr uint32_t kSvoSmall = 8192;
__global__ void __launch_bounds__(1024) k_inner_level_small(const uint64_t* __restrict__ ccode, uint32_t n,
                                                            uint32_t* __restrict__ rec, uint64_t* __restrict__ pcode) {
  __shared__ uint32_t pidx[kSvoSmall];
  __shared__ uint32_t lds[1024 / 64 + 1];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t h = (i < n && head64(ccode, i)) ? 1u : 0u;
    uint32_t agg;
    const uint32_t ex = block_exclusive_sum(h, lds, agg) + carry;
    if (i < n) pidx[i] = ex;
    carry += agg;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += 1024) {
    if (!head64(ccode, i)) continue;
    const uint64_t p = ccode[i] >> 3;
    uint32_t w[kRecWords];
    w[0] = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) w[1 + c] = kNull;
    for (uint32_t j = i; j < n && j < i + 8 && (ccode[j] >> 3) == p; ++j) {
      const uint32_t sl = (uint32_t)(ccode[j] & 7u);
      w[0] |= 1u << sl;
      w[1 + sl] = j;
    }
    uint32_t* o = rec + (uint64_t)pidx[i] * kRecWords;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
    if (pcode) pcode[pidx[i]] = p;
  }
}

// Every inner SVO level from the first with at most kSvoSmall child codes up
// to the root in one workgroup: per level the heads scan and the records as
// k_inner_level_small, the parent codes of one level the child codes of the
// next (ping-pong between ca and cb in global memory, ordered by barriers).
struct SvoTop {
  uint32_t* rec[kMaxLevels];
  int32_t top;                 // first level handled (levels top .. 0)
};
__global__ void __launch_bounds__(1024) k_inner_levels_small(uint64_t* ca, uint64_t* cb, uint32_t n, SvoTop lv) {
  __shared__ uint32_t pidx[kSvoSmall];
  __shared__ uint32_t lds[1024 / 64 + 1];
  for (int lev = lv.top; lev >= 0; --lev) {
    const uint64_t* ccode = ca;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
      const uint32_t i = base + threadIdx.x;
      const uint32_t h = (i < n && head64(ccode, i)) ? 1u : 0u;
      uint32_t agg;
      const uint32_t ex = block_exclusive_sum(h, lds, agg) + carry;
      if (i < n) pidx[i] = ex;
      carry += agg;
    }
    __syncthreads();
    uint32_t* rec = lv.rec[lev];
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
      if (!head64(ccode, i)) continue;
      const uint64_t p = ccode[i] >> 3;
      uint32_t w[kRecWords];
      w[0] = 0;
#pragma unroll
      for (int c = 0; c < 8; ++c) w[1 + c] = kNull;
      for (uint32_t j = i; j < n && j < i + 8 && (ccode[j] >> 3) == p; ++j) {
        const uint32_t sl = (uint32_t)(ccode[j] & 7u);
        w[0] |= 1u << sl;
        w[1 + sl] = j;
      }
      uint32_t* o = rec + (uint64_t)pidx[i] * kRecWords;
#pragma unroll
      for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
      cb[pidx[i]] = p;
    }
    __threadfence_block();
    __syncthreads();
    n = carry;
    uint64_t* t = ca;
    ca = cb;
    cb = t;
  }
}

// ---------------------------------------------------------------- hash dedup
__device__ __forceinline__ void load_rec(const uint32_t* __restrict__ rec, uint64_t i, uint32_t w[kRecWords]) {
  const uint32_t* r = rec + i * kRecWords;
#pragma unroll
  for (int k = 0; k < kRecWords; ++k) w[k] = r[k];
}

// Record j is loaded whole before it is compared: a fingerprint match is
// almost always an equal key, and nine loads issued together cost one memory
// round trip where an early exit per word would chain nine.
__device__ __forceinline__ bool rec_eq(const uint32_t* __restrict__ rec, uint32_t j, const uint32_t w[kRecWords]) {
  uint32_t r[kRecWords];
  load_rec(rec, j, r);
  uint32_t diff = 0;
#pragma unroll
  for (int k = 0; k < kRecWords; ++k) diff |= r[k] ^ w[k];
  return diff == 0;
}

// Open-addressing table of node indices.  An entry is (fingerprint << 32 |
// index), the fingerprint being the key hash's upper 32 bits, published by one
// 64-bit CAS: a probe compares fingerprints in registers and reads the 36-byte
// key only when they agree.  Empty slots are all ones (two kNone words).
using TEntry = unsigned long long;
constexpr TEntry kEmpty = ~0ull;

// Probe the table for node i's key (record w, hash hv) and insert it; the
// slot ends holding the smallest index of all nodes with an equal key (equal
// keys share the fingerprint).  eq(j) compares record j with w.  Returns the
// slot.
// The device table goes straight to the CAS (it returns the entry it meets,
// one round trip); the LDS table first loads.
template <bool CasFirst, typename Eq>
__device__ __forceinline__ uint32_t insert_key(TEntry* table, uint32_t cmask, uint64_t hv, uint32_t i, Eq eq) {
  const uint32_t fp = (uint32_t)(hv >> 32);
  const TEntry mine = ((TEntry)fp << 32) | (TEntry)i;
  uint32_t h = (uint32_t)hv & cmask;
  while (true) {
    TEntry e = CasFirst ? kEmpty : __atomic_load_n(&table[h], __ATOMIC_RELAXED);
    if (e == kEmpty) {
      const TEntry old = atomicCAS(&table[h], kEmpty, mine);
      if (old == kEmpty) return h;
      e = old;
    }
    if ((uint32_t)(e >> 32) == fp && eq((uint32_t)e)) {
      // the slot only decreases, so a reader above the value it saw has
      // nothing to add: a key costs one atomic per decrease, not per node
      if (mine < e) atomicMin(&table[h], mine);
      return h;
    }
    h = (h + 1) & cmask;
  }
}

// Insert the nodes of a level; slot[i] = the table slot of node i's key, whose
// entry ends holding the smallest index of all nodes with that key.  Empty
// nodes (mask 0) are skipped, like the reference's `if (!n.hasChildren())
// continue;`.  A common key (a full brick, say) occurs in most workgroups many
// times, and atomics on one address serialize, so each workgroup first
// dedups its nodes in an LDS table (keys compared with the records staged in
// LDS) and only the smallest index of each of its keys goes to the device
// table; the others take its slot.  PER nodes per thread: 4 on the big levels
// (fewer workgroups per common key), 1 on the others, where the chain of
// memory round trips per node sets the time.
constexpr int kInsT = 256;
constexpr uint64_t kInsWide = 1u << 19;   // levels from this size use 4 nodes per thread
template <int PER>
__global__ void __launch_bounds__(kInsT) k_insert(const uint32_t* __restrict__ rec, uint64_t n,
                                                  TEntry* __restrict__ table, uint32_t cmask,
                                                  uint32_t* __restrict__ slot) {
  constexpr uint32_t kNodes = kInsT * PER, kCap = 2 * kNodes;   // LDS table load <= 1/2
  __shared__ TEntry lt[kCap];
  __shared__ uint32_t gs[kCap];
  __shared__ uint16_t ls[kNodes];
  __shared__ uint32_t lrec[kNodes * kRecWords];
  const uint32_t tid = threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kNodes;
  for (uint32_t j = tid; j < kCap; j += kInsT) lt[j] = kEmpty;
  uint32_t w[PER][kRecWords];
  bool live[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t loc = k * kInsT + tid;
    const uint64_t i = base + loc;
    live[k] = false;
    ls[loc] = 0xFFFFu;
    if (i >= n) continue;
    load_rec(rec, i, w[k]);
#pragma unroll
    for (int q = 0; q < kRecWords; ++q) lrec[loc * kRecWords + q] = w[k][q];
    if ((w[k][0] & 0xFFu) == 0) slot[i] = kNone;
    else live[k] = true;
  }
  __syncthreads();
  uint64_t hv[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if (!live[k]) continue;
    const uint32_t loc = k * kInsT + tid;
    hv[k] = hash_rec(w[k]);
    const uint32_t* wk = w[k];
    auto eqLds = [&](uint32_t j) {
      const uint32_t* r = lrec + (j - (uint32_t)base) * kRecWords;
      uint32_t diff = 0;
#pragma unroll
      for (int q = 0; q < kRecWords; ++q) diff |= r[q] ^ wk[q];
      return diff == 0;
    };
    ls[loc] = (uint16_t)insert_key<false>(lt, kCap - 1, hv[k], (uint32_t)(base + loc), eqLds);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t loc = k * kInsT + tid;
    const uint32_t l = ls[loc];
    if (!live[k] || (uint32_t)lt[l] != (uint32_t)(base + loc)) continue;
    const uint32_t* wk = w[k];
    auto eqDev = [&](uint32_t j) { return rec_eq(rec, j, wk); };
    gs[l] = insert_key<true>(table, cmask, hv[k], (uint32_t)(base + loc), eqDev);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t loc = k * kInsT + tid;
    if (live[k]) slot[base + loc] = gs[ls[loc]];
  }
}

void launch_insert(svd_ctx* c, const uint32_t* rec, uint64_t n, TEntry* table, uint32_t cmask, uint32_t* slot) {
  if (n >= kInsWide) k_insert<4><<<div_up(n, 4 * kInsT), kInsT, 0, c->stream>>>(rec, n, table, cmask, slot);
  else k_insert<1><<<div_up(n, kInsT), kInsT, 0, c->stream>>>(rec, n, table, cmask, slot);
}

__device__ __forceinline__ uint32_t entry_index(TEntry e) { return (uint32_t)e; }

// Probe for a key without inserting; returns the stored (minimum) index.
__device__ __forceinline__ uint32_t lookup(const uint32_t* __restrict__ rec, const TEntry* __restrict__ table,
                                           uint32_t cmask, const uint32_t w[kRecWords]) {
  const uint64_t hv = hash_rec(w);
  const uint32_t fp = (uint32_t)(hv >> 32);
  uint32_t h = (uint32_t)hv & cmask;
  while (true) {
    const TEntry e = table[h];
    if (e == kEmpty) return kNone;
    if ((uint32_t)(e >> 32) == fp && rec_eq(rec, (uint32_t)e, w)) return (uint32_t)e;
    h = (h + 1) & cmask;
  }
}

// A small inner level of toDAG (n <= kSmallLevel) in one workgroup: the
// children remap, the first-occurrence insert (table in LDS), the
// representative flags, their scan and the compaction of k_remap_children,
// k_insert, the representative flags and their scan, and k_compact, with workgroup
// barriers instead of launches.  The table size does not change the result:
// each key's slot ends with its smallest index whatever the probe sequence.
constexpr uint32_t kSmallLevel = 4096;
constexpr uint32_t kSmallCap = 8192;        // power of two >= 2 * kSmallLevel
// toSDAG's one-workgroup path pays a long serial chain per node it handles:
// measured on the 2048^3 terrain it wins up to one node per thread only
constexpr uint32_t kSdagSmall = 1024;
constexpr int kSmallThreads = 1024;
struct SmallLds {
  TEntry table[kSmallCap];
  uint16_t slotOf[kSmallLevel];
  uint32_t posOf[kSmallLevel];
  uint32_t lds[kSmallThreads / 64 + 1];
};
__device__ void dedup_small_level(const uint32_t* __restrict__ src, uint32_t n, const uint32_t* __restrict__ remapIn,
                                  uint32_t* __restrict__ keyed, uint32_t* __restrict__ out,
                                  uint32_t* __restrict__ remapOut, uint32_t* __restrict__ total, SmallLds& sh) {
  TEntry* table = sh.table;
  uint16_t* slotOf = sh.slotOf;
  uint32_t* posOf = sh.posOf;
  uint32_t* lds = sh.lds;
  const uint32_t tid = threadIdx.x;
  for (uint32_t j = tid; j < kSmallCap; j += kSmallThreads) table[j] = kEmpty;
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    uint32_t w[kRecWords];
    load_rec(src, i, w);
    const uint32_t m = w[0] & 0xFFu;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if ((m >> c) & 1u) w[1 + c] = remapIn[w[1 + c]];
    uint32_t* o = keyed + (uint64_t)i * kRecWords;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    uint32_t w[kRecWords];
    load_rec(keyed, i, w);
    if ((w[0] & 0xFFu) == 0) { slotOf[i] = 0xFFFFu; continue; }
    const uint64_t hv = hash_rec(w);
    const uint32_t fp = (uint32_t)(hv >> 32);
    const TEntry mine = ((TEntry)fp << 32) | (TEntry)i;
    uint32_t h = (uint32_t)hv & (kSmallCap - 1);
    while (true) {
      TEntry e = table[h];
      if (e == kEmpty) {
        const TEntry old = atomicCAS(&table[h], kEmpty, mine);
        if (old == kEmpty) break;
        e = old;
      }
      if ((uint32_t)(e >> 32) == fp && rec_eq(keyed, (uint32_t)e, w)) {
        if (mine < e) atomicMin(&table[h], mine);
        break;
      }
      h = (h + 1) & (kSmallCap - 1);
    }
    slotOf[i] = (uint16_t)h;
  }
  __syncthreads();
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kSmallThreads) {
    const uint32_t i = base + tid;
    uint32_t f = 0;
    if (i < n && slotOf[i] != 0xFFFFu) f = entry_index(table[slotOf[i]]) == i ? 1u : 0u;
    uint32_t agg;
    const uint32_t ex = block_exclusive_sum(f, lds, agg) + carry;
    if (i < n) posOf[i] = ex;
    carry += agg;
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    const uint32_t sl = slotOf[i];
    // an empty (skipped) node keeps correspondence 0 (geom_octree.cpp:489-495)
    if (sl == 0xFFFFu) { remapOut[i] = 0; continue; }
    const uint32_t r = entry_index(table[sl]);
    remapOut[i] = posOf[r];
    if (r == i) {
      const uint32_t* a = keyed + (uint64_t)i * kRecWords;
      uint32_t* b = out + (uint64_t)posOf[i] * kRecWords;
#pragma unroll
      for (int k = 0; k < kRecWords; ++k) b[k] = a[k];
    }
  }
  if (tid == 0) *total = carry;
}


// toDAG of every level from `top` up to the root in one workgroup, each of
// at most kSmallLevel nodes: per level dedup_small_level (children remapped
// by the level below's correspondence, then the first-occurrence dedup); the
// root only has its children remapped (it is never reduced,
// geom_octree.cpp:483).  A level's correspondence goes to global memory and
// is read by the next level after a barrier.
struct DagTop {
  const uint32_t* src[kMaxLevels];
  uint32_t* out[kMaxLevels];
  uint32_t n[kMaxLevels];
  const uint32_t* remapIn;   // correspondence of level top + 1
  uint32_t* remap[2];        // ping-pong correspondences of the levels handled
  uint32_t* keyed;           // scratch: n[top] records
  uint32_t* counts;          // counts[lev] = unique nodes of level lev
  int32_t top;
};
__global__ void __launch_bounds__(kSmallThreads) k_dedup_levels_small(DagTop a) {
  __shared__ SmallLds sh;
  const uint32_t* remapIn = a.remapIn;
  for (int lev = a.top; lev >= 1; --lev) {
    uint32_t* remapOut = a.remap[(a.top - lev) & 1];
    dedup_small_level(a.src[lev], a.n[lev], remapIn, a.keyed, a.out[lev], remapOut, a.counts + lev, sh);
    __threadfence_block();
    __syncthreads();
    remapIn = remapOut;
  }
  for (uint32_t i = threadIdx.x; i < a.n[0]; i += kSmallThreads) {
    uint32_t w[kRecWords];
    load_rec(a.src[0], i, w);
    const uint32_t m = w[0] & 0xFFu;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if ((m >> c) & 1u) w[1 + c] = remapIn[w[1 + c]];
    uint32_t* o = a.out[0] + (uint64_t)i * kRecWords;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
  }
}

// pos[i] = number of representatives before node i (scan_op): a node is its
// key's representative when the table slot holds its own index.
struct RepFlagOp {
  const uint32_t* __restrict__ slot;
  const TEntry* __restrict__ table;
  uint32_t* __restrict__ pos;
  __device__ uint32_t value(uint64_t i, uint32_t&) const {
    const uint32_t s = slot[i];
    return (s != kNone && entry_index(table[s]) == (uint32_t)i) ? 1u : 0u;
  }
  __device__ void apply(uint64_t i, uint32_t ex) const { pos[i] = ex; }
  __device__ void report(uint32_t) const {}
};

__global__ void k_compact(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ slot,
                          const TEntry* __restrict__ table, const uint32_t* __restrict__ pos, uint64_t n,
                          uint32_t* __restrict__ out, uint32_t* __restrict__ remap) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = slot[i];
  // an empty (skipped) node keeps correspondence 0, like the reference's
  // zero-initialised correspondences vector (geom_octree.cpp:489-495)
  if (s == kNone) { remap[i] = 0; return; }
  const uint32_t r = entry_index(table[s]);
  remap[i] = pos[r];
  if (r == (uint32_t)i) {
    const uint32_t* src = rec + i * kRecWords;
    uint32_t* dst = out + (uint64_t)pos[i] * kRecWords;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) dst[k] = src[k];
  }
}

// children[c] <- remap[children[c]] where mask bit c is set (geom_octree.cpp:519-538)
// clear (optional): a hash table of nClear entries set to empty on the way,
// grid-stride, in place of a fill launch before the next level's insert.
__global__ void k_remap_children(const uint32_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ remap,
                                 uint32_t* __restrict__ out, unsigned long long* __restrict__ clear = nullptr,
                                 uint64_t nClear = 0) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t j = i; j < nClear; j += (uint64_t)gridDim.x * blockDim.x) clear[j] = ~0ull;
  if (i >= n) return;
  uint32_t w[kRecWords];
  load_rec(in, i, w);
  const uint32_t m = w[0] & 0xFFu;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if ((m >> c) & 1u) w[1 + c] = remap[w[1 + c]];
  uint32_t* o = out + i * kRecWords;
#pragma unroll
  for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
}

// ---------------------------------------------------------------- leaf level
// Each block folds a contiguous chunk of masks into an LDS table first, so the
// global table sees at most 255 atomics per block instead of one per node
// (all nodes hit the same 255 addresses).
constexpr int kFoBlocks = 1024;
__global__ void __launch_bounds__(256) k_leaf_firstocc(const uint32_t* __restrict__ mask, uint64_t n,
                                                       uint32_t* __restrict__ firstocc) {
  __shared__ uint32_t fo[256];
  fo[threadIdx.x] = kNone;
  __syncthreads();
  const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const uint32_t m = mask[i] & 0xFFu;
    if (m && fo[m] > (uint32_t)i) atomicMin(&fo[m], (uint32_t)i);
  }
  __syncthreads();
  const uint32_t v = fo[threadIdx.x];
  if (v != kNone) atomicMin(&firstocc[threadIdx.x], v);
}

// Distinct masks ranked by first occurrence -> DAG ids.
__global__ void k_leaf_rank(const uint32_t* __restrict__ firstocc, uint32_t* __restrict__ idOfMask,
                            uint32_t* __restrict__ dagMask, uint32_t* __restrict__ count) {
  __shared__ uint32_t fo[256];
  const int m = threadIdx.x;
  fo[m] = firstocc[m];
  __syncthreads();
  uint32_t id = kNone;
  if (fo[m] != kNone) {
    uint32_t r = 0;
    for (int k = 0; k < 256; ++k) r += (fo[k] < fo[m]) ? 1u : 0u;
    id = r;
    dagMask[r] = (uint32_t)m;
  }
  idOfMask[m] = id;
  if (m == 0) {
    uint32_t cnt = 0;
    for (int k = 0; k < 256; ++k) cnt += fo[k] != kNone;
    *count = cnt;
  }
}

__global__ void k_leaf_remap(const uint32_t* __restrict__ mask, uint64_t n, const uint32_t* __restrict__ idOfMask,
                             uint32_t* __restrict__ remap) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = mask[i] & 0xFFu;
  remap[i] = m ? idOfMask[m] : 0u;
}

// ---------------------------------------------------------------- symmetry
// Transform order of the reference lookup: I, X, Y, Z, XY, XZ, YZ, XYZ, as
// child-index xor masks {0, 4, 2, 1, 6, 5, 3, 7}, nibble t of one constant
// (no table load in a lane's dependency chain).
__device__ __forceinline__ uint32_t flip_of(int t) { return (0x73561240u >> (4 * t)) & 0xFu; }

// Bit c of every byte of w moves to bit c ^ flip of that byte: one
// conditional swap of bit groups per flip bit.
__device__ __forceinline__ uint32_t perm_bytes(uint32_t w, uint32_t flip) {
  if (flip & 1u) w = ((w & 0x55555555u) << 1) | ((w >> 1) & 0x55555555u);
  if (flip & 2u) w = ((w & 0x33333333u) << 2) | ((w >> 2) & 0x33333333u);
  if (flip & 4u) w = ((w & 0x0F0F0F0Fu) << 4) | ((w >> 4) & 0x0F0F0F0Fu);
  return w;
}
__device__ __forceinline__ uint32_t perm8(uint32_t m, uint32_t flip) { return perm_bytes(m & 0xFFu, flip); }

// Invariance bits of a node's children, one 8-bit mask per axis (bit c =
// child c in the node's own order), loaded once per node; zero without a
// child level (childInv == nullptr).
struct ChildInv {
  uint32_t x, y, z;
};
__device__ __forceinline__ ChildInv child_inv(const uint32_t k0[kRecWords], const uint8_t* __restrict__ childInv) {
  ChildInv r{0u, 0u, 0u};
  if (!childInv) return r;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (k0[1 + c] == kNull) continue;
    const uint32_t ci = childInv[k0[1 + c]];
    r.x |= (ci & 1u) << c;
    r.y |= ((ci >> 1) & 1u) << c;
    r.z |= ((ci >> 2) & 1u) << c;
  }
  return r;
}

// Node::mirror with applyToChildren (octree_node.cpp:88-195) by transform t,
// then GeomOctree::invertInvs (geom_octree.cpp:822-831): a mirrored child
// that is invariant under the axis keeps its mirror bit clear.  Child c of
// the variant is child c ^ flip of the node, so the invariance masks are
// permuted like the children.  The children move by conditional swaps (one
// per flip bit) rather than by an index computed at run time, and the four
// mask bytes of word 0 are permuted together.
__device__ __forceinline__ void make_variant(const uint32_t k0[kRecWords], int t, const ChildInv& inv,
                                             uint32_t v[kRecWords]) {
  const uint32_t flip = flip_of(t);
  const uint32_t tx = (flip >> 2) & 1u, ty = (flip >> 1) & 1u, tz = flip & 1u;
  const uint32_t pw = perm_bytes(k0[0], flip);
  const uint32_t mask = pw & 0xFFu;
  uint32_t mx = (pw >> 8) & 0xFFu, my = (pw >> 16) & 0xFFu, mz = pw >> 24;
  uint32_t ch[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) ch[c] = k0[1 + c];
#pragma unroll
  for (int b = 1; b <= 4; b <<= 1) {
    const bool sw = (flip & (uint32_t)b) != 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (c & b) continue;
      const uint32_t lo = ch[c], hi = ch[c | b];
      ch[c] = sw ? hi : lo;
      ch[c | b] = sw ? lo : hi;
    }
  }
  uint32_t ex = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    v[1 + c] = ch[c];
    ex |= (ch[c] != kNull ? 1u : 0u) << c;
  }
  if (tx) mx ^= ex;
  if (ty) my ^= ex;
  if (tz) mz ^= ex;
  const uint32_t pinv = perm_bytes(inv.x | (inv.y << 8) | (inv.z << 16), flip);
  if (tx) mx &= ~(pinv & 0xFFu);
  if (ty) my &= ~((pinv >> 8) & 0xFFu);
  if (tz) mz &= ~((pinv >> 16) & 0xFFu);
  v[0] = mask | (mx << 8) | (my << 16) | (mz << 24);
}

__device__ __forceinline__ bool same_rec(const uint32_t a[kRecWords], const uint32_t b[kRecWords]) {
#pragma unroll
  for (int k = 0; k < kRecWords; ++k)
    if (a[k] != b[k]) return false;
  return true;
}

// Identity key K0 of each DAG node of a level: children moved to the SDAG
// ids of the level below, mirror bits set from their correspondence
// (geom_octree.cpp:666-679); plus raw invariance bits (:603-606).
__device__ __forceinline__ void sdag_key(const uint32_t* __restrict__ dag, uint64_t i,
                                         const uint32_t* __restrict__ childMap, uint32_t* __restrict__ key,
                                         uint8_t* __restrict__ inv) {
  uint32_t w[kRecWords];
  load_rec(dag, i, w);
  uint32_t mx = (w[0] >> 8) & 0xFFu, my = (w[0] >> 16) & 0xFFu, mz = (w[0] >> 24) & 0xFFu;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (w[1 + c] == kNull) continue;
    const uint32_t m = childMap[w[1 + c]];
    w[1 + c] = m & 0x1FFFFFFFu;
    mx |= ((m >> 29) & 1u) << c;
    my |= ((m >> 30) & 1u) << c;
    mz |= ((m >> 31) & 1u) << c;
  }
  w[0] = (w[0] & 0xFFu) | (mx << 8) | (my << 16) | (mz << 24);
  uint32_t* o = key + i * kRecWords;
#pragma unroll
  for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
  if (inv) {
    uint32_t b = 0, v[kRecWords];
    const ChildInv none{0u, 0u, 0u};
    make_variant(w, 1, none, v); b |= same_rec(v, w) ? 1u : 0u;
    make_variant(w, 2, none, v); b |= same_rec(v, w) ? 2u : 0u;
    make_variant(w, 3, none, v); b |= same_rec(v, w) ? 4u : 0u;
    inv[i] = (uint8_t)b;
  }
}

// Also clears the level's hash table (nt entries), which would otherwise take
// a fill launch of its own.
__global__ void k_sdag_keys(const uint32_t* __restrict__ dag, uint64_t n, const uint32_t* __restrict__ childMap,
                            uint32_t* __restrict__ key, uint8_t* __restrict__ inv, TEntry* __restrict__ table,
                            uint32_t nt) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = i; j < nt; j += stride) table[j] = kEmpty;
  if (i < n) sdag_key(dag, i, childMap, key, inv);
}

enum : uint32_t { kUndecided = 0, kRep = 1, kNonRep = 2 };

// First-occurrence nodes look up their 7 mirrored variants; an earlier node
// carrying that key is a candidate that, if it is a representative, takes
// this node.  Later duplicates of a key are never representatives.  Eight
// lanes per node: lane t of the group looks up variant t, lane 0 sets the
// state from the group's ballot, so a node's chain is one lookup long.
__global__ void k_sdag_cands(const uint32_t* __restrict__ key, uint64_t n, const TEntry* __restrict__ table,
                             uint32_t cmask, const uint32_t* __restrict__ slot, const uint8_t* __restrict__ childInv,
                             uint32_t* __restrict__ cand, uint32_t* __restrict__ state) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i = g >> 3;
  const int t = (int)(g & 7u);
  const bool live = i < n;
  // a node without children was never inserted (slot kNone): it is no
  // representative and k_sdag_outcome reports it, as sdag_small_level does
  bool first = false;
  if (live) {
    const uint32_t si = slot[i];
    first = si != kNone && entry_index(table[si]) == (uint32_t)i;
  }
  bool found = false;
  uint32_t r = kNone;
  if (first && t > 0) {
    uint32_t k0[kRecWords], v[kRecWords];
    load_rec(key, i, k0);
    const ChildInv ci = child_inv(k0, childInv);
    make_variant(k0, t, ci, v);
    // a variant equal to the node's own key finds the node itself
    bool self = true;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) self = self && v[k] == k0[k];
    r = self ? (uint32_t)i : lookup(key, table, cmask, v);
    found = r != kNone && r < (uint32_t)i;
  }
  const uint64_t b = __ballot(found);
  if (!live) return;
  uint32_t* cd = cand + i * 7;
  if (t > 0) cd[t - 1] = found ? r : kNone;
  else state[i] = !first ? kNonRep : (((b >> (threadIdx.x & 63)) & 0xFEull) ? kUndecided : kRep);
}

// One resolution round.  state only moves Undecided -> final, so a stale
// read only delays a decision.
__global__ void k_sdag_resolve(const uint32_t* __restrict__ cand, uint64_t n, uint32_t* state,
                               uint32_t* __restrict__ remaining) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (__atomic_load_n(&state[i], __ATOMIC_RELAXED) != kUndecided) return;
  const uint32_t* cd = cand + i * 7;
  bool allNon = true, anyRep = false;
  for (int t = 0; t < 7; ++t) {
    const uint32_t c = cd[t];
    if (c == kNone) continue;
    const uint32_t s = __atomic_load_n(&state[c], __ATOMIC_RELAXED);
    if (s == kRep) { anyRep = true; break; }
    if (s != kNonRep) allNon = false;
  }
  if (anyRep) __atomic_store_n(&state[i], kNonRep, __ATOMIC_RELAXED);
  else if (allNon) __atomic_store_n(&state[i], kRep, __ATOMIC_RELAXED);
  else atomicAdd(remaining, 1u);
}

// pos[i] = number of representatives before node i (scan_op).
struct StateFlagOp {
  const uint32_t* __restrict__ state;
  uint32_t* __restrict__ pos;
  __device__ uint32_t value(uint64_t i, uint32_t&) const { return state[i] == kRep ? 1u : 0u; }
  __device__ void apply(uint64_t i, uint32_t ex) const { pos[i] = ex; }
  __device__ void report(uint32_t) const {}
};

// Correspondence of every DAG node: itself if a representative, else the
// first transform (in reference order) whose key is an earlier
// representative's identity key.  Representatives are compacted in order.
__global__ void k_sdag_outcome(const uint32_t* __restrict__ key, uint64_t n, const TEntry* __restrict__ table,
                               uint32_t cmask, const uint32_t* __restrict__ slot, const uint32_t* __restrict__ state,
                               const uint32_t* __restrict__ pos, const uint32_t* __restrict__ cand,
                               const uint8_t* __restrict__ childInv, const uint8_t* __restrict__ invIn,
                               uint32_t* __restrict__ map, uint32_t* __restrict__ out, uint8_t* __restrict__ invOut,
                               uint32_t* __restrict__ err) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t st = state[i];
  const uint32_t si = slot[i];
  if (st != kRep && si != kNone && entry_index(table[si]) == (uint32_t)i) {
    // the first node of its key: k_sdag_cands already looked up transforms
    // 1..7 and kept the earlier nodes found; the identity finds itself
    const uint32_t* cd = cand + i * 7;
    for (int t = 1; t < 8; ++t) {
      const uint32_t r = cd[t - 1];
      if (r != kNone && state[r] == kRep) {
        const uint32_t f = flip_of(t);
        map[i] = pos[r] | (((f >> 2) & 1u) << 29) | (((f >> 1) & 1u) << 30) | ((f & 1u) << 31);
        return;
      }
    }
    map[i] = kNone;
    atomicAdd(err, 1u);
    return;
  }
  uint32_t k0[kRecWords];
  load_rec(key, i, k0);
  if (st == kRep) {
    const uint32_t p = pos[i];
    map[i] = p;
    uint32_t* o = out + (uint64_t)p * kRecWords;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) o[k] = k0[k];
    if (invOut) invOut[p] = invIn ? invIn[i] : 0;
    return;
  }
  uint32_t v[kRecWords];
  const ChildInv ci = child_inv(k0, childInv);
  for (int t = 0; t < 8; ++t) {
    uint32_t r;
    if (t == 0) r = si == kNone ? kNone : entry_index(table[si]);
    else { make_variant(k0, t, ci, v); r = lookup(key, table, cmask, v); }
    if (r != kNone && r < (uint32_t)i && state[r] == kRep) {
      const uint32_t f = flip_of(t);
      map[i] = pos[r] | (((f >> 2) & 1u) << 29) | (((f >> 1) & 1u) << 30) | ((f & 1u) << 31);
      return;
    }
  }
  map[i] = kNone;
  atomicAdd(err, 1u);
}

// A small level of toSDAG (n <= kSdagSmall) in one workgroup: keys,
// first-occurrence insert (table and node states in LDS), candidates, the
// resolution rounds until every node is decided, representative flags and
// their scan, and the outcome, as k_sdag_keys ... k_sdag_outcome do over
// several launches.  Every thread reads the round's undecided count after
// a barrier, so all leave the loop together; each round decides at least
// the earliest undecided node, so it ends within n rounds.
// The level's keys, candidates and invariance bits live in LDS beside the
// table (91 KB): every probe's record compare and every resolution step is
// then an LDS access instead of a global memory round trip.
constexpr uint32_t kSdagCap = 2 * kSdagSmall;   // table slots, power of two
struct SdagLds {
  TEntry table[kSdagCap];
  uint32_t key[kSdagSmall * kRecWords];
  uint32_t cand[kSdagSmall * 7];
  uint16_t slotOf[kSdagSmall];
  uint32_t state[kSdagSmall];
  uint32_t posOf[kSdagSmall];
  uint8_t inv[kSdagSmall];
  uint32_t lds[kSmallThreads / 64 + 1];
  uint32_t remaining;
};
__device__ void sdag_small_level(const uint32_t* __restrict__ dag, uint32_t n, const uint32_t* __restrict__ childMap,
                                 const uint8_t* __restrict__ childInv, uint32_t* __restrict__ map,
                                 uint32_t* __restrict__ out, uint8_t* __restrict__ invOut,
                                 uint32_t* __restrict__ total, uint32_t* __restrict__ err, SdagLds& sh) {
  TEntry* table = sh.table;
  uint32_t* key = sh.key;
  uint32_t* cand = sh.cand;
  uint8_t* inv = sh.inv;
  uint16_t* slotOf = sh.slotOf;
  uint32_t* state = sh.state;
  uint32_t* posOf = sh.posOf;
  uint32_t* lds = sh.lds;
  uint32_t& remaining = sh.remaining;
  const uint32_t tid = threadIdx.x;
  for (uint32_t j = tid; j < kSdagCap; j += kSmallThreads) table[j] = kEmpty;
  for (uint32_t i = tid; i < n; i += kSmallThreads) sdag_key(dag, i, childMap, key, inv);
  __syncthreads();
  // insert (as k_insert)
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    uint32_t w[kRecWords];
    load_rec(key, i, w);
    if ((w[0] & 0xFFu) == 0) { slotOf[i] = 0xFFFFu; continue; }
    const uint64_t hv = hash_rec(w);
    const uint32_t fp = (uint32_t)(hv >> 32);
    const TEntry mine = ((TEntry)fp << 32) | (TEntry)i;
    uint32_t h = (uint32_t)hv & (kSdagCap - 1);
    while (true) {
      TEntry e = table[h];
      if (e == kEmpty) {
        const TEntry old = atomicCAS(&table[h], kEmpty, mine);
        if (old == kEmpty) break;
        e = old;
      }
      if ((uint32_t)(e >> 32) == fp && rec_eq(key, (uint32_t)e, w)) {
        if (mine < e) atomicMin(&table[h], mine);
        break;
      }
      h = (h + 1) & (kSdagCap - 1);
    }
    slotOf[i] = (uint16_t)h;
  }
  __syncthreads();
  auto find = [&](const uint32_t w[kRecWords]) -> uint32_t {
    const uint64_t hv = hash_rec(w);
    const uint32_t fp = (uint32_t)(hv >> 32);
    uint32_t h = (uint32_t)hv & (kSdagCap - 1);
    while (true) {
      const TEntry e = table[h];
      if (e == kEmpty) return kNone;
      if ((uint32_t)(e >> 32) == fp && rec_eq(key, (uint32_t)e, w)) return (uint32_t)e;
      h = (h + 1) & (kSdagCap - 1);
    }
  };
  // candidates (as k_sdag_cands): P lanes per node (8 when the level is
  // small enough), lane q trying the variants t = 1 + q, 1 + q + P, ...; a
  // variant is a hash and a probe chain, so spreading them shortens the
  // longest chain of a thread
  const uint32_t P = n * 8 <= kSmallThreads ? 8u : n * 4 <= kSmallThreads ? 4u : n * 2 <= kSmallThreads ? 2u : 1u;
  for (uint32_t x = tid; x < n * P; x += kSmallThreads) {
    const uint32_t i = x / P, q = x % P;
    uint32_t* cd = cand + (uint64_t)i * 7;
    const bool first = slotOf[i] != 0xFFFFu && entry_index(table[slotOf[i]]) == i;
    if (!first) {
      for (uint32_t t = 1 + q; t < 8; t += P) cd[t - 1] = kNone;
      continue;
    }
    uint32_t k0[kRecWords], v[kRecWords];
    load_rec(key, i, k0);
    const ChildInv ci = child_inv(k0, childInv);
    for (uint32_t t = 1 + q; t < 8; t += P) {
      make_variant(k0, (int)t, ci, v);
      const uint32_t r = same_rec(v, k0) ? i : find(v);
      cd[t - 1] = (r != kNone && r < i) ? r : kNone;
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    const bool first = slotOf[i] != 0xFFFFu && entry_index(table[slotOf[i]]) == i;
    const uint32_t* cd = cand + (uint64_t)i * 7;
    bool any = false;
#pragma unroll
    for (int t = 0; t < 7; ++t) any |= cd[t] != kNone;
    state[i] = !first ? kNonRep : (any ? kUndecided : kRep);
  }
  __syncthreads();
  // resolution rounds (as k_sdag_resolve), to convergence
  for (uint32_t round = 0; round <= n; ++round) {
    if (tid == 0) remaining = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kSmallThreads) {
      if (__hip_atomic_load(&state[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != kUndecided) continue;
      const uint32_t* cd = cand + (uint64_t)i * 7;
      bool allNon = true, anyRep = false;
      for (int t = 0; t < 7; ++t) {
        const uint32_t cc = cd[t];
        if (cc == kNone) continue;
        const uint32_t st = __hip_atomic_load(&state[cc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (st == kRep) { anyRep = true; break; }
        if (st != kNonRep) allNon = false;
      }
      if (anyRep) __hip_atomic_store(&state[i], kNonRep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else if (allNon) __hip_atomic_store(&state[i], kRep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else atomicAdd(&remaining, 1u);
    }
    __syncthreads();
    const uint32_t left = remaining;
    __syncthreads();
    if (left == 0) break;
  }
  // flags + scan (as StateFlagOp)
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kSmallThreads) {
    const uint32_t i = base + tid;
    const uint32_t f = (i < n && state[i] == kRep) ? 1u : 0u;
    uint32_t agg;
    const uint32_t ex = block_exclusive_sum(f, lds, agg) + carry;
    if (i < n) posOf[i] = ex;
    carry += agg;
  }
  __syncthreads();
  // outcome (as k_sdag_outcome)
  for (uint32_t i = tid; i < n; i += kSmallThreads) {
    const uint32_t st = state[i];
    const bool first = slotOf[i] != 0xFFFFu && entry_index(table[slotOf[i]]) == i;
    if (st != kRep && first) {
      const uint32_t* cd = cand + (uint64_t)i * 7;
      bool found = false;
      for (int t = 1; t < 8 && !found; ++t) {
        const uint32_t r = cd[t - 1];
        if (r != kNone && state[r] == kRep) {
          const uint32_t f = flip_of(t);
          map[i] = posOf[r] | (((f >> 2) & 1u) << 29) | (((f >> 1) & 1u) << 30) | ((f & 1u) << 31);
          found = true;
        }
      }
      if (!found) { map[i] = kNone; atomicAdd(err, 1u); }
      continue;
    }
    uint32_t k0[kRecWords];
    load_rec(key, i, k0);
    if (st == kRep) {
      const uint32_t p = posOf[i];
      map[i] = p;
      uint32_t* o = out + (uint64_t)p * kRecWords;
#pragma unroll
      for (int k = 0; k < kRecWords; ++k) o[k] = k0[k];
      invOut[p] = inv[i];
      continue;
    }
    uint32_t v[kRecWords];
    const ChildInv ci = child_inv(k0, childInv);
    bool found = false;
    for (int t = 0; t < 8 && !found; ++t) {
      uint32_t r;
      if (t == 0) r = slotOf[i] != 0xFFFFu ? entry_index(table[slotOf[i]]) : kNone;
      else { make_variant(k0, t, ci, v); r = find(v); }
      if (r != kNone && r < i && state[r] == kRep) {
        const uint32_t f = flip_of(t);
        map[i] = posOf[r] | (((f >> 2) & 1u) << 29) | (((f >> 1) & 1u) << 30) | ((f & 1u) << 31);
        found = true;
      }
    }
    if (!found) { map[i] = kNone; atomicAdd(err, 1u); }
  }
  if (tid == 0) *total = carry;
}

// toSDAG of every level from `top` up to the root in one workgroup, each of
// at most kSdagSmall nodes: per level sdag_small_level, whose
// correspondence and invariance bits the next level reads after a barrier;
// the root only gets its identity key (children moved to the SDAG ids below,
// mirror bits from their correspondence) and clear invariance bits: it is
// never reduced.
struct SdagTop {
  const uint32_t* dag[kMaxLevels];
  uint32_t n[kMaxLevels];
  uint32_t* map[kMaxLevels + 1];      // map[lev + 1]: correspondence of the level below lev
  uint8_t* invOut[kMaxLevels + 1];    // SDAG invariance bits per level (inv of level lev + 1 = childInv)
  uint32_t* out[kMaxLevels];
  uint32_t* counts;
  uint32_t* err;
  int32_t top;
};
__global__ void __launch_bounds__(kSmallThreads) k_sdag_levels_small(SdagTop a) {
  __shared__ SdagLds sh;
  for (int lev = a.top; lev >= 1; --lev) {
    sdag_small_level(a.dag[lev], a.n[lev], a.map[lev + 1], a.invOut[lev + 1], a.map[lev], a.out[lev], a.invOut[lev],
                     a.counts + lev, a.err, sh);
    __threadfence_block();
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < a.n[0]; i += kSmallThreads) {
    sdag_key(a.dag[0], i, a.map[1], a.out[0], nullptr);
    a.invOut[0][i] = 0;
  }
}

// Finest level.  The reference loop walks the nodes in order and makes a node
// a representative unless one of its eight mirrored masks (kFlip order) is an
// earlier representative's.  Mirroring a leaf mask only permutes its bits by
// c -> c ^ f, and these permutations form a group, so the mirrored masks of a
// node are its orbit: a node is a representative exactly when it is the first
// node of its orbit, and any other node maps to that first node through the
// first transform, in kFlip order, that turns its mask into the
// representative's.  Representatives are numbered in node order.  One
// workgroup of 256 threads, any n.
__global__ void __launch_bounds__(256) k_sdag_leaves(const uint32_t* __restrict__ dagMask, uint32_t n,
                                                     uint32_t* __restrict__ map, uint32_t* __restrict__ sdagMask,
                                                     uint8_t* __restrict__ inv, uint32_t* __restrict__ count) {
  __shared__ uint32_t firstOf[256], rankOf[256];
  const uint32_t t = threadIdx.x;
  firstOf[t] = kNone;
  __syncthreads();
  auto canon = [](uint32_t m) {
    uint32_t c = m;
#pragma unroll
    for (uint32_t f = 1; f < 8; ++f) c = min(c, perm8(m, f));
    return c;
  };
  for (uint32_t i = t; i < n; i += 256) atomicMin(&firstOf[canon(dagMask[i] & 0xFFu)], i);
  __syncthreads();
  // rank of orbit t's first node among all orbits' first nodes
  uint32_t r = 0;
  const uint32_t me = firstOf[t];
  for (int k = 0; k < 256; ++k) r += firstOf[k] < me ? 1u : 0u;
  rankOf[t] = r;
  __syncthreads();
  if (t == 0) {
    uint32_t cnt = 0;
    for (int k = 0; k < 256; ++k) cnt += firstOf[k] != kNone;
    *count = cnt;
  }
  for (uint32_t i = t; i < n; i += 256) {
    const uint32_t m = dagMask[i] & 0xFFu;
    const uint32_t c = canon(m);
    const uint32_t first = firstOf[c], rank = rankOf[c];
    if (first == i) {
      uint32_t b = 0;
      if (perm8(m, 4) == m) b |= 1;
      if (perm8(m, 2) == m) b |= 2;
      if (perm8(m, 1) == m) b |= 4;
      sdagMask[rank] = m;
      inv[rank] = (uint8_t)b;
      map[i] = rank;
      continue;
    }
    const uint32_t mr = dagMask[first] & 0xFFu;
    for (int k = 0; k < 8; ++k) {
      const uint32_t f = flip_of(k);
      if (perm8(m, f) == mr) {
        map[i] = rank | (((f >> 2) & 1u) << 29) | (((f >> 1) & 1u) << 30) | ((f & 1u) << 31);
        break;
      }
    }
  }
}

// The scene s of code t * kLeafTile (the last s with off[s] <= it; every
// scene holds at least one code, so off ascends strictly), as 8 words per
// tile: off[s], off[s + 1], off[s + 2] (64-bit, low word first; off[S] past
// the last scene), s, 3 * levels[s] | 3 * levels[s + 1] << 8.
__global__ void k_tile_scenes(const uint64_t* __restrict__ off, const uint32_t* __restrict__ levels, uint32_t S,
                              uint64_t tiles, uint32_t* __restrict__ tileScene) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  const uint64_t i = t * kLeafTile;
  uint32_t lo = 0, hi = S;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  const uint64_t b = off[lo], e = off[lo + 1], e1 = lo + 1 < S ? off[lo + 2] : e;
  const uint32_t sh1 = lo + 1 < S ? 3 * levels[lo + 1] : 0u;
  uint4* r = (uint4*)(tileScene + 8 * t);
  r[0] = make_uint4((uint32_t)b, (uint32_t)(b >> 32), (uint32_t)e, (uint32_t)(e >> 32));
  r[1] = make_uint4((uint32_t)e1, (uint32_t)(e1 >> 32), lo, 3 * levels[lo] | sh1 << 8);
}

// Scene id of every node of a level of scene-tagged codes (codes >> shift).
__global__ void k_code_scenes(const uint64_t* __restrict__ codes, uint64_t n, uint32_t shift,
                              uint32_t* __restrict__ sid) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sid[i] = (uint32_t)(codes[i] >> shift);
}

uint32_t table_cap(uint64_t n) {
  uint64_t c = 1024;
  while (c < 2 * n) c <<= 1;
  return (uint32_t)c;
}

}  // namespace

// ============================================================================
void build_svo_from_morton(svd_octree* o, const uint64_t* dcodes, uint64_t n) {
  svd_ctx* c = o->ctx;
  const uint32_t L = o->levels;
  o->svo.assign(L, Level());
  o->nVoxels = n;
  if (n == 0) {
    // Empty scene: a single empty root (buildSVO on no geometry).
    o->svo[0].rec = dalloc<uint32_t>(c, kRecWords);
    uint32_t h[kRecWords] = {0, kNull, kNull, kNull, kNull, kNull, kNull, kNull, kNull};
    SVD_HIP(hipMemcpyAsync(o->svo[0].rec, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    o->svo[0].n = 1;
    for (uint32_t l = 1; l < L; ++l) o->svo[l].leaf = (l == L - 1);
    if (L == 1) o->svo[0].leaf = true;
    o->nNodesSVO = 0;
    o->state = SVD_S_SVO;
    SVD_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  o->nNodesSVO = svo_levels_from_codes(c, dcodes, n, L, o->svo, nullptr, ~((1ull << (3 * L)) - 1));
  o->state = SVD_S_SVO;
}

// Inner levels L-2 .. 0 from the finest level's parent codes (codeA, nc of
// them, taken over and freed).  Returns the non-root nodes created.
uint64_t svo_inner_levels(svd_ctx* c, uint64_t* codeA, uint64_t nc, uint32_t L, std::vector<Level>& svo,
                          std::vector<uint32_t*>* levelScenes, const uint64_t* sizes) {
  uint64_t* codeB = dalloc<uint64_t>(c, nc ? nc : 1);
  auto keep = [&](uint32_t lev, const uint64_t* codes, uint64_t m) {
    if (!levelScenes) return;
    (*levelScenes)[lev] = dalloc<uint32_t>(c, m + 1);
    if (m)
      k_code_scenes<<<div_up(m, kB), kB, 0, c->stream>>>(codes, m, kSceneTagShift - 3 * (L - lev),
                                                         (*levelScenes)[lev]);
  };
  uint64_t nNodes = 0;
  for (int lev = (int)L - 2; lev >= 0; --lev) {
    Level& lv = svo[lev];
    if (nc <= small_limit(c->smallSvo, kSvoSmall) && !levelScenes) {
      // this level and all above it in one launch
      SvoTop top{};
      top.top = lev;
      for (int l = lev; l >= 0; --l) {
        svo[l].n = sizes[l];
        svo[l].rec = dalloc<uint32_t>(c, (sizes[l] ? sizes[l] : 1) * kRecWords);
        top.rec[l] = svo[l].rec;
        if (l > 0) nNodes += sizes[l];
      }
      k_inner_levels_small<<<1, 1024, 0, c->stream>>>(codeA, codeB, (uint32_t)nc, top);
      check_launch("k_inner_levels_small");
      break;
    }
    const uint64_t np = sizes[lev];
    lv.n = np;
    lv.rec = dalloc<uint32_t>(c, np * kRecWords);
    if (nc <= small_limit(c->smallSvo, kSvoSmall)) {
      k_inner_level_small<<<1, 1024, 0, c->stream>>>(codeA, (uint32_t)nc, lv.rec, codeB);
      check_launch("k_inner_level_small");
    } else {
      const uint64_t tiles = div_up(nc, kInnerTile);
      k_inner_build<<<(uint32_t)tiles, kRunT, 0, c->stream>>>(codeA, nc, lb_pass(c, tiles), lv.rec, codeB);
      check_launch("k_inner_build");
    }
    keep((uint32_t)lev, codeB, lv.n);
    std::swap(codeA, codeB);
    nc = lv.n;
    if (lev > 0) nNodes += lv.n;
  }
  dfree(c, codeA); dfree(c, codeB);
  return nNodes;
}

// SVO levels from n > 0 strictly ascending codes with no bit of hiMask set,
// in one pass over the codes for the finest level (k_leaf_build: the input
// check, every level's node count, the leaf masks and parent codes) and one
// per inner level.  Node order per level = order of first arrival.
// tags (optional): the codes are a batch's scenes one after another, each
// ascending below 2^(3 * its levels); they are tagged with their scene id as
// they are read (hiMask is then ignored).
// levelScenes (optional, with tags) receives, per level, every node's scene
// code (code >> 3*(L-lev)).  Returns the number of non-root nodes; throws
// SVD_ERR_ARG on bad input.
uint64_t svo_levels_from_codes(svd_ctx* c, const uint64_t* dcodes, uint64_t n, uint32_t L, std::vector<Level>& svo,
                               std::vector<uint32_t*>* levelScenes, uint64_t hiMask, const SceneTags* tags) {
  svo.assign(L, Level());
  if (levelScenes) levelScenes->assign(L, nullptr);
  const uint64_t tiles = div_up(n, kLeafTile);
  uint32_t* part = dalloc<uint32_t>(c, tiles * (L + 1));
  uint32_t* hist = dalloc<uint32_t>(c, kMaxLevels + 1);
  uint32_t* leafMask = dalloc<uint32_t>(c, n);
  uint64_t* codeA = dalloc<uint64_t>(c, n);
  uint32_t* tileScene = nullptr;
  if (tags) {
    tileScene = dalloc<uint32_t>(c, 8 * tiles);
    k_tile_scenes<<<div_up(tiles, kB), kB, 0, c->stream>>>(tags->off, tags->levels, tags->S, tiles, tileScene);
    SceneTags t = *tags;
    t.tileScene = tileScene;
    k_leaf_build<<<(uint32_t)tiles, kRunT, 0, c->stream>>>(TaggedCodes{dcodes, t, kLeafTile}, n, L, 1ull << 63,
                                                           lb_pass(c, tiles), part, hist, leafMask, codeA);
  } else {
    k_leaf_build<<<(uint32_t)tiles, kRunT, 0, c->stream>>>(PlainCodes{dcodes}, n, L, hiMask, lb_pass(c, tiles), part,
                                                           hist, leafMask, codeA);
  }
  k_level_counts<<<(uint32_t)div_up(tiles, kCountTilesPerWg), 1024, 0, c->stream>>>(part, (uint32_t)tiles, L, hist);
  check_launch("k_leaf_build");
  uint32_t h[kMaxLevels + 1];
  read_u32s(c, hist, h, L + 1);
  dfree(c, hist);
  dfree(c, part);
  dfree(c, tileScene);
  if (h[0]) {
    dfree(c, leafMask);
    dfree(c, codeA);
    throw Error{SVD_ERR_ARG, tags ? "each scene's Morton codes must be strictly ascending and below 2^(3*levels)"
                                  : "Morton codes must be strictly ascending and below 2^(3*levels)"};
  }
  // level lev is L - lev steps above the voxels
  std::vector<uint64_t> sizes(L);
  for (uint32_t lev = 0; lev < L; ++lev) sizes[lev] = h[L - lev];
  Level& leaf = svo[L - 1];
  leaf.leaf = true;
  leaf.n = sizes[L - 1];
  leaf.rec = leafMask;
  if (levelScenes) {
    (*levelScenes)[L - 1] = dalloc<uint32_t>(c, leaf.n + 1);
    k_code_scenes<<<div_up(leaf.n, kB), kB, 0, c->stream>>>(codeA, leaf.n, kSceneTagShift - 3, (*levelScenes)[L - 1]);
  }
  const uint64_t nNodes = (L >= 2) ? leaf.n : 0;
  return nNodes + svo_inner_levels(c, codeA, leaf.n, L, svo, levelScenes, sizes.data());
}

void remap_children(svd_ctx* c, const uint32_t* in, uint64_t n, const uint32_t* remap, uint32_t* out) {
  if (n) k_remap_children<<<div_up(n, kB), kB, 0, c->stream>>>(in, n, remap, out);
  check_launch("k_remap_children");
}

// Generic dedup of one inner level's records -> remap + compacted records.
// The unique count goes to the device word *total; the output holds n
// records (the upper bound), so no host read is needed.
// cleared: a table of table_cap(n) entries already set to empty (null: one is
// allocated and cleared here); it is freed here either way.
void dedup_level_dev(svd_ctx* c, const uint32_t* rec, uint64_t n, uint32_t* remap, uint32_t** outRec,
                     uint32_t* total, TEntry* cleared = nullptr) {
  const uint32_t cap = table_cap(n);
  TEntry* table = cleared ? cleared : dalloc<TEntry>(c, cap);
  uint32_t* slot = dalloc<uint32_t>(c, n);
  uint32_t* pos = dalloc<uint32_t>(c, n + 1);
  if (!cleared) fill_u32(c, (uint32_t*)table, kNone, 2ull * cap);
  launch_insert(c, rec, n, table, cap - 1, slot);
  check_launch("k_insert");
  scan_op<4>(c, n, RepFlagOp{slot, table, pos}, total);
  uint32_t* out = dalloc<uint32_t>(c, (n ? n : 1) * kRecWords);
  k_compact<<<div_up(n, kB), kB, 0, c->stream>>>(rec, slot, table, pos, n, out, remap);
  check_launch("k_compact");
  dfree(c, table); dfree(c, slot); dfree(c, pos);
  *outRec = out;
}

uint64_t dedup_level(svd_ctx* c, const uint32_t* rec, uint64_t n, uint32_t* remap, uint32_t** outRec) {
  uint32_t* total = dalloc<uint32_t>(c, 1);
  dedup_level_dev(c, rec, n, remap, outRec, total);
  const uint64_t d = read_u32(c, total);
  dfree(c, total);
  return d;
}

// toDAG over any set of levels (the SVO, or a lossy DAG before its final
// dedup, geom_octree_extension.cpp:1176): dst gets the reduced levels.
uint64_t dedup_all(svd_ctx* c, uint32_t L, const std::vector<Level>& src, std::vector<Level>& dst) {
  dst.assign(L, Level());
  if (L == 1) {
    dst[0] = src[0];
    dst[0].rec = dalloc<uint32_t>(c, 1);
    SVD_HIP(hipMemcpyAsync(dst[0].rec, src[0].rec, 4, hipMemcpyDeviceToDevice, c->stream));
    return 1;
  }
  uint64_t nNodes = 1;
  // unique counts of every reduced level, read back once at the end
  uint32_t* counts = dalloc<uint32_t>(c, L);
  // finest level: dedup of 8-bit masks by first occurrence
  const Level& sl = src[L - 1];
  uint32_t* remap = dalloc<uint32_t>(c, sl.n + 1);
  {
    uint32_t* firstocc = dalloc<uint32_t>(c, 256);
    uint32_t* idOfMask = dalloc<uint32_t>(c, 256);
    Level& dl = dst[L - 1];
    dl.leaf = true;
    dl.rec = dalloc<uint32_t>(c, 256);
    fill_u32(c, firstocc, kNone, 256);
    if (sl.n) {
      const uint64_t nb = div_up(sl.n, 256) < (uint64_t)kFoBlocks ? div_up(sl.n, 256) : (uint64_t)kFoBlocks;
      k_leaf_firstocc<<<nb, 256, 0, c->stream>>>(sl.rec, sl.n, firstocc);
    }
    k_leaf_rank<<<1, 256, 0, c->stream>>>(firstocc, idOfMask, dl.rec, counts + (L - 1));
    if (sl.n) k_leaf_remap<<<div_up(sl.n, kB), kB, 0, c->stream>>>(sl.rec, sl.n, idOfMask, remap);
    check_launch("leaf dedup");
    dfree(c, firstocc); dfree(c, idOfMask);
  }
  for (int lev = (int)L - 2; lev >= 0; --lev) {
    const Level& s = src[lev];
    // this level and all above it at most kSmallLevel nodes: one launch
    uint64_t maxTop = 0;
    for (int l = lev; l >= 1; --l) maxTop = std::max<uint64_t>(maxTop, src[l].n);
    if (maxTop <= small_limit(c->smallDag, kSmallLevel)) {
      DagTop top{};
      top.top = lev;
      top.remapIn = remap;
      for (int l = lev; l >= 0; --l) {
        top.src[l] = src[l].rec;
        top.n[l] = (uint32_t)src[l].n;
        dst[l].rec = dalloc<uint32_t>(c, (src[l].n ? src[l].n : 1) * kRecWords);
        top.out[l] = dst[l].rec;
      }
      dst[0].n = src[0].n;
      top.remap[0] = dalloc<uint32_t>(c, maxTop + 1);
      top.remap[1] = dalloc<uint32_t>(c, maxTop + 1);
      top.keyed = dalloc<uint32_t>(c, (maxTop ? maxTop : 1) * kRecWords);
      top.counts = counts;
      k_dedup_levels_small<<<1, kSmallThreads, 0, c->stream>>>(top);
      check_launch("k_dedup_levels_small");
      dfree(c, top.remap[0]); dfree(c, top.remap[1]); dfree(c, top.keyed);
      break;
    }
    uint32_t* keyed = dalloc<uint32_t>(c, (s.n ? s.n : 1) * kRecWords);
    // the level's dedup table is cleared by the same launch
    const uint32_t cap = table_cap(s.n);
    TEntry* table = dalloc<TEntry>(c, cap);
    if (s.n) k_remap_children<<<div_up(s.n, kB), kB, 0, c->stream>>>(s.rec, s.n, remap, keyed, table, cap);
    else fill_u32(c, (uint32_t*)table, kNone, 2ull * cap);
    check_launch("k_remap_children");
    Level& d = dst[lev];
    uint32_t* nremap = dalloc<uint32_t>(c, s.n + 1);
    dedup_level_dev(c, keyed, s.n, nremap, &d.rec, counts + lev, table);
    dfree(c, keyed);
    dfree(c, remap);
    remap = nremap;
  }
  dfree(c, remap);
  std::vector<uint32_t> h(L);
  read_u32s(c, counts, h.data(), L);
  dfree(c, counts);
  for (uint32_t lev = 1; lev < L; ++lev) {
    dst[lev].n = h[lev];
    nNodes += h[lev];
  }
  return nNodes;
}

void reduce_dag(svd_octree* o) {
  svd_ctx* c = o->ctx;
  if (!(o->state == SVD_S_SVO)) throw Error{SVD_ERR_STATE, "toDAG: octree is not an SVO"};
  for (auto& lv : o->dag) { dfree(c, lv.rec); }
  o->dag.clear();
  o->nNodesDAG = dedup_all(c, o->levels, o->svo, o->dag);
  o->state = SVD_S_DAG;
}

namespace {

// One toSDAG pass.  speculative: every level runs two resolution rounds
// without looking at the host, and the nodes still undecided after the
// second are counted per level and read back at the end with the level
// sizes; the pass returns false (the octree's SDAG levels are dropped) when
// any level had not converged, and the caller reruns it with a convergence
// check after every round.
bool reduce_sym_pass(svd_octree* o, bool speculative) {
  svd_ctx* c = o->ctx;
  const uint32_t L = o->levels;
  for (auto& lv : o->sdag) { dfree(c, lv.rec); dfree(c, lv.inv); }
  for (auto* m : o->sdagMap) dfree(c, m);
  o->sdag.assign(L, Level());
  o->sdagMap.assign(L, nullptr);
  uint64_t nNodes = 0;
  uint32_t* count = dalloc<uint32_t>(c, 1);
  // [0, L): representative counts of every level; [L, 2L): nodes undecided
  // after the speculative rounds; [2L]: nodes without correspondence
  uint32_t* dev = dalloc<uint32_t>(c, 2 * L + 1);
  uint32_t* counts = dev;
  uint32_t* undecided = dev + L;
  uint32_t* err = dev + 2 * L;
  fill_u32(c, dev, 0u, 2 * L + 1);
  {
    const Level& dl = o->dag[L - 1];
    Level& sl = o->sdag[L - 1];
    sl.leaf = true;
    sl.rec = dalloc<uint32_t>(c, 256);
    sl.inv = dalloc<uint8_t>(c, 256);
    o->sdagMap[L - 1] = dalloc<uint32_t>(c, dl.n + 1);
    k_sdag_leaves<<<1, 256, 0, c->stream>>>(dl.rec, (uint32_t)dl.n, o->sdagMap[L - 1], sl.rec, sl.inv,
                                            counts + (L - 1));
    check_launch("k_sdag_leaves");
  }
  for (int lev = (int)L - 2; lev >= 0; --lev) {
    const Level& d = o->dag[lev];
    Level& s = o->sdag[lev];
    const uint64_t n = d.n;
    const uint8_t* childInv = (lev < (int)L - 1) ? o->sdag[lev + 1].inv : nullptr;
    uint32_t* key = dalloc<uint32_t>(c, (n ? n : 1) * kRecWords);
    uint8_t* inv = dalloc<uint8_t>(c, n + 1);
    // this level and all above it at most kSdagSmall nodes: one launch
    uint64_t maxTop = 0;
    for (int l = lev; l >= 1; --l) maxTop = std::max<uint64_t>(maxTop, o->dag[l].n);
    if (maxTop <= small_limit(c->smallSdag, kSdagSmall)) {
      dfree(c, key);
      dfree(c, inv);
      SdagTop top{};
      top.top = lev;
      for (int l = lev; l >= 0; --l) {
        const uint64_t m = o->dag[l].n;
        top.dag[l] = o->dag[l].rec;
        top.n[l] = (uint32_t)m;
        Level& sl = o->sdag[l];
        sl.rec = dalloc<uint32_t>(c, (m ? m : 1) * kRecWords);
        sl.inv = dalloc<uint8_t>(c, m + 1);
        top.out[l] = sl.rec;
        top.invOut[l] = sl.inv;
        if (l > 0) o->sdagMap[l] = dalloc<uint32_t>(c, m + 1);
        top.map[l] = o->sdagMap[l];
      }
      top.map[lev + 1] = o->sdagMap[lev + 1];
      top.invOut[lev + 1] = o->sdag[lev + 1].inv;
      o->sdag[0].n = o->dag[0].n;
      o->sdagMap[0] = nullptr;
      top.counts = counts;
      top.err = err;
      k_sdag_levels_small<<<1, kSmallThreads, 0, c->stream>>>(top);
      check_launch("k_sdag_levels_small");
      break;
    }
    // the root level is never reduced and needs no table
    const uint32_t cap = lev == 0 ? 0u : table_cap(n);
    TEntry* table = cap ? dalloc<TEntry>(c, cap) : nullptr;
    if (n || cap)
      k_sdag_keys<<<std::max<uint64_t>(div_up(n, kB), 1), kB, 0, c->stream>>>(d.rec, n, o->sdagMap[lev + 1], key, inv,
                                                                            table, cap);
    check_launch("k_sdag_keys");
    if (lev == 0) {
      // root: children remapped, never reduced, invariant bits stay clear
      s.rec = key;
      s.n = n;
      s.inv = inv;
      fill_u8(c, inv, 0, n);
      o->sdagMap[0] = nullptr;
      break;
    }
    uint32_t* slot = dalloc<uint32_t>(c, n);
    uint32_t* cand = dalloc<uint32_t>(c, n * 7);
    uint32_t* state = dalloc<uint32_t>(c, n);
    uint32_t* pos = dalloc<uint32_t>(c, n + 1);
    launch_insert(c, key, n, table, cap - 1, slot);
    k_sdag_cands<<<div_up(8 * n, kB), kB, 0, c->stream>>>(key, n, table, cap - 1, slot, childInv, cand, state);
    check_launch("k_sdag_cands");
    if (speculative) {
      k_sdag_resolve<<<div_up(n, kB), kB, 0, c->stream>>>(cand, n, state, count);
      k_sdag_resolve<<<div_up(n, kB), kB, 0, c->stream>>>(cand, n, state, undecided + lev);
      check_launch("k_sdag_resolve");
    } else {
      for (int round = 0;; ++round) {
        fill_u32(c, count, 0u, 1);
        k_sdag_resolve<<<div_up(n, kB), kB, 0, c->stream>>>(cand, n, state, count);
        check_launch("k_sdag_resolve");
        if (read_u32(c, count) == 0) break;
        if (round > (int)n + 2) throw Error{SVD_ERR_DEVICE, "toSDAG: resolution did not converge"};
      }
    }
    scan_op<4>(c, n, StateFlagOp{state, pos}, counts + lev);
    // n records: the upper bound of the representative count, known on the host
    s.rec = dalloc<uint32_t>(c, (n ? n : 1) * kRecWords);
    s.inv = dalloc<uint8_t>(c, n + 1);
    o->sdagMap[lev] = dalloc<uint32_t>(c, n + 1);
    k_sdag_outcome<<<div_up(n, kB), kB, 0, c->stream>>>(key, n, table, cap - 1, slot, state, pos, cand, childInv,
                                                         inv, o->sdagMap[lev], s.rec, s.inv, err);
    check_launch("k_sdag_outcome");
    dfree(c, key); dfree(c, inv); dfree(c, table); dfree(c, slot); dfree(c, cand); dfree(c, state); dfree(c, pos);
  }
  std::vector<uint32_t> h(2 * L + 1);
  read_u32s(c, dev, h.data(), 2 * L + 1);
  dfree(c, count); dfree(c, dev);
  for (uint32_t lev = 0; lev < L; ++lev)
    if (h[L + lev]) return false;
  const uint64_t nerr = h[2 * L];
  for (uint32_t lev = 1; lev < L; ++lev) {
    o->sdag[lev].n = h[lev];
    nNodes += h[lev];
  }
  if (nerr) throw Error{SVD_ERR_DEVICE, "toSDAG: node without correspondence"};
  o->nNodesSDAG = nNodes;
  o->state = SVD_S_SDAG;
  return true;
}

}  // namespace

namespace {

// toSDAG with skipSymmetry (geom_octree.cpp:591-596): a node with children
// becomes its own representative with no mirror bits; a node without children
// is dropped.  flag[i] = 1 for the nodes kept.
__global__ void k_nonempty(const uint32_t* __restrict__ rec, uint64_t n, uint32_t leaf, uint32_t* __restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t w0 = leaf ? rec[i] : rec[i * kRecWords];
  flag[i] = (w0 & 0xFFu) ? 1u : 0u;
}

// Node i of the level moves to pos[i] (kept nodes only; the root always
// stays), its children to childMap[child]; map[i] is its new index or kNone.
__global__ void k_pointer_level(const uint32_t* __restrict__ rec, uint64_t n, uint32_t leaf, uint32_t keepAll,
                                const uint32_t* __restrict__ pos, const uint32_t* __restrict__ childMap,
                                uint32_t* __restrict__ out, uint32_t* __restrict__ map) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t w0 = leaf ? rec[i] : rec[i * kRecWords];
  const uint32_t m = w0 & 0xFFu;
  if (!m && !keepAll) {
    if (map) map[i] = kNone;
    return;
  }
  const uint32_t p = keepAll ? (uint32_t)i : pos[i];
  if (map) map[i] = p;
  if (leaf) {
    out[p] = m;
    return;
  }
  uint32_t w[kRecWords];
  load_rec(rec, i, w);
  w[0] = m;
#pragma unroll
  for (int c = 0; c < 8; ++c) w[1 + c] = ((m >> c) & 1u) ? childMap[w[1 + c]] : kNull;
  uint32_t* o = out + (uint64_t)p * kRecWords;
#pragma unroll
  for (int k = 0; k < kRecWords; ++k) o[k] = w[k];
}

}  // namespace

void reduce_sym_pointers(svd_octree* o) {
  if (o->state != SVD_S_SVO && o->state != SVD_S_DAG)
    throw Error{SVD_ERR_STATE, "toSDAG: octree is not an SVO or a DAG"};
  svd_ctx* c = o->ctx;
  const uint32_t L = o->levels;
  const std::vector<Level>& src = o->state == SVD_S_DAG ? o->dag : o->svo;
  for (auto& lv : o->sdag) { dfree(c, lv.rec); dfree(c, lv.inv); }
  for (auto* m : o->sdagMap) dfree(c, m);
  o->sdag.assign(L, Level());
  o->sdagMap.assign(L, nullptr);
  uint32_t* counts = dalloc<uint32_t>(c, L);
  fill_u32(c, counts, 0u, L);
  for (int lev = (int)L - 1; lev >= 0; --lev) {
    const Level& d = src[lev];
    Level& s = o->sdag[lev];
    const uint64_t n = d.n;
    s.leaf = d.leaf;
    s.rec = dalloc<uint32_t>(c, (n ? n : 1) * (d.leaf ? 1 : kRecWords));
    s.inv = dalloc<uint8_t>(c, n + 1);
    fill_u8(c, s.inv, 0, n + 1);
    const uint32_t* childMap = lev + 1 < (int)L ? o->sdagMap[lev + 1] : nullptr;
    if (lev == 0) {
      // the root is never reduced (the reference loops lev > 0)
      if (n) k_pointer_level<<<div_up(n, kB), kB, 0, c->stream>>>(d.rec, n, d.leaf ? 1u : 0u, 1u, nullptr,
                                                                    childMap, s.rec, nullptr);
      check_launch("k_pointer_level");
      s.n = n;
      break;
    }
    uint32_t* pos = dalloc<uint32_t>(c, n + 1);
    o->sdagMap[lev] = dalloc<uint32_t>(c, n + 1);
    if (n) {
      k_nonempty<<<div_up(n, kB), kB, 0, c->stream>>>(d.rec, n, d.leaf ? 1u : 0u, pos);
      scan_exclusive(c, pos, pos, n, counts + lev);
      k_pointer_level<<<div_up(n, kB), kB, 0, c->stream>>>(d.rec, n, d.leaf ? 1u : 0u, 0u, pos, childMap, s.rec,
                                                            o->sdagMap[lev]);
      check_launch("k_pointer_level");
    }
    dfree(c, pos);
  }
  std::vector<uint32_t> h(L);
  read_u32s(c, counts, h.data(), L);
  dfree(c, counts);
  uint64_t nNodes = 0;
  for (uint32_t lev = 1; lev < L; ++lev) {
    o->sdag[lev].n = h[lev];
    nNodes += h[lev];
  }
  o->nNodesSDAG = nNodes;
  o->state = SVD_S_SDAG;
}

void reduce_sym(svd_octree* o) {
  if (o->state != SVD_S_DAG && o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "toSDAG: octree is not a DAG"};
  if (o->levels == 1) throw Error{SVD_ERR_LEVELS, "toSDAG needs at least 2 levels"};
  if (o->ctx->sdagChecked || !reduce_sym_pass(o, true)) reduce_sym_pass(o, false);
}

}  // namespace svd
