// Gray-code attribute bit-tree (the reference names the feature in readme.md:8
// and has no code for it; the encoding is defined in DESIGN.md).
//
// Attributes are given in voxel (Morton / DFS) order.  Code = Gray b^(b>>1)
// or binary.  Per bit plane, a pruned binary tree over the sequence padded to
// P = 2^ceil(log2 V) with its last value is emitted in pre-order:
// single element -> its bit; uniform node -> 0, bit; mixed node -> 1, left,
// right.  Bits are packed LSB-first in u32 words; each plane starts on a word.
//
// Parallel form: a node is emitted iff its parent is mixed (mixedness is
// monotone upward, so the parent alone decides).  The emitted nodes that
// start at position i form a chain (i, t), (i, t-1), ... with t = tz(i); one
// lane walks it, so per plane: change flags -> exclusive scan E (uniformity
// of [lo,hi) is E[hi] == E[lo+1]) -> per-position bit counts -> exclusive
// scan (pre-order offsets) -> emit.  Everything streams through HBM.
#include "svd_internal.hpp"

namespace svd {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint32_t code_of(const uint8_t* __restrict__ a, uint64_t V, uint64_t i, int gray) {
  const uint32_t b = a[i < V ? i : V - 1];
  return gray ? (b ^ (b >> 1)) : b;
}

// flags[i] = 1 iff bit k of element i differs from element i-1 (i < P)
__global__ void k_change_flags(const uint8_t* __restrict__ a, uint64_t V, uint64_t P, int gray, int k,
                               uint32_t* __restrict__ flags) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > P) return;
  if (i == 0 || i >= P) { flags[i] = 0; return; }
  const uint32_t b0 = (code_of(a, V, i - 1, gray) >> k) & 1u, b1 = (code_of(a, V, i, gray) >> k) & 1u;
  flags[i] = b0 != b1 ? 1u : 0u;
}

__device__ __forceinline__ bool mixed(const uint32_t* __restrict__ E, uint64_t lo, uint32_t h) {
  if (h == 0) return false;
  return E[lo + (1ull << h)] != E[lo + 1];
}

__device__ __forceinline__ uint32_t top_height(uint64_t i, uint32_t H) {
  if (i == 0) return H;
  const uint32_t t = (uint32_t)__builtin_ctzll(i);
  return t < H ? t : H;
}

// number of bits emitted by the nodes that start at position i
__global__ void k_bit_counts(const uint32_t* __restrict__ E, uint64_t P, uint32_t H, uint32_t* __restrict__ cnt) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > P) return;
  if (i == P) { cnt[i] = 0; return; }
  uint32_t h = top_height(i, H);
  if (i > 0 && !mixed(E, i - (1ull << h), h + 1)) { cnt[i] = 0; return; }   // parent uniform: not emitted
  uint32_t n = 0;
  while (true) {
    if (h == 0) { n += 1; break; }
    if (!mixed(E, i, h)) { n += 2; break; }
    n += 1;
    --h;
  }
  cnt[i] = n;
}

__global__ void k_emit(const uint8_t* __restrict__ a, uint64_t V, int gray, int k, const uint32_t* __restrict__ E,
                       uint64_t P, uint32_t H, const uint32_t* __restrict__ off, uint32_t* __restrict__ words) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  uint32_t h = top_height(i, H);
  if (i > 0 && !mixed(E, i - (1ull << h), h + 1)) return;
  uint64_t pos = off[i];
  const uint32_t bit = (code_of(a, V, i, gray) >> k) & 1u;
  auto put = [&](uint32_t b) {
    if (b) atomicOr(&words[pos >> 5], 1u << (pos & 31));
    ++pos;
  };
  while (true) {
    if (h == 0) { put(bit); break; }
    if (!mixed(E, i, h)) { put(0); put(bit); break; }
    put(1);
    --h;
  }
}

}  // namespace

// Returns the blob size; fills dst (host) when cap suffices.
uint64_t attr_bittree(svd_ctx* c, const uint8_t* dattrs, uint64_t V, int gray, uint8_t* dst, uint64_t cap) {
  uint64_t P = 1;
  uint32_t H = 0;
  while (P < V) { P <<= 1; ++H; }
  uint32_t bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<uint32_t*> E(8, nullptr), O(8, nullptr);
  uint32_t* total = dalloc<uint32_t>(c, 8);
  if (V) {
    for (int k = 0; k < 8; ++k) {
      E[k] = dalloc<uint32_t>(c, P + 1);
      O[k] = dalloc<uint32_t>(c, P + 1);
      k_change_flags<<<div_up(P + 1, kB), kB, 0, c->stream>>>(dattrs, V, P, gray, k, E[k]);
      scan_exclusive(c, E[k], E[k], P + 1, nullptr);
      k_bit_counts<<<div_up(P + 1, kB), kB, 0, c->stream>>>(E[k], P, H, O[k]);
      scan_exclusive(c, O[k], O[k], P + 1, total + k);
    }
    check_launch("bit-tree counts");
    read_u32s(c, total, bits, 8);
  }
  uint64_t nwords = 0;
  std::vector<uint64_t> planeWord(8);
  for (int k = 0; k < 8; ++k) { planeWord[k] = nwords; nwords += (bits[k] + 31) / 32; }
  const uint64_t bytes = 48 + 4 * nwords;
  if (dst && cap >= bytes) {
    uint32_t* words = dalloc<uint32_t>(c, nwords + 1);
    fill_u32(c, words, 0u, nwords + 1);
    for (int k = 0; k < 8 && V; ++k)
      k_emit<<<div_up(P, kB), kB, 0, c->stream>>>(dattrs, V, gray, k, E[k], P, H, O[k], words + planeWord[k]);
    check_launch("k_emit");
    const uint32_t hdr[4] = {0x42415653u, (uint32_t)V, (uint32_t)gray, 8u};
    memcpy(dst, hdr, 16);
    memcpy(dst + 16, bits, 32);
    if (nwords) {
      SVD_HIP(hipMemcpyAsync(dst + 48, words, 4 * nwords, hipMemcpyDeviceToHost, c->stream));
    }
    SVD_HIP(hipStreamSynchronize(c->stream));
    dfree(c, words);
  }
  for (int k = 0; k < 8; ++k) { dfree(c, E[k]); dfree(c, O[k]); }
  dfree(c, total);
  return bytes;
}

}  // namespace svd
