// On-device encoders producing the exact file images of
//   EncodedSVDAG   ::encode+save (encoded_svdag.cpp:76-199)
//   EncodedUSSVDAG ::encode+save (encoded_ussvdag.cpp:60-170)
//   EncodedSSVDAG  ::encode+save (encoded_ssvdag.cpp:84-466), on the SDAG
//                  (.ssvdag) or on the DAG (.esvdag, svbuilder/main.cpp:243-244).
#include "svd_internal.hpp"
#include <algorithm>

namespace svd {

namespace {

constexpr int kB = 256;

// ---------------------------------------------------------------- SVDAG / USSVDAG
__global__ void k_word_counts(const uint32_t* __restrict__ rec, uint64_t n, int leaf, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = leaf ? 1u : 1u + (uint32_t)__popc(rec[i * kRecWords] & 0xFFu);
}

// ussv = 0: header word = mask; child pointers follow childLevels
//           (offset = first node of that level), c = 7..0.
// ussv = 1: header word = mZ<<24 | mY<<16 | mX<<8 | mask; pointers into the
//           next level (childLevels ignored, encoded_ussvdag.cpp:137).
__global__ void k_write_words(const uint32_t* __restrict__ rec, uint64_t n, int leaf, int ussv,
                              const uint32_t* __restrict__ childLv, uint32_t lev, const uint32_t* __restrict__ levStart,
                              const uint32_t* __restrict__ truePtrs, uint32_t base, uint64_t totalNodes,
                              uint32_t* __restrict__ words, uint32_t* __restrict__ err) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t at = truePtrs[base + i];
  if (leaf) { words[at] = rec[i] & 0xFFu; return; }
  const uint32_t* r = rec + i * kRecWords;
  const uint32_t w0 = r[0];
  words[at] = ussv ? w0 : (w0 & 0xFFu);
  uint32_t k = at + 1;
  const uint32_t kEnd = k + (uint32_t)__popc(w0 & 0xFFu);   // the words the size pass reserved
#pragma unroll
  for (int c = 7; c >= 0; --c) {
    const uint32_t ch = r[1 + c];
    if (ch == kNull) continue;
    if (k >= kEnd) { atomicAdd(err, 1u); break; }
    const uint32_t cl = (!ussv && childLv) ? childLv[i * 8 + c] : lev + 1;
    const uint64_t g = (uint64_t)ch + levStart[cl];
    if (g >= totalNodes) { atomicAdd(err, 1u); words[k++] = 0u; continue; }
    words[k++] = truePtrs[g];
  }
}

// ---------------------------------------------------------------- SSVDAG
// Reference counts of the children of a level.  A few children (common
// bricks) take most references, so each block first aggregates in an LDS
// hash table (1024 slots, at most 8 probes; a child that finds no slot goes
// straight to the global histogram) and flushes one add per distinct child.
constexpr int kRcSlots = 1024;
constexpr int kRcBlocks = 2048;
__global__ void __launch_bounds__(256) k_refcount(const uint32_t* __restrict__ parent, uint64_t np,
                                                  uint32_t* __restrict__ hist, uint64_t nChild,
                                                  uint32_t* __restrict__ err) {
  __shared__ uint32_t key[kRcSlots], cnt[kRcSlots];
  for (int t = threadIdx.x; t < kRcSlots; t += blockDim.x) { key[t] = kNone; cnt[t] = 0; }
  __syncthreads();
  const uint64_t chunk = (np + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < np ? lo + chunk : np;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const uint32_t* r = parent + i * kRecWords;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t ch = r[1 + c];
      if (ch == kNull) continue;
      if (ch >= nChild) { atomicAdd(err, 1u); continue; }
      uint32_t h = (ch * 2654435761u) >> 22;                 // 10-bit slot
      bool done = false;
#pragma unroll 1
      for (int p = 0; p < 8 && !done; ++p, h = (h + 1) & (kRcSlots - 1)) {
        uint32_t k = key[h];
        if (k == kNone) {
          const uint32_t old = atomicCAS(&key[h], kNone, ch);
          k = old == kNone ? ch : old;
        }
        if (k == ch) { atomicAdd(&cnt[h], 1u); done = true; }
      }
      if (!done) atomicAdd(&hist[ch], 1u);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kRcSlots; t += blockDim.x)
    if (key[t] != kNone && cnt[t]) atomicAdd(&hist[key[t]], cnt[t]);
}

__global__ void k_scatter_rank(const uint32_t* __restrict__ perm, uint64_t n, uint32_t* __restrict__ ind) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ind[perm[i]] = (uint32_t)i;
}

// Encoded size (u16 words) of the node at sorted position i.
__global__ void k_inner_sizes(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ perm, uint64_t n,
                              const uint32_t* __restrict__ indChild, uint32_t* __restrict__ sz,
                              uint32_t* __restrict__ err) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* r = rec + (uint64_t)perm[i] * kRecWords;
  const uint32_t m = r[0] & 0xFFu;
  uint32_t s = 1;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (!((m >> c) & 1u)) continue;
    const uint32_t a = indChild[r[1 + c]];
    if (a < (1u << 13)) s += 1;
    else if (a < (1u << 30)) s += 2;
    else atomicAdd(err, 1u);
  }
  sz[i] = s;
}

__global__ void k_offsets_to_ind(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ offs, uint64_t n,
                                 uint32_t* __restrict__ ind) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ind[perm[i]] = offs[i];
}

// Variable-length inner node (encoded_ssvdag.cpp:352-430).
__global__ void k_inner_write(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ perm, uint64_t n,
                              const uint32_t* __restrict__ indChild, const uint32_t* __restrict__ offs,
                              uint16_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* r = rec + (uint64_t)perm[i] * kRecWords;
  const uint32_t w0 = r[0];
  const uint32_t m = w0 & 0xFFu, mx = (w0 >> 8) & 0xFFu, my = (w0 >> 16) & 0xFFu, mz = (w0 >> 24) & 0xFFu;
  uint16_t* o = out + offs[i];
  uint32_t hdr = 0;
  uint32_t k = 1;
#pragma unroll
  for (int c = 7; c >= 0; --c) {
    if (!((m >> c) & 1u)) continue;
    const uint32_t a = indChild[r[1 + c]];
    const uint32_t bx = (mx >> c) & 1u, by = (my >> c) & 1u, bz = (mz >> c) & 1u;
    if (a < (1u << 13)) {
      hdr |= 1u << (2 * c);
      o[k++] = (uint16_t)(a | (bx << 13) | (by << 14) | (bz << 15));
    } else {
      uint32_t p = a;
      if (p & (1u << 29)) { hdr |= 3u << (2 * c); p &= ~(1u << 29); }
      else hdr |= 2u << (2 * c);
      p |= (bx << 29) | (by << 30) | (bz << 31);
      o[k++] = (uint16_t)(p >> 16);
      o[k++] = (uint16_t)(p & 0xFFFFu);
    }
  }
  o[0] = (uint16_t)hdr;
}

__device__ __forceinline__ uint32_t perm8(uint32_t m, uint32_t flip) {
  uint32_t r = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) r |= ((m >> c) & 1u) << (c ^ flip);
  return r;
}

// 4^3 brick in matrix order (encoded_ssvdag.cpp:303-346): bit x+4y+16z of the
// output <- byte (x>>1)+2(y>>1)+4(z>>1), bit (x&1)+2(y&1)+4(z&1) of the
// children masks (each mirrored by the parent's mirror bits).
__global__ void k_brick_write(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ perm, uint64_t n,
                              const uint32_t* __restrict__ leafMask, uint16_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* r = rec + (uint64_t)perm[i] * kRecWords;
  const uint32_t w0 = r[0];
  uint32_t buf[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t ch = r[1 + c];
    if (ch == kNull) { buf[c] = 0; continue; }
    const uint32_t flip = (((w0 >> (8 + c)) & 1u) << 2) | (((w0 >> (16 + c)) & 1u) << 1) | ((w0 >> (24 + c)) & 1u);
    buf[c] = perm8(leafMask[ch] & 0xFFu, flip);
  }
  uint64_t bits = 0;
#pragma unroll
  for (int z = 0; z < 4; ++z)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int byte = (x >> 1) + 2 * (y >> 1) + 4 * (z >> 1);
        const int bit = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
        bits |= (uint64_t)((buf[byte] >> bit) & 1u) << (x + 4 * y + 16 * z);
      }
  uint16_t* o = out + 4 * i;
  o[0] = (uint16_t)bits; o[1] = (uint16_t)(bits >> 16); o[2] = (uint16_t)(bits >> 32); o[3] = (uint16_t)(bits >> 48);
}

void put(std::vector<uint8_t>& b, const void* p, size_t n) {
  const uint8_t* s = (const uint8_t*)p;
  b.insert(b.end(), s, s + n);
}

std::vector<uint8_t> base_header(const svd_octree* o, uint64_t nNodes) {
  std::vector<uint8_t> h;
  put(h, o->bbox, 12);
  put(h, o->bbox + 3, 12);
  const float rs = (float)o->rootSide;
  put(h, &rs, 4);
  const uint32_t lv = o->levels;
  put(h, &lv, 4);
  const uint32_t nn = (uint32_t)nNodes;   // size_t truncated to 4 bytes on save
  put(h, &nn, 4);
  return h;
}

void upload(svd_ctx* c, void* dst, const std::vector<uint8_t>& src) {
  if (src.empty()) return;
  SVD_HIP(hipMemcpyAsync(dst, src.data(), src.size(), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));   // src is pageable host memory
}

void encode_words(svd_octree* o, int fmt) {
  svd_ctx* c = o->ctx;
  const bool ussv = fmt == SVD_FMT_USSVDAG;
  const std::vector<Level>& lv = ussv ? o->sdag : (o->hasX ? o->xdag : o->dag);
  const uint32_t L = o->levels;
  std::vector<uint32_t> levStart(L + 1, 0);
  for (uint32_t l = 0; l < L; ++l) levStart[l + 1] = levStart[l] + (uint32_t)lv[l].n;
  const uint64_t N = levStart[L];
  uint32_t* dLevStart = dalloc<uint32_t>(c, L + 1);
  SVD_HIP(hipMemcpyAsync(dLevStart, levStart.data(), 4 * (L + 1), hipMemcpyHostToDevice, c->stream));
  uint32_t* ptrs = dalloc<uint32_t>(c, N + 1);
  uint32_t* total = dalloc<uint32_t>(c, 1);
  uint32_t* err = dalloc<uint32_t>(c, 1);
  fill_u32(c, err, 0u, 1);
  for (uint32_t l = 0; l < L; ++l)
    if (lv[l].n) k_word_counts<<<div_up(lv[l].n, kB), kB, 0, c->stream>>>(lv[l].rec, lv[l].n, lv[l].leaf, ptrs + levStart[l]);
  check_launch("k_word_counts");
  scan_exclusive(c, ptrs, ptrs, N, total);
  uint32_t h2[2];
  read_u32s(c, total, h2, 1);
  const uint64_t count = h2[0];
  // The reference assigns _firstLeafPtr after the loop over ALL levels
  // (encoded_svdag.cpp:129-137), so the stored value is the total word count.
  const uint32_t firstLeaf = (uint32_t)count;
  const uint64_t nNodes = ussv ? o->nNodesSDAG : (o->hasX ? o->nNodesX : o->nNodesDAG);
  std::vector<uint8_t> hdr = base_header(o, nNodes);
  put(hdr, &firstLeaf, 4);
  const uint32_t cnt32 = (uint32_t)count;
  put(hdr, &cnt32, 4);
  Encoded& e = o->enc[fmt];
  dfree(c, e.d);
  e.bytes = hdr.size() + 4 * count;
  e.d = dalloc<uint8_t>(c, e.bytes);
  upload(c, e.d, hdr);
  uint32_t* words = (uint32_t*)(e.d + hdr.size());
  for (uint32_t l = 0; l < L; ++l) {
    if (!lv[l].n) continue;
    k_write_words<<<div_up(lv[l].n, kB), kB, 0, c->stream>>>(lv[l].rec, lv[l].n, lv[l].leaf, ussv ? 1 : 0,
                                                              lv[l].childLevels, l, dLevStart, ptrs, levStart[l], N,
                                                              words, err);
  }
  check_launch("k_write_words");
  const uint64_t nerr = read_u32(c, err);
  dfree(c, dLevStart); dfree(c, ptrs); dfree(c, total); dfree(c, err);
  if (nerr) throw Error{SVD_ERR_POINTER, "encode: child pointer without a target node"};
  e.valid = true;
}

void encode_ssvdag(svd_octree* o, int fmt) {
  svd_ctx* c = o->ctx;
  const std::vector<Level>& lv = (fmt == SVD_FMT_SSVDAG) ? o->sdag : o->dag;
  const int L = (int)o->levels;
  if (L < 3) throw Error{SVD_ERR_LEVELS, "SSVDAG encoding needs at least 3 levels"};
  for (int l = 0; l < L; ++l)
    if (lv[l].n > (1ull << 30)) throw Error{SVD_ERR_CAPACITY, "level too big for encoding in 30 bits"};
  std::vector<uint32_t*> perm(L, nullptr), ind(L, nullptr), offs(L, nullptr);
  std::vector<uint32_t> innerSize(L, 0);
  // dev[0..L-1]: inner-size totals per level, dev[L]: pointer-size error flag
  uint32_t* dev = dalloc<uint32_t>(c, L + 1);
  uint32_t* err = dev + L;
  fill_u32(c, dev, 0u, L + 1);
  // Phase 1a: the reference counts of every level land in one array, and one
  // segmented sort orders all levels (each level is its own std::sort call).
  std::vector<uint64_t> off(L, 0), rfirst, rlast;
  uint64_t tot = 0;
  for (int lev = 0; lev <= L - 2; ++lev) { off[lev] = tot; tot += lv[lev].n; }
  uint32_t* permAll = dalloc<uint32_t>(c, tot + 1);
  uint32_t* hist = dalloc<uint32_t>(c, tot + 1);
  fill_u32(c, hist, 0u, tot);
  for (int lev = 0; lev <= L - 2; ++lev) {
    const uint64_t n = lv[lev].n;
    perm[lev] = permAll + off[lev];
    iota_u32(c, perm[lev], n);
    if (lev > 0 && n > 0) {
      if (lv[lev - 1].n)
        k_refcount<<<std::min<uint64_t>(div_up(lv[lev - 1].n, kB), kRcBlocks), kB, 0, c->stream>>>(
            lv[lev - 1].rec, lv[lev - 1].n, hist + off[lev], n, err);
      check_launch("k_refcount");
      rfirst.push_back(off[lev]);
      rlast.push_back(off[lev] + n);
    }
  }
  refsort_ranges(c, permAll, hist, tot, rfirst.data(), rlast.data(), rfirst.size());
  dfree(c, hist);
  // Phase 1b, bottom-up: children addresses of each level in its order.
  for (int lev = L - 2; lev >= 0; --lev) {
    const uint64_t n = lv[lev].n;
    ind[lev] = dalloc<uint32_t>(c, n + 1);
    if (lev == L - 2) {
      if (n) k_scatter_rank<<<div_up(n, kB), kB, 0, c->stream>>>(perm[lev], n, ind[lev]);
      check_launch("k_scatter_rank");
    } else {
      offs[lev] = dalloc<uint32_t>(c, n + 1);
      if (n) {
        k_inner_sizes<<<div_up(n, kB), kB, 0, c->stream>>>(lv[lev].rec, perm[lev], n, ind[lev + 1], offs[lev], err);
        check_launch("k_inner_sizes");
      }
      scan_exclusive(c, offs[lev], offs[lev], n, dev + lev);
      if (n) k_offsets_to_ind<<<div_up(n, kB), kB, 0, c->stream>>>(perm[lev], offs[lev], n, ind[lev]);
      check_launch("k_offsets_to_ind");
    }
  }
  {
    std::vector<uint32_t> h(L + 1);
    read_u32s(c, dev, h.data(), L + 1);
    if (h[L]) throw Error{SVD_ERR_CAPACITY, "unsupported pointer size (>= 2^30)"};
    for (int lev = 0; lev < L - 2; ++lev) innerSize[lev] = h[lev];
  }
  // Layout of the file image.
  std::vector<uint32_t> levelOffsets(L - 2, 0);
  for (int i = 1; i < L - 2; ++i) levelOffsets[i] = innerSize[i - 1] + levelOffsets[i - 1];
  uint64_t nInner = 0;
  for (int i = 0; i < L - 2; ++i) nInner += innerSize[i];
  const uint64_t nBricks = lv[L - 2].n;
  const uint64_t nNodes = (fmt == SVD_FMT_SSVDAG) ? o->nNodesSDAG : o->nNodesDAG;
  std::vector<uint8_t> head = base_header(o, nNodes);
  const uint32_t nInner32 = (uint32_t)nInner;
  put(head, &nInner32, 4);
  const uint64_t innerAt = head.size();
  const uint64_t leafCntAt = innerAt + 2 * nInner;
  const uint64_t leafAt = leafCntAt + 4;
  const uint64_t tailAt = leafAt + 8 * nBricks;
  std::vector<uint8_t> mid, tail;
  const uint32_t nLeafBytes = (uint32_t)(8 * nBricks);
  put(mid, &nLeafBytes, 4);
  const uint32_t nOff = (uint32_t)levelOffsets.size();
  put(tail, &nOff, 4);
  put(tail, levelOffsets.data(), 4 * levelOffsets.size());
  Encoded& e = o->enc[fmt];
  dfree(c, e.d);
  e.bytes = tailAt + tail.size();
  e.d = dalloc<uint8_t>(c, e.bytes);
  SVD_HIP(hipMemcpyAsync(e.d, head.data(), head.size(), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(e.d + leafCntAt, mid.data(), mid.size(), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(e.d + tailAt, tail.data(), tail.size(), hipMemcpyHostToDevice, c->stream));
  // Phase 2: write every level into the image.
  uint16_t* inner = (uint16_t*)(e.d + innerAt);
  for (int lev = 0; lev <= L - 3; ++lev) {
    const uint64_t n = lv[lev].n;
    if (!n) continue;
    k_inner_write<<<div_up(n, kB), kB, 0, c->stream>>>(lv[lev].rec, perm[lev], n, ind[lev + 1], offs[lev],
                                                        inner + levelOffsets[lev]);
  }
  if (nBricks)
    k_brick_write<<<div_up(nBricks, kB), kB, 0, c->stream>>>(lv[L - 2].rec, perm[L - 2], nBricks, lv[L - 1].rec,
                                                              (uint16_t*)(e.d + leafAt));
  check_launch("ssvdag write");
  SVD_HIP(hipStreamSynchronize(c->stream));   // host staging vectors go out of scope
  for (int l = 0; l < L; ++l) { dfree(c, ind[l]); dfree(c, offs[l]); }
  dfree(c, permAll); dfree(c, dev);
  e.valid = true;
}

}  // namespace

void encode(svd_octree* o, int fmt) {
  switch (fmt) {
    case SVD_FMT_SVDAG:
      if (o->state < SVD_S_DAG) throw Error{SVD_ERR_STATE, "SVDAG encoding needs a DAG"};
      encode_words(o, fmt);
      break;
    case SVD_FMT_USSVDAG:
      if (o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "USSVDAG encoding needs an SDAG"};
      encode_words(o, fmt);
      break;
    case SVD_FMT_SSVDAG:
      if (o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "SSVDAG encoding needs an SDAG"};
      encode_ssvdag(o, fmt);
      break;
    case SVD_FMT_ESVDAG:
      if (o->state < SVD_S_DAG) throw Error{SVD_ERR_STATE, "ESVDAG encoding needs a DAG"};
      encode_ssvdag(o, fmt);
      break;
    default:
      throw Error{SVD_ERR_ARG, "unknown format"};
  }
}

}  // namespace svd
