// On-device encoders producing the exact file images of
//   EncodedSVDAG   ::encode+save (encoded_svdag.cpp:76-199)
//   EncodedUSSVDAG ::encode+save (encoded_ussvdag.cpp:60-170)
//   EncodedSSVDAG  ::encode+save (encoded_ssvdag.cpp:84-466), on the SDAG
//                  (.ssvdag) or on the DAG (.esvdag, svbuilder/main.cpp:243-244).
#include "svd_internal.hpp"
#include "scan_op.hpp"
#include "block_scan.hpp"
#include <algorithm>

namespace svd {

namespace {

constexpr int kB = 256;

// ---------------------------------------------------------------- level sets
// All levels of an octree in one launch: node j of level l is element
// start[l] + j of the concatenation (start[L] = total), and per-level arrays
// travel in the kernel arguments.
struct LevelSet {
  const uint32_t* rec[kMaxLevels];
  const uint32_t* aux[kMaxLevels];   // childLevels (SVDAG), or a per-level array of the kernel
  uint32_t start[kMaxLevels + 1];
  uint32_t L;
  uint32_t leafBits;                 // bit l: level l holds leaf masks
};

__device__ __forceinline__ uint32_t level_of(const LevelSet& a, uint64_t i) {
  uint32_t l = 0;
  while (l + 1 < a.L && i >= a.start[l + 1]) ++l;
  return l;
}

// File header pieces (up to kPutMax of up to 128 bytes) written in one
// launch, one workgroup per piece, from the kernel arguments: no host staging
// buffer has to outlive the launch.
constexpr uint32_t kPutMax = 8;
struct PutBatch {
  uint8_t* dst[kPutMax];
  uint32_t n[kPutMax];
  uint8_t b[kPutMax][128];
};
__global__ void k_put_batch(PutBatch p) {
  const uint32_t k = blockIdx.x, t = threadIdx.x;
  if (t < p.n[k]) p.dst[k][t] = p.b[k][t];
}

// ---------------------------------------------------------------- SVDAG / USSVDAG
__global__ void k_word_counts(LevelSet a, uint64_t n, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t l = level_of(a, i);
  const uint64_t j = i - a.start[l];
  out[i] = ((a.leafBits >> l) & 1u) ? 1u : 1u + (uint32_t)__popc(a.rec[l][j * kRecWords] & 0xFFu);
}

// ussv = 0: header word = mask; child pointers follow childLevels
//           (offset = first node of that level), c = 7..0.
// ussv = 1: header word = mZ<<24 | mY<<16 | mX<<8 | mask; pointers into the
//           next level (childLevels ignored, encoded_ussvdag.cpp:137).
// The header's two count words (_firstLeafPtr and the array length, both =
// the word count *total, see encode_words) go to cnt from thread 0.
__global__ void k_write_words(LevelSet a, int ussv, const uint32_t* __restrict__ truePtrs, uint64_t totalNodes,
                              uint32_t* __restrict__ words, uint32_t* __restrict__ err,
                              const uint32_t* __restrict__ total, uint32_t* __restrict__ cnt) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) cnt[0] = cnt[1] = *total;
  if (i >= totalNodes) return;
  const uint32_t lev = level_of(a, i);
  const uint64_t j = i - a.start[lev];
  const uint32_t at = truePtrs[i];
  if ((a.leafBits >> lev) & 1u) { words[at] = a.rec[lev][j] & 0xFFu; return; }
  const uint32_t* r = a.rec[lev] + j * kRecWords;
  const uint32_t* childLv = a.aux[lev];
  const uint32_t w0 = r[0];
  words[at] = ussv ? w0 : (w0 & 0xFFu);
  uint32_t k = at + 1;
  const uint32_t kEnd = k + (uint32_t)__popc(w0 & 0xFFu);   // the words the size pass reserved
#pragma unroll
  for (int c = 7; c >= 0; --c) {
    const uint32_t ch = r[1 + c];
    if (ch == kNull) continue;
    if (k >= kEnd) { atomicAdd(err, 1u); break; }
    const uint32_t cl = (!ussv && childLv) ? childLv[j * 8 + c] : lev + 1;
    const uint64_t g = cl < a.L ? (uint64_t)ch + a.start[cl] : totalNodes;
    if (g >= totalNodes) { atomicAdd(err, 1u); words[k++] = 0u; continue; }
    words[k++] = truePtrs[g];
  }
}

// ---------------------------------------------------------------- SSVDAG
// The levels 0..L-2 of one or two octree views (.ssvdag of the SDAG, .esvdag
// of the DAG) concatenated: set s level l is span level s * P + l, P = L - 1.
struct Span {
  const uint32_t* rec[2 * kMaxLevels];
  uint32_t start[2 * kMaxLevels + 1];
  uint32_t L;     // span levels (K * P)
  uint32_t P;     // levels per set
};

__device__ __forceinline__ uint32_t span_level(const Span& a, uint64_t i) {
  uint32_t l = 0;
  while (l + 1 < a.L && i >= a.start[l + 1]) ++l;
  return l;
}

// Reference counts of the children of a level.  Most references go to a few
// children (common bricks), whose counters share cache lines, and device
// atomics on one line serialize; so each workgroup counts one range of
// kRcRange children over one chunk of parents in LDS and adds each nonzero
// counter to the histogram once: a child gets at most one device atomic per
// parent chunk, and those of a workgroup are coalesced.
constexpr uint32_t kRcRange = 32768;
constexpr uint32_t kRcParents = 4096;   // parents per workgroup, about
constexpr int kRcT = 1024;
// Levels 1..L-2 of every set in one launch.  Parent level job j is span
// level lev[j], whose children are the next span level (element start[l+1] +
// child of the concatenated histogram); it owns workgroups blk0[j] ..
// blk0[j+1]-1: nr[j] child ranges times (blk0[j+1]-blk0[j]) / nr[j] parent
// chunks.
struct RcJobs {
  uint32_t lev[2 * kMaxLevels];
  uint32_t nr[2 * kMaxLevels];
  uint32_t blk0[2 * kMaxLevels + 1];
  uint32_t n;
};
__global__ void __launch_bounds__(kRcT) k_refcount(Span a, RcJobs jb, uint32_t* __restrict__ histAll,
                                                   uint32_t* __restrict__ err) {
  __shared__ uint32_t cnt[kRcRange];
  uint32_t j = 0;
  while (j + 1 < jb.n && blockIdx.x >= jb.blk0[j + 1]) ++j;
  const uint32_t l = jb.lev[j], nr = jb.nr[j];
  const uint32_t b = blockIdx.x - jb.blk0[j], nch = (jb.blk0[j + 1] - jb.blk0[j]) / nr;
  const uint32_t range = b % nr, chunk = b / nr;
  const uint32_t* __restrict__ parent = a.rec[l];
  const uint64_t np = a.start[l + 1] - a.start[l];
  uint32_t* __restrict__ hist = histAll + a.start[l + 1];
  const uint32_t nChild = a.start[l + 2] - a.start[l + 1];
  const uint32_t r0 = range * kRcRange, rn = min(kRcRange, nChild - r0);
  for (uint32_t t = threadIdx.x; t < rn; t += kRcT) cnt[t] = 0;
  __syncthreads();
  const uint64_t lo = np * chunk / nch, hi = np * (chunk + 1) / nch;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kRcT) {
    const uint32_t* r = parent + i * kRecWords;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t ch = r[1 + c];
      if (ch == kNull) continue;
      if (ch >= nChild) {
        if (range == 0) atomicAdd(err, 1u);
        continue;
      }
      const uint32_t o = ch - r0;
      if (o < rn) atomicAdd(&cnt[o], 1u);
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < rn; t += kRcT) {
    const uint32_t v = cnt[t];
    if (v) atomicAdd(&hist[r0 + t], v);
  }
}

// perm[i] = position of element i within its level (the identity order of
// every level).  The same launch clears the refcount histogram (n words) and
// the nz counters in zero, which would otherwise take two fill launches.
__global__ void k_level_iota(Span a, uint64_t n, uint32_t* __restrict__ perm, uint32_t* __restrict__ hist,
                             uint32_t* __restrict__ zero, uint32_t nz) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    perm[i] = (uint32_t)(i - a.start[span_level(a, i)]);
    hist[i] = 0u;
  }
  if (i < nz) zero[i] = 0u;
}

// The rank of every brick in its set's order, ind[perm[i]] = i, for one or two
// sets' brick levels in one launch (elements n0.. are the second set's).
__global__ void k_scatter_rank2(const uint32_t* __restrict__ perm0, uint64_t n0, uint32_t* __restrict__ ind0,
                                const uint32_t* __restrict__ perm1, uint64_t n1, uint32_t* __restrict__ ind1) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) ind0[perm0[i]] = (uint32_t)i;
  else if (i < n0 + n1) ind1[perm1[i - n0]] = (uint32_t)(i - n0);
}

// Encoded size (u16 words) of the node at sorted position i.
// The layout of an inner level in its refcount order (scan_op): the encoded
// size in u16 words of the node at sorted position i (one word of mask plus
// one or two per child by the child's address), its offset = the sizes
// before it, and ind[node] = that offset.
struct LayoutOp {
  const uint32_t* __restrict__ rec;
  const uint32_t* __restrict__ perm;
  const uint32_t* __restrict__ indChild;
  uint32_t* __restrict__ offs;
  uint32_t* __restrict__ ind;
  uint32_t* __restrict__ err;
  __device__ uint32_t value(uint64_t i, uint32_t& flags) const {
    const uint32_t* r = rec + (uint64_t)perm[i] * kRecWords;
    const uint32_t m = r[0] & 0xFFu;
    uint32_t s = 1;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (!((m >> c) & 1u)) continue;
      const uint32_t a = indChild[r[1 + c]];
      if (a < (1u << 13)) s += 1;
      else if (a < (1u << 30)) s += 2;
      else flags = 1;
    }
    return s;
  }
  __device__ void report(uint32_t) const { atomicAdd(err, 1u); }
  __device__ void apply(uint64_t i, uint32_t ex) const {
    offs[i] = ex;
    ind[perm[i]] = ex;
  }
};

// The layout of a small level (n <= kLayoutSmall) in one workgroup: the
// sizes, their exclusive scan and the node addresses of LayoutOp with
// workgroup barriers.
constexpr uint32_t kLayoutSmall = 4096;   // above this the three launches are faster (measured)
constexpr int kLayoutThreads = 1024;
__device__ void inner_layout_small(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ perm, uint32_t n,
                                   const uint32_t* __restrict__ indChild, uint32_t* __restrict__ offs,
                                   uint32_t* __restrict__ ind, uint32_t* __restrict__ total,
                                   uint32_t* __restrict__ err, uint32_t* lds) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kLayoutThreads) {
    const uint32_t i = base + threadIdx.x;
    uint32_t sz = 0, p = 0;
    if (i < n) {
      p = perm[i];
      const uint32_t* r = rec + (uint64_t)p * kRecWords;
      const uint32_t m = r[0] & 0xFFu;
      sz = 1;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (!((m >> c) & 1u)) continue;
        const uint32_t a = indChild[r[1 + c]];
        if (a < (1u << 13)) sz += 1;
        else if (a < (1u << 30)) sz += 2;
        else atomicAdd(err, 1u);
      }
    }
    uint32_t agg;
    const uint32_t ex = block_exclusive_sum(sz, lds, agg) + carry;
    if (i < n) {
      offs[i] = ex;
      ind[p] = ex;
    }
    carry += agg;
  }
  if (threadIdx.x == 0) *total = carry;
}

// The layouts of every level from top[s] up to the root of set s in
// workgroup s, each level of at most kLayoutSmall nodes: inner_layout_small
// per level, a level's child addresses read after a barrier.
struct LayoutTop {
  const uint32_t* rec[2][kMaxLevels];
  const uint32_t* perm[2][kMaxLevels];
  uint32_t* offs[2][kMaxLevels];
  uint32_t* ind[2][kMaxLevels + 1];
  uint32_t* total[2][kMaxLevels];
  uint32_t n[2][kMaxLevels];
  int32_t top[2];
  uint32_t* err;
};
__global__ void __launch_bounds__(kLayoutThreads) k_inner_layouts_small(LayoutTop a) {
  __shared__ uint32_t lds[kLayoutThreads / 64 + 1];
  const int s = blockIdx.x;
  for (int lev = a.top[s]; lev >= 0; --lev) {
    inner_layout_small(a.rec[s][lev], a.perm[s][lev], a.n[s][lev], a.ind[s][lev + 1], a.offs[s][lev],
                       a.ind[s][lev], a.total[s][lev], a.err, lds);
    __threadfence_block();
    __syncthreads();
  }
}

// Inner levels 0..L-3 of the SSVDAG image in one launch.
struct InnerLevels {
  const uint32_t* rec[kMaxLevels];
  const uint32_t* indChild[kMaxLevels];   // addresses of the next level's nodes
  const uint32_t* offs[kMaxLevels];       // u16 offset of each sorted node in its level
  uint32_t outOff[kMaxLevels];            // levelOffsets
  uint32_t start[kMaxLevels + 1];         // sorted positions, concatenated (= permAll layout)
  uint32_t L;
};

// Variable-length inner node (encoded_ssvdag.cpp:352-430): sorted position g.
__device__ __forceinline__ void inner_write(const InnerLevels& a, const uint32_t* __restrict__ permAll, uint64_t g,
                                            uint16_t* __restrict__ inner) {
  uint32_t l = 0;
  while (l + 1 < a.L && g >= a.start[l + 1]) ++l;
  const uint64_t i = g - a.start[l];
  const uint32_t* __restrict__ indChild = a.indChild[l];
  uint16_t* out = inner + a.outOff[l];
  const uint32_t* r = a.rec[l] + (uint64_t)permAll[g] * kRecWords;
  const uint32_t w0 = r[0];
  const uint32_t m = w0 & 0xFFu, mx = (w0 >> 8) & 0xFFu, my = (w0 >> 16) & 0xFFu, mz = (w0 >> 24) & 0xFFu;
  uint16_t* o = out + a.offs[l][i];
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
// children masks (each mirrored by the parent's mirror bits): brick i.
__device__ __forceinline__ void brick_write(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ perm,
                                            uint64_t i, const uint32_t* __restrict__ leafMask,
                                            uint16_t* __restrict__ out) {
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

// The write phase of one or two SSVDAG-layout images in one launch: the
// blocks before end[0] write set 0's inner nodes, then set 0's bricks (to
// end[1]), set 1's inner nodes (end[2]) and set 1's bricks (end[3]).
struct SetWrite {
  InnerLevels ia;
  const uint32_t* permAll;        // the set's sorted inner positions
  uint16_t* inner;
  const uint32_t* brickRec;
  const uint32_t* brickPerm;
  const uint32_t* leafMask;
  uint16_t* bricks;
  uint64_t nInner, nBricks;
};
struct SetWrites {
  SetWrite s[2];
  uint32_t end[4];
};
__device__ __forceinline__ void set_write(const SetWrite& w, bool bricks, uint64_t g) {
  if (bricks) {
    if (g < w.nBricks) brick_write(w.brickRec, w.brickPerm, g, w.leafMask, w.bricks);
  } else if (g < w.nInner) {
    inner_write(w.ia, w.permAll, g, w.inner);
  }
}
__global__ void k_ssvdag_write(SetWrites w) {
  const uint32_t b = blockIdx.x;
  const uint32_t part = b < w.end[0] ? 0u : b < w.end[1] ? 1u : b < w.end[2] ? 2u : 3u;
  const uint32_t b0 = part ? w.end[part - 1] : 0u;
  const uint64_t g = (uint64_t)(b - b0) * blockDim.x + threadIdx.x;
  if (part < 2) set_write(w.s[0], part == 1, g);
  else set_write(w.s[1], part == 3, g);
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

LevelSet level_set(const std::vector<Level>& lv, uint32_t L, bool withChildLevels) {
  if (L > kMaxLevels) throw Error{SVD_ERR_LEVELS, "too many levels"};
  LevelSet a{};
  a.L = L;
  uint64_t at = 0;
  for (uint32_t l = 0; l < L; ++l) {
    a.start[l] = (uint32_t)at;
    a.rec[l] = lv[l].rec;
    a.aux[l] = withChildLevels ? lv[l].childLevels : nullptr;
    if (lv[l].leaf) a.leafBits |= 1u << l;
    at += lv[l].n;
  }
  if (at >= 0xFFFFFFFFull) throw Error{SVD_ERR_CAPACITY, "too many nodes for 32-bit word pointers"};
  a.start[L] = (uint32_t)at;
  return a;
}

// Header pieces collected on the host and written by flush() in one launch
// (or earlier, when the batch is full).
struct Puts {
  PutBatch b{};
  uint32_t count = 0;
  void add(svd_ctx* c, uint8_t* dst, const std::vector<uint8_t>& v) {
    if (v.empty()) return;
    if (v.size() > sizeof(b.b[0])) throw Error{SVD_ERR_DEVICE, "header too long"};
    if (count == kPutMax) flush(c);
    b.dst[count] = dst;
    b.n[count] = (uint32_t)v.size();
    memcpy(b.b[count], v.data(), v.size());
    ++count;
  }
  void flush(svd_ctx* c) {
    if (!count) return;
    k_put_batch<<<count, 128, 0, c->stream>>>(b);
    check_launch("k_put_batch");
    count = 0;
  }
};

// .svdag (of the DAG, or its cross-level view) and .ussvdag (of the SDAG):
// each format's passes are enqueued, and the word counts of all are read back
// once at the end (or, when SSVDAG-layout images follow, together with their
// layout sizes).
struct WordsPending {
  int K = 0;
  int fmts[2] = {0, 0};
  uint32_t* dev = nullptr;        // per format: word count, error flags
  uint32_t* ptrs[2] = {nullptr, nullptr};
  uint64_t hdrBytes[2] = {0, 0};
  Puts puts;                      // the images' header bytes, written with the next flush
};

void encode_words_enqueue(svd_octree* o, const int* fmts, int K, WordsPending& wp) {
  svd_ctx* c = o->ctx;
  const uint32_t L = o->levels;
  uint32_t* dev = dalloc<uint32_t>(c, 2 * K);
  fill_u32(c, dev, 0u, 2 * K);
  wp.K = K;
  wp.dev = dev;
  uint32_t** ptrsOf = wp.ptrs;
  uint64_t* hdrBytesOf = wp.hdrBytes;
  for (int s = 0; s < K; ++s) {
    wp.fmts[s] = fmts[s];
    const int fmt = fmts[s];
    const bool ussv = fmt == SVD_FMT_USSVDAG;
    const std::vector<Level>& lv = ussv ? o->sdag : (o->hasX ? o->xdag : o->dag);
    const LevelSet a = level_set(lv, L, !ussv);
    const uint64_t N = a.start[L];
    uint32_t* ptrs = dalloc<uint32_t>(c, N + 1);
    ptrsOf[s] = ptrs;
    uint32_t* d = dev + 2 * s;
    if (N) k_word_counts<<<div_up(N, kB), kB, 0, c->stream>>>(a, N, ptrs);
    check_launch("k_word_counts");
    scan_exclusive(c, ptrs, ptrs, N, d);
    // The reference assigns _firstLeafPtr after the loop over ALL levels
    // (encoded_svdag.cpp:129-137), so the stored value is the total word
    // count.  Every node takes at most 9 words: the image is allocated for
    // that bound, the header's two count words are written on the device and
    // the count is read back once, with the error flags, at the end.
    uint64_t maxWords = 0;
    for (uint32_t l = 0; l < L; ++l) maxWords += lv[l].leaf ? lv[l].n : lv[l].n * kRecWords;
    const uint64_t nNodes = ussv ? o->nNodesSDAG : (o->hasX ? o->nNodesX : o->nNodesDAG);
    std::vector<uint8_t> hdr = base_header(o, nNodes);
    const uint64_t hdrBytes = hdr.size() + 8;   // + firstLeafPtr + count
    hdrBytesOf[s] = hdrBytes;
    Encoded& e = o->enc[fmt];
    dfree(c, e.d);
    e.d = dalloc<uint8_t>(c, hdrBytes + 4 * maxWords);
    e.valid = false;
    if (!N) hdr.resize(hdr.size() + 8, 0);   // no words: both count words are 0
    wp.puts.add(c, e.d, hdr);
    uint32_t* words = (uint32_t*)(e.d + hdrBytes);
    if (N)
      k_write_words<<<div_up(N, kB), kB, 0, c->stream>>>(a, ussv ? 1 : 0, ptrs, N, words, d + 1, d,
                                                         (uint32_t*)(e.d + hdrBytes - 8));
    check_launch("k_write_words");
  }
}

// h: the 2K words of wp.dev, read back.
void encode_words_finish(svd_octree* o, WordsPending& wp, const uint32_t* h) {
  svd_ctx* c = o->ctx;
  for (int s = 0; s < wp.K; ++s) dfree(c, wp.ptrs[s]);
  dfree(c, wp.dev);
  wp.dev = nullptr;
  for (int s = 0; s < wp.K; ++s)
    if (h[2 * s + 1]) throw Error{SVD_ERR_POINTER, "encode: child pointer without a target node"};
  for (int s = 0; s < wp.K; ++s) {
    Encoded& e = o->enc[wp.fmts[s]];
    e.bytes = wp.hdrBytes[s] + 4ull * h[2 * s];
    e.valid = true;
  }
}

void encode_words_sets(svd_octree* o, const int* fmts, int K) {
  WordsPending wp;
  encode_words_enqueue(o, fmts, K, wp);
  wp.puts.flush(o->ctx);
  uint32_t h[4];
  read_u32s(o->ctx, wp.dev, h, 2 * K);
  encode_words_finish(o, wp, h);
}

// One or two SSVDAG-layout images (.ssvdag of the SDAG, .esvdag of the DAG)
// with every pass shared: one launch clears and numbers all their levels, one
// counts the references of all, one segmented sort orders every level of
// both (each level is its own std::sort call), one host read returns the
// layout sizes, and the writes follow.
// checked: the refcount order reads its list size back until it is done
// (otherwise it reports an unfinished order with the layout sizes, and the
// whole encoding runs again checked).
void encode_ssvdag_sets(svd_octree* o, const int* fmts, int K, WordsPending* wp = nullptr, bool checked = false) {
  if (K < 1 || K > 2) throw Error{SVD_ERR_ARG, "SSVDAG-layout encodings: one or two sets"};
  svd_ctx* c = o->ctx;
  const int L = (int)o->levels;
  if (L < 3) throw Error{SVD_ERR_LEVELS, "SSVDAG encoding needs at least 3 levels"};
  const int P = L - 1;                       // levels 0..L-2 of each set in the span
  std::vector<const std::vector<Level>*> lvs(K);
  for (int s = 0; s < K; ++s) {
    lvs[s] = fmts[s] == SVD_FMT_SSVDAG ? &o->sdag : &o->dag;
    for (int l = 0; l < L; ++l)
      if ((*lvs[s])[l].n > (1ull << 30)) throw Error{SVD_ERR_CAPACITY, "level too big for encoding in 30 bits"};
  }
  // perm / ind / offs of set s, level lev at [s * L + lev]
  std::vector<uint32_t*> perm(K * L, nullptr), ind(K * L, nullptr), offs(K * L, nullptr);
  // dev[s * L + lev]: inner-size totals per level, dev[K * L]: pointer-size
  // error flag, dev[K * L + 1]: unfinished refcount order
  uint32_t* dev = dalloc<uint32_t>(c, K * L + 2);
  uint32_t* err = dev + K * L;   // dev is cleared by k_level_iota below
  // Phase 1a: the reference counts of every level land in one array, and one
  // segmented sort orders all levels of all sets.
  Span ua{};
  ua.L = (uint32_t)(K * P);
  ua.P = (uint32_t)P;
  std::vector<uint64_t> off(K * L, 0), rfirst, rlast;
  uint64_t tot = 0;
  for (int s = 0; s < K; ++s)
    for (int lev = 0; lev <= L - 2; ++lev) {
      const Level& v = (*lvs[s])[lev];
      off[s * L + lev] = tot;
      ua.rec[s * P + lev] = v.rec;
      ua.start[s * P + lev] = (uint32_t)tot;
      if (lev > 0 && v.n > 0) {
        rfirst.push_back(tot);
        rlast.push_back(tot + v.n);
      }
      tot += v.n;
    }
  if (tot >= 0xFFFFFFFFull) throw Error{SVD_ERR_CAPACITY, "too many nodes for one refcount order"};
  ua.start[K * P] = (uint32_t)tot;
  uint32_t* permAll = dalloc<uint32_t>(c, tot + 1);
  uint32_t* hist = dalloc<uint32_t>(c, tot + 1);   // cleared by k_level_iota
  for (int s = 0; s < K; ++s)
    for (int lev = 0; lev <= L - 2; ++lev) perm[s * L + lev] = permAll + off[s * L + lev];
  k_level_iota<<<div_up(std::max<uint64_t>(tot, (uint64_t)(K * L + 2)), kB), kB, 0, c->stream>>>(
      ua, tot, permAll, hist, dev, (uint32_t)(K * L + 2));
  check_launch("k_level_iota");
  RcJobs jb{};
  for (int s = 0; s < K; ++s)
    for (int lev = 0; lev <= L - 3; ++lev) {
      const uint64_t np = (*lvs[s])[lev].n, nc = (*lvs[s])[lev + 1].n;
      if (!np) continue;
      const uint32_t nr = (uint32_t)std::max<uint64_t>(1, div_up(nc, kRcRange));
      const uint32_t nch = (uint32_t)std::max<uint64_t>(1, div_up(np, kRcParents));
      jb.lev[jb.n] = (uint32_t)(s * P + lev);
      jb.nr[jb.n] = nr;
      jb.blk0[jb.n + 1] = jb.blk0[jb.n] + nr * nch;
      ++jb.n;
    }
  if (jb.n) k_refcount<<<jb.blk0[jb.n], kRcT, 0, c->stream>>>(ua, jb, hist, err);
  check_launch("k_refcount");
  refsort_ranges(c, permAll, hist, tot, rfirst.data(), rlast.data(), rfirst.size(),
                 checked ? nullptr : dev + K * L + 1);
  dfree(c, hist);
  // Phase 1b, bottom-up: children addresses of each level in its order.  A
  // set's levels from the first whose level and all above are small go to
  // one workgroup of one launch at the end.
  LayoutTop lt{};
  lt.err = err;
  lt.top[0] = lt.top[1] = -1;
  for (int lev = L - 2; lev >= 0; --lev) {
    // the level's passes of both sets go out together: one launch each
    int nj = 0;
    LayoutOp jop[2];
    uint64_t jn[2] = {0, 0};
    uint32_t* jtot[2] = {nullptr, nullptr};
    for (int s = 0; s < K; ++s) {
      if (lt.top[s] >= 0) continue;   // handled by the launch at the end
      const Level& v = (*lvs[s])[lev];
      const uint64_t n = v.n;
      const int q = s * L + lev;
      ind[q] = dalloc<uint32_t>(c, n + 1);
      if (lev == L - 2) {
        jn[nj] = n;
        jop[nj].perm = perm[q];
        jop[nj].ind = ind[q];
        ++nj;
        continue;
      }
      offs[q] = dalloc<uint32_t>(c, n + 1);
      uint64_t maxTop = 0;
      for (int l = lev; l >= 0; --l) maxTop = std::max<uint64_t>(maxTop, (*lvs[s])[l].n);
      if (maxTop <= small_limit(c->smallLayout, kLayoutSmall)) {
        lt.top[s] = lev;
        for (int l = lev; l >= 0; --l) {
          const int ql = s * L + l;
          if (l < lev) {
            ind[ql] = dalloc<uint32_t>(c, (*lvs[s])[l].n + 1);
            offs[ql] = dalloc<uint32_t>(c, (*lvs[s])[l].n + 1);
          }
          lt.rec[s][l] = (*lvs[s])[l].rec;
          lt.perm[s][l] = perm[ql];
          lt.offs[s][l] = offs[ql];
          lt.ind[s][l] = ind[ql];
          lt.total[s][l] = dev + ql;
          lt.n[s][l] = (uint32_t)(*lvs[s])[l].n;
        }
        lt.ind[s][lev + 1] = ind[q + 1];
        continue;
      }
      jop[nj] = LayoutOp{v.rec, perm[q], ind[q + 1], offs[q], ind[q], err};
      jn[nj] = n;
      jtot[nj] = dev + q;
      ++nj;
    }
    if (lev == L - 2) {
      const uint64_t nt = jn[0] + (nj > 1 ? jn[1] : 0);
      if (nt) k_scatter_rank2<<<div_up(nt, kB), kB, 0, c->stream>>>(jop[0].perm, jn[0], jop[0].ind,
                                                                      nj > 1 ? jop[1].perm : nullptr,
                                                                      nj > 1 ? jn[1] : 0,
                                                                      nj > 1 ? jop[1].ind : nullptr);
      check_launch("k_scatter_rank2");
    } else if (nj == 2) {
      scan_op2<1>(c, jn[0], jop[0], jtot[0], jn[1], jop[1], jtot[1]);
    } else if (nj == 1) {
      scan_op<1>(c, jn[0], jop[0], jtot[0]);
    }
  }
  {
    // sets whose every level was big (never here for a DAG: the root level has one node)
    int order[2], nb = 0;
    for (int s = 0; s < K; ++s)
      if (lt.top[s] >= 0) order[nb++] = s;
    if (nb) {
      LayoutTop ltc = lt;
      for (int k = 0; k < nb; ++k) {
        const int s = order[k];
        for (int l = 0; l < kMaxLevels; ++l) {
          ltc.rec[k][l] = lt.rec[s][l]; ltc.perm[k][l] = lt.perm[s][l]; ltc.offs[k][l] = lt.offs[s][l];
          ltc.total[k][l] = lt.total[s][l]; ltc.n[k][l] = lt.n[s][l];
        }
        for (int l = 0; l <= kMaxLevels; ++l) ltc.ind[k][l] = lt.ind[s][l];
        ltc.top[k] = lt.top[s];
      }
      k_inner_layouts_small<<<nb, kLayoutThreads, 0, c->stream>>>(ltc);
      check_launch("k_inner_layouts_small");
    }
  }
  std::vector<uint32_t> h(K * L + 2);
  if (wp) {
    // the pending word counts ride on the same host round trip
    uint32_t hw[4];
    read_u32s_2(c, dev, h.data(), K * L + 2, wp->dev, hw, 2 * wp->K);
    wp->puts.flush(c);
    encode_words_finish(o, *wp, hw);
  } else {
    read_u32s(c, dev, h.data(), K * L + 2);
  }
  Puts puts;
  if (h[K * L + 1]) {   // the refcount order needed more rounds: again, checked
    for (auto* p : ind) dfree(c, p);
    for (auto* p : offs) dfree(c, p);
    dfree(c, permAll); dfree(c, dev);
    encode_ssvdag_sets(o, fmts, K, nullptr, true);
    return;
  }
  if (h[K * L]) throw Error{SVD_ERR_CAPACITY, "unsupported pointer size (>= 2^30)"};
  SetWrites sw{};
  uint32_t blocks = 0;
  for (int s = 0; s < K; ++s) {
    const std::vector<Level>& lv = *lvs[s];
    const int fmt = fmts[s];
    // Layout of the file image.
    std::vector<uint32_t> levelOffsets(L - 2, 0);
    for (int i = 1; i < L - 2; ++i) levelOffsets[i] = h[s * L + i - 1] + levelOffsets[i - 1];
    uint64_t nInner = 0;
    for (int i = 0; i < L - 2; ++i) nInner += h[s * L + i];
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
    e.d = nullptr;
    e.valid = false;
    e.bytes = tailAt + tail.size();
    e.d = dalloc<uint8_t>(c, e.bytes);
    puts.add(c, e.d, head);
    puts.add(c, e.d + leafCntAt, mid);
    puts.add(c, e.d + tailAt, tail);
    // Phase 2: every level of every set written in one launch.
    SetWrite& w = sw.s[s];
    const uint64_t base = off[s * L];
    const uint64_t nInnerNodes = off[s * L + L - 2] - base;
    w.ia.L = (uint32_t)(L - 2);
    for (int lev = 0; lev <= L - 3; ++lev) {
      w.ia.rec[lev] = lv[lev].rec;
      w.ia.indChild[lev] = ind[s * L + lev + 1];
      w.ia.offs[lev] = offs[s * L + lev];
      w.ia.outOff[lev] = levelOffsets[lev];
      w.ia.start[lev] = (uint32_t)(off[s * L + lev] - base);
    }
    w.ia.start[L - 2] = (uint32_t)nInnerNodes;
    w.permAll = permAll + base;
    w.inner = (uint16_t*)(e.d + innerAt);
    w.nInner = nInnerNodes;
    w.brickRec = lv[L - 2].rec;
    w.brickPerm = perm[s * L + L - 2];
    w.leafMask = lv[L - 1].rec;
    w.bricks = (uint16_t*)(e.d + leafAt);
    w.nBricks = nBricks;
    blocks += (uint32_t)div_up(nInnerNodes, kB);
    sw.end[2 * s] = blocks;
    blocks += (uint32_t)div_up(nBricks, kB);
    sw.end[2 * s + 1] = blocks;
  }
  for (int p = 2 * K; p < 4; ++p) sw.end[p] = blocks;
  if (blocks) k_ssvdag_write<<<blocks, kB, 0, c->stream>>>(sw);
  check_launch("k_ssvdag_write");
  puts.flush(c);
  for (int s = 0; s < K; ++s) o->enc[fmts[s]].valid = true;
  for (auto* p : ind) dfree(c, p);
  for (auto* p : offs) dfree(c, p);
  dfree(c, permAll); dfree(c, dev);
}

}  // namespace

namespace {
void check_state(const svd_octree* o, int fmt) {
  switch (fmt) {
    case SVD_FMT_SVDAG:
      if (o->state < SVD_S_DAG) throw Error{SVD_ERR_STATE, "SVDAG encoding needs a DAG"};
      break;
    case SVD_FMT_USSVDAG:
      if (o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "USSVDAG encoding needs an SDAG"};
      break;
    case SVD_FMT_SSVDAG:
      if (o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "SSVDAG encoding needs an SDAG"};
      break;
    case SVD_FMT_ESVDAG:
      if (o->state < SVD_S_DAG) throw Error{SVD_ERR_STATE, "ESVDAG encoding needs a DAG"};
      break;
    default:
      throw Error{SVD_ERR_ARG, "unknown format"};
  }
}
}  // namespace

void encode(svd_octree* o, int fmt) { encode_formats(o, &fmt, 1); }

void encode_formats(svd_octree* o, const int* fmts, int n) {
  int words[2], ss[2], nw = 0, nss = 0;
  bool seen[4] = {false, false, false, false};
  for (int k = 0; k < n; ++k) {
    check_state(o, fmts[k]);
    if (seen[fmts[k]]) throw Error{SVD_ERR_ARG, "a format is listed twice"};
    seen[fmts[k]] = true;
    if (fmts[k] == SVD_FMT_SVDAG || fmts[k] == SVD_FMT_USSVDAG) words[nw++] = fmts[k];
    else ss[nss++] = fmts[k];
  }
  if (nw && nss) {
    WordsPending wp;
    encode_words_enqueue(o, words, nw, wp);
    encode_ssvdag_sets(o, ss, nss, &wp);
    return;
  }
  if (nw) encode_words_sets(o, words, nw);
  if (nss) encode_ssvdag_sets(o, ss, nss);
}

}  // namespace svd
