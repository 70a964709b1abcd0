// Exact emulation of the reference-count ordering of the SSVDAG encoder:
//   std::sort(hist.begin(), hist.end(),
//             [](pair a, pair b) { return a.second > b.second; });
// (encoded_ssvdag.cpp:273-275, libstdc++ introsort: median-of-3 pivot moved
// to the front, Hoare "unguarded partition", 16-element threshold, heapsort
// once the depth budget 2*floor(log2 n) is spent, final insertion sort).
//
// The comparator looks at the count only, so the permutation of equal counts
// depends on every swap the library makes.  One partition step is restated
// data-parallel: the k-th stop of the left scanner (value <= pivot) is swapped
// with the k-th stop of the right scanner (value >= pivot) for every k with
// L_k < R_k, and the cut is L_K if L_K < R_{K-1} else R_{K-1}.  Partitions of
// disjoint segments are independent, so all segments of one depth run
// together.  The final insertion sort never moves an element across a
// partition boundary, so it reduces to a stable insertion sort of each
// <=16-element leaf segment.
#include <algorithm>
#include "svd_internal.hpp"
#include "block_scan.hpp"

namespace svd {

namespace {

constexpr int kThreshold = 16;
constexpr int kLdsN = 2048;           // segments up to this size finish in one workgroup
constexpr int kBT = 256;              // threads per workgroup
constexpr int kPer = kLdsN / kBT;     // elements per thread in the workgroup phase
constexpr int kTile = 2048;           // elements per workgroup in the device-wide rounds
constexpr uint32_t kLeafTag = 0x8000u; // workgroup phase: segOf tags
constexpr uint32_t kEqTag = 0xC000u;
constexpr uint32_t kTagMask = 0xC000u;
constexpr uint32_t kDone = 0xFFFFu;

struct Seg { uint32_t first, last; int32_t depth; uint32_t pad; };

__device__ __forceinline__ bool comp(uint32_t a, uint32_t b) { return a > b; }

// ---- libstdc++ heap routines (bits/stl_heap.h) on a generic array view -----
template <typename Arr>
__device__ void adjust_heap(Arr& a, int64_t hole, int64_t len, uint32_t vv, uint32_t vi) {
  const int64_t top = hole;
  int64_t sc = hole;
  while (sc < (len - 1) / 2) {
    sc = 2 * (sc + 1);
    if (comp(a.v(sc), a.v(sc - 1))) sc--;
    a.set(hole, a.v(sc), a.i(sc));
    hole = sc;
  }
  if ((len & 1) == 0 && sc == (len - 2) / 2) {
    sc = 2 * (sc + 1);
    a.set(hole, a.v(sc - 1), a.i(sc - 1));
    hole = sc - 1;
  }
  int64_t parent = (hole - 1) / 2;
  while (hole > top && comp(a.v(parent), vv)) {
    a.set(hole, a.v(parent), a.i(parent));
    hole = parent;
    parent = (hole - 1) / 2;
  }
  a.set(hole, vv, vi);
}

// std::__partial_sort(first, last, last) == make_heap + sort_heap
template <typename Arr>
__device__ void heap_sort(Arr& a, int64_t len) {
  if (len >= 2) {
    int64_t parent = (len - 2) / 2;
    while (true) {
      uint32_t vv = a.v(parent), vi = a.i(parent);
      adjust_heap(a, parent, len, vv, vi);
      if (parent == 0) break;
      parent--;
    }
  }
  int64_t last = len;
  while (last > 1) {
    --last;
    uint32_t vv = a.v(last), vi = a.i(last);
    a.set(last, a.v(0), a.i(0));
    adjust_heap(a, 0, last, vv, vi);
  }
}

// Stable insertion sort (the effect of __final_insertion_sort on one leaf).
template <typename Arr>
__device__ void insertion_sort(Arr& a, int64_t len) {
  for (int64_t i = 1; i < len; ++i) {
    uint32_t vv = a.v(i), vi = a.i(i);
    int64_t j = i;
    while (j > 0 && comp(vv, a.v(j - 1))) { a.set(j, a.v(j - 1), a.i(j - 1)); --j; }
    a.set(j, vv, vi);
  }
}

// __move_median_to_first(first, first+1, mid, last-1): returns the position
// whose element is swapped to `first`.
template <typename V>
__device__ __forceinline__ int64_t median_pos(const V& val, int64_t first, int64_t last) {
  const int64_t mid = first + (last - first) / 2;
  const int64_t A = first + 1, B = mid, C = last - 1;
  const uint32_t a = val(A), b = val(B), c = val(C);
  if (comp(a, b)) {
    if (comp(b, c)) return B;
    if (comp(a, c)) return C;
    return A;
  }
  if (comp(a, c)) return A;
  if (comp(b, c)) return C;
  return B;
}

struct GArr {   // global-memory view of one segment
  uint32_t* vp; uint32_t* ip;
  __device__ uint32_t v(int64_t k) const { return vp[k]; }
  __device__ uint32_t i(int64_t k) const { return ip[k]; }
  __device__ void set(int64_t k, uint32_t vv, uint32_t ii) { vp[k] = vv; ip[k] = ii; }
};
struct LArr {   // LDS view of one segment
  uint32_t* vp; uint32_t* ip;
  __device__ uint32_t v(int64_t k) const { return vp[k]; }
  __device__ uint32_t i(int64_t k) const { return ip[k]; }
  __device__ void set(int64_t k, uint32_t vv, uint32_t ii) { vp[k] = vv; ip[k] = ii; }
};

struct Lists {
  Seg* big; uint32_t* nbig;
  Seg* small; uint32_t* nsmall;
  Seg* heap; uint32_t* nheap;
  Seg* leaf; uint32_t* nleaf;
};

// ---------------------------------------------------------------- device-wide rounds
struct SegState { uint32_t pivot, K, nL, nR, blk0, nblk, cut, pad; };

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t atomic_read(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The cut of a partitioned segment: L_K if L_K < R_{K-1}, else R_{K-1}
// (R_{-1} = last).
__device__ __forceinline__ uint32_t seg_cut(const Seg& g, const SegState& z, const uint32_t* __restrict__ Lpos,
                                            const uint32_t* __restrict__ Rpos) {
  const uint32_t rprev = z.K ? Rpos[g.first + z.K - 1] : g.last;
  return (z.K < z.nL && Lpos[g.first + z.K] < rprev) ? Lpos[g.first + z.K] : rprev;
}

// One workgroup between two device-wide rounds.  Phase A splits the segments
// of the round just finished (prev, when there is one) into the four lists;
// the big children form the new list cur.  Phase B moves the median of three
// of every segment of cur to its front (its pivot) and lays the segments'
// tiles out one after another: blk0[s] = first tile of segment s, blk0[ns] =
// the tile count.
constexpr int kPlanT = 1024;
__global__ void __launch_bounds__(kPlanT) k_g_plan(const Seg* __restrict__ prev, const uint32_t* __restrict__ nprev,
                                                   const uint32_t* __restrict__ Lpos, const uint32_t* __restrict__ Rpos,
                                                   Lists L, uint32_t* val, uint32_t* idx, SegState* __restrict__ st,
                                                   uint32_t* __restrict__ blk0) {
  __shared__ uint32_t lds[kPlanT / 64 + 1];
  __shared__ uint32_t nBig;
  const uint32_t tid = threadIdx.x;
  if (prev) {
    if (tid == 0) nBig = 0;
    __syncthreads();
    const uint32_t np = *nprev;
    for (uint32_t s = tid; s < np; s += kPlanT) {
      const Seg g = prev[s];
      const uint32_t cut = seg_cut(g, st[s], Lpos, Rpos);
      const uint32_t ff[2] = {cut, g.first}, ll[2] = {g.last, cut};
      for (int h = 0; h < 2; ++h) {
        const uint32_t size = ll[h] - ff[h];
        const Seg c{ff[h], ll[h], g.depth - 1, 0};
        if (size <= (uint32_t)kThreshold) { if (size > 1) L.leaf[atomicAdd(L.nleaf, 1u)] = c; }
        else if (size <= (uint32_t)kLdsN) L.small[atomicAdd(L.nsmall, 1u)] = c;
        else if (c.depth == 0) L.heap[atomicAdd(L.nheap, 1u)] = c;
        else L.big[atomicAdd(&nBig, 1u)] = c;
      }
    }
    __syncthreads();
    if (tid == 0) *L.nbig = nBig;
  }
  const uint32_t ns = prev ? nBig : atomic_read(L.nbig);
  uint32_t carry = 0;
  for (uint32_t s0 = 0; s0 < ns; s0 += kPlanT) {
    const uint32_t s = s0 + tid;
    uint32_t nb = 0;
    if (s < ns) {
      const Seg g = L.big[s];
      auto vf = [&](int64_t k) { return val[k]; };
      const int64_t m = median_pos(vf, g.first, g.last);
      const uint32_t tv = val[g.first], ti = idx[g.first];
      val[g.first] = val[m]; idx[g.first] = idx[m];
      val[m] = tv; idx[m] = ti;
      SegState z{};
      z.pivot = val[g.first];
      st[s] = z;
      nb = (g.last - g.first + kTile - 1) / kTile;
    }
    uint32_t agg;
    const uint32_t ex = block_exclusive_sum(nb, lds, agg);
    if (s < ns) blk0[s] = carry + ex;
    carry += agg;
  }
  if (tid == 0) blk0[ns] = carry;
}

__device__ __forceinline__ int find_seg(const uint32_t* __restrict__ blk0, uint32_t ns, uint32_t b) {
  // largest s with blk0[s] <= b (blk0 has ns+1 entries, blk0[ns] = total)
  if (b >= blk0[ns]) return -1;
  uint32_t lo = 0, hi = ns;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (blk0[mid] <= b) lo = mid; else hi = mid;
  }
  return (int)lo;
}

// Position k of a tile for thread t: wave w covers 64 * (kTile / kBT)
// consecutive elements, 64 at a time, so ballots give in-order ranks.
__device__ __forceinline__ uint32_t tile_pos(int t, int k) {
  return (uint32_t)((t >> 6) * (64 * (kTile / kBT)) + 64 * k + (t & 63));
}

// Per-tile counts of left stoppers (value <= pivot, past the front) and
// right stoppers (value >= pivot).
__global__ void __launch_bounds__(kBT) k_g_count(const Seg* __restrict__ segs, const uint32_t* __restrict__ nsp,
                                                 const uint32_t* __restrict__ blk0, const SegState* __restrict__ st,
                                                 const uint32_t* __restrict__ val, uint32_t* __restrict__ cntL,
                                                 uint32_t* __restrict__ cntR) {
  __shared__ uint32_t wl[kBT / 64], wr[kBT / 64];
  const int s = find_seg(blk0, *nsp, blockIdx.x);
  if (s < 0) return;
  const Seg g = segs[s];
  const uint32_t p = st[s].pivot;
  const uint64_t base = (uint64_t)g.first + (uint64_t)(blockIdx.x - blk0[s]) * kTile;
  uint32_t l = 0, r = 0;
#pragma unroll
  for (int k = 0; k < kTile / kBT; ++k) {
    const uint64_t pos = base + tile_pos(threadIdx.x, k);
    bool fl = false, fr = false;
    if (pos < g.last) {
      const uint32_t v = val[pos];
      fl = pos > g.first && v <= p;
      fr = v >= p;
    }
    l += (uint32_t)__popcll(__ballot(fl));
    r += (uint32_t)__popcll(__ballot(fr));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { wl[w] = l; wr[w] = r; }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tl = 0, tr = 0;
    for (int v = 0; v < kBT / 64; ++v) { tl += wl[v]; tr += wr[v]; }
    cntL[blockIdx.x] = tl;
    cntR[blockIdx.x] = tr;
  }
}

// Rank every stopper inside its segment; Lpos[first+k] / Rpos[first+k] = k-th
// stop.  A tile's offset is the sum of the counts of the tiles before it in
// its segment, and the segment's totals the sum of all of them (a segment has
// at most n / kTile tiles).
__global__ void __launch_bounds__(kBT) k_g_scatter(const Seg* __restrict__ segs, const uint32_t* __restrict__ nsp,
                                                   const uint32_t* __restrict__ blk0, SegState* __restrict__ st,
                                                   const uint32_t* __restrict__ val, const uint32_t* __restrict__ cntL,
                                                   const uint32_t* __restrict__ cntR, uint32_t* __restrict__ Lpos,
                                                   uint32_t* __restrict__ Rpos) {
  __shared__ uint32_t red[4][kBT / 64];
  __shared__ uint32_t wl[kBT / 64], wr[kBT / 64];
  const int s = find_seg(blk0, *nsp, blockIdx.x);
  if (s < 0) return;
  const Seg g = segs[s];
  const uint32_t p = st[s].pivot;
  const uint32_t b0 = blk0[s], b1 = blk0[s + 1];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  uint32_t q[4] = {0, 0, 0, 0};   // left before, left total, right before, right total
  for (uint32_t j = b0 + tid; j < b1; j += kBT) {
    const uint32_t cl = cntL[j], cr = cntR[j];
    q[1] += cl; q[3] += cr;
    if (j < blockIdx.x) { q[0] += cl; q[2] += cr; }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    q[i] = wave_sum(q[i]);
    if (lane == 0) red[i][w] = q[i];
  }
  const uint64_t base = (uint64_t)g.first + (uint64_t)(blockIdx.x - b0) * kTile;
  uint64_t bl[kTile / kBT], br[kTile / kBT];
  uint32_t cl = 0, cr = 0;
#pragma unroll
  for (int k = 0; k < kTile / kBT; ++k) {
    const uint64_t pos = base + tile_pos(tid, k);
    bool fl = false, fr = false;
    if (pos < g.last) {
      const uint32_t v = val[pos];
      fl = pos > g.first && v <= p;
      fr = v >= p;
    }
    bl[k] = __ballot(fl);
    br[k] = __ballot(fr);
  }
#pragma unroll
  for (int k = 0; k < kTile / kBT; ++k) { cl += (uint32_t)__popcll(bl[k]); cr += (uint32_t)__popcll(br[k]); }
  if (lane == 0) { wl[w] = cl; wr[w] = cr; }
  __syncthreads();
  uint32_t t[4] = {0, 0, 0, 0}, offL = 0, offR = 0;
  for (int v = 0; v < kBT / 64; ++v) {
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] += red[i][v];
    if (v < w) { offL += wl[v]; offR += wr[v]; }
  }
  const uint32_t nR = t[3];
  if (blockIdx.x == b0 && tid == 0) { st[s].nL = t[1]; st[s].nR = nR; }
  offL += t[0];
  offR += t[2];
#pragma unroll
  for (int k = 0; k < kTile / kBT; ++k) {
    const uint64_t pos = base + tile_pos(tid, k);
    const uint64_t below = (1ull << lane) - 1ull;
    if ((bl[k] >> lane) & 1ull) Lpos[g.first + offL + (uint32_t)__popcll(bl[k] & below)] = (uint32_t)pos;
    if ((br[k] >> lane) & 1ull) Rpos[g.first + (nR - 1 - (offR + (uint32_t)__popcll(br[k] & below)))] = (uint32_t)pos;
    offL += (uint32_t)__popcll(bl[k]);
    offR += (uint32_t)__popcll(br[k]);
  }
}

// Swap the k-th left stop with the k-th right stop while L_k < R_k; record K.
// kTile / kBT workgroups per tile, one pair per thread.
__global__ void __launch_bounds__(kBT) k_g_swap(const Seg* __restrict__ segs, const uint32_t* __restrict__ nsp,
                                                const uint32_t* __restrict__ blk0, SegState* __restrict__ st,
                                                uint32_t* val, uint32_t* idx, const uint32_t* __restrict__ Lpos,
                                                const uint32_t* __restrict__ Rpos) {
  constexpr uint32_t kSub = kTile / kBT;
  const uint32_t tile = blockIdx.x / kSub;
  const int s = find_seg(blk0, *nsp, tile);
  if (s < 0) return;
  const Seg g = segs[s];
  const uint32_t m = min(st[s].nL, st[s].nR);
  const uint64_t k = (uint64_t)(tile - blk0[s]) * kTile + (uint64_t)(blockIdx.x % kSub) * kBT + threadIdx.x;
  if (k >= m) return;
  const uint32_t l = Lpos[g.first + k], r = Rpos[g.first + k];
  if (l < r) {
    const uint32_t tv = val[l], ti = idx[l];
    val[l] = val[r]; idx[l] = idx[r];
    val[r] = tv; idx[r] = ti;
    const bool nextOk = (k + 1 < m) && (Lpos[g.first + k + 1] < Rpos[g.first + k + 1]);
    if (!nextOk) st[s].K = (uint32_t)(k + 1);
  }
}

// ---------------------------------------------------------------- per-thread finishers
// The finishers take their list length from the device (grid-stride), so
// the host launches them without reading it.
__global__ void k_leaf_sort(const Seg* __restrict__ segs, const uint32_t* __restrict__ ns, uint32_t* val, uint32_t* idx) {
  const uint32_t m = *ns;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < m; s += gridDim.x * blockDim.x) {
    const Seg g = segs[s];
    GArr a{val + g.first, idx + g.first};
    insertion_sort(a, (int64_t)(g.last - g.first));
  }
}

__global__ void k_heap_sort(const Seg* __restrict__ segs, const uint32_t* __restrict__ ns, uint32_t* val, uint32_t* idx) {
  const uint32_t m = *ns;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < m; s += gridDim.x * blockDim.x) {
    const Seg g = segs[s];
    GArr a{val + g.first, idx + g.first};
    heap_sort(a, (int64_t)(g.last - g.first));
    insertion_sort(a, (int64_t)(g.last - g.first));   // no-op on sorted input (kept for exactness)
  }
}

// ---------------------------------------------------------------- workgroup phase
// One workgroup runs the complete introsort recursion of one segment of
// 17..2048 elements in LDS, all subsegments of one depth together.
// Introsort of a segment of s equal values with depth budget d.  The median
// of three is the middle element B = f + s/2 (no comparison holds), so the
// front and B swap; every element stops both scanners, so the partition
// swaps f+1+k with l-1-k for k < K = floor((s-1)/2) and cuts at f+1+K; the
// halves recurse, and the final insertion sort moves no equal key.
// eq_depth_ok: the recursion reaches no heap sort (a segment of more than
// kThreshold elements at depth 0); the larger half has ceil(s/2) elements.
__device__ __forceinline__ bool eq_depth_ok(uint32_t s, int d) {
  while (s > (uint32_t)kThreshold) {
    if (d <= 0) return false;
    s = (s + 1) / 2;
    --d;
  }
  return true;
}
// Final place (0..s-1) of the element at place x of such a segment.
__device__ __forceinline__ uint32_t eq_dest(uint32_t x, uint32_t s) {
  uint32_t f = 0, l = s;
  while (l - f > (uint32_t)kThreshold) {
    const uint32_t sz = l - f, mid = f + sz / 2, K = (sz - 1) / 2;
    if (x == f) x = mid;
    else if (x == mid) x = f;
    if (x > f) {
      if (x - (f + 1) < K) x = l - 1 - (x - (f + 1));
      else if (l - 1 - x < K) x = f + 1 + (l - 1 - x);
    }
    const uint32_t cut = f + 1 + K;
    if (x < cut) l = cut; else f = cut;
  }
  return x;
}

// Element k of thread (wave w, lane j) is w * 64 * kPer + 64 * k + j: every
// LDS access of a per-element loop hits 64 consecutive words, and a wave's
// stopper prefix comes from ballots instead of a workgroup scan.
__device__ __forceinline__ uint32_t blk_elem(int tid, int k) {
  return (uint32_t)((tid >> 6) * (64 * kPer) + 64 * k + (tid & 63));
}

__global__ void __launch_bounds__(kBT) k_block_sort(const Seg* __restrict__ segs, const uint32_t* __restrict__ nsp,
                                                    uint32_t* val, uint32_t* idx) {
  constexpr int kWaves = kBT / 64;
  __shared__ uint32_t sv[kLdsN], si[kLdsN];
  __shared__ uint16_t segOf[kLdsN];
  __shared__ uint16_t exL[kLdsN + 1], exR[kLdsN + 1];
  __shared__ uint16_t Lp[kLdsN], Rp[kLdsN];
  __shared__ uint16_t aFirst[2][128], aLast[2][128];
  __shared__ int8_t aDepth[2][128];
  __shared__ uint32_t aPivot[128];
  __shared__ uint16_t aK[128], aCut[128], aLeft[128], aRight[128];
  __shared__ uint16_t lfFirst[kLdsN / 2], lfLast[kLdsN / 2];   // leaves of 2..16 elements
  __shared__ uint16_t eqFirst[128], eqLast[128];               // equal-value segments
  __shared__ uint32_t wTot[2][kWaves];
  __shared__ uint32_t nActive[2];
  __shared__ uint32_t nLeaf, nEq;
  // segOf: < kLeafTag active segment slot, kLeafTag | q leaf q, kEqTag | q
  // equal-value segment q, kDone finished

  const uint32_t nSeg = *nsp;
  for (uint32_t sIdx = blockIdx.x; sIdx < nSeg; sIdx += gridDim.x) {
  const Seg g = segs[sIdx];
  const uint32_t n = g.last - g.first;
  const int tid = threadIdx.x, wv = tid >> 6;
  const uint32_t v0 = val[g.first];
  bool differs = false;
  for (uint32_t k = tid; k < n; k += kBT) {
    sv[k] = val[g.first + k];
    si[k] = idx[g.first + k];
    segOf[k] = 0;
    differs |= sv[k] != v0;
  }
  if (!__syncthreads_or(differs) && eq_depth_ok(n, g.depth)) {
    // all values equal: only the indices move, each straight to its place
    for (uint32_t k = tid; k < n; k += kBT) idx[g.first + eq_dest(k, n)] = si[k];
    __syncthreads();   // the next segment reuses the LDS arrays
    continue;
  }
  if (tid == 0) {
    nActive[0] = 0;
    nActive[1] = 0;
    nLeaf = 0;
    nEq = 0;
    if (g.depth == 0) {
      LArr a{sv, si};
      heap_sort(a, n);
    } else {
      aFirst[0][0] = 0; aLast[0][0] = (uint16_t)n; aDepth[0][0] = (int8_t)g.depth;
      nActive[0] = 1;
    }
  }
  __syncthreads();
  // 1. median-of-3 to the front of every active segment, its pivot
  auto pivots = [&](int c) {
    if (tid < (int)nActive[c]) {
      const int64_t f = aFirst[c][tid], l = aLast[c][tid];
      auto vf = [&](int64_t k) { return sv[k]; };
      const int64_t m = median_pos(vf, f, l);
      const uint32_t tv = sv[f], ti = si[f];
      sv[f] = sv[m]; si[f] = si[m]; sv[m] = tv; si[m] = ti;
      aPivot[tid] = sv[f];
      aK[tid] = 0;
    }
  };
  pivots(0);
  __syncthreads();
  // Every phase below first issues all the LDS loads of its kPer elements and
  // only then uses them: one LDS round trip per dependent step instead of one
  // per element and step.
  int cur = 0;
  while (nActive[cur] > 0) {
    const uint32_t na = nActive[cur];
    const int nxt = cur ^ 1;
    // 2. stopper flags and their exclusive prefix over all kLdsN positions
    uint32_t sOf[kPer], v[kPer], fst[kPer], lst[kPer], piv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t e = blk_elem(tid, k);
      sOf[k] = e < n ? (uint32_t)segOf[e] : (uint32_t)kDone;
      v[k] = sv[e];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t q = sOf[k] < kLeafTag ? sOf[k] : 0u;
      piv[k] = aPivot[q];
      fst[k] = aFirst[cur][q];
      lst[k] = aLast[cur][q];
    }
    bool fl[kPer], fr[kPer];
    uint32_t pl[kPer], pr[kPer], cl = 0, cr = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t e = blk_elem(tid, k);
      const bool act = sOf[k] < kLeafTag;
      fl[k] = act && e > fst[k] && v[k] <= piv[k];
      fr[k] = act && v[k] >= piv[k];
      const uint64_t bl = __ballot(fl[k]), br = __ballot(fr[k]);
      pl[k] = cl + lanes_below(bl);
      pr[k] = cr + lanes_below(br);
      cl += (uint32_t)__popcll(bl);
      cr += (uint32_t)__popcll(br);
    }
    if ((tid & 63) == 0) { wTot[0][wv] = cl; wTot[1][wv] = cr; }
    __syncthreads();
    uint32_t offL = 0, offR = 0, totL = 0, totR = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t a = wTot[0][w], b = wTot[1][w];
      if (w < wv) { offL += a; offR += b; }
      totL += a; totR += b;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t e = blk_elem(tid, k);
      exL[e] = (uint16_t)(offL + pl[k]);
      exR[e] = (uint16_t)(offR + pr[k]);
    }
    if (tid == 0) { exL[kLdsN] = (uint16_t)totL; exR[kLdsN] = (uint16_t)totR; }
    __syncthreads();
    // 3. ranks -> Lp / Rp; the segment's stopper counts stay in registers
    uint32_t nLk[kPer], nRk[kPer];
    {
      uint32_t eLf[kPer], eLl[kPer], eRf[kPer], eRl[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        eLf[k] = exL[fst[k]]; eLl[k] = exL[lst[k]];
        eRf[k] = exR[fst[k]]; eRl[k] = exR[lst[k]];
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const uint32_t e = blk_elem(tid, k);
        nLk[k] = eLl[k] - eLf[k];
        nRk[k] = eRl[k] - eRf[k];
        // exL[e] - exL[f] = offL + pl[k] - exL[f]; exR[l] - 1 - exR[e] likewise
        if (fl[k]) Lp[fst[k] + (offL + pl[k] - eLf[k])] = (uint16_t)e;
        if (fr[k]) Rp[fst[k] + (eRl[k] - 1 - (offR + pr[k]))] = (uint16_t)e;
      }
    }
    __syncthreads();
    // 4. paired swaps, K
    {
      uint32_t la[kPer], rb[kPer], la1[kPer], rb1[kPer];
      bool sw[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const uint32_t e = blk_elem(tid, k);
        const uint32_t m = min(nLk[k], nRk[k]);
        sw[k] = sOf[k] < kLeafTag && e - fst[k] < m;
        const uint32_t q = sw[k] ? e : 0u;
        const uint32_t q1 = sw[k] && e - fst[k] + 1 < m ? e + 1 : 0u;
        la[k] = Lp[q]; rb[k] = Rp[q];
        la1[k] = Lp[q1]; rb1[k] = Rp[q1];
      }
      uint32_t va[kPer], ia[kPer], vb[kPer], ib[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const bool go = sw[k] && la[k] < rb[k];
        const uint32_t a = go ? la[k] : 0u, b = go ? rb[k] : 0u;
        va[k] = sv[a]; ia[k] = si[a]; vb[k] = sv[b]; ib[k] = si[b];
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        if (!(sw[k] && la[k] < rb[k])) continue;
        const uint32_t e = blk_elem(tid, k);
        const uint32_t kk = e - fst[k], m = min(nLk[k], nRk[k]);
        sv[la[k]] = vb[k]; si[la[k]] = ib[k]; sv[rb[k]] = va[k]; si[rb[k]] = ia[k];
        const bool nextOk = (kk + 1 < m) && (la1[k] < rb1[k]);
        if (!nextOk) aK[sOf[k]] = (uint16_t)(kk + 1);
      }
    }
    __syncthreads();
    // 5. cut, children: <=16 -> insertion sort now, depth 0 -> heap sort now, else next round
    if (tid < (int)na) {
      const uint32_t f = aFirst[cur][tid], l = aLast[cur][tid], K = aK[tid];
      const int d = aDepth[cur][tid] - 1;
      const uint32_t eLf = exL[f], eLl = exL[l], eRf = exR[f], eRl = exR[l];
      const uint32_t rK1 = Rp[K ? f + K - 1 : f], lK = Lp[f + K];
      const uint32_t nL = eLl - eLf, nR = eRl - eRf;
      // no value above the pivot: the left half is all pivots; none below:
      // the right half is
      const bool eqc[2] = {nL == l - f - 1, nR == l - f};
      const uint32_t rprev = K ? rK1 : l;
      const uint32_t cut = (K < nL && lK < rprev) ? lK : rprev;
      aCut[tid] = (uint16_t)cut;
      const uint32_t cf[2] = {f, cut}, cl2[2] = {cut, l};
      // the halves' kinds first, then their list slots (independent atomics)
      int kind[2];   // 0 nothing, 1 leaf, 2 equal values, 3 heap sort now, 4 next round
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t sz = cl2[h] - cf[h];
        kind[h] = sz <= (uint32_t)kThreshold ? (sz > 1 ? 1 : 0)
                  : (eqc[h] && eq_depth_ok(sz, d)) ? 2 : (d == 0 ? 3 : 4);
      }
      uint32_t slot[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
        slot[h] = kind[h] == 1 ? atomicAdd(&nLeaf, 1u)
                  : kind[h] == 2 ? atomicAdd(&nEq, 1u) : kind[h] == 4 ? atomicAdd(&nActive[nxt], 1u) : 0u;
      uint16_t ids[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        ids[h] = kDone;
        if (kind[h] == 1) {   // sorted after the rounds, all leaves at once
          lfFirst[slot[h]] = (uint16_t)cf[h];
          lfLast[slot[h]] = (uint16_t)cl2[h];
          ids[h] = (uint16_t)(kLeafTag | slot[h]);
        } else if (kind[h] == 2) {   // placed after the rounds
          eqFirst[slot[h]] = (uint16_t)cf[h];
          eqLast[slot[h]] = (uint16_t)cl2[h];
          ids[h] = (uint16_t)(kEqTag | slot[h]);
        } else if (kind[h] == 4) {
          aFirst[nxt][slot[h]] = (uint16_t)cf[h]; aLast[nxt][slot[h]] = (uint16_t)cl2[h];
          aDepth[nxt][slot[h]] = (int8_t)d;
          ids[h] = (uint16_t)slot[h];
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (kind[h] == 3) {
          LArr a{sv + cf[h], si + cf[h]};
          heap_sort(a, cl2[h] - cf[h]);
        }
      aLeft[tid] = ids[0]; aRight[tid] = ids[1];
    }
    __syncthreads();
    // 6. new segment ids, and step 1 of the next round (it moves values, not ids)
    {
      uint32_t ct[kPer], lf[kPer], rt[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const uint32_t q = sOf[k] < kLeafTag ? sOf[k] : 0u;
        ct[k] = aCut[q]; lf[k] = aLeft[q]; rt[k] = aRight[q];
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        if (sOf[k] >= kLeafTag) continue;
        const uint32_t e = blk_elem(tid, k);
        segOf[e] = (uint16_t)(e < ct[k] ? lf[k] : rt[k]);
      }
    }
    pivots(nxt);
    if (tid == 0) nActive[cur] = 0;   // the round after next appends here
    cur = nxt;
    __syncthreads();
  }
  // The final insertion sort only acts inside leaves and is stable: every
  // element's place is the number of larger elements of its leaf plus the
  // equal ones before it.
  // (loads first, as in the rounds: a leaf has at most kThreshold elements,
  // all of them read in one unrolled window)
  uint32_t rv[kPer], ri[kPer], sO[kPer], fq[kPer], lq[kPer];
  uint16_t rp[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t e = blk_elem(tid, k);
    sO[k] = e < n ? (uint32_t)segOf[e] : (uint32_t)kDone;
    rv[k] = sv[e];
    ri[k] = si[e];
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t s = sO[k], q = s & ~kTagMask;
    const bool tagged = s >= kLeafTag && s != kDone;
    const bool eq = (s & kTagMask) == kEqTag;
    fq[k] = tagged ? (eq ? eqFirst[q] : lfFirst[q]) : 0u;
    lq[k] = tagged ? (eq ? eqLast[q] : lfLast[q]) : 0u;
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const uint32_t e = blk_elem(tid, k);
    const uint32_t s = sO[k];
    rp[k] = kDone;
    if (s < kLeafTag || s == kDone) continue;
    const uint32_t f = fq[k], l = lq[k], v = rv[k];
    uint32_t r;
    if ((s & kTagMask) == kEqTag) {
      r = f + eq_dest(e - f, l - f);
    } else {
      r = f;
#pragma unroll
      for (uint32_t j = 0; j < (uint32_t)kThreshold; ++j) {
        const uint32_t pos = f + j;
        const uint32_t u = sv[pos < l ? pos : f];
        r += (pos < l && (comp(u, v) || (u == v && pos < e))) ? 1u : 0u;
      }
    }
    rp[k] = (uint16_t)r;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (rp[k] != kDone) { sv[rp[k]] = rv[k]; si[rp[k]] = ri[k]; }
  __syncthreads();
  for (uint32_t k = tid; k < n; k += kBT) { val[g.first + k] = sv[k]; idx[g.first + k] = si[k]; }
  __syncthreads();   // the next segment reuses the LDS arrays
  }
}

// The initial segment lists and their counters, written from the kernel
// arguments in one launch instead of one copy per list: roots of the big,
// small and leaf lists in that order, n = the four list sizes (heap empty).
constexpr int kSeedMax = 48;
struct Seeds {
  Seg seg[kSeedMax];
  uint32_t n[4];
};
__global__ void k_seed_lists(Seeds sd, Seg* __restrict__ big, Seg* __restrict__ small, Seg* __restrict__ leaf,
                             uint32_t* __restrict__ counters) {
  const uint32_t t = threadIdx.x;
  const uint32_t n0 = sd.n[0], n1 = sd.n[1], n3 = sd.n[3];
  if (t < n0) big[t] = sd.seg[t];
  else if (t < n0 + n1) small[t - n0] = sd.seg[t];
  else if (t < n0 + n1 + n3) leaf[t - n0 - n1] = sd.seg[t];
  if (t < 8) counters[t] = t < 4 ? sd.n[t] : 0u;
}

}  // namespace

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) { *dst = *src; }

void refsort_ranges(svd_ctx* c, uint32_t* idx, uint32_t* val, uint64_t n, const uint64_t* rfirst,
                    const uint64_t* rlast, size_t nr, uint32_t* unfinished) {
  if (n > 0xFFFFFFF0ull) throw Error{SVD_ERR_CAPACITY, "refsort: array too large"};
  // Each range is its own std::sort call: own depth budget 2*floor(log2 size).
  std::vector<Seg> roots[4];   // big, small, (heap), leaf
  std::vector<std::pair<uint64_t, uint64_t>> chk(nr);
  for (size_t r = 0; r < nr; ++r) chk[r] = {rfirst[r], rlast[r]};
  std::sort(chk.begin(), chk.end());
  for (size_t r = 0; r < nr; ++r)
    if (chk[r].first > chk[r].second || chk[r].second > n || (r && chk[r].first < chk[r - 1].second))
      throw Error{SVD_ERR_ARG, "refsort: ranges must be disjoint and inside the array"};
  for (size_t r = 0; r < nr; ++r) {
    const uint64_t size = rlast[r] - rfirst[r];
    if (size <= 1) continue;
    const int lg = 63 - __builtin_clzll(size);
    const Seg s{(uint32_t)rfirst[r], (uint32_t)rlast[r], 2 * lg, 0};
    roots[size <= (uint64_t)kThreshold ? 3 : size <= (uint64_t)kLdsN ? 1 : 0].push_back(s);
  }
  if (roots[0].empty() && roots[1].empty() && roots[3].empty()) return;
  const uint64_t capBig = n / kLdsN + nr + 2;
  const uint64_t capSeg = n + nr + 2;
  Lists L;
  Seg* bigA = dalloc<Seg>(c, capBig);
  Seg* bigB = dalloc<Seg>(c, capBig);
  L.small = dalloc<Seg>(c, capSeg);
  L.heap = dalloc<Seg>(c, capSeg);
  L.leaf = dalloc<Seg>(c, capSeg);
  uint32_t* counters = dalloc<uint32_t>(c, 8);
  L.nbig = counters + 0; L.nsmall = counters + 1; L.nheap = counters + 2; L.nleaf = counters + 3;
  uint32_t hcnt[8] = {(uint32_t)roots[0].size(), (uint32_t)roots[1].size(), 0, (uint32_t)roots[3].size(), 0, 0, 0, 0};
  if (roots[0].size() + roots[1].size() + roots[3].size() <= (size_t)kSeedMax) {
    Seeds sd{};
    size_t k = 0;
    for (int l : {0, 1, 3})
      for (const Seg& g : roots[l]) sd.seg[k++] = g;
    for (int l = 0; l < 4; ++l) sd.n[l] = hcnt[l];
    k_seed_lists<<<1, 64, 0, c->stream>>>(sd, bigA, L.small, L.leaf, counters);
    check_launch("k_seed_lists");
  } else {
    Seg* dstOf[4] = {bigA, L.small, L.heap, L.leaf};
    for (int k = 0; k < 4; ++k)
      if (!roots[k].empty())
        SVD_HIP(hipMemcpyAsync(dstOf[k], roots[k].data(), sizeof(Seg) * roots[k].size(), hipMemcpyHostToDevice,
                               c->stream));
    SVD_HIP(hipMemcpyAsync(counters, hcnt, sizeof(hcnt), hipMemcpyHostToDevice, c->stream));
  }
  if (hcnt[0]) {
    uint32_t* Lpos = dalloc<uint32_t>(c, n);
    uint32_t* Rpos = dalloc<uint32_t>(c, n);
    SegState* st = dalloc<SegState>(c, capBig);
    uint32_t* blk0 = dalloc<uint32_t>(c, capBig + 1);
    // tiles of one round: sum of ceil(size / kTile) <= n / kTile + live segments
    const uint32_t G = (uint32_t)(n / kTile + capBig + 1);
    uint32_t* cntL = dalloc<uint32_t>(c, G);
    uint32_t* cntR = dalloc<uint32_t>(c, G);
    Seg* cur = bigA;
    Seg* nxt = bigB;
    uint32_t* ncur = counters + 0;
    uint32_t* nnxt = counters + 4;
    // A round is plan (split the last round's segments, pick the pivots of
    // the new ones), count, scatter and swap.  Every round leaves a piece of
    // at least half the largest segment, so the first ceil(log2(max /
    // kLdsN)) rounds are all needed.  Refcount arrays hold long runs of equal
    // values, which split unevenly: the lossless builds measured need up to
    // four rounds more, so the first batch has those too (the context's
    // refsortExtra; a round over an empty list changes nothing and costs four
    // short launches, a look at the list size a host round trip); after it
    // the size is checked every second round.
    uint64_t maxRoot = 0;
    for (const Seg& r : roots[0]) maxRoot = std::max<uint64_t>(maxRoot, r.last - r.first);
    int batch = c->refsortExtra;
    while (((uint64_t)kLdsN << (batch - c->refsortExtra)) < maxRoot) ++batch;
    L.big = cur;
    L.nbig = ncur;
    k_g_plan<<<1, kPlanT, 0, c->stream>>>(nullptr, nullptr, Lpos, Rpos, L, val, idx, st, blk0);
    while (true) {
      for (int r = 0; r < batch; ++r) {
        k_g_count<<<G, kBT, 0, c->stream>>>(cur, ncur, blk0, st, val, cntL, cntR);
        k_g_scatter<<<G, kBT, 0, c->stream>>>(cur, ncur, blk0, st, val, cntL, cntR, Lpos, Rpos);
        k_g_swap<<<G * (kTile / kBT), kBT, 0, c->stream>>>(cur, ncur, blk0, st, val, idx, Lpos, Rpos);
        L.big = nxt;
        L.nbig = nnxt;
        k_g_plan<<<1, kPlanT, 0, c->stream>>>(cur, ncur, Lpos, Rpos, L, val, idx, st, blk0);
        check_launch("refsort round");
        std::swap(cur, nxt);
        std::swap(ncur, nnxt);
      }
      if (unfinished) {   // the caller reads the list size later, with its own read-back
        k_copy_u32<<<1, 1, 0, c->stream>>>(ncur, unfinished);
        check_launch("k_copy_u32");
        break;
      }
      if (!read_u32(c, ncur)) break;
      batch = 2;
    }
    dfree(c, Lpos); dfree(c, Rpos); dfree(c, st); dfree(c, blk0); dfree(c, cntL); dfree(c, cntR);
  }
  // finishers, sized without reading the list lengths (grid-stride)
  k_block_sort<<<(uint32_t)std::max(c->numCU, 1) * 3, kBT, 0, c->stream>>>(L.small, L.nsmall, val, idx);
  k_heap_sort<<<64, 64, 0, c->stream>>>(L.heap, L.nheap, val, idx);
  k_leaf_sort<<<512, 64, 0, c->stream>>>(L.leaf, L.nleaf, val, idx);
  check_launch("refsort finish");
  dfree(c, bigA); dfree(c, bigB); dfree(c, L.small); dfree(c, L.heap); dfree(c, L.leaf); dfree(c, counters);
}

void refsort(svd_ctx* c, uint32_t* idx, uint32_t* val, uint64_t n) {
  const uint64_t f = 0;
  refsort_ranges(c, idx, val, n, &f, &n, 1);
}

}  // namespace svd
