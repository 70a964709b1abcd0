// Lossy DAG (LSVDAG) on the device: GeomOctree::toLossyDag
// (geom_octree_extension.cpp:797-1179) with a deterministic clustering rule in
// place of the reference's external MCL run (cluster.hpp:134-192).
//
// Per merged level (lev_hi down to lev_lo, bottom-up as in :882):
//   1. signatures: 512-bit occupancy of each node's cells three levels down
//      (bit c1*64 + c2*8 + c3 for the child-slot path; voxels at level L-3);
//   2. lists: for each candidate j (effective reference count <= max_refs,
//      :945) the first list_cap earlier nodes i < j with
//      popcount(sig_i ^ sig_j) <= bound -- exhaustively (an LDS-tiled
//      all-pairs XOR-popcount scan), or the list_cap nearest of them by
//      (popcount, i) (ranked, the same scan to the end), or among the learned
//      proposals (kNN of the MLP embeddings, sorted and de-duplicated);
//   3. verify: exact voxel difference of the two subtrees by a joint descent
//      with early abort (the Hamming distance above is a lower bound of it);
//   4. resolve: j walks its verified entries in order; a representative
//      becomes its target, an entry that merged earlier offers its own target
//      (taken when the exact difference to j is within the bound); resolved
//      in monotone rounds (each round settles at least the smallest
//      unresolved node, so the loop ends);
//   5. compact the level, remap the parent level's child pointers.
// Finally the whole DAG is re-deduplicated exactly (:1176).
#include "svd_internal.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace svd {

namespace {

constexpr int kB = 256;
constexpr uint8_t kUnres = 0, kRep = 1, kMerged = 2;

struct LevelView {
  const uint32_t* rec[kMaxLevels];
  const uint64_t* vox[kMaxLevels];
  uint32_t L;
};

__device__ __forceinline__ uint32_t mask_of(const LevelView& v, uint32_t lev, uint32_t n) {
  return (lev == v.L - 1 ? v.rec[lev][n] : v.rec[lev][(uint64_t)n * kRecWords]) & 0xFFu;
}
__device__ __forceinline__ uint32_t child_of(const LevelView& v, uint32_t lev, uint32_t n, int c) {
  return v.rec[lev][(uint64_t)n * kRecWords + 1 + c];
}

__global__ void k_fill_u64(uint64_t* d, uint64_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

// refs[child] += refs[parent] for every existing child (sortByEffectiveRefCount, :527-537)
__global__ void k_refs_level(const uint32_t* __restrict__ prec, uint64_t np, const uint64_t* __restrict__ pref,
                             unsigned long long* __restrict__ cref) {
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const uint32_t* r = prec + p * kRecWords;
  const uint32_t m = r[0] & 0xFFu;
  const uint64_t v = pref[p];
  if (!v) return;
  for (int c = 0; c < 8; ++c)
    if (((m >> c) & 1u) && r[1 + c] != kNull) atomicAdd(&cref[r[1 + c]], (unsigned long long)v);
}

__global__ void k_vox_level(const uint32_t* __restrict__ rec, uint64_t n, int leaf, const uint64_t* __restrict__ cvox,
                            uint64_t* __restrict__ vox) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (leaf) { vox[i] = (uint64_t)__popc(rec[i] & 0xFFu); return; }
  const uint32_t* r = rec + i * kRecWords;
  const uint32_t m = r[0] & 0xFFu;
  uint64_t s = 0;
  for (int c = 0; c < 8; ++c)
    if ((m >> c) & 1u) s += cvox[r[1 + c]];
  vox[i] = s;
}

__global__ void k_signatures(LevelView v, uint32_t lev, uint64_t n, uint32_t* __restrict__ sig) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s[16];
#pragma unroll
  for (int w = 0; w < 16; ++w) s[w] = 0;
  const uint32_t L = v.L;
  const uint32_t ma = mask_of(v, lev, (uint32_t)i);
  for (int c1 = 0; c1 < 8; ++c1) {
    if (!((ma >> c1) & 1u)) continue;
    if (lev + 1 >= L) { s[0] |= 1u << c1; continue; }
    const uint32_t b = child_of(v, lev, (uint32_t)i, c1);
    const uint32_t mb = mask_of(v, lev + 1, b);
    for (int c2 = 0; c2 < 8; ++c2) {
      if (!((mb >> c2) & 1u)) continue;
      if (lev + 2 >= L) { const int bit = c1 * 8 + c2; s[bit >> 5] |= 1u << (bit & 31); continue; }
      const uint32_t d = child_of(v, lev + 1, b, c2);
      const uint32_t md = mask_of(v, lev + 2, d);
      const int bit = c1 * 64 + c2 * 8;          // the 8 bits of md land in one word
      s[bit >> 5] |= md << (bit & 31);
    }
  }
  uint4* o = reinterpret_cast<uint4*>(sig + i * 16);
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = make_uint4(s[4 * q], s[4 * q + 1], s[4 * q + 2], s[4 * q + 3]);
}

__device__ __forceinline__ uint32_t hamming512(const uint4* a, const uint32_t s[16]) {
  uint32_t h = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 t = a[q];
    h += __popc(t.x ^ s[4 * q]) + __popc(t.y ^ s[4 * q + 1]) + __popc(t.z ^ s[4 * q + 2]) + __popc(t.w ^ s[4 * q + 3]);
  }
  return h;
}

// Exhaustive pre-filter lists: block = 256 consecutive j; the earlier nodes are
// streamed through LDS in tiles of 256 signatures (16 KB) that every lane
// reads in the same order (LDS broadcast).  Stops once every lane holds cap.
__global__ void __launch_bounds__(256) k_exhaustive_lists(const uint32_t* __restrict__ sig, uint64_t n,
                                                          const uint64_t* __restrict__ refs, uint64_t maxRefs,
                                                          uint32_t bound, uint32_t cap, uint32_t* __restrict__ lists,
                                                          uint32_t* __restrict__ counts) {
  __shared__ uint4 tile[256 * 4];
  const uint64_t j0 = (uint64_t)blockIdx.x * 256;
  const uint64_t j = j0 + threadIdx.x;
  const bool active = j < n && refs[j] <= maxRefs;
  uint32_t s[16];
  if (j < n) {
    const uint4* p = reinterpret_cast<const uint4*>(sig + j * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q) { const uint4 t = p[q]; s[4 * q] = t.x; s[4 * q + 1] = t.y; s[4 * q + 2] = t.z; s[4 * q + 3] = t.w; }
  }
  uint32_t cnt = 0;
  const uint64_t jend = j0 + 256 < n ? j0 + 256 : n;
  for (uint64_t i0 = 0; i0 < jend; i0 += 256) {
    if (!__syncthreads_or(active && cnt < cap)) break;
    for (int q = threadIdx.x; q < 256 * 4; q += 256) {
      const uint64_t i = i0 + (q >> 2);
      tile[q] = i < n ? reinterpret_cast<const uint4*>(sig + i * 16)[q & 3] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    if (active && cnt < cap) {
      const uint64_t iend = i0 + 256 < j ? i0 + 256 : j;
      for (uint64_t i = i0; i < iend; ++i) {
        if (hamming512(&tile[(i - i0) * 4], s) <= bound) {
          lists[j * cap + cnt] = (uint32_t)i;
          if (++cnt == cap) break;
        }
      }
    }
    __syncthreads();
  }
  if (j < n) counts[j] = active ? cnt : 0u;
}

// Ranked pre-filter lists: every earlier node i < j within the bound, the K
// nearest kept by (hamming << 32 | i) in a sorted register list (an insertion
// bubbles the new key in and drops the largest); the same LDS tiles as
// k_exhaustive_lists, scanned to the end (no early stop).
template <int K>
__global__ void __launch_bounds__(256) k_ranked_lists(const uint32_t* __restrict__ sig, uint64_t n,
                                                      const uint64_t* __restrict__ refs, uint64_t maxRefs,
                                                      uint32_t bound, uint32_t cap, uint32_t* __restrict__ lists,
                                                      uint32_t* __restrict__ counts) {
  __shared__ uint4 tile[256 * 4];
  const uint64_t j0 = (uint64_t)blockIdx.x * 256;
  const uint64_t j = j0 + threadIdx.x;
  const bool active = j < n && refs[j] <= maxRefs;
  uint32_t s[16];
  if (j < n) {
    const uint4* p = reinterpret_cast<const uint4*>(sig + j * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q) { const uint4 t = p[q]; s[4 * q] = t.x; s[4 * q + 1] = t.y; s[4 * q + 2] = t.z; s[4 * q + 3] = t.w; }
  }
  unsigned long long top[K];
#pragma unroll
  for (int q = 0; q < K; ++q) top[q] = ~0ull;
  uint32_t cnt = 0;
  const uint64_t jend = j0 + 256 < n ? j0 + 256 : n;
  for (uint64_t i0 = 0; i0 < jend; i0 += 256) {
    if (!__syncthreads_or(active)) break;
    for (int q = threadIdx.x; q < 256 * 4; q += 256) {
      const uint64_t i = i0 + (q >> 2);
      tile[q] = i < n ? reinterpret_cast<const uint4*>(sig + i * 16)[q & 3] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    if (active) {
      const uint64_t iend = i0 + 256 < j ? i0 + 256 : j;
      for (uint64_t i = i0; i < iend; ++i) {
        const uint32_t h = hamming512(&tile[(i - i0) * 4], s);
        if (h > bound) continue;
        unsigned long long x = ((unsigned long long)h << 32) | (unsigned long long)i;
        if (x >= top[K - 1]) continue;
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const unsigned long long t = top[q];
          const bool lt = x < t;
          top[q] = lt ? x : t;
          x = lt ? t : x;
        }
        ++cnt;
      }
    }
    __syncthreads();
  }
  if (j >= n) return;
  const uint32_t m = active ? min(cnt, cap) : 0u;
#pragma unroll
  for (int q = 0; q < K; ++q)
    if ((uint32_t)q < m) lists[j * cap + q] = (uint32_t)top[q];
  counts[j] = m;
}

// Learned proposals -> (hi << 32 | lo) keys; invalid entries get ~0.
__global__ void k_proposal_keys(const uint32_t* __restrict__ knn, uint64_t n, uint32_t k,
                                const uint32_t* __restrict__ sig, const uint64_t* __restrict__ refs, uint64_t maxRefs,
                                uint32_t bound, uint64_t* __restrict__ keys) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const uint64_t x = e / k;
  const uint32_t y = knn[e];
  uint64_t key = ~0ull;
  if (y != kNone && (uint64_t)y != x && (uint64_t)y < n) {
    const uint64_t lo = x < y ? x : y, hi = x < y ? y : x;
    if (refs[hi] <= maxRefs) {
      uint32_t s[16];
      const uint4* p = reinterpret_cast<const uint4*>(sig + hi * 16);
#pragma unroll
      for (int q = 0; q < 4; ++q) { const uint4 t = p[q]; s[4 * q] = t.x; s[4 * q + 1] = t.y; s[4 * q + 2] = t.z; s[4 * q + 3] = t.w; }
      if (hamming512(reinterpret_cast<const uint4*>(sig + lo * 16), s) <= bound) key = (hi << 32) | lo;
    }
  }
  keys[e] = key;
}

// Sorted keys -> per hi node its first cap distinct lo (ascending).
__global__ void k_lists_from_sorted(const uint64_t* __restrict__ keys, uint64_t m, uint32_t cap,
                                    uint32_t* __restrict__ lists, uint32_t* __restrict__ counts) {
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const uint64_t key = keys[p];
  if (key == ~0ull) return;
  const uint64_t hi = key >> 32;
  if (p > 0 && (keys[p - 1] >> 32) == hi) return;
  uint32_t cnt = 0;
  uint64_t prev = ~0ull;
  for (uint64_t q = p; q < m && cnt < cap; ++q) {
    const uint64_t kq = keys[q];
    if (kq == ~0ull || (kq >> 32) != hi) break;
    if (kq != prev) { lists[hi * cap + cnt++] = (uint32_t)kq; prev = kq; }
  }
  counts[hi] = cnt;
}

// Exact voxel difference of the subtrees a0, b0 at level lev0, capped at
// bound + 1 (the running total only grows, so the capped value does not depend
// on the visiting order).
__device__ uint32_t subtree_diff(const LevelView& v, uint32_t lev0, uint32_t a0, uint32_t b0, uint32_t bound) {
  if (a0 == b0) return 0;
  const uint32_t L = v.L;
  if (lev0 == L - 1) {
    const uint32_t d = __popc(mask_of(v, lev0, a0) ^ mask_of(v, lev0, b0));
    return d > bound ? bound + 1 : d;
  }
  uint32_t sa[kMaxLevels], sb[kMaxLevels];
  uint8_t sc[kMaxLevels];
  int top = 0;
  sa[0] = a0; sb[0] = b0; sc[0] = 0;
  uint64_t acc = 0;
  while (top >= 0) {
    const uint32_t lev = lev0 + (uint32_t)top;     // < L - 1 for every frame on the stack
    const uint32_t a = sa[top], b = sb[top];
    const uint32_t ma = mask_of(v, lev, a), mb = mask_of(v, lev, b);
    int c = sc[top];
    bool pushed = false;
    for (; c < 8; ++c) {
      const bool ea = (ma >> c) & 1u, eb = (mb >> c) & 1u;
      if (ea && eb) {
        const uint32_t ca = child_of(v, lev, a, c), cb = child_of(v, lev, b, c);
        if (ca != cb) {
          if (lev + 1 == L - 1) {
            acc += __popc(mask_of(v, lev + 1, ca) ^ mask_of(v, lev + 1, cb));
          } else {
            sc[top] = (uint8_t)(c + 1);
            ++top;
            sa[top] = ca; sb[top] = cb; sc[top] = 0;
            pushed = true;
            break;
          }
        }
      } else if (ea) {
        acc += v.vox[lev + 1][child_of(v, lev, a, c)];
      } else if (eb) {
        acc += v.vox[lev + 1][child_of(v, lev, b, c)];
      }
      if (acc > bound) return bound + 1;
    }
    if (!pushed) --top;
  }
  return (uint32_t)acc;
}

__global__ void k_verify(LevelView v, uint32_t lev, const uint32_t* __restrict__ lists,
                         const uint32_t* __restrict__ counts, uint64_t n, uint32_t cap, uint32_t bound,
                         uint32_t* __restrict__ dv, unsigned long long* __restrict__ stats) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * cap) return;
  const uint64_t j = e / cap;
  const uint32_t t = (uint32_t)(e - j * cap);
  if (t >= counts[j]) return;
  const uint32_t d = subtree_diff(v, lev, lists[e], (uint32_t)j, bound);
  dv[e] = d;
  atomicAdd(&stats[0], 1ull);
  if (d <= bound) atomicAdd(&stats[1], 1ull);
}

// Node state word: state << 32 | target node (one 64-bit word, so a reader
// that sees "merged" also sees the target).
constexpr uint64_t kStRep = (uint64_t)kRep << 32, kStMerged = (uint64_t)kMerged << 32;

__global__ void k_resolve_init(uint64_t n, const uint64_t* __restrict__ refs, uint64_t maxRefs,
                               unsigned long long* __restrict__ st, unsigned long long* __restrict__ stats) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const bool cand = refs[j] <= maxRefs;
  st[j] = cand ? (unsigned long long)kUnres << 32 : (unsigned long long)(kStRep | j);
  if (cand) atomicAdd(&stats[2], 1ull);
}

// One round: an unresolved j walks its verified entries in order.  A
// representative becomes its target; an entry that merged offers its own
// target (a representative), taken when the exact difference to j is within
// the bound; an unresolved entry makes j wait.  States only move away from
// "unresolved", so a stale read delays a node by one round and never changes
// its outcome.
__global__ void k_resolve(LevelView v, uint32_t lev, const uint32_t* __restrict__ lists,
                          const uint32_t* __restrict__ dv, const uint32_t* __restrict__ counts, uint64_t n,
                          uint32_t cap, uint32_t bound, unsigned long long* st, uint32_t* __restrict__ tdist,
                          uint32_t* __restrict__ unresolved) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if ((__atomic_load_n(&st[j], __ATOMIC_RELAXED) >> 32) != kUnres) return;
  const uint32_t cnt = counts[j];
  for (uint32_t t = 0; t < cnt; ++t) {
    const uint64_t e = j * cap + t;
    if (dv[e] > bound) continue;
    const uint32_t i = lists[e];
    const unsigned long long w = __atomic_load_n(&st[i], __ATOMIC_RELAXED);
    const uint32_t s = (uint32_t)(w >> 32);
    if (s == kRep) {
      tdist[j] = dv[e];
      __atomic_store_n(&st[j], (unsigned long long)(kStMerged | i), __ATOMIC_RELAXED);
      return;
    }
    if (s == kMerged) {
      const uint32_t r = (uint32_t)w;
      const uint32_t d = subtree_diff(v, lev, r, (uint32_t)j, bound);
      if (d <= bound) {
        tdist[j] = d;
        __atomic_store_n(&st[j], (unsigned long long)(kStMerged | r), __ATOMIC_RELAXED);
        return;
      }
      continue;
    }
    atomicAdd(unresolved, 1u);
    return;
  }
  __atomic_store_n(&st[j], (unsigned long long)(kStRep | j), __ATOMIC_RELAXED);
}

// Also clears hit[], which k_lossy_compact sets on every representative that
// takes in a merged node.
__global__ void k_rep_flag(const unsigned long long* __restrict__ st, uint64_t n, uint32_t* __restrict__ flag,
                           uint32_t* __restrict__ hit) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  flag[j] = (st[j] >> 32) == kRep ? 1u : 0u;
  hit[j] = 0u;
}

__global__ void k_lossy_compact(const uint32_t* __restrict__ rec, const uint64_t* __restrict__ vox,
                                const unsigned long long* __restrict__ st, const uint32_t* __restrict__ tdist,
                                const uint32_t* __restrict__ pos, const uint64_t* __restrict__ refs, uint64_t n,
                                uint32_t* __restrict__ out, uint64_t* __restrict__ outVox,
                                uint32_t* __restrict__ corr, uint32_t* __restrict__ hit,
                                unsigned long long* __restrict__ stats) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned long long w = st[j];
  if ((w >> 32) == kRep) {
    const uint32_t p = pos[j];
    const uint32_t* s = rec + j * kRecWords;
    uint32_t* d = out + (uint64_t)p * kRecWords;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) d[k] = s[k];
    outVox[p] = vox[j];
    corr[j] = p;
  } else {
    corr[j] = pos[(uint32_t)w];
    atomicAdd(&stats[3], 1ull);
    atomicAdd(&stats[4], (unsigned long long)tdist[j] * refs[j]);
    // clusters: representatives with at least one member
    if (atomicExch(&hit[(uint32_t)w], 1u) == 0u) atomicAdd(&stats[8], 1ull);
  }
}

__global__ void k_remap_inplace(uint32_t* __restrict__ rec, uint64_t n, const uint32_t* __restrict__ corr) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* r = rec + i * kRecWords;
  const uint32_t m = r[0] & 0xFFu;
  for (int c = 0; c < 8; ++c)
    if ((m >> c) & 1u) r[1 + c] = corr[r[1 + c]];
}

LevelView make_view(const svd_octree* o, const std::vector<uint64_t*>& vox) {
  LevelView v = {};
  v.L = o->levels;
  for (uint32_t l = 0; l < o->levels; ++l) { v.rec[l] = o->dag[l].rec; v.vox[l] = vox[l]; }
  return v;
}

struct PhaseClock {
  svd_ctx* c;
  std::vector<hipEvent_t> ev;
  explicit PhaseClock(svd_ctx* ctx, int n) : c(ctx), ev(n, nullptr) {
    for (auto& e : ev) SVD_HIP(hipEventCreate(&e));
  }
  ~PhaseClock() { for (auto& e : ev) if (e) (void)hipEventDestroy(e); }
  void mark(int i) { SVD_HIP(hipEventRecord(ev[i], c->stream)); }
  double ms(int a, int b) {
    float t = 0.f;
    SVD_HIP(hipEventSynchronize(ev[b]));
    SVD_HIP(hipEventElapsedTime(&t, ev[a], ev[b]));
    return (double)t;
  }
};

}  // namespace

// Sort n u64 keys ascending in place (rocPRIM radix sort through a scratch copy).
void sort_u64(svd_ctx* c, uint64_t* keys, uint64_t n) {
  if (n < 2) return;
  uint64_t* tmpKeys = dalloc<uint64_t>(c, n);
  size_t bytes = 0;
  SVD_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys, tmpKeys, (size_t)n, 0, 64, c->stream));
  void* tmp = c->arena->get(bytes + 16);
  SVD_HIP(rocprim::radix_sort_keys(tmp, bytes, keys, tmpKeys, (size_t)n, 0, 64, c->stream));
  SVD_HIP(hipMemcpyAsync(keys, tmpKeys, n * 8, hipMemcpyDeviceToDevice, c->stream));
  c->arena->put(tmp);
  dfree(c, tmpKeys);
}

void lossy_signatures(svd_octree* o, uint32_t lev, uint32_t* sig) {
  if (o->state != SVD_S_DAG && o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "signatures need the DAG"};
  if (lev >= o->levels) throw Error{SVD_ERR_LEVELS, "level out of range"};
  std::vector<uint64_t*> vox(o->levels, nullptr);
  const LevelView v = make_view(o, vox);
  const uint64_t n = o->dag[lev].n;
  if (n) k_signatures<<<div_up(n, kB), kB, 0, o->ctx->stream>>>(v, lev, n, sig);
  check_launch("k_signatures");
}

void lossy_merge(svd_octree* o, const svd_lossy_params* p, std::vector<svd_lossy_level_stats>& out) {
  svd_ctx* c = o->ctx;
  const uint32_t L = o->levels;
  if (o->state != SVD_S_DAG) throw Error{SVD_ERR_STATE, "toLossyDag: octree is not a DAG (geom_octree_extension.cpp:886-891)"};
  if (L < 3 || p->lev_lo < 1 || p->lev_lo > p->lev_hi || p->lev_hi > L - 2)
    throw Error{SVD_ERR_ARG, "lossy levels must satisfy 1 <= lev_lo <= lev_hi <= L-2"};
  if (p->list_cap < 1 || p->list_cap > 64) throw Error{SVD_ERR_ARG, "list_cap must be 1..64"};
  const bool learned = p->proposer == SVD_PROPOSER_LEARNED;
  const bool ranked = p->proposer == SVD_PROPOSER_RANKED;
  if (p->proposer > SVD_PROPOSER_RANKED) throw Error{SVD_ERR_ARG, "unknown proposer"};
  if (learned && (!p->model || p->knn_k < 1 || p->knn_k > 32 || p->knn_tables < 1 || p->knn_bits < 1 || p->knn_refine > 16 ||
                  p->knn_bits > 24 || p->knn_window < 1 || p->knn_window > 96))
    throw Error{SVD_ERR_ARG, "learned proposer needs a model, 1<=k<=32, tables>=1, 1<=bits<=24, 1<=window<=96"};
  const uint32_t cap = p->list_cap;
  const uint64_t maxRefs = p->max_refs;
  uint64_t recUsed = 0;

  // effective reference counts of levels 0..lev_hi on the input DAG
  std::vector<uint64_t*> refs(L, nullptr);
  refs[0] = dalloc<uint64_t>(c, o->dag[0].n + 1);
  k_fill_u64<<<div_up(o->dag[0].n + 1, kB), kB, 0, c->stream>>>(refs[0], 1ull, o->dag[0].n);
  for (uint32_t lev = 1; lev <= p->lev_hi; ++lev) {
    const uint64_t n = o->dag[lev].n;
    refs[lev] = dalloc<uint64_t>(c, n + 1);
    k_fill_u64<<<div_up(n + 1, kB), kB, 0, c->stream>>>(refs[lev], 0ull, n + 1);
    const uint64_t np = o->dag[lev - 1].n;
    if (np) k_refs_level<<<div_up(np, kB), kB, 0, c->stream>>>(o->dag[lev - 1].rec, np, refs[lev - 1],
                                                              (unsigned long long*)refs[lev]);
  }
  check_launch("effective refcounts");
  // voxel counts below the merged range
  std::vector<uint64_t*> vox(L, nullptr);
  for (int lev = (int)L - 1; lev > (int)p->lev_hi; --lev) {
    const uint64_t n = o->dag[lev].n;
    vox[lev] = dalloc<uint64_t>(c, n + 1);
    if (n) k_vox_level<<<div_up(n, kB), kB, 0, c->stream>>>(o->dag[lev].rec, n, lev == (int)L - 1,
                                                          lev + 1 < (int)L ? vox[lev + 1] : nullptr, vox[lev]);
  }
  check_launch("voxel counts");
  unsigned long long* dstats = dalloc<unsigned long long>(c, 9);
  uint32_t* unresolved = dalloc<uint32_t>(c, 1);
  PhaseClock clk(c, 7);

  for (int lev = (int)p->lev_hi; lev >= (int)p->lev_lo; --lev) {
    Level& lv = o->dag[lev];
    const uint64_t n = lv.n;
    svd_lossy_level_stats st = {};
    st.level = (uint64_t)lev;
    if (p->max_diff) st.bound = p->max_diff;
    else {
      const double d = (double)(L - (uint32_t)lev);
      st.bound = (uint64_t)std::llround(d * d * (double)p->diff_factor);
    }
    const uint32_t bound = st.bound > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)st.bound;
    st.nodes_before = n;
    vox[lev] = dalloc<uint64_t>(c, n + 1);
    if (n) k_vox_level<<<div_up(n, kB), kB, 0, c->stream>>>(lv.rec, n, 0, vox[lev + 1], vox[lev]);
    k_fill_u64<<<1, 64, 0, c->stream>>>((uint64_t*)dstats, 0ull, 9);
    uint32_t* sig = dalloc<uint32_t>(c, (n + 1) * 16);
    uint32_t* lists = dalloc<uint32_t>(c, (n + 1) * cap);
    uint32_t* counts = dalloc<uint32_t>(c, n + 1);
    uint32_t* dv = dalloc<uint32_t>(c, (n + 1) * cap);
    unsigned long long* state = dalloc<unsigned long long>(c, n + 1);
    uint32_t* tdist = dalloc<uint32_t>(c, n + 1);
    uint32_t* pos = dalloc<uint32_t>(c, n + 1);
    uint32_t* total = dalloc<uint32_t>(c, 1);
    uint32_t* corr = dalloc<uint32_t>(c, n + 1);
    uint32_t* hit = dalloc<uint32_t>(c, n + 1);
    LevelView v = make_view(o, vox);
    clk.mark(0);
    if (n) k_signatures<<<div_up(n, kB), kB, 0, c->stream>>>(v, (uint32_t)lev, n, sig);
    check_launch("k_signatures");
    clk.mark(1);
    if (!learned) {
      clk.mark(2);
      if (n && !ranked)
        k_exhaustive_lists<<<div_up(n, 256), 256, 0, c->stream>>>(sig, n, refs[lev], maxRefs, bound, cap, lists,
                                                                  counts);
      else if (n && cap <= 16)
        k_ranked_lists<16><<<div_up(n, 256), 256, 0, c->stream>>>(sig, n, refs[lev], maxRefs, bound, cap, lists,
                                                                   counts);
      else if (n && cap <= 32)
        k_ranked_lists<32><<<div_up(n, 256), 256, 0, c->stream>>>(sig, n, refs[lev], maxRefs, bound, cap, lists,
                                                                   counts);
      else if (n)
        k_ranked_lists<64><<<div_up(n, 256), 256, 0, c->stream>>>(sig, n, refs[lev], maxRefs, bound, cap, lists,
                                                                   counts);
      check_launch("pre-filter lists");
    } else {
      const uint32_t k = p->knn_k;
      float* emb = dalloc<float>(c, (n + 32) * 32);
      uint32_t* knn = dalloc<uint32_t>(c, (n + 1) * k);
      if (n) mlp_embed(p->model, sig, n, emb);
      clk.mark(2);
      if (n) knn_lsh(c, emb, n, k, p->knn_tables, p->knn_bits, p->knn_window, p->knn_seed, knn, sig);
      if (n) knn_refine(c, sig, n, k, p->knn_refine, knn);
      if (p->knn_record) {
        if (recUsed + n * k > p->knn_record_cap) throw Error{SVD_ERR_CAPACITY, "knn_record buffer too small"};
        if (n) SVD_HIP(hipMemcpyAsync(p->knn_record + recUsed, knn, n * k * 4, hipMemcpyDeviceToHost, c->stream));
        recUsed += n * k;
      }
      clk.mark(3);
      uint64_t* keys = dalloc<uint64_t>(c, n * k + 1);
      fill_u32(c, counts, 0u, n + 1);
      if (n) {
        k_proposal_keys<<<div_up(n * k, kB), kB, 0, c->stream>>>(knn, n, k, sig, refs[lev], maxRefs, bound, keys);
        check_launch("k_proposal_keys");
        sort_u64(c, keys, n * k);
        k_lists_from_sorted<<<div_up(n * k, kB), kB, 0, c->stream>>>(keys, n * k, cap, lists, counts);
        check_launch("k_lists_from_sorted");
      }
      dfree(c, keys); dfree(c, knn); dfree(c, emb);
    }
    if (!learned) clk.mark(3);
    clk.mark(4);
    if (n) k_verify<<<div_up(n * cap, kB), kB, 0, c->stream>>>(v, (uint32_t)lev, lists, counts, n, cap, bound, dv,
                                                             dstats + 5);
    check_launch("k_verify");
    clk.mark(5);
    if (n) k_resolve_init<<<div_up(n, kB), kB, 0, c->stream>>>(n, refs[lev], maxRefs, state, dstats + 5);
    check_launch("k_resolve_init");
    for (uint64_t round = 0; n; ++round) {
      fill_u32(c, unresolved, 0u, 1);
      k_resolve<<<div_up(n, kB), kB, 0, c->stream>>>(v, (uint32_t)lev, lists, dv, counts, n, cap, bound, state, tdist,
                                                     unresolved);
      check_launch("k_resolve");
      if (read_u32(c, unresolved) == 0) break;
      if (round > n) throw Error{SVD_ERR_DEVICE, "lossy resolution did not converge"};
    }
    if (n) k_rep_flag<<<div_up(n, kB), kB, 0, c->stream>>>(state, n, pos, hit);
    scan_exclusive(c, pos, pos, n, total);
    const uint64_t kept = n ? read_u32(c, total) : 0;
    uint32_t* newRec = dalloc<uint32_t>(c, (kept + 1) * kRecWords);
    uint64_t* newVox = dalloc<uint64_t>(c, kept + 1);
    if (n) k_lossy_compact<<<div_up(n, kB), kB, 0, c->stream>>>(lv.rec, vox[lev], state, tdist, pos, refs[lev], n,
                                                              newRec, newVox, corr, hit, dstats);
    check_launch("k_lossy_compact");
    const uint64_t np = o->dag[lev - 1].n;
    if (np) k_remap_inplace<<<div_up(np, kB), kB, 0, c->stream>>>(o->dag[lev - 1].rec, np, corr);
    check_launch("k_remap_inplace");
    clk.mark(6);
    uint64_t hs[9];
    SVD_HIP(hipMemcpyAsync(hs, dstats, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
    st.merged = hs[3];
    st.error_bound = hs[4];
    st.listed = hs[5];
    st.verified = hs[6];
    st.candidates = hs[7];
    st.clusters = hs[8];
    st.nodes_after = kept;
    st.ms_embed = learned ? clk.ms(1, 2) : 0.0;
    st.ms_knn = learned ? clk.ms(2, 3) : 0.0;
    st.ms_lists = learned ? clk.ms(3, 4) : clk.ms(2, 3);
    st.ms_verify = clk.ms(4, 5);
    st.ms_merge = clk.ms(5, 6);
    st.ms_total = clk.ms(0, 6);
    out.push_back(st);
    dfree(c, lv.rec);
    lv.rec = newRec;
    lv.n = kept;
    dfree(c, vox[lev]);
    vox[lev] = newVox;
    dfree(c, sig); dfree(c, lists); dfree(c, counts); dfree(c, dv); dfree(c, state); dfree(c, tdist);
    dfree(c, pos); dfree(c, total); dfree(c, corr); dfree(c, hit);
  }
  for (auto* r : refs) dfree(c, r);
  for (auto* x : vox) dfree(c, x);
  dfree(c, dstats); dfree(c, unresolved);
  // final exact dedup of every level (geom_octree_extension.cpp:1176)
  std::vector<Level> reduced;
  const uint64_t nn = dedup_all(c, L, o->dag, reduced);
  for (auto& lv : o->dag) dfree(c, lv.rec);
  o->dag = reduced;
  o->nNodesDAG = nn;
  SVD_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace svd
