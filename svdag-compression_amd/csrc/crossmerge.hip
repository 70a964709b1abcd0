// CSVDAG cross-level merge on the device (geom_octree_extension.cpp:1192-1544).
//
// The reference walks a top-down frontier: every frontier node A of level levA
// is compared (compareSubtrees, :238-290) with the nodes of the levels above
// that share its hash (computeNodeHash, :34-74); the first equal one in
// (levB, index) order wins, A's whole subtree is recorded as corresponding to
// B's, and unmatched nodes send their children to the next frontier.  Then the
// levels are compacted bottom-up and pointers redirected (:1428-1498).
//
// compareSubtrees(A, B) is equality of the two subtrees cut at the depth of
// A's leaves: every pair of visited nodes has equal child masks.  Here that
// equality is a class number: cls_0(x) = mask(x), and cls_d(x) = the hash-table
// slot of (mask(x), cls_{d-1}(children)) over all nodes of levels 1..L-1-d, so
// cls_d(A) == cls_d(B) exactly when the depth-d subtrees are equal.  Equal
// subtrees have equal hashes (the hash depth never exceeds the compared
// depth), so the hash pre-filter drops nothing and is not needed.  The first
// match is then an atomicMin of (levB << 32 | index) per class.
//
// Recording: inside one comparison the reference's map keeps the last
// pre-order visit of a node; between matches of one round the earliest
// frontier node wins.  The last pre-order visit of x is its lexicographically
// largest child-slot path from A, so a top-down pass per level keeps, per
// node, the smallest matching frontier index and then the largest path
// (path * 8 + slot), and the winning parent writes B's descendant on the same
// path.  No subtree is walked per thread.
#include "svd_internal.hpp"

namespace svd {

namespace {

constexpr int kB = 256;
constexpr uint32_t kMaxL = 21;

struct LevelPtrs {
  const uint32_t* rec[kMaxL];   // 9-word records (leaf level: 1 word)
  uint32_t n[kMaxL];
  uint32_t leafLevel;
};

__device__ __forceinline__ uint32_t node_mask(const LevelPtrs& P, uint32_t lev, uint32_t i) {
  return lev == P.leafLevel ? (P.rec[lev][i] & 0xFFu) : (P.rec[lev][(uint64_t)i * kRecWords] & 0xFFu);
}
__device__ __forceinline__ uint32_t node_child(const LevelPtrs& P, uint32_t lev, uint32_t i, int c) {
  return lev == P.leafLevel ? kNull : P.rec[lev][(uint64_t)i * kRecWords + 1 + c];
}

// ---- truncated-subtree classes ---------------------------------------------
// keys of level lev for depth d >= 1: mask, then the children's cls_{d-1}
// (clsChild == nullptr: d == 1, the children's masks)
__global__ void k_class_keys(LevelPtrs P, uint32_t lev, const uint32_t* __restrict__ clsChild,
                             uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n[lev]) return;
  const uint32_t m = node_mask(P, lev, i);
  uint32_t* k = keys + (uint64_t)i * kRecWords;
  k[0] = m;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    uint32_t v = kNull;
    if ((m >> c) & 1u) {
      const uint32_t ch = node_child(P, lev, i, c);
      v = clsChild ? clsChild[ch] : node_mask(P, lev + 1, ch);
    }
    k[1 + c] = v;
  }
}

// cls[i] = the slot that holds key i's value (one slot per distinct key)
__global__ void k_class_insert(const uint32_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ table,
                               uint32_t cmask, uint32_t* __restrict__ cls) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t w[kRecWords];
#pragma unroll
  for (int k = 0; k < kRecWords; ++k) w[k] = keys[i * kRecWords + k];
  uint32_t h = (uint32_t)hash_rec(w) & cmask;
  while (true) {
    uint32_t v = __atomic_load_n(&table[h], __ATOMIC_RELAXED);
    if (v == kNone) {
      const uint32_t old = atomicCAS(&table[h], kNone, (uint32_t)i);
      if (old == kNone) { cls[i] = h; return; }
      v = old;
    }
    bool eq = true;
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) eq = eq && keys[(uint64_t)v * kRecWords + k] == w[k];
    if (eq) { cls[i] = h; return; }
    h = (h + 1) & cmask;
  }
}

__device__ __forceinline__ uint32_t class_of(const LevelPtrs& P, uint32_t lev, uint32_t i,
                                             const uint32_t* __restrict__ cls) {
  return cls ? cls[i] : node_mask(P, lev, i);
}

// bestOf[class] = min over the nodes of level levB of (levB << 32 | index)
__global__ void k_class_best(LevelPtrs P, uint32_t levB, const uint32_t* __restrict__ cls,
                             unsigned long long* __restrict__ bestOf) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.n[levB]) return;
  const unsigned long long key = ((unsigned long long)levB << 32) | j;
  unsigned long long* b = &bestOf[class_of(P, levB, j, cls)];
  if (key < __atomic_load_n(b, __ATOMIC_RELAXED)) atomicMin(b, key);
}

// Frontier nodes of levA take their class's first match; matched ones start
// the recording state (frontier index, empty path, B).
__global__ void k_frontier_best(LevelPtrs P, uint32_t levA, const uint8_t* __restrict__ front,
                                const uint32_t* __restrict__ cls, const unsigned long long* __restrict__ bestOf,
                                unsigned long long* __restrict__ best, uint32_t* __restrict__ sa,
                                uint64_t* __restrict__ sp, uint64_t* __restrict__ sb, uint32_t* __restrict__ nMatched) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= P.n[levA]) return;
  unsigned long long b = ~0ull;
  if (front[a]) b = bestOf[class_of(P, levA, a, cls)];
  best[a] = b;
  sa[a] = b != ~0ull ? a : kNone;
  sp[a] = 0;
  sb[a] = b;
  if (b != ~0ull) atomicAdd(nMatched, 1u);
}

// One level of the recording pass: children (level lev+1) of recorded nodes.
__global__ void k_prop_min(LevelPtrs P, uint32_t lev, const uint32_t* __restrict__ sa, uint32_t* __restrict__ na) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= P.n[lev]) return;
  const uint32_t a = sa[x];
  if (a == kNone) return;
  const uint32_t m = node_mask(P, lev, x);
  for (int c = 0; c < 8; ++c) {
    if (!((m >> c) & 1u)) continue;
    uint32_t* t = &na[node_child(P, lev, x, c)];
    if (a < __atomic_load_n(t, __ATOMIC_RELAXED)) atomicMin(t, a);
  }
}

__global__ void k_prop_path(LevelPtrs P, uint32_t lev, const uint32_t* __restrict__ sa, const uint64_t* __restrict__ sp,
                            const uint32_t* __restrict__ na, unsigned long long* __restrict__ np) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= P.n[lev]) return;
  const uint32_t a = sa[x];
  if (a == kNone) return;
  const uint32_t m = node_mask(P, lev, x);
  const unsigned long long base = (unsigned long long)sp[x] << 3;
  for (int c = 0; c < 8; ++c) {
    if (!((m >> c) & 1u)) continue;
    const uint32_t y = node_child(P, lev, x, c);
    if (na[y] != a) continue;
    const unsigned long long v = base | (unsigned long long)c;
    if (v > __atomic_load_n(&np[y], __ATOMIC_RELAXED)) atomicMax(&np[y], v);
  }
}

__global__ void k_prop_write(LevelPtrs P, uint32_t lev, const uint32_t* __restrict__ sa, const uint64_t* __restrict__ sp,
                             const uint64_t* __restrict__ sb, const uint32_t* __restrict__ na,
                             const uint64_t* __restrict__ np, uint64_t* __restrict__ nb) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= P.n[lev]) return;
  const uint32_t a = sa[x];
  if (a == kNone) return;
  const uint32_t m = node_mask(P, lev, x);
  const uint64_t base = sp[x] << 3;
  const uint32_t lb = (uint32_t)(sb[x] >> 32), b = (uint32_t)sb[x];
  for (int c = 0; c < 8; ++c) {
    if (!((m >> c) & 1u)) continue;
    const uint32_t y = node_child(P, lev, x, c);
    if (na[y] == a && np[y] == (base | (uint64_t)c))
      nb[y] = ((uint64_t)(lb + 1) << 32) | node_child(P, lb, b, c);
  }
}

// Merge this round into the global correspondence map: frontier matches of
// levA overwrite (operator[]), subtree entries only fill empty slots (insert).
__global__ void k_merge_self(uint32_t n, const uint8_t* __restrict__ front, const unsigned long long* __restrict__ best,
                             uint64_t* __restrict__ mlc, uint8_t* __restrict__ nextFront, const uint32_t* __restrict__ rec,
                             int leafLevel) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n || !front[a]) return;
  if (best[a] != ~0ull) { mlc[a] = best[a]; return; }
  if (leafLevel || !nextFront) return;
  const uint32_t* r = rec + (uint64_t)a * kRecWords;
  for (int c = 0; c < 8; ++c)
    if ((r[0] >> c) & 1u) nextFront[r[1 + c]] = 1;
}

__global__ void k_merge_sub(uint32_t n, const uint64_t* __restrict__ cand, uint64_t* __restrict__ mlc) {
  uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  if (cand[x] != ~0ull && mlc[x] == ~0ull) mlc[x] = cand[x];
}

__global__ void k_fill64(uint64_t* d, uint64_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}
__global__ void k_fill8(uint8_t* d, uint8_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

// ---- compaction (geom_octree_extension.cpp:1428-1498) ------------------------
__global__ void k_keep_flags(const uint64_t* __restrict__ mlc, uint32_t n, uint32_t* __restrict__ flag) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = mlc[i] == ~0ull ? 1u : 0u;
}

__global__ void k_compact_level(const uint32_t* __restrict__ rec, const uint32_t* __restrict__ cl, uint32_t n, int leaf,
                                const uint64_t* __restrict__ mlc, const uint32_t* __restrict__ pos,
                                uint32_t* __restrict__ outRec, uint32_t* __restrict__ outCl, uint32_t* __restrict__ corr) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mlc[i] != ~0ull) { corr[i] = kNone; return; }
  const uint32_t p = pos[i];
  corr[i] = p;
  if (leaf) { outRec[p] = rec[i]; return; }
  for (int k = 0; k < kRecWords; ++k) outRec[(uint64_t)p * kRecWords + k] = rec[(uint64_t)i * kRecWords + k];
  for (int k = 0; k < 8; ++k) outCl[(uint64_t)p * 8 + k] = cl[(uint64_t)i * 8 + k];
}

// Parents (level lev-1) of the nodes of level lev: redirect or renumber.
__global__ void k_fix_parents(uint32_t* __restrict__ rec, uint32_t* __restrict__ cl, uint32_t n,
                              const uint64_t* __restrict__ mlc, const uint32_t* __restrict__ corr) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* r = rec + (uint64_t)i * kRecWords;
  uint32_t* l = cl + (uint64_t)i * 8;
  for (int c = 0; c < 8; ++c) {
    if (!((r[0] >> c) & 1u)) continue;
    const uint32_t ch = r[1 + c];
    const uint64_t m = mlc[ch];
    if (m != ~0ull) { l[c] = (uint32_t)(m >> 32); r[1 + c] = (uint32_t)m; }
    else r[1 + c] = corr[ch];
  }
}

// Nodes of a lower level whose pointers were redirected into level lev.
__global__ void k_fix_redirected(uint32_t* __restrict__ rec, const uint32_t* __restrict__ cl, uint32_t n, uint32_t lev,
                                 const uint32_t* __restrict__ corr, uint32_t* __restrict__ missing) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* r = rec + (uint64_t)i * kRecWords;
  const uint32_t* l = cl + (uint64_t)i * 8;
  for (int c = 0; c < 8; ++c) {
    if (!((r[0] >> c) & 1u) || l[c] != lev) continue;
    const uint32_t v = corr[r[1 + c]];
    if (v == kNone) atomicAdd(missing, 1u);
    r[1 + c] = v;
  }
}

__global__ void k_init_cl(uint32_t* __restrict__ cl, uint32_t n, uint32_t lev) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * 8) cl[i] = lev + 1;
}

}  // namespace

void cross_level_merge(svd_octree* o) {
  svd_ctx* c = o->ctx;
  if (o->state != SVD_S_DAG) throw Error{SVD_ERR_STATE, "cross-level merge needs a DAG"};
  const uint32_t L = o->levels;
  if (L > kMaxL) throw Error{SVD_ERR_LEVELS, "too many levels"};
  std::vector<Level>& D = o->dag;
  // working copy of the DAG view with childLevels (the merge rewrites pointers)
  std::vector<Level> W(L);
  LevelPtrs P{};
  P.leafLevel = L - 1;
  for (uint32_t l = 0; l < L; ++l) {
    W[l].n = D[l].n;
    W[l].leaf = D[l].leaf;
    const uint64_t words = D[l].leaf ? D[l].n : D[l].n * kRecWords;
    W[l].rec = dalloc<uint32_t>(c, words + 1);
    if (words) SVD_HIP(hipMemcpyAsync(W[l].rec, D[l].rec, 4 * words, hipMemcpyDeviceToDevice, c->stream));
    if (!D[l].leaf) {
      W[l].childLevels = dalloc<uint32_t>(c, 8 * D[l].n + 1);
      if (D[l].n) k_init_cl<<<div_up(8 * D[l].n, kB), kB, 0, c->stream>>>(W[l].childLevels, (uint32_t)D[l].n, l);
    }
    P.rec[l] = W[l].rec;
    P.n[l] = (uint32_t)W[l].n;
  }
  // per-level state
  std::vector<uint64_t*> mlc(L, nullptr), cand(L, nullptr), sp(L, nullptr);
  std::vector<uint32_t*> sa(L, nullptr);
  std::vector<uint8_t*> front(L, nullptr);
  for (uint32_t l = 0; l < L; ++l) {
    const uint64_t n = W[l].n + 1;
    mlc[l] = dalloc<uint64_t>(c, n);
    cand[l] = dalloc<uint64_t>(c, n);
    sp[l] = dalloc<uint64_t>(c, n);
    sa[l] = dalloc<uint32_t>(c, n);
    front[l] = dalloc<uint8_t>(c, n);
    k_fill64<<<div_up(n, kB), kB, 0, c->stream>>>(mlc[l], ~0ull, n);
    k_fill8<<<div_up(n, kB), kB, 0, c->stream>>>(front[l], 0, n);
  }
  // classes: cls[d] covers levels 1..L-1-d, level l at base[l]
  std::vector<uint64_t> base(L + 1, 0);
  for (uint32_t l = 1; l < L; ++l) base[l + 1] = base[l] + P.n[l];
  std::vector<uint32_t*> cls(L, nullptr);
  std::vector<uint32_t> capOf(L, 256);
  if (L >= 3) {
    uint32_t* keys = dalloc<uint32_t>(c, (base[L - 1] + 1) * kRecWords);
    for (uint32_t d = 1; d + 2 <= L; ++d) {
      const uint32_t top = L - 1 - d;   // deepest level with d levels below it
      const uint64_t nd = base[top + 1];
      uint64_t cap = 1024;
      while (cap < 2 * nd) cap <<= 1;
      if (cap > (1ull << 31)) throw Error{SVD_ERR_CAPACITY, "cross-level merge: too many nodes"};
      capOf[d] = (uint32_t)cap;
      cls[d] = dalloc<uint32_t>(c, nd + 1);
      for (uint32_t l = 1; l <= top; ++l)
        if (P.n[l])
          k_class_keys<<<div_up(P.n[l], kB), kB, 0, c->stream>>>(P, l, d == 1 ? nullptr : cls[d - 1] + base[l + 1],
                                                                 keys + base[l] * kRecWords);
      uint32_t* table = dalloc<uint32_t>(c, cap);
      fill_u32(c, table, kNone, cap);
      if (nd) k_class_insert<<<div_up(nd, kB), kB, 0, c->stream>>>(keys, nd, table, (uint32_t)(cap - 1), cls[d]);
      check_launch("k_class");
      dfree(c, table);
    }
    dfree(c, keys);
  }
  auto clsAt = [&](uint32_t d, uint32_t l) -> const uint32_t* { return d == 0 ? nullptr : cls[d] + base[l]; };

  if (P.n[1]) k_fill8<<<div_up(P.n[1], kB), kB, 0, c->stream>>>(front[1], 1, P.n[1]);
  uint32_t* nMatched = dalloc<uint32_t>(c, 1);
  for (uint32_t levA = 1; levA < L; ++levA) {
    const uint32_t D = L - 1 - levA;   // levels below A
    const uint32_t nA = P.n[levA];
    if (!nA) continue;
    unsigned long long* best = dalloc<unsigned long long>(c, nA);
    unsigned long long* bestOf = dalloc<unsigned long long>(c, capOf[D]);
    k_fill64<<<div_up(capOf[D], kB), kB, 0, c->stream>>>((uint64_t*)bestOf, ~0ull, capOf[D]);
    for (uint32_t levB = 1; levB < levA; ++levB)
      if (P.n[levB]) k_class_best<<<div_up(P.n[levB], kB), kB, 0, c->stream>>>(P, levB, clsAt(D, levB), bestOf);
    fill_u32(c, nMatched, 0u, 1);
    k_frontier_best<<<div_up(nA, kB), kB, 0, c->stream>>>(P, levA, front[levA], clsAt(D, levA), bestOf, best,
                                                            sa[levA], sp[levA], cand[levA], nMatched);
    check_launch("k_frontier_best");
    if (levA + 1 < L && read_u32(c, nMatched)) {
      for (uint32_t l = levA; l + 1 < L; ++l) {
        const uint32_t nc = P.n[l + 1];
        fill_u32(c, sa[l + 1], kNone, nc);
        k_fill64<<<div_up(nc + 1, kB), kB, 0, c->stream>>>(sp[l + 1], 0ull, nc + 1);
        k_fill64<<<div_up(nc + 1, kB), kB, 0, c->stream>>>(cand[l + 1], ~0ull, nc + 1);
        if (!P.n[l]) continue;
        const uint32_t g = div_up(P.n[l], kB);
        k_prop_min<<<g, kB, 0, c->stream>>>(P, l, sa[l], sa[l + 1]);
        k_prop_path<<<g, kB, 0, c->stream>>>(P, l, sa[l], sp[l], sa[l + 1], (unsigned long long*)sp[l + 1]);
        k_prop_write<<<g, kB, 0, c->stream>>>(P, l, sa[l], sp[l], cand[l], sa[l + 1], sp[l + 1], cand[l + 1]);
        check_launch("cross-level record");
      }
      for (uint32_t l = levA + 1; l < L; ++l)
        if (P.n[l]) k_merge_sub<<<div_up(P.n[l], kB), kB, 0, c->stream>>>(P.n[l], cand[l], mlc[l]);
    }
    k_merge_self<<<div_up(nA, kB), kB, 0, c->stream>>>(nA, front[levA], best, mlc[levA],
                                                        levA + 1 < L ? front[levA + 1] : nullptr, W[levA].rec,
                                                        levA == L - 1 ? 1 : 0);
    check_launch("cross-level round");
    dfree(c, best); dfree(c, bestOf);
  }
  dfree(c, nMatched);
  for (auto* p : cls) dfree(c, p);

  // compaction, bottom-up
  uint32_t* missing = dalloc<uint32_t>(c, 1);
  uint32_t* total = dalloc<uint32_t>(c, 1);
  fill_u32(c, missing, 0u, 1);
  std::vector<Level> X(L);
  X[0] = W[0];
  uint64_t nNodes = 1, merged = 0;
  for (uint32_t lev = L - 1; lev > 0; --lev) {
    const uint32_t n = P.n[lev];
    uint32_t* pos = dalloc<uint32_t>(c, n + 1);
    uint32_t* corr = dalloc<uint32_t>(c, n + 1);
    if (n) k_keep_flags<<<div_up(n, kB), kB, 0, c->stream>>>(mlc[lev], n, pos);
    scan_exclusive(c, pos, pos, n, total);
    const uint64_t keep = read_u32(c, total);
    merged += n - keep;
    Level& x = X[lev];
    x.n = keep;
    x.leaf = W[lev].leaf;
    x.rec = dalloc<uint32_t>(c, (x.leaf ? keep : keep * kRecWords) + 1);
    if (!x.leaf) x.childLevels = dalloc<uint32_t>(c, 8 * keep + 1);
    if (n) k_compact_level<<<div_up(n, kB), kB, 0, c->stream>>>(W[lev].rec, W[lev].childLevels, n, x.leaf ? 1 : 0,
                                                                mlc[lev], pos, x.rec, x.childLevels, corr);
    check_launch("k_compact_level");
    nNodes += keep;
    // parents in level lev-1: for lev-1 >= 1 they live in W (not yet compacted)
    const uint32_t np = P.n[lev - 1];
    if (np) k_fix_parents<<<div_up(np, kB), kB, 0, c->stream>>>(W[lev - 1].rec, W[lev - 1].childLevels, np,
                                                                mlc[lev], corr);
    // already compacted levels [lev, L-2] that were redirected into level lev
    for (uint32_t low = lev; low + 1 < L; ++low) {
      if (X[low].n) k_fix_redirected<<<div_up(X[low].n, kB), kB, 0, c->stream>>>(X[low].rec, X[low].childLevels,
                                                                                   (uint32_t)X[low].n, lev, corr, missing);
    }
    check_launch("k_fix");
    dfree(c, pos); dfree(c, corr);
  }
  X[0] = W[0];
  const uint64_t nMissing = read_u32(c, missing);
  // replace the cross-level view
  for (auto& lv : o->xdag) { dfree(c, lv.rec); dfree(c, lv.childLevels); }
  o->xdag = X;
  for (uint32_t l = 1; l < L; ++l) { dfree(c, W[l].rec); dfree(c, W[l].childLevels); }
  o->hasX = true;
  o->nNodesX = nNodes;
  o->crossMerged = merged;
  for (uint32_t l = 0; l < L; ++l) {
    dfree(c, mlc[l]); dfree(c, cand[l]); dfree(c, sp[l]); dfree(c, sa[l]); dfree(c, front[l]);
  }
  dfree(c, missing); dfree(c, total);
  for (auto& e : o->enc) e.valid = false;
  if (nMissing)
    throw Error{SVD_ERR_POINTER, "cross-level merge: " + std::to_string(nMissing) +
                                     " pointer(s) to a removed node (the reference prints 'Missing correspondence')"};
}

}  // namespace svd
