// CSVDAG cross-level merge on the device (geom_octree_extension.cpp:1192-1544).
//
// The reference walks a top-down frontier: every frontier node A of level levA
// is compared (compareSubtrees, :238-290) with the nodes of the levels above
// that share its depth-d hash (computeNodeHash, :34-74); the first equal one in
// (levB, index) order wins, A's whole subtree is recorded as corresponding to
// B's, and unmatched nodes send their children to the next frontier.  Then the
// levels are compacted bottom-up and pointers redirected (:1428-1498).
//
// Here, per level: all (frontier node, candidate) pairs with equal hashes are
// compared in parallel and the first match is an atomicMin over
// (levB << 32 | index); subtree correspondences get the priority the
// reference's map semantics give them (earlier match wins between matches,
// last pre-order visit wins inside one comparison); the compaction runs the
// reference's two pointer-update loops level by level.
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

// computeNodeHash with childLevels = lev+1 (std::hash<uint64_t> is the identity).
__global__ void k_hash(LevelPtrs P, uint32_t lev, uint32_t depth, const uint64_t* __restrict__ childHash,
                       uint64_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n[lev]) return;
  uint64_t key = 0;
  if (depth == 0) {
    key = node_mask(P, lev, i);
  } else if (depth == 1) {
    for (int c = 0; c < 8; ++c) {
      key <<= 8;
      const uint32_t ch = node_child(P, lev, i, c);
      if (ch != kNull) key += node_mask(P, lev + 1, ch);
    }
  } else {
    for (int c = 0; c < 8; ++c) {
      const uint64_t orig = key;
      uint64_t h = orig << (c + 1);
      const uint32_t ch = node_child(P, lev, i, c);
      if (ch != kNull) h = childHash[ch];
      key = (orig << 1) ^ h;
      key = key ^ orig;
    }
  }
  out[i] = key;
}

// Chained table of the frontier nodes of level levA, keyed by hash.
__global__ void k_frontier_table(const uint8_t* __restrict__ front, uint32_t n, const uint64_t* __restrict__ hash,
                                 uint32_t* __restrict__ head, uint32_t cmask, uint32_t* __restrict__ next) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !front[i]) return;
  const uint32_t b = (uint32_t)mix64(hash[i]) & cmask;
  next[i] = atomicExch(&head[b], i);
}

// compareSubtrees (geom_octree_extension.cpp:238-290) as an iterative DFS.
// record = 0: only decide equality.  record = 1: also emit, in pre-order,
// every (A-descendant -> B-descendant) correspondence with its visit number.
struct Frame { uint32_t a, b, c; };

template <bool kRecord>
__device__ bool compare_subtrees(const LevelPtrs& P, uint32_t L, uint32_t levA, uint32_t a, uint32_t levB, uint32_t b,
                                 uint64_t idA, unsigned long long* const* prio, uint32_t round) {
  // depth of the walk is at most L-1-levA <= 20
  Frame st[kMaxL];
  int sp = 0;
  uint32_t visit = 0;
  auto check_node = [&](uint32_t la, uint32_t x, uint32_t lb, uint32_t y) -> int {
    // returns 0 = unequal, 1 = equal leaf (stop), 2 = descend
    const uint32_t ma = node_mask(P, la, x), mb = node_mask(P, lb, y);
    if (ma != 0 && mb == 0) return 0;
    if (la == L - 1) return ma == mb ? 1 : 0;
    return 2;
  };
  int r = check_node(levA, a, levB, b);
  if (r != 2) return r == 1;
  st[sp++] = Frame{a, b, 0};
  uint32_t la = levA, lb = levB;
  while (sp > 0) {
    Frame& f = st[sp - 1];
    if (f.c >= 8) { --sp; --la; --lb; continue; }
    const int i = (int)f.c++;
    const uint32_t ma = node_mask(P, la, f.a), mb = node_mask(P, lb, f.b);
    const bool ea = (ma >> i) & 1u, eb = (mb >> i) & 1u;
    if (ea != eb) return false;
    if (!ea) continue;
    const uint32_t ca = node_child(P, la, f.a, i), cb = node_child(P, lb, f.b, i);
    if (kRecord) {
      // operator[] inside one comparison: the last visit wins -> larger visit number wins
      const unsigned long long pr = ((unsigned long long)idA << 32) | (0xFFFFFFFFu - visit);
      atomicMin(&prio[la + 1][ca], pr);
      (void)round;
    }
    ++visit;
    r = check_node(la + 1, ca, lb + 1, cb);
    if (r == 0) return false;
    if (r == 2) { st[sp++] = Frame{ca, cb, 0}; ++la; ++lb; }
  }
  return true;
}

// For every node B of level levB: walk the chain of frontier nodes with the
// same hash; an equal subtree lowers best[A] to (levB, B).
__global__ void k_match(LevelPtrs P, uint32_t L, uint32_t levA, uint32_t levB, const uint64_t* __restrict__ hashB,
                        const uint64_t* __restrict__ hashA, const uint32_t* __restrict__ head, uint32_t cmask,
                        const uint32_t* __restrict__ next, unsigned long long* __restrict__ best) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.n[levB]) return;
  const uint64_t k = hashB[j];
  uint32_t a = head[(uint32_t)mix64(k) & cmask];
  while (a != kNone) {
    if (hashA[a] == k) {
      const unsigned long long cand = ((unsigned long long)levB << 32) | j;
      if (cand < best[a] && compare_subtrees<false>(P, L, levA, a, levB, j, 0, nullptr, 0)) atomicMin(&best[a], cand);
    }
    a = next[a];
  }
}

// Matched frontier nodes record their subtree correspondences (priority pass).
__global__ void k_record_prio(LevelPtrs P, uint32_t L, uint32_t levA, const uint8_t* __restrict__ front,
                              const unsigned long long* __restrict__ best, unsigned long long* const* prio) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= P.n[levA] || !front[a] || best[a] == ~0ull) return;
  const uint32_t levB = (uint32_t)(best[a] >> 32), b = (uint32_t)best[a];
  compare_subtrees<true>(P, L, levA, a, levB, b, a, prio, 0);
}

// Second pass: the winning visit writes its correspondence (levB', index).
__global__ void k_record_write(LevelPtrs P, uint32_t L, uint32_t levA, const uint8_t* __restrict__ front,
                               const unsigned long long* __restrict__ best, unsigned long long* const* prio,
                               uint64_t* const* cand) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= P.n[levA] || !front[a] || best[a] == ~0ull) return;
  const uint32_t levB = (uint32_t)(best[a] >> 32), b = (uint32_t)best[a];
  Frame st[kMaxL];
  int sp = 0;
  uint32_t visit = 0, la = levA, lb = levB;
  if (levA == L - 1) return;
  st[sp++] = Frame{a, b, 0};
  while (sp > 0) {
    Frame& f = st[sp - 1];
    if (f.c >= 8) { --sp; --la; --lb; continue; }
    const int i = (int)f.c++;
    if (!((node_mask(P, la, f.a) >> i) & 1u)) continue;
    const uint32_t ca = node_child(P, la, f.a, i), cb = node_child(P, lb, f.b, i);
    const unsigned long long pr = ((unsigned long long)a << 32) | (0xFFFFFFFFu - visit);
    if (prio[la + 1][ca] == pr) cand[la + 1][ca] = ((uint64_t)(lb + 1) << 32) | cb;
    ++visit;
    if (la + 1 < L - 1) { st[sp++] = Frame{ca, cb, 0}; ++la; ++lb; }
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
  std::vector<uint64_t*> hash(L, nullptr), hashTmp(L, nullptr), mlc(L, nullptr), cand(L, nullptr);
  std::vector<unsigned long long*> prio(L, nullptr);
  std::vector<uint8_t*> front(L, nullptr);
  for (uint32_t l = 0; l < L; ++l) {
    const uint64_t n = W[l].n + 1;
    hash[l] = dalloc<uint64_t>(c, n);
    hashTmp[l] = dalloc<uint64_t>(c, n);
    mlc[l] = dalloc<uint64_t>(c, n);
    cand[l] = dalloc<uint64_t>(c, n);
    prio[l] = dalloc<unsigned long long>(c, n);
    front[l] = dalloc<uint8_t>(c, n);
    k_fill64<<<div_up(n, kB), kB, 0, c->stream>>>(mlc[l], ~0ull, n);
    k_fill8<<<div_up(n, kB), kB, 0, c->stream>>>(front[l], 0, n);
  }
  unsigned long long** dPrio = dalloc<unsigned long long*>(c, L);
  uint64_t** dCand = dalloc<uint64_t*>(c, L);
  SVD_HIP(hipMemcpyAsync(dPrio, prio.data(), sizeof(void*) * L, hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(dCand, cand.data(), sizeof(void*) * L, hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));   // pointer tables come from host vectors

  auto compute_hashes = [&](uint32_t depth) {
    // hash_d(level l) needs hash_{d-1}(level l+1): sweep d = 0..depth
    for (uint32_t d = 0; d <= depth; ++d) {
      for (uint32_t l = 1; l < L; ++l) {
        if (!P.n[l]) continue;
        const uint64_t* ch = (l + 1 < L) ? hash[l + 1] : nullptr;
        k_hash<<<div_up(P.n[l], kB), kB, 0, c->stream>>>(P, l, d, ch, hashTmp[l]);
      }
      std::swap(hash, hashTmp);
    }
    check_launch("k_hash");
  };

  uint32_t curDepth = L / 2;
  compute_hashes(curDepth);
  if (P.n[1]) k_fill8<<<div_up(P.n[1], kB), kB, 0, c->stream>>>(front[1], 1, P.n[1]);
  unsigned long long* best = nullptr;
  uint32_t* head = nullptr;
  uint32_t* nextA = nullptr;
  for (uint32_t levA = 1; levA < L; ++levA) {
    const uint32_t maxDepth = L - levA - 1;
    if (maxDepth < curDepth) { curDepth = maxDepth; compute_hashes(curDepth); }
    const uint32_t nA = P.n[levA];
    if (!nA) continue;
    uint32_t cap = 1024;
    while (cap < 2 * nA) cap <<= 1;
    best = dalloc<unsigned long long>(c, nA);
    head = dalloc<uint32_t>(c, cap);
    nextA = dalloc<uint32_t>(c, nA);
    k_fill64<<<div_up(nA, kB), kB, 0, c->stream>>>((uint64_t*)best, ~0ull, nA);
    fill_u32(c, head, kNone, cap);
    k_frontier_table<<<div_up(nA, kB), kB, 0, c->stream>>>(front[levA], nA, hash[levA], head, cap - 1, nextA);
    for (uint32_t levB = 1; levB < levA; ++levB) {
      if (!P.n[levB]) continue;
      k_match<<<div_up(P.n[levB], kB), kB, 0, c->stream>>>(P, L, levA, levB, hash[levB], hash[levA], head, cap - 1,
                                                             nextA, best);
    }
    check_launch("k_match");
    if (levA + 1 < L) {
      for (uint32_t l = levA + 1; l < L; ++l) {
        const uint64_t n = P.n[l] + 1;
        k_fill64<<<div_up(n, kB), kB, 0, c->stream>>>((uint64_t*)prio[l], ~0ull, n);
        k_fill64<<<div_up(n, kB), kB, 0, c->stream>>>(cand[l], ~0ull, n);
      }
      k_record_prio<<<div_up(nA, kB), kB, 0, c->stream>>>(P, L, levA, front[levA], best, dPrio);
      k_record_write<<<div_up(nA, kB), kB, 0, c->stream>>>(P, L, levA, front[levA], best, dPrio, dCand);
      for (uint32_t l = levA + 1; l < L; ++l)
        if (P.n[l]) k_merge_sub<<<div_up(P.n[l], kB), kB, 0, c->stream>>>(P.n[l], cand[l], mlc[l]);
    }
    k_merge_self<<<div_up(nA, kB), kB, 0, c->stream>>>(nA, front[levA], best, mlc[levA],
                                                        levA + 1 < L ? front[levA + 1] : nullptr, W[levA].rec,
                                                        levA == L - 1 ? 1 : 0);
    check_launch("cross-level round");
    dfree(c, best); dfree(c, head); dfree(c, nextA);
  }

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
    dfree(c, hash[l]); dfree(c, hashTmp[l]); dfree(c, mlc[l]); dfree(c, cand[l]); dfree(c, prio[l]); dfree(c, front[l]);
  }
  dfree(c, dPrio); dfree(c, dCand); dfree(c, missing); dfree(c, total);
  for (auto& e : o->enc) e.valid = false;
  if (nMissing)
    throw Error{SVD_ERR_POINTER, "cross-level merge: " + std::to_string(nMissing) +
                                     " pointer(s) to a removed node (the reference prints 'Missing correspondence')"};
}

}  // namespace svd
