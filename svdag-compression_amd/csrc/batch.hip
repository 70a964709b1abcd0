// Batched build of many scenes in shared launches (uneven depths), plus the
// SVDAG images of all of them, and multi-camera traversal over the batch.
//
// Every scene's sorted Morton codes are tagged with the scene id above bit
// 40 (as the leaf pass reads them) and the tagged list (ascending: scene-major) is built as ONE octree of
// depth Lmax = max depth: a node's code carries its scene id, so runs never
// cross scenes and every level is scene-major with each scene's nodes in its
// own first-arrival order.  A scene of depth L_s < Lmax occupies batched
// levels Lmax-L_s .. Lmax-1 (its root is the single node of its id at level
// Lmax-L_s); the levels above are a chain of one-child nodes that belong to
// no scene and are ignored.  Dedup per level: inner nodes of different
// scenes never share a key (their children lie in disjoint id ranges), so the
// single-scene hash dedup applies unchanged; leaf masks are keyed by
// (scene, mask).  The result for every scene equals its own build bit for bit.
#include "svd_internal.hpp"
#include <algorithm>

namespace svd {

namespace {

constexpr int kB = 256;

// start[s] = first index of scene s in a scene-major array (kNone if absent)
__global__ void k_scene_starts(const uint32_t* __restrict__ sid, uint64_t n, uint32_t* __restrict__ start) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || sid[i] != sid[i - 1]) start[sid[i]] = (uint32_t)i;
}

// Each block folds a contiguous chunk into an LDS table for the chunk's first
// scene (the array is scene-major, so other scenes appear only at chunk
// boundaries and go straight to the global table).
constexpr int kFoBlocks = 4096;
__global__ void __launch_bounds__(256) k_leaf_firstocc_scene(const uint32_t* __restrict__ mask,
                                                             const uint32_t* __restrict__ sid, uint64_t n,
                                                             uint32_t* __restrict__ firstocc) {
  __shared__ uint32_t fo[256];
  fo[threadIdx.x] = kNone;
  __syncthreads();
  const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  const uint32_t s0 = lo < n ? sid[lo] : 0u;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const uint32_t m = mask[i] & 0xFFu, s = sid[i];
    if (s == s0) {
      if (fo[m] > (uint32_t)i) atomicMin(&fo[m], (uint32_t)i);
    } else {
      atomicMin(&firstocc[(uint64_t)s * 256 + m], (uint32_t)i);
    }
  }
  __syncthreads();
  const uint32_t v = fo[threadIdx.x];
  if (lo < n && v != kNone) atomicMin(&firstocc[(uint64_t)s0 * 256 + threadIdx.x], v);
}

// Per scene (one workgroup each): its distinct leaf masks ranked by first
// occurrence, rank[s][m] (kNone for a mask the scene lacks), and their count.
__global__ void __launch_bounds__(256) k_leaf_rank_scene(const uint32_t* __restrict__ firstocc,
                                                         uint32_t* __restrict__ rank, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t fo[256];
  __shared__ uint32_t total;
  const uint32_t m = threadIdx.x;
  const uint64_t row = (uint64_t)blockIdx.x * 256;
  fo[m] = firstocc[row + m];
  if (m == 0) total = 0;
  __syncthreads();
  uint32_t id = kNone;
  if (fo[m] != kNone) {
    uint32_t r = 0;
    for (int k = 0; k < 256; ++k) r += fo[k] < fo[m] ? 1u : 0u;
    id = r;
    atomicAdd(&total, 1u);
  }
  rank[row + m] = id;
  __syncthreads();
  if (m == 0) cnt[blockIdx.x] = total;
}

// The batched leaf level's dedup: leaf i of scene s with mask m becomes DAG
// leaf base[s] + rank[s][m] (scene-major, first occurrence within the scene:
// the order of first occurrence over the whole level); the first occurrence
// writes the mask.
__global__ void k_leaf_remap_scene(const uint32_t* __restrict__ mask, const uint32_t* __restrict__ sid, uint64_t n,
                                   const uint32_t* __restrict__ firstocc, const uint32_t* __restrict__ rank,
                                   const uint32_t* __restrict__ base, uint32_t* __restrict__ out,
                                   uint32_t* __restrict__ remap) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = mask[i] & 0xFFu, s = sid[i];
  const uint64_t e = (uint64_t)s * 256 + m;
  const uint32_t id = base[s] + rank[e];
  remap[i] = id;
  if (firstocc[e] == (uint32_t)i) out[id] = m;
}

// Scene of each DAG node from the per-scene start table of its level.
__global__ void k_sid_from_starts(const uint32_t* __restrict__ start, uint32_t S, uint64_t n, uint32_t* __restrict__ sid) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = S;                   // last s with start[s] <= i (absent scenes hold kNone)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] != kNone && start[mid] <= i) lo = mid; else hi = mid;
  }
  sid[i] = lo;
}

// SVDAG word counts of one batched level (0 outside the node's scene).
__global__ void k_batch_word_counts(const uint32_t* __restrict__ rec, uint64_t n, int leaf,
                                    const uint32_t* __restrict__ sid, const uint32_t* __restrict__ firstLev,
                                    uint32_t lev, uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (lev < firstLev[sid[i]]) { out[i] = 0; return; }
  out[i] = leaf ? 1u : 1u + (uint32_t)__popc(rec[i * kRecWords] & 0xFFu);
}

// EncodedSVDAG words (encoded_svdag.cpp:119-186) of every scene at once:
// node i of batched level lev (scene s) goes to word
//   fileWord[s] + base[lev][s] + scan[lev][i] - scanStart[lev][s]
// and its child c (level lev+1) is addressed by the scene-local word index
//   base[lev+1][s] + scan[lev+1][c] - scanStart[lev+1][s].
__global__ void k_batch_write_words(const uint32_t* __restrict__ rec, uint64_t n, int leaf, uint32_t lev, uint32_t S,
                                    const uint32_t* __restrict__ sid, const uint32_t* __restrict__ firstLev,
                                    const uint32_t* __restrict__ scan, const uint32_t* __restrict__ scanNext,
                                    const uint64_t* __restrict__ base, const uint32_t* __restrict__ scanStart,
                                    const uint64_t* __restrict__ fileWord, uint32_t* __restrict__ words) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = sid[i];
  if (lev < firstLev[s]) return;
  const uint64_t at = fileWord[s] + base[(uint64_t)lev * S + s] + scan[i] - scanStart[(uint64_t)lev * S + s];
  if (leaf) { words[at] = rec[i] & 0xFFu; return; }
  const uint32_t* r = rec + i * kRecWords;
  const uint32_t m = r[0] & 0xFFu;
  words[at] = m;
  uint64_t k = at + 1;
  const uint64_t nb = base[(uint64_t)(lev + 1) * S + s];
  const uint32_t ns = scanStart[(uint64_t)(lev + 1) * S + s];
#pragma unroll
  for (int c = 7; c >= 0; --c) {
    if (!((m >> c) & 1u)) continue;
    words[k++] = (uint32_t)(nb + scanNext[r[1 + c]] - ns);
  }
}

__global__ void k_gather_starts(const uint32_t* __restrict__ remap, const uint32_t* __restrict__ svoStart, uint32_t S,
                                uint32_t* __restrict__ dagStart) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) dagStart[s] = svoStart[s] == kNone ? kNone : remap[svoStart[s]];
}

// scan value at each scene's first node and one past its last node
__global__ void k_gather_scan(const uint32_t* __restrict__ scan, const uint32_t* __restrict__ total, uint64_t m,
                              const uint32_t* __restrict__ start, uint32_t S, uint32_t* __restrict__ outStart,
                              uint32_t* __restrict__ outEnd) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t st = start[s];
  const uint64_t en = s + 1 < S ? start[s + 1] : m;
  outStart[s] = st < m ? scan[st] : *total;
  outEnd[s] = en < m ? scan[en] : *total;
}

__global__ void k_scatter_headers(const uint32_t* __restrict__ hdr, uint32_t S, const uint64_t* __restrict__ fileWord,
                                  uint32_t* __restrict__ words) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S * 11) return;
  const uint32_t s = t / 11, w = t % 11;
  words[fileWord[s] - 11 + w] = hdr[t];
}

}  // namespace

void batch_build(svd_batch* b, const uint64_t* dcodes, uint64_t n, const std::vector<uint64_t>& off) {
  svd_ctx* c = b->ctx;
  const uint32_t S = b->S, Lm = b->Lmax;
  // ---- SVO of the whole batch: the codes are tagged with their scene id as
  // the leaf pass reads them, and checked there
  uint64_t* dOff = dalloc<uint64_t>(c, S + 1);
  uint32_t* dLev = dalloc<uint32_t>(c, S);
  SVD_HIP(hipMemcpyAsync(dOff, off.data(), 8 * (S + 1), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(dLev, b->L.data(), 4 * S, hipMemcpyHostToDevice, c->stream));
  std::vector<Level> svo;
  std::vector<uint32_t*> sid;   // scene of every SVO node, per level
  const SceneTags tags{dOff, dLev, nullptr, S};
  try {
    svo_levels_from_codes(c, dcodes, n, Lm, svo, &sid, 0, &tags);
  } catch (...) {
    dfree(c, dOff); dfree(c, dLev);
    throw;
  }
  // first node of every scene, per level
  uint32_t* startTab = dalloc<uint32_t>(c, (uint64_t)Lm * S);
  fill_u32(c, startTab, kNone, (uint64_t)Lm * S);
  for (uint32_t lev = 0; lev < Lm; ++lev) {
    const uint64_t m = svo[lev].n;
    k_scene_starts<<<div_up(m, kB), kB, 0, c->stream>>>(sid[lev], m, startTab + (uint64_t)lev * S);
  }
  check_launch("scene tables");
  // ---- DAG: leaf masks keyed by (scene, mask), inner levels by the plain key
  std::vector<Level>& dag = b->dag;
  dag.assign(Lm, Level());
  uint32_t* remap = nullptr;
  {
    const Level& sl = svo[Lm - 1];
    const uint64_t m = sl.n;
    // first occurrence of every (scene, mask), the masks of each scene ranked,
    // the scenes' counts scanned; the level's output holds at most 255 masks
    // per scene, so it is allocated before the count is read
    uint32_t* fo = dalloc<uint32_t>(c, (uint64_t)S * 256);
    uint32_t* rank = dalloc<uint32_t>(c, (uint64_t)S * 256);
    uint32_t* base = dalloc<uint32_t>(c, S + 1);
    uint32_t* total = dalloc<uint32_t>(c, 1);
    fill_u32(c, fo, kNone, (uint64_t)S * 256);
    if (m) {
      const uint64_t nb = div_up(m, 256) < (uint64_t)kFoBlocks ? div_up(m, 256) : (uint64_t)kFoBlocks;
      k_leaf_firstocc_scene<<<nb, 256, 0, c->stream>>>(sl.rec, sid[Lm - 1], m, fo);
    }
    k_leaf_rank_scene<<<S, 256, 0, c->stream>>>(fo, rank, base);
    scan_exclusive(c, base, base, S, total);
    Level& dl = dag[Lm - 1];
    dl.leaf = true;
    dl.rec = dalloc<uint32_t>(c, std::min<uint64_t>(m, 255ull * S) + 1);
    remap = dalloc<uint32_t>(c, m + 1);
    if (m)
      k_leaf_remap_scene<<<div_up(m, kB), kB, 0, c->stream>>>(sl.rec, sid[Lm - 1], m, fo, rank, base, dl.rec, remap);
    check_launch("batched leaf dedup");
    dl.n = read_u32(c, total);
    dfree(c, fo); dfree(c, rank); dfree(c, base); dfree(c, total);
  }
  // per-scene first DAG node per level = remap of the scene's first SVO node
  // (every scene has nodes on every batched level: its chain above the root)
  uint32_t* dStart = dalloc<uint32_t>(c, (uint64_t)Lm * S);
  auto scene_starts = [&](uint32_t lev, const uint32_t* rmap) {
    k_gather_starts<<<div_up(S, kB), kB, 0, c->stream>>>(rmap, startTab + (uint64_t)lev * S, S,
                                                         dStart + (uint64_t)lev * S);
  };
  scene_starts(Lm - 1, remap);
  for (int lev = (int)Lm - 2; lev >= 0; --lev) {
    const Level& s = svo[lev];
    uint32_t* keyed = dalloc<uint32_t>(c, (s.n + 1) * kRecWords);
    remap_children(c, s.rec, s.n, remap, keyed);
    uint32_t* nremap = dalloc<uint32_t>(c, s.n + 1);
    dag[lev].n = dedup_level(c, keyed, s.n, nremap, &dag[lev].rec);
    dfree(c, keyed);
    dfree(c, remap);
    remap = nremap;
    scene_starts((uint32_t)lev, remap);
  }
  check_launch("scene starts");
  dfree(c, remap);
  for (auto& lv : svo) dfree(c, lv.rec);
  for (auto* p : sid) dfree(c, p);
  dfree(c, startTab);
  std::vector<uint32_t> dagStart((uint64_t)Lm * S);
  read_u32s(c, dStart, dagStart.data(), (uint64_t)Lm * S);
  // ---- per-scene counts (levels of scene s: batched Lm-L_s .. Lm-1)
  std::vector<uint32_t> firstLev(S);
  for (uint32_t s = 0; s < S; ++s) firstLev[s] = Lm - b->L[s];
  uint32_t* dFirst = dalloc<uint32_t>(c, S);
  SVD_HIP(hipMemcpyAsync(dFirst, firstLev.data(), 4 * S, hipMemcpyHostToDevice, c->stream));
  b->nodes.assign(S, 1);
  for (uint32_t lev = 0; lev < Lm; ++lev)
    for (uint32_t s = 0; s < S; ++s) {
      const uint64_t st = dagStart[(uint64_t)lev * S + s];
      const uint64_t en = s + 1 < S ? dagStart[(uint64_t)lev * S + s + 1] : dag[lev].n;
      if (st == kNone || en < st) throw Error{SVD_ERR_DEVICE, "batched build: scene table out of order"};
      if (lev > firstLev[s]) b->nodes[s] += en - st;     // 1 + sum of the non-root levels
    }
  // ---- SVDAG images: word counts and per-level scans
  std::vector<uint32_t*> scans(Lm, nullptr), sidDag(Lm, nullptr);
  uint32_t* totals = dalloc<uint32_t>(c, Lm);
  uint32_t* dScanStart = dalloc<uint32_t>(c, (uint64_t)Lm * S);
  uint32_t* dScanEnd = dalloc<uint32_t>(c, (uint64_t)Lm * S);
  for (uint32_t lev = 0; lev < Lm; ++lev) {
    const uint64_t m = dag[lev].n;
    scans[lev] = dalloc<uint32_t>(c, m + 1);
    sidDag[lev] = dalloc<uint32_t>(c, m + 1);
    k_sid_from_starts<<<div_up(m, kB), kB, 0, c->stream>>>(dStart + (uint64_t)lev * S, S, m, sidDag[lev]);
    k_batch_word_counts<<<div_up(m, kB), kB, 0, c->stream>>>(dag[lev].rec, m, lev == Lm - 1, sidDag[lev], dFirst, lev,
                                                             scans[lev]);
    check_launch("k_batch_word_counts");
    scan_exclusive(c, scans[lev], scans[lev], m, totals + lev);
    k_gather_scan<<<div_up(S, kB), kB, 0, c->stream>>>(scans[lev], totals + lev, m, dStart + (uint64_t)lev * S, S,
                                                       dScanStart + (uint64_t)lev * S, dScanEnd + (uint64_t)lev * S);
  }
  check_launch("k_gather_scan");
  std::vector<uint32_t> scanStart((uint64_t)Lm * S), scanEnd((uint64_t)Lm * S);
  read_u32s(c, dScanStart, scanStart.data(), scanStart.size());
  read_u32s(c, dScanEnd, scanEnd.data(), scanEnd.size());
  // scene file layout: 11 header words, then levels in order
  std::vector<uint64_t> base((uint64_t)(Lm + 1) * S, 0), fileWord(S);
  std::vector<uint32_t> hdr((uint64_t)S * 11);
  b->fileOff.assign(S, 0);
  b->fileBytes.assign(S, 0);
  uint64_t at = 0;
  for (uint32_t s = 0; s < S; ++s) {
    uint64_t w = 0;
    for (uint32_t lev = firstLev[s]; lev < Lm; ++lev) {
      base[(uint64_t)lev * S + s] = w;
      w += scanEnd[(uint64_t)lev * S + s] - scanStart[(uint64_t)lev * S + s];
    }
    b->fileOff[s] = at;
    b->fileBytes[s] = 44 + 4 * w;
    fileWord[s] = at / 4 + 11;
    at += (44 + 4 * w + 255) & ~255ull;
    // header (encoded_svdag.cpp:88-98): bbox, rootSide, levels, nNodes, firstLeafPtr, nWords
    const float bb[7] = {0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 1.f};
    uint32_t* h = &hdr[(uint64_t)s * 11];
    memcpy(h, bb, 28);
    h[7] = b->L[s];
    h[8] = (uint32_t)b->nodes[s];
    h[9] = (uint32_t)w;          // _firstLeafPtr = the total word count (assigned after the level loop)
    h[10] = (uint32_t)w;
  }
  b->files = dalloc<uint8_t>(c, at + 256);
  uint32_t* words = (uint32_t*)b->files;
  uint64_t* dBase = dalloc<uint64_t>(c, base.size());
  uint64_t* dFileWord = dalloc<uint64_t>(c, S);
  uint32_t* dHdr = dalloc<uint32_t>(c, hdr.size());
  SVD_HIP(hipMemcpyAsync(dBase, base.data(), 8 * base.size(), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(dFileWord, fileWord.data(), 8 * S, hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(dHdr, hdr.data(), 4 * hdr.size(), hipMemcpyHostToDevice, c->stream));
  k_scatter_headers<<<div_up((uint64_t)S * 11, kB), kB, 0, c->stream>>>(dHdr, S, dFileWord, words);
  for (uint32_t lev = 0; lev < Lm; ++lev) {
    const uint64_t m = dag[lev].n;
    if (!m) continue;
    k_batch_write_words<<<div_up(m, kB), kB, 0, c->stream>>>(dag[lev].rec, m, lev == Lm - 1, lev, S, sidDag[lev],
                                                             dFirst, scans[lev], lev + 1 < Lm ? scans[lev + 1] : nullptr,
                                                             dBase, dScanStart, dFileWord, words);
  }
  check_launch("k_batch_write_words");
  SVD_HIP(hipStreamSynchronize(c->stream));
  for (auto* p : scans) dfree(c, p);
  for (auto* p : sidDag) dfree(c, p);
  dfree(c, totals); dfree(c, dStart); dfree(c, dFirst); dfree(c, dBase); dfree(c, dScanStart); dfree(c, dScanEnd);
  dfree(c, dFileWord);
  dfree(c, dHdr); dfree(c, dOff); dfree(c, dLev);
  // scenes for traversal
  b->scenes.assign(S, svd_scene());
  for (uint32_t s = 0; s < S; ++s) {
    svd_scene& sc = b->scenes[s];
    sc.ctx = c;
    sc.fmt = SVD_FMT_SVDAG;
    sc.levels = b->L[s];
    for (int a = 0; a < 3; ++a) { sc.bmin[a] = 0.f; sc.bmax[a] = 1.f; sc.center[a] = 0.5f; }
    sc.rootHalfSide = 0.5f;
    sc.nodes = (const uint32_t*)(b->files + b->fileOff[s] + 44);
    sc.nWords = (uint32_t)((b->fileBytes[s] - 44) / 4);
  }
}

}  // namespace svd
