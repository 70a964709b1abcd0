// .svdag input: EncodedSVDAG::load + decode (encoded_svdag.cpp:43-66, 200-281).
//
// File: f32 min[3], f32 max[3], f32 rootSide, u32 levels, u32 nNodes,
// u32 firstLeafPtr, u32 count, u32 data[count].  Above the leaf level a node
// is its mask word followed by one absolute word index per child, child 7
// first; a leaf is its mask word.  The reference parses the words in order:
// a level begins at the smallest pointer seen in the level above, and a
// pointer becomes the target's node index in its level (words before the
// target minus pointer words before it).
//
// Device form, one level at a time from the root: the nodes of level l+1 are
// exactly the targets of level l's pointers (in a file written by the
// encoder every node is referenced), so marking the targets in the word range
// [smallest pointer, largest pointer] and scanning the marks numbers the nodes
// in file order.  The decoded nodes must tile every level's word range and the
// leaf level must end at the last word; a file that breaks this (unreferenced
// or overlapping words, where a sequential parse could read differently) is
// rejected with SVD_ERR_FORMAT instead of being decoded differently.
#include <algorithm>
#include "svd_internal.hpp"

namespace svd {

namespace {

constexpr int kB = 256;

// Node i of an inner level at word starts[i]: the 9-word record with raw
// (absolute word) child pointers; marks the targets; folds min/max pointer.
// Below the root a node without children never comes out of the encoder
// (buildSVO's cleanEmptyNodes and toDAG drop them), so it flags the file.
__global__ void k_dec_inner(const uint32_t* __restrict__ data, uint32_t count, const uint32_t* __restrict__ starts,
                            uint32_t n, uint32_t belowRoot, uint32_t* __restrict__ rec, uint32_t* __restrict__ mark,
                            uint32_t* __restrict__ range, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t w = starts[i];
  const uint32_t m = data[w] & 0xFFu;
  const uint32_t nc = (uint32_t)__popc(m);
  uint32_t* r = rec + (uint64_t)i * kRecWords;
  r[0] = m;
  if (belowRoot && m == 0) atomicOr(err, 4u);
  // tiling: this node ends where the next one starts
  const uint32_t end = w + 1 + nc;
  if (end > count || (i + 1 < n && starts[i + 1] != end)) { atomicOr(err, 1u); return; }
  if (i + 1 == n) range[2] = end;   // end of this level's words
  uint32_t lo = 0xFFFFFFFFu, hi = 0;
  uint32_t k = 0;
  for (int c = 7; c >= 0; --c) {
    uint32_t v = kNull;
    if ((m >> c) & 1u) {
      v = data[w + 1 + k++];
      if (v >= count) { atomicOr(err, 2u); v = kNull; }
      else {
        mark[v] = 1u;
        lo = min(lo, v);
        hi = max(hi, v);
      }
    }
    r[1 + c] = v;
  }
  if (nc) { atomicMin(&range[0], lo); atomicMax(&range[1], hi); }
}

__global__ void k_dec_leaf(const uint32_t* __restrict__ data, const uint32_t* __restrict__ starts, uint32_t n,
                           uint32_t count, uint32_t* __restrict__ rec, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t w = starts[i];
  rec[i] = data[w] & 0xFFu;
  if ((data[w] & 0xFFu) == 0) atomicOr(err, 4u);
  const uint32_t next = i + 1 < n ? starts[i + 1] : count;
  if (next != w + 1) atomicOr(err, 1u);
}

// marks[lo..hi] -> 0/1 flags for the scan
__global__ void k_dec_flags(const uint32_t* __restrict__ mark, uint32_t lo, uint32_t len, uint32_t* __restrict__ flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < len) flag[j] = mark[lo + j];
}

__global__ void k_dec_starts(const uint32_t* __restrict__ mark, const uint32_t* __restrict__ pos, uint32_t lo,
                             uint32_t len, uint32_t* __restrict__ starts) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < len && mark[lo + j]) starts[pos[j]] = lo + j;
}

// raw pointers of level l -> node indices of level l+1
__global__ void k_dec_remap(uint32_t* __restrict__ rec, uint32_t n, const uint32_t* __restrict__ pos, uint32_t lo) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* r = rec + (uint64_t)i * kRecWords;
  for (int c = 0; c < 8; ++c)
    if (r[1 + c] != kNull) r[1 + c] = pos[r[1 + c] - lo];
}

}  // namespace

void octree_from_svdag(svd_octree* o, const uint8_t* file, uint64_t bytes, bool onDevice) {
  svd_ctx* c = o->ctx;
  constexpr uint64_t kHead = 40;
  if (bytes < kHead) throw Error{SVD_ERR_FORMAT, ".svdag: file shorter than its header"};
  uint8_t head[kHead + 4];
  if (bytes < kHead + 4) throw Error{SVD_ERR_FORMAT, ".svdag: file shorter than its header"};
  if (onDevice) {
    SVD_HIP(hipMemcpyAsync(head, file, kHead + 4, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
  } else {
    memcpy(head, file, kHead + 4);
  }
  float bb[6], rootSide;
  uint32_t levels, nNodes, firstLeaf, count;
  memcpy(bb, head, 24);
  memcpy(&rootSide, head + 24, 4);
  memcpy(&levels, head + 28, 4);
  memcpy(&nNodes, head + 32, 4);
  memcpy(&firstLeaf, head + 36, 4);
  (void)firstLeaf;
  memcpy(&count, head + kHead, 4);
  if (levels < 2 || levels > (uint32_t)kMaxLevels) throw Error{SVD_ERR_LEVELS, ".svdag: unsupported level count"};
  if (bytes < kHead + 4 + 4ull * count || count == 0) throw Error{SVD_ERR_FORMAT, ".svdag: truncated word array"};
  uint32_t* data = dalloc<uint32_t>(c, (uint64_t)count + 1);
  SVD_HIP(hipMemcpyAsync(data, file + kHead + 4, 4ull * count, onDevice ? hipMemcpyDeviceToDevice
                                                                          : hipMemcpyHostToDevice, c->stream));
  uint32_t* mark = dalloc<uint32_t>(c, (uint64_t)count + 1);
  fill_u32(c, mark, 0u, count + 1);
  uint32_t* dev = dalloc<uint32_t>(c, 4);   // range lo, hi, end of level, error flags
  std::vector<Level> lv(levels);
  uint32_t* starts = dalloc<uint32_t>(c, 2);
  fill_u32(c, starts, 0u, 1);               // level 0: the root at word 0
  uint64_t n = 1;
  auto fail = [&](const std::string& what) {
    for (auto& x : lv) dfree(c, x.rec);
    dfree(c, data); dfree(c, mark); dfree(c, dev); dfree(c, starts);
    throw Error{SVD_ERR_FORMAT, ".svdag: " + what};
  };
  for (uint32_t l = 0; l < levels; ++l) {
    lv[l].n = n;
    lv[l].leaf = (l == levels - 1);
    uint32_t h[4] = {0xFFFFFFFFu, 0u, 0u, 0u};
    SVD_HIP(hipMemcpyAsync(dev, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    if (lv[l].leaf) {
      lv[l].rec = dalloc<uint32_t>(c, n + 1);
      k_dec_leaf<<<div_up(n, kB), kB, 0, c->stream>>>(data, starts, (uint32_t)n, count, lv[l].rec, dev + 3);
      check_launch("k_dec_leaf");
      read_u32s(c, dev, h, 4);
      if (h[3] & 4u) fail("empty leaf node in level " + std::to_string(l));
      if (h[3]) fail("leaf words do not end the file");
      break;
    }
    lv[l].rec = dalloc<uint32_t>(c, n * kRecWords + 1);
    k_dec_inner<<<div_up(n, kB), kB, 0, c->stream>>>(data, count, starts, (uint32_t)n, l > 0 ? 1u : 0u, lv[l].rec,
                                                     mark, dev, dev + 3);
    check_launch("k_dec_inner");
    read_u32s(c, dev, h, 4);
    if (h[3] & 4u) fail("node without children below the root in level " + std::to_string(l));
    if (h[3] & 2u) fail("child pointer outside the word array");
    if (h[3]) fail("nodes do not tile level " + std::to_string(l));
    if (h[0] == 0xFFFFFFFFu) fail("level " + std::to_string(l) + " has no children");
    if (h[0] != h[2]) fail("level " + std::to_string(l + 1) + " does not start where level " + std::to_string(l) +
                           " ends");
    const uint32_t lo = h[0], len = h[1] - h[0] + 1;
    uint32_t* pos = dalloc<uint32_t>(c, (uint64_t)len + 1);
    uint32_t* total = dalloc<uint32_t>(c, 1);
    k_dec_flags<<<div_up(len, kB), kB, 0, c->stream>>>(mark, lo, len, pos);
    scan_exclusive(c, pos, pos, len, total);
    const uint64_t nNext = read_u32(c, total);
    dfree(c, starts);
    starts = dalloc<uint32_t>(c, nNext + 1);
    k_dec_starts<<<div_up(len, kB), kB, 0, c->stream>>>(mark, pos, lo, len, starts);
    k_dec_remap<<<div_up(n, kB), kB, 0, c->stream>>>(lv[l].rec, (uint32_t)n, pos, lo);
    check_launch("k_dec_remap");
    dfree(c, pos);
    dfree(c, total);
    n = nNext;
  }
  SVD_HIP(hipStreamSynchronize(c->stream));
  dfree(c, data); dfree(c, mark); dfree(c, dev); dfree(c, starts);
  o->levels = levels;
  for (int i = 0; i < 6; ++i) o->bbox[i] = bb[i];
  o->rootSide = rootSide;
  o->dag = lv;
  o->nNodesDAG = nNodes;   // the header's count (GeomOctree(NodeData, ...) takes stats.nNodesDAG)
  o->nVoxels = 0;          // EncodedSVDAG::load leaves _nVoxels at 0 and decode copies it
  o->state = SVD_S_DAG;
}

}  // namespace svd
