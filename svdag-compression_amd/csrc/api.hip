// C ABI of libsvdag_amd: handle management, error mapping, host <-> device
// transfers.  Every entry point converts internal svd::Error exceptions into
// an int status plus a message stored in the context.
#include <cstdlib>
#include "svd_internal.hpp"
#include <algorithm>
#include <chrono>
#include <new>

namespace svd {

static size_t size_class(size_t b) {
  if (b <= 4096) return (b + 255) & ~size_t(255);
  // 8 classes per power of two above 4 KiB
  size_t p = 1;
  while ((p << 1) <= b) p <<= 1;
  const size_t step = p / 8;
  return (b + step - 1) / step * step;
}

Arena::~Arena() {
  trim();
  for (auto& kv : live_) (void)hipFree(kv.first);
  live_.clear();
}

void* Arena::get(size_t bytes) {
  const size_t cls = size_class(bytes);
  auto it = free_.find(cls);
  if (it != free_.end()) {
    void* p = it->second;
    free_.erase(it);
    live_[p] = cls;
    return p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, cls) != hipSuccess) {
    (void)hipGetLastError();
    trim();
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      throw Error{SVD_ERR_NOMEM, "device allocation of " + std::to_string(cls) + " bytes failed"};
    }
  }
  live_[p] = cls;
  return p;
}

void Arena::put(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.insert(std::make_pair(it->second, p));
  live_.erase(it);
}

void Arena::trim() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
}

// Two small arrays (together at most 2048 words) in one host round trip.
void read_u32s_2(svd_ctx* c, const uint32_t* d1, uint32_t* h1, size_t n1, const uint32_t* d2, uint32_t* h2,
                 size_t n2) {
  if (n1 + n2 > 2048) {
    read_u32s(c, d1, h1, n1);
    read_u32s(c, d2, h2, n2);
    return;
  }
  uint32_t* stage = (uint32_t*)c->pinned;
  if (n1) SVD_HIP(hipMemcpyAsync(stage, d1, 4 * n1, hipMemcpyDeviceToHost, c->stream));
  if (n2) SVD_HIP(hipMemcpyAsync(stage + n1, d2, 4 * n2, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
  memcpy(h1, stage, 4 * n1);
  memcpy(h2, stage + n1, 4 * n2);
}

double timer_ms(svd_ctx* c, const std::string& name) {
  auto it = c->timers.find(name);
  if (it == c->timers.end() || !it->second.second) return -1.0;
  if (hipEventSynchronize(it->second.second) != hipSuccess) return -1.0;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, it->second.first, it->second.second) != hipSuccess) return -1.0;
  return ms;
}

uint64_t read_u32(svd_ctx* c, const uint32_t* d) {
  uint32_t v = 0;
  read_u32s(c, d, &v, 1);
  return v;
}

void read_u32s(svd_ctx* c, const uint32_t* d, uint32_t* h, size_t n) {
  uint32_t* stage = (uint32_t*)c->pinned;
  size_t done = 0;
  while (done < n) {
    const size_t k = std::min<size_t>(n - done, 1024);
    SVD_HIP(hipMemcpyAsync(stage, d + done, 4 * k, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
    memcpy(h + done, stage, 4 * k);
    done += k;
  }
}

}  // namespace svd

using namespace svd;

namespace {

template <typename F>
int guarded(svd_ctx* c, F&& f) {
  if (!c) return SVD_ERR_ARG;
  try {
    SVD_HIP(hipSetDevice(c->device));
    f();
    c->err.clear();
    return SVD_OK;
  } catch (const Error& e) {
    c->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    c->err = "host allocation failed";
    return SVD_ERR_NOMEM;
  } catch (const std::exception& e) {
    c->err = e.what();
    return SVD_ERR_DEVICE;
  }
}

void free_levels(svd_ctx* c, std::vector<Level>& v) {
  for (auto& l : v) { dfree(c, l.rec); dfree(c, l.inv); dfree(c, l.childLevels); }
  v.clear();
}

const std::vector<Level>* view_of(svd_octree* o, int view) {
  switch (view) {
    case SVD_VIEW_SVO: return &o->svo;
    case SVD_VIEW_DAG: return o->hasX ? &o->xdag : &o->dag;
    case SVD_VIEW_SDAG: return &o->sdag;
    default: return nullptr;
  }
}

}  // namespace

extern "C" {

int svd_ctx_create(int device, svd_ctx** out) {
  if (!out) return SVD_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SVD_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return SVD_ERR_DEVICE;
  svd_ctx* c = new (std::nothrow) svd_ctx();
  if (!c) return SVD_ERR_NOMEM;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&c->pinned, 8192, hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess) {
    delete c;
    return SVD_ERR_DEVICE;
  }
  c->arena = new Arena(device);
  if (hipDeviceGetAttribute(&c->numCU, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->numCU = 256;
  // svd_render_noalloc's static texture: its device address on this device
  render_scratch_init(c);
  {
    const char* e = std::getenv("SVD_TOSDAG_CHECKED");
    c->sdagChecked = e && e[0] && !(e[0] == '0' && e[1] == 0);
    // SVD_REFSORT_EXTRA_ROUNDS: device-wide refcount-order rounds beyond the
    // size bound before the list size is first looked at (default 4; 0..16)
    const char* r = std::getenv("SVD_REFSORT_EXTRA_ROUNDS");
    if (r && r[0]) c->refsortExtra = std::min(16, std::max(0, std::atoi(r)));
    const char* sl = std::getenv("SVD_SMALL_LEVELS");
    if (sl && sl[0]) {
      unsigned v[4] = {0, 0, 0, 0};
      if (std::sscanf(sl, "%u,%u,%u,%u", &v[0], &v[1], &v[2], &v[3]) == 4) {
        c->smallSvo = v[0]; c->smallDag = v[1]; c->smallSdag = v[2]; c->smallLayout = v[3];
      }
    }
  }
  *out = c;
  return SVD_OK;
}

int svd_ctx_destroy(svd_ctx* c) {
  if (!c) return SVD_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete c->arena;
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
  if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
  for (auto& kv : c->timers) {
    if (kv.second.first) (void)hipEventDestroy(kv.second.first);
    if (kv.second.second) (void)hipEventDestroy(kv.second.second);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SVD_OK;
}

const char* svd_last_error(const svd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int svd_ctx_synchronize(svd_ctx* c) {
  return guarded(c, [&] { SVD_HIP(hipStreamSynchronize(c->stream)); });
}

void* svd_ctx_stream(svd_ctx* c) { return c ? (void*)c->stream : nullptr; }

int svd_octree_from_morton(svd_ctx* c, const uint64_t* codes, uint64_t n, uint32_t levels, int on_device,
                           svd_octree** out) {
  if (!out || (n && !codes)) return SVD_ERR_ARG;
  if (levels < 2 || levels > kMaxLevels) return SVD_ERR_LEVELS;
  *out = nullptr;
  svd_octree* o = new svd_octree();
  o->ctx = c;
  o->levels = levels;
  int rc = guarded(c, [&] {
    // device codes: return once the work is enqueued (host codes: their copy
    // is waited for)
    Timer t(c, "build_svo", !on_device);
    const uint64_t* d = codes;
    uint64_t* tmp = nullptr;
    if (!on_device && n) {
      tmp = dalloc<uint64_t>(c, n);
      SVD_HIP(hipMemcpyAsync(tmp, codes, 8 * n, hipMemcpyHostToDevice, c->stream));
      d = tmp;
    }
    build_svo_from_morton(o, d, n);
    SVD_HIP(hipStreamSynchronize(c->stream));
    dfree(c, tmp);
  });
  if (rc) { free_levels(c, o->svo); delete o; return rc; }
  *out = o;
  return SVD_OK;
}

namespace {
int mesh_octree(svd_ctx* c, const void* prims, uint64_t ntris, uint32_t levels, int stepLevel, const double bbox[6],
                int on_device, svd_octree** out);
}

int svd_octree_from_mesh(svd_ctx* c, const float* tris, uint64_t ntris, uint32_t levels, const double bbox[6],
                         int on_device, svd_octree** out) {
  return mesh_octree(c, tris, ntris, levels, -1, bbox, on_device, out);
}

int svd_octree_from_mesh_subtrees(svd_ctx* c, const float* tris, uint64_t ntris, uint32_t levels,
                                  uint32_t step_level, const double bbox[6], int on_device, svd_octree** out) {
  if (step_level == 0 || step_level >= 64) return SVD_ERR_ARG;
  return mesh_octree(c, tris, ntris, levels, (int)step_level, bbox, on_device, out);
}

int svd_octree_from_points(svd_ctx* c, const double* xyz, uint64_t npts, uint32_t levels, const double bbox[6],
                           int on_device, svd_octree** out) {
  return mesh_octree(c, xyz, npts, levels, -2, bbox, on_device, out);
}

}  // extern "C"

namespace {
// stepLevel: -1 buildSVO, -2 buildSVOFromPoints (prims = points, 3 doubles
// each), else buildDAG's joined SVO with that step level.
int mesh_octree(svd_ctx* c, const void* prims, uint64_t ntris, uint32_t levels, int stepLevel, const double bbox[6],
                int on_device, svd_octree** out) {
  if (!c || !out || !bbox || (ntris && !prims)) return SVD_ERR_ARG;
  if (levels < 2 || levels > kMaxLevels) return SVD_ERR_LEVELS;
  *out = nullptr;
  svd_octree* o = new svd_octree();
  o->ctx = c;
  o->levels = levels;
  int rc = guarded(c, [&] {
    Timer t(c, "build_svo");
    const uint64_t primBytes = stepLevel == -2 ? 24 : 36;
    const void* d = prims;
    uint8_t* tmp = nullptr;
    if (!on_device && ntris) {
      tmp = dalloc<uint8_t>(c, primBytes * ntris);
      SVD_HIP(hipMemcpyAsync(tmp, prims, primBytes * ntris, hipMemcpyHostToDevice, c->stream));
      d = tmp;
    }
    if (stepLevel == -2) build_svo_from_points(o, (const double*)d, ntris, bbox, bbox + 3);
    else if (stepLevel < 0) build_svo_from_mesh(o, (const float*)d, ntris, bbox, bbox + 3);
    else build_svo_from_mesh_subtrees(o, (const float*)d, ntris, bbox, bbox + 3, (uint32_t)stepLevel);
    dfree(c, tmp);
    // nVoxels: set bits of the leaf level
    const Level& leaf = o->svo[levels - 1];
    std::vector<uint32_t> m(leaf.n);
    if (leaf.n) SVD_HIP(hipMemcpyAsync(m.data(), leaf.rec, 4 * leaf.n, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
    uint64_t v = 0;
    for (uint32_t x : m) v += (uint64_t)__builtin_popcount(x & 0xFFu);
    o->nVoxels = v;
    o->state = SVD_S_SVO;
  });
  if (rc) { free_levels(c, o->svo); delete o; return rc; }
  *out = o;
  return SVD_OK;
}

}  // namespace

extern "C" {

namespace {

// Host level arrays (9 words per node; the finest level's mask in word 0) into
// device levels; child pointers are checked against their masks first.
void upload_levels(svd_ctx* c, uint32_t levels, const uint64_t* sizes, const uint32_t* const* words,
                   std::vector<Level>& dst, uint64_t* nNodes, uint64_t* nVox) {
  dst.assign(levels, Level());
  *nNodes = 0;
  *nVox = 0;
  if (sizes[0] != 1) throw Error{SVD_ERR_ARG, "level 0 must hold exactly the root"};
  // validate child pointers on the host: every set mask bit of an inner node
  // points into the next level, cleared bits hold the null pointer
  for (uint32_t l = 0; l + 1 < levels; ++l) {
    if (!sizes[l]) continue;
    if (!words[l]) throw Error{SVD_ERR_ARG, "missing level words"};
    for (uint64_t i = 0; i < sizes[l]; ++i) {
      const uint32_t* w = words[l] + 9 * i;
      for (int c = 0; c < 8; ++c) {
        const bool bit = (w[0] >> c) & 1u;
        if (bit ? (w[1 + c] >= sizes[l + 1]) : (w[1 + c] != kNull))
          throw Error{SVD_ERR_POINTER, "level " + std::to_string(l) + " node " + std::to_string(i) +
                                           ": child pointer inconsistent with its mask"};
      }
    }
  }
  for (uint32_t l = 0; l < levels; ++l) {
    Level& lv = dst[l];
    lv.n = sizes[l];
    lv.leaf = (l == levels - 1);
    if (l > 0) *nNodes += lv.n;
    if (!lv.n) continue;
    if (!words[l]) throw Error{SVD_ERR_ARG, "missing level words"};
    if (lv.leaf) {
      std::vector<uint32_t> m(lv.n);
      for (uint64_t i = 0; i < lv.n; ++i) { m[i] = words[l][9 * i] & 0xFFu; *nVox += __builtin_popcount(m[i]); }
      lv.rec = dalloc<uint32_t>(c, lv.n);
      SVD_HIP(hipMemcpyAsync(lv.rec, m.data(), 4 * lv.n, hipMemcpyHostToDevice, c->stream));
      SVD_HIP(hipStreamSynchronize(c->stream));
    } else {
      std::vector<uint32_t> w(words[l], words[l] + kRecWords * lv.n);
      for (uint64_t i = 0; i < lv.n; ++i) w[kRecWords * i] &= 0xFFu;   // masks only: no mirror bits
      lv.rec = dalloc<uint32_t>(c, lv.n * kRecWords);
      SVD_HIP(hipMemcpyAsync(lv.rec, w.data(), 4 * kRecWords * lv.n, hipMemcpyHostToDevice, c->stream));
      SVD_HIP(hipStreamSynchronize(c->stream));
    }
  }
}

// Deep copy of a level view: records (9 words, 1 per leaf), invariance bytes
// and child levels where present.
std::vector<Level> copy_levels(svd_ctx* c, const std::vector<Level>& src) {
  std::vector<Level> dst(src.size());
  try {
    for (size_t l = 0; l < src.size(); ++l) {
      const Level& a = src[l];
      Level& b = dst[l];
      b.n = a.n;
      b.leaf = a.leaf;
      const uint64_t words = a.n * (a.leaf ? 1 : kRecWords);
      if (a.rec) {
        b.rec = dalloc<uint32_t>(c, words ? words : 1);
        if (words) SVD_HIP(hipMemcpyAsync(b.rec, a.rec, 4 * words, hipMemcpyDeviceToDevice, c->stream));
      }
      if (a.inv) {
        b.inv = dalloc<uint8_t>(c, a.n + 1);
        if (a.n) SVD_HIP(hipMemcpyAsync(b.inv, a.inv, a.n, hipMemcpyDeviceToDevice, c->stream));
      }
      if (a.childLevels) {
        b.childLevels = dalloc<uint32_t>(c, 8 * a.n + 1);
        if (a.n) SVD_HIP(hipMemcpyAsync(b.childLevels, a.childLevels, 32 * a.n, hipMemcpyDeviceToDevice, c->stream));
      }
    }
  } catch (...) {
    free_levels(c, dst);
    throw;
  }
  return dst;
}

void destroy_octree(svd_octree* o) {
  svd_ctx* c = o->ctx;
  free_levels(c, o->svo);
  free_levels(c, o->dag);
  free_levels(c, o->sdag);
  free_levels(c, o->xdag);
  for (auto* m : o->sdagMap) dfree(c, m);
  for (auto& e : o->enc) dfree(c, e.d);
  delete o;
}

}  // namespace

int svd_octree_from_svo(svd_ctx* c, uint32_t levels, const uint64_t* sizes, const uint32_t* const* words,
                        const float bbox[6], double rootSide, svd_octree** out) {
  if (!out || !sizes || !words || !bbox) return SVD_ERR_ARG;
  if (levels < 2 || levels > kMaxLevels) return SVD_ERR_LEVELS;
  *out = nullptr;
  svd_octree* o = new svd_octree();
  o->ctx = c;
  o->levels = levels;
  for (int i = 0; i < 6; ++i) o->bbox[i] = bbox[i];
  o->rootSide = rootSide;
  int rc = guarded(c, [&] {
    uint64_t nNodes = 0, nVox = 0;
    upload_levels(c, levels, sizes, words, o->svo, &nNodes, &nVox);
    o->nVoxels = nVox;
    o->nNodesSVO = nNodes;
    o->state = SVD_S_SVO;
  });
  if (rc) { free_levels(c, o->svo); delete o; return rc; }
  *out = o;
  return SVD_OK;
}

int svd_octree_from_levels(svd_ctx* c, uint32_t levels, const uint64_t* sizes, const uint32_t* const* words,
                           const float bbox[6], double rootSide, uint64_t n_nodes, uint64_t n_voxels,
                           svd_octree** out) {
  if (!out || !sizes || !words || !bbox) return SVD_ERR_ARG;
  if (levels < 2 || levels > kMaxLevels) return SVD_ERR_LEVELS;
  *out = nullptr;
  svd_octree* o = new svd_octree();
  o->ctx = c;
  o->levels = levels;
  for (int i = 0; i < 6; ++i) o->bbox[i] = bbox[i];
  o->rootSide = rootSide;
  int rc = guarded(c, [&] {
    uint64_t nNodes = 0, nVox = 0;
    upload_levels(c, levels, sizes, words, o->dag, &nNodes, &nVox);
    o->nVoxels = n_voxels;     // GeomOctree(NodeData, ...) takes the counts from the Stats it is given
    o->nNodesDAG = n_nodes;
    o->state = SVD_S_DAG;
  });
  if (rc) { free_levels(c, o->dag); delete o; return rc; }
  *out = o;
  return SVD_OK;
}

int svd_octree_copy(const svd_octree* src, svd_octree** out) {
  if (!src || !out) return SVD_ERR_ARG;
  *out = nullptr;
  svd_ctx* c = src->ctx;
  svd_octree* o = new svd_octree();
  o->ctx = c;
  o->levels = src->levels;
  o->state = src->state;
  for (int i = 0; i < 6; ++i) o->bbox[i] = src->bbox[i];
  o->rootSide = src->rootSide;
  o->nVoxels = src->nVoxels;
  o->nNodesSVO = src->nNodesSVO;
  o->nNodesDAG = src->nNodesDAG;
  o->nNodesSDAG = src->nNodesSDAG;
  o->nNodesX = src->nNodesX;
  o->crossMerged = src->crossMerged;
  o->hasX = src->hasX;
  int rc = guarded(c, [&] {
    o->svo = copy_levels(c, src->svo);
    o->dag = copy_levels(c, src->dag);
    o->sdag = copy_levels(c, src->sdag);
    o->xdag = copy_levels(c, src->xdag);
    o->sdagMap.assign(src->sdagMap.size(), nullptr);
    for (size_t l = 0; l < src->sdagMap.size(); ++l) {
      if (!src->sdagMap[l]) continue;
      const uint64_t n = l < src->dag.size() ? src->dag[l].n : 0;
      o->sdagMap[l] = dalloc<uint32_t>(c, n + 1);
      if (n) SVD_HIP(hipMemcpyAsync(o->sdagMap[l], src->sdagMap[l], 4 * n, hipMemcpyDeviceToDevice, c->stream));
    }
    for (int f = 0; f < 4; ++f) {
      const Encoded& e = src->enc[f];
      if (!e.d) continue;
      o->enc[f].d = dalloc<uint8_t>(c, e.bytes + 1);
      if (e.bytes) SVD_HIP(hipMemcpyAsync(o->enc[f].d, e.d, e.bytes, hipMemcpyDeviceToDevice, c->stream));
      o->enc[f].bytes = e.bytes;
      o->enc[f].valid = e.valid;
    }
    SVD_HIP(hipStreamSynchronize(c->stream));
  });
  if (rc) { destroy_octree(o); return rc; }
  *out = o;
  return SVD_OK;
}

int svd_octree_from_svdag(svd_ctx* c, const void* file, uint64_t bytes, int on_device, svd_octree** out) {
  if (!c || !file || !out) return SVD_ERR_ARG;
  *out = nullptr;
  svd_octree* o = new svd_octree();
  o->ctx = c;
  int rc = guarded(c, [&] {
    Timer t(c, "decode");
    octree_from_svdag(o, (const uint8_t*)file, bytes, on_device != 0);
  });
  if (rc) { delete o; return rc; }
  *out = o;
  return SVD_OK;
}

int svd_octree_free(svd_octree* o) {
  if (!o) return SVD_ERR_ARG;
  svd_ctx* c = o->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  destroy_octree(o);
  return SVD_OK;
}

int svd_reduce_dag(svd_octree* o) {
  if (!o) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    Timer t(o->ctx, "to_dag", false);
    reduce_dag(o);
    free_levels(o->ctx, o->xdag);
    o->hasX = false;
    for (auto& e : o->enc) e.valid = false;
  });
}

int svd_reduce_sym(svd_octree* o) {
  if (!o) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    Timer t(o->ctx, "to_sdag", false);
    reduce_sym(o);
  });
}

int svd_reduce_sym_pointers(svd_octree* o) {
  if (!o) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    Timer t(o->ctx, "to_sdag");
    reduce_sym_pointers(o);
  });
}

int svd_cross_level_merge(svd_octree* o, uint64_t* eliminated) {
  if (!o) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    Timer t(o->ctx, "cross_merge");
    cross_level_merge(o);
    if (eliminated) *eliminated = o->crossMerged;
  });
}

int svd_octree_state(svd_octree* o, int* s) {
  if (!o || !s) return SVD_ERR_ARG;
  *s = o->state;
  return SVD_OK;
}

int svd_octree_levels(svd_octree* o, uint32_t* l) {
  if (!o || !l) return SVD_ERR_ARG;
  *l = o->levels;
  return SVD_OK;
}

int svd_octree_counts(svd_octree* o, uint64_t* nn, uint64_t* nv) {
  if (!o) return SVD_ERR_ARG;
  if (nn) {
    switch (o->state) {
      case SVD_S_SVO: *nn = o->nNodesSVO; break;
      case SVD_S_DAG: *nn = o->hasX ? o->nNodesX : o->nNodesDAG; break;
      case SVD_S_SDAG: *nn = o->nNodesSDAG; break;
      default: *nn = 0;
    }
  }
  if (nv) *nv = o->nVoxels;
  return SVD_OK;
}

int svd_level_size(svd_octree* o, int view, uint32_t lev, uint64_t* n) {
  if (!o || !n) return SVD_ERR_ARG;
  const std::vector<Level>* v = view_of(o, view);
  if (!v || lev >= v->size()) return SVD_ERR_ARG;
  *n = (*v)[lev].n;
  return SVD_OK;
}

int svd_get_level(svd_octree* o, int view, uint32_t lev, uint32_t* words, uint8_t* inv, uint32_t* childLevels) {
  if (!o || !words) return SVD_ERR_ARG;
  const std::vector<Level>* v = view_of(o, view);
  if (!v || lev >= v->size()) return SVD_ERR_ARG;
  svd_ctx* c = o->ctx;
  return guarded(c, [&] {
    const Level& lv = (*v)[lev];
    if (!lv.n) return;
    if (lv.leaf) {
      std::vector<uint32_t> m(lv.n);
      SVD_HIP(hipMemcpyAsync(m.data(), lv.rec, 4 * lv.n, hipMemcpyDeviceToHost, c->stream));
      SVD_HIP(hipStreamSynchronize(c->stream));
      for (uint64_t i = 0; i < lv.n; ++i) {
        words[9 * i] = m[i];
        for (int k = 1; k < 9; ++k) words[9 * i + k] = kNull;
      }
    } else {
      SVD_HIP(hipMemcpyAsync(words, lv.rec, 4 * kRecWords * lv.n, hipMemcpyDeviceToHost, c->stream));
    }
    if (inv) {
      if (lv.inv) SVD_HIP(hipMemcpyAsync(inv, lv.inv, lv.n, hipMemcpyDeviceToHost, c->stream));
      else memset(inv, 0, lv.n);
    }
    if (childLevels) {
      if (lv.childLevels) SVD_HIP(hipMemcpyAsync(childLevels, lv.childLevels, 32 * lv.n, hipMemcpyDeviceToHost, c->stream));
      else for (uint64_t i = 0; i < 8 * lv.n; ++i) childLevels[i] = lev + 1;
    }
    SVD_HIP(hipStreamSynchronize(c->stream));
  });
}

int svd_encode(svd_octree* o, int fmt, uint64_t* bytes) {
  if (!o || fmt < 0 || fmt > 3) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    Timer t(o->ctx, "encode");
    encode(o, fmt);
    if (bytes) *bytes = o->enc[fmt].bytes;
  });
}

int svd_encode_formats(svd_octree* o, const int* fmts, uint32_t n, uint64_t* bytes) {
  if (!o || (n && !fmts) || n > 4) return SVD_ERR_ARG;
  for (uint32_t k = 0; k < n; ++k)
    if (fmts[k] < 0 || fmts[k] > 3) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    Timer t(o->ctx, "encode", false);
    encode_formats(o, fmts, (int)n);
    if (bytes)
      for (uint32_t k = 0; k < n; ++k) bytes[k] = o->enc[fmts[k]].bytes;
  });
}

int svd_encoded_copy(svd_octree* o, int fmt, void* dst, uint64_t cap) {
  if (!o || fmt < 0 || fmt > 3 || !dst) return SVD_ERR_ARG;
  if (!o->enc[fmt].valid) return SVD_ERR_STATE;
  if (cap < o->enc[fmt].bytes) return SVD_ERR_ARG;
  svd_ctx* c = o->ctx;
  return guarded(c, [&] {
    SVD_HIP(hipMemcpyAsync(dst, o->enc[fmt].d, o->enc[fmt].bytes, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
  });
}

int svd_encoded_device(svd_octree* o, int fmt, const void** dptr, uint64_t* bytes) {
  if (!o || fmt < 0 || fmt > 3 || !dptr) return SVD_ERR_ARG;
  if (!o->enc[fmt].valid) return SVD_ERR_STATE;
  // the image may still be being written on the context stream: a pointer
  // handed out is to finished bytes
  if (hipStreamSynchronize(o->ctx->stream) != hipSuccess) return SVD_ERR_DEVICE;
  *dptr = o->enc[fmt].d;
  if (bytes) *bytes = o->enc[fmt].bytes;
  return SVD_OK;
}

// ---------------------------------------------------------------- scenes
static int scene_setup(svd_ctx* c, int fmt, const uint8_t* dimg, uint64_t bytes, svd_scene* s) {
  return guarded(c, [&] {
    auto rd = [&](uint64_t off, void* h, size_t n) {
      if (off + n > bytes) throw Error{SVD_ERR_FORMAT, "encoded image truncated"};
      SVD_HIP(hipMemcpyAsync(h, dimg + off, n, hipMemcpyDeviceToHost, c->stream));
      SVD_HIP(hipStreamSynchronize(c->stream));
    };
    uint8_t hdr[36];
    rd(0, hdr, 36);
    float rs;
    memcpy(s->bmin, hdr, 12);
    memcpy(s->bmax, hdr + 12, 12);
    memcpy(&rs, hdr + 24, 4);
    memcpy(&s->levels, hdr + 28, 4);
    if (s->levels < 2 || s->levels > kMaxLevels) throw Error{SVD_ERR_FORMAT, "bad level count"};
    s->rootHalfSide = rs / 2.0f;   // EncodedOctree::getHalfSide() (encoded_octree.hpp:68-70)
    for (int a = 0; a < 3; ++a) s->center[a] = (s->bmin[a] + s->bmax[a]) * 0.5f;
    s->fmt = fmt;
    if (fmt == SVD_FMT_SVDAG || fmt == SVD_FMT_USSVDAG) {
      uint32_t cnt;
      rd(40, &cnt, 4);
      if (44 + 4ull * cnt > bytes) throw Error{SVD_ERR_FORMAT, "node array truncated"};
      s->nodes = (const uint32_t*)(dimg + 44);
      s->nWords = cnt;
    } else {
      if (s->levels < 3) throw Error{SVD_ERR_FORMAT, "SSVDAG needs 3 levels"};
      uint32_t nInner, nLeafBytes, nOff;
      rd(36, &nInner, 4);
      const uint64_t leafCnt = 40 + 2ull * nInner;
      rd(leafCnt, &nLeafBytes, 4);
      const uint64_t offCnt = leafCnt + 4 + nLeafBytes;
      rd(offCnt, &nOff, 4);
      if (nOff != s->levels - 2 || offCnt + 4 + 4ull * nOff > bytes || (nLeafBytes & 7))
        throw Error{SVD_ERR_FORMAT, "bad SSVDAG layout"};
      // aligned private copies for 8-byte brick loads
      const uint64_t innerBytes = 2ull * nInner;
      const uint64_t own = ((innerBytes + 15) & ~15ull) + nLeafBytes + 4ull * nOff + 64;
      s->owned = dalloc<uint8_t>(c, own);
      uint8_t* p = s->owned;
      SVD_HIP(hipMemcpyAsync(p, dimg + 40, innerBytes, hipMemcpyDeviceToDevice, c->stream));
      s->inner = (const uint16_t*)p;
      p += (innerBytes + 15) & ~15ull;
      SVD_HIP(hipMemcpyAsync(p, dimg + leafCnt + 4, nLeafBytes, hipMemcpyDeviceToDevice, c->stream));
      s->leaves = p;
      p += (nLeafBytes + 15) & ~15ull;
      SVD_HIP(hipMemcpyAsync(p, dimg + offCnt + 4, 4ull * nOff, hipMemcpyDeviceToDevice, c->stream));
      s->levelOffsets = (uint32_t*)p;
      s->nLevelOffsets = nOff;
      s->nInner = nInner;
      s->nLeaves = nLeafBytes / 8;
      SVD_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int svd_scene_create(svd_ctx* c, int fmt, const void* file, uint64_t bytes, int on_device, svd_scene** out) {
  if (!c || !file || !out || fmt < 0 || fmt > 3) return SVD_ERR_ARG;
  *out = nullptr;
  svd_scene* s = new svd_scene();
  s->ctx = c;
  uint8_t* img = nullptr;
  int rc = guarded(c, [&] {
    img = dalloc<uint8_t>(c, bytes);
    SVD_HIP(hipMemcpyAsync(img, file, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
  });
  if (rc) { delete s; return rc; }
  rc = scene_setup(c, fmt, img, bytes, s);
  if (rc) { dfree(c, img); dfree(c, s->owned); delete s; return rc; }
  if (fmt == SVD_FMT_SVDAG || fmt == SVD_FMT_USSVDAG) {
    // keep the image alive: it backs s->nodes
    s->owned = img;
  } else {
    dfree(c, img);
  }
  *out = s;
  return SVD_OK;
}

int svd_scene_from_octree(svd_octree* o, int fmt, svd_scene** out) {
  if (!o || !out || fmt < 0 || fmt > 3) return SVD_ERR_ARG;
  if (!o->enc[fmt].valid) return SVD_ERR_STATE;
  *out = nullptr;
  svd_scene* s = new svd_scene();
  s->ctx = o->ctx;
  int rc = scene_setup(o->ctx, fmt, o->enc[fmt].d, o->enc[fmt].bytes, s);
  if (rc) { dfree(o->ctx, s->owned); delete s; return rc; }
  *out = s;
  return SVD_OK;
}

int svd_scene_free(svd_scene* s) {
  if (!s) return SVD_ERR_ARG;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  dfree(s->ctx, s->owned);
  delete s;
  return SVD_OK;
}

int svd_trace(svd_scene* s, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad) {
  if (!s || !p || !prim) return SVD_ERR_ARG;
  return guarded(s->ctx, [&] {
    Timer t(s->ctx, "trace");
    trace_frame(s, p, prim, shad);
  });
}

int svd_trace_hits(svd_scene* s, const svd_trace_params* p, float* prim, float* shad) {
  if (!s || !p || !prim) return SVD_ERR_ARG;
  return guarded(s->ctx, [&] {
    Timer t(s->ctx, "trace");
    trace_frame_hits(s, p, prim, shad);
  });
}

int svd_trace_beam(svd_scene* s, const svd_trace_params* p, float depth_projection_factor, svd_ray_result* prim,
                   svd_ray_result* shad, float* min_depth) {
  if (!s || !p || !prim) return SVD_ERR_ARG;
  return guarded(s->ctx, [&] {
    Timer t(s->ctx, "trace");
    trace_frame_beam(s, p, depth_projection_factor, prim, shad, min_depth);
  });
}

int svd_render(svd_scene* s, const svd_trace_params* p, const svd_view_params* v, float beam_projection_factor,
               float* rgb, svd_ray_result* primary) {
  if (!s || !p || !v || !rgb) return SVD_ERR_ARG;
  return guarded(s->ctx, [&] {
    Timer t(s->ctx, "trace");
    render_frame(s, p, v, beam_projection_factor, rgb, primary);
  });
}

int svd_render_noalloc(svd_scene* s, const svd_trace_params* p, const svd_view_params* v,
                       float beam_projection_factor, float* rgb, svd_ray_result* primary) {
  if (!s || !p || !v || !rgb) return SVD_ERR_ARG;
  return guarded(s->ctx, [&] { render_frame_noalloc(s, p, v, beam_projection_factor, rgb, primary); });
}

int svd_trace_ao(svd_scene* s, const svd_trace_params* p, uint32_t num_rays, float* visibility,
                 svd_ray_result* primary) {
  if (!s || !p || !visibility) return SVD_ERR_ARG;
  return guarded(s->ctx, [&] {
    Timer t(s->ctx, "trace");
    ao_frame(s, p, num_rays, visibility, primary);
  });
}

int svd_refcount_order(svd_ctx* c, uint32_t* idx, uint32_t* val, uint64_t n) {
  if (!c || (n && (!idx || !val))) return SVD_ERR_ARG;
  return guarded(c, [&] {
    if (!n) return;
    uint32_t* di = dalloc<uint32_t>(c, n);
    uint32_t* dv = dalloc<uint32_t>(c, n);
    SVD_HIP(hipMemcpyAsync(di, idx, 4 * n, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(dv, val, 4 * n, hipMemcpyHostToDevice, c->stream));
    {
      Timer t(c, "refsort");
      refsort(c, di, dv, n);
    }
    SVD_HIP(hipMemcpyAsync(idx, di, 4 * n, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipMemcpyAsync(val, dv, 4 * n, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
    dfree(c, di);
    dfree(c, dv);
  });
}

int svd_refcount_order_ranges(svd_ctx* c, uint32_t* idx, uint32_t* val, uint64_t n, const uint64_t* first,
                              const uint64_t* last, uint64_t nr) {
  if (!c || (n && (!idx || !val)) || (nr && (!first || !last))) return SVD_ERR_ARG;
  return guarded(c, [&] {
    if (!n) return;
    uint32_t* di = dalloc<uint32_t>(c, n);
    uint32_t* dv = dalloc<uint32_t>(c, n);
    SVD_HIP(hipMemcpyAsync(di, idx, 4 * n, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(dv, val, 4 * n, hipMemcpyHostToDevice, c->stream));
    refsort_ranges(c, di, dv, n, first, last, (size_t)nr);
    SVD_HIP(hipMemcpyAsync(idx, di, 4 * n, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipMemcpyAsync(val, dv, 4 * n, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
    dfree(c, di);
    dfree(c, dv);
  });
}

int svd_attr_bittree_encode(svd_ctx* c, const uint8_t* attrs, uint64_t n, int gray, int on_device, void* out,
                            uint64_t cap, uint64_t* bytes) {
  if (!c || (n && !attrs) || !bytes) return SVD_ERR_ARG;
  return guarded(c, [&] {
    Timer t(c, "attr");
    const uint8_t* d = attrs;
    uint8_t* tmp = nullptr;
    if (!on_device && n) {
      tmp = dalloc<uint8_t>(c, n);
      SVD_HIP(hipMemcpyAsync(tmp, attrs, n, hipMemcpyHostToDevice, c->stream));
      d = tmp;
    }
    *bytes = attr_bittree(c, d, n, gray ? 1 : 0, (uint8_t*)out, out ? cap : 0);
    SVD_HIP(hipStreamSynchronize(c->stream));
    dfree(c, tmp);
  });
}

int svd_last_kernel_ms(svd_ctx* c, const char* name, double* ms) {
  if (!c || !name || !ms) return SVD_ERR_ARG;
  const double v = timer_ms(c, name);
  if (v < 0) return SVD_ERR_ARG;
  *ms = v;
  return SVD_OK;
}

}  // extern "C"

// ---- lossy DAG ---------------------------------------------------------------
int svd_lossy_merge(svd_octree* o, const svd_lossy_params* p, svd_lossy_level_stats* stats, uint32_t cap) {
  if (!o || !p) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    std::vector<svd_lossy_level_stats> st;
    {
      Timer t(o->ctx, "lossy");
      lossy_merge(o, p, st);
    }
    free_levels(o->ctx, o->xdag);
    o->hasX = false;
    free_levels(o->ctx, o->sdag);
    for (auto* m : o->sdagMap) dfree(o->ctx, m);
    o->sdagMap.clear();
    for (auto& e : o->enc) e.valid = false;
    if (stats)
      for (size_t i = 0; i < st.size() && i < cap; ++i) stats[i] = st[i];
  });
}

int svd_lossy_signatures(svd_octree* o, uint32_t lev, uint32_t* sig) {
  if (!o || !sig) return SVD_ERR_ARG;
  return guarded(o->ctx, [&] {
    lossy_signatures(o, lev, sig);
    SVD_HIP(hipStreamSynchronize(o->ctx->stream));
  });
}

// ---- learned encoder ---------------------------------------------------------
int svd_mlp_create(svd_ctx* c, uint64_t seed, svd_mlp** out) {
  if (!c || !out) return SVD_ERR_ARG;
  svd_mlp* m = new svd_mlp();
  m->ctx = c;
  const int rc = guarded(c, [&] {
    mlp_alloc(m);
    mlp_init(m, seed);
  });
  if (rc != SVD_OK) { (void)guarded(c, [&] { mlp_release(m); }); delete m; return rc; }
  *out = m;
  return SVD_OK;
}

int svd_mlp_free(svd_mlp* m) {
  if (!m) return SVD_ERR_ARG;
  const int rc = guarded(m->ctx, [&] { mlp_release(m); });
  delete m;
  return rc;
}

int svd_mlp_get_params(svd_mlp* m, float* host, uint64_t count) {
  if (!m || !host || count != SVD_MLP_PARAMS) return SVD_ERR_ARG;
  return guarded(m->ctx, [&] { mlp_get_params(m, host); });
}

int svd_mlp_set_params(svd_mlp* m, const float* host, uint64_t count) {
  if (!m || !host || count != SVD_MLP_PARAMS) return SVD_ERR_ARG;
  return guarded(m->ctx, [&] { mlp_set_params(m, host); });
}

int svd_mlp_embed(svd_mlp* m, const uint32_t* sig, uint64_t n, float* emb) {
  if (!m || (n && (!sig || !emb))) return SVD_ERR_ARG;
  return guarded(m->ctx, [&] {
    Timer t(m->ctx, "mlp_embed");
    if (n) mlp_embed(m, sig, n, emb);
  });
}

int svd_mlp_loss_grad(svd_mlp* m, const uint32_t* sig, const uint32_t* labels, uint64_t npairs, float* loss,
                      float* grads) {
  if (!m || !sig || !labels || !loss || !grads || npairs < 16 || npairs % 16) return SVD_ERR_ARG;
  return guarded(m->ctx, [&] { mlp_loss_grad(m, sig, labels, npairs, loss, grads); });
}

int svd_mlp_train(svd_mlp* m, svd_octree* o, const svd_mlp_train_params* p, svd_mlp_train_stats* st) {
  if (!m || !o || !p || o->ctx != m->ctx) return SVD_ERR_ARG;
  return guarded(m->ctx, [&] { mlp_train(m, o, p, st); });
}

int svd_knn(svd_ctx* c, const float* emb, uint64_t n, uint32_t k, uint32_t tables, uint32_t bits, uint32_t window,
            uint64_t seed, uint32_t* out, const uint32_t* sig) {
  if (!c || (n && (!emb || !out))) return SVD_ERR_ARG;
  return guarded(c, [&] {
    Timer t(c, "knn");
    knn_lsh(c, emb, n, k, tables, bits, window, seed, out, sig);
  });
}

// ---- synthetic scenes ----------------------------------------------------------
int svd_workload_caves(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double threshold,
                       double base_cells, uint64_t** out, uint64_t* n) {
  if (!c || !out || !n || octaves < 1 || octaves > 16) return SVD_ERR_ARG;
  return guarded(c, [&] {
    Timer t(c, "workload");
    *n = gen_caves(c, levels, seed, octaves, threshold, base_cells, out);
  });
}

int svd_device_free(svd_ctx* c, void* p) {
  if (!c) return SVD_ERR_ARG;
  return guarded(c, [&] { if (p) SVD_HIP(hipFree(p)); });
}

int svd_device_read(svd_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return SVD_ERR_ARG;
  return guarded(c, [&] {
    if (bytes) SVD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
  });
}

// ---- batches -------------------------------------------------------------------
int svd_batch_build(svd_ctx* c, uint32_t S, const uint64_t* codes, const uint64_t* offsets, const uint32_t* levels,
                    int on_device, svd_batch** out) {
  if (!c || !offsets || !levels || !out || S == 0 || S > (1u << 20)) return SVD_ERR_ARG;
  *out = nullptr;
  const uint64_t n = offsets[S];
  if (n && !codes) return SVD_ERR_ARG;
  svd_batch* b = new svd_batch();
  b->ctx = c;
  b->S = S;
  const int rc = guarded(c, [&] {
    Timer t(c, "batch_build");
    b->L.assign(levels, levels + S);
    b->voxels.resize(S);
    std::vector<uint64_t> off(offsets, offsets + S + 1);
    if (off[0] != 0) throw Error{SVD_ERR_ARG, "offsets[0] must be 0"};
    for (uint32_t s = 0; s < S; ++s) {
      if (levels[s] < 2 || levels[s] > 13) throw Error{SVD_ERR_LEVELS, "batched scenes need 2..13 levels"};
      if (off[s + 1] <= off[s]) throw Error{SVD_ERR_ARG, "every batched scene needs at least one voxel"};
      b->voxels[s] = off[s + 1] - off[s];
      b->Lmax = std::max(b->Lmax, levels[s]);
    }
    const uint64_t* d = codes;
    uint64_t* tmp = nullptr;
    if (!on_device) {
      tmp = dalloc<uint64_t>(c, n);
      SVD_HIP(hipMemcpyAsync(tmp, codes, 8 * n, hipMemcpyHostToDevice, c->stream));
      d = tmp;
    }
    batch_build(b, d, n, off);
    dfree(c, tmp);
  });
  if (rc != SVD_OK) { (void)svd_batch_free(b); return rc; }
  *out = b;
  return SVD_OK;
}

int svd_batch_free(svd_batch* b) {
  if (!b) return SVD_ERR_ARG;
  const int rc = guarded(b->ctx, [&] {
    for (auto& lv : b->dag) dfree(b->ctx, lv.rec);
    dfree(b->ctx, b->files);
  });
  delete b;
  return rc;
}

int svd_batch_scene_info(svd_batch* b, uint32_t s, uint32_t* levels, uint64_t* nv, uint64_t* nn, uint64_t* bytes) {
  if (!b || s >= b->S) return SVD_ERR_ARG;
  if (levels) *levels = b->L[s];
  if (nv) *nv = b->voxels[s];
  if (nn) *nn = b->nodes[s];
  if (bytes) *bytes = b->fileBytes[s];
  return SVD_OK;
}

int svd_batch_svdag(svd_batch* b, uint32_t s, const void** dptr, uint64_t* bytes) {
  if (!b || s >= b->S || !dptr) return SVD_ERR_ARG;
  *dptr = b->files + b->fileOff[s];
  if (bytes) *bytes = b->fileBytes[s];
  return SVD_OK;
}

int svd_batch_svdag_copy(svd_batch* b, uint32_t s, void* dst, uint64_t cap) {
  if (!b || s >= b->S || !dst || cap < b->fileBytes[s]) return SVD_ERR_ARG;
  return guarded(b->ctx, [&] {
    SVD_HIP(hipMemcpyAsync(dst, b->files + b->fileOff[s], b->fileBytes[s], hipMemcpyDeviceToHost, b->ctx->stream));
    SVD_HIP(hipStreamSynchronize(b->ctx->stream));
  });
}

int svd_batch_trace(svd_batch* b, uint32_t n_cams, const svd_trace_params* cams, const uint32_t* cam_scene,
                    svd_ray_result* prim, svd_ray_result* shad) {
  if (!b || (n_cams && (!cams || !cam_scene || !prim))) return SVD_ERR_ARG;
  return guarded(b->ctx, [&] {
    Timer t(b->ctx, "trace");
    std::vector<svd_trace_params> cv(cams, cams + n_cams);
    std::vector<uint32_t> cs(cam_scene, cam_scene + n_cams);
    trace_batch_svdag(b->ctx, b->scenes, cv, cs, prim, shad);
  });
}

int svd_workload_balls(svd_ctx* c, uint32_t levels, const double* balls, uint32_t nb, uint64_t** out, uint64_t* n) {
  if (!c || !out || !n || (nb && !balls)) return SVD_ERR_ARG;
  return guarded(c, [&] {
    Timer t(c, "workload");
    *n = gen_balls(c, levels, balls, nb, out);
  });
}

int svd_workload_menger(svd_ctx* c, uint32_t levels, uint32_t order, uint64_t** out, uint64_t* n) {
  if (!c || !out || !n) return SVD_ERR_ARG;
  return guarded(c, [&] {
    Timer t(c, "workload");
    *n = gen_menger(c, levels, order, out);
  });
}

int svd_workload_terrain(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double hmin, double hmax,
                         double base_cells, uint64_t** out, uint64_t* n) {
  if (!c || !out || !n || octaves < 1 || octaves > 16) return SVD_ERR_ARG;
  return guarded(c, [&] {
    Timer t(c, "workload");
    *n = gen_terrain(c, levels, seed, octaves, hmin, hmax, base_cells, out);
  });
}

int svd_device_copy(svd_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return SVD_ERR_ARG;
  return guarded(c, [&] {
    if (bytes) SVD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));
  });
}

// ---- autotuning --------------------------------------------------------------
int svd_set_trace_tile(svd_ctx* c, uint32_t w) {
  if (!c || (w != 8 && w != 16 && w != 32 && w != 64)) return SVD_ERR_ARG;
  c->traceTileW = w;
  return SVD_OK;
}

int svd_get_trace_tile(svd_ctx* c, uint32_t* w) {
  if (!c || !w) return SVD_ERR_ARG;
  *w = c->traceTileW;
  return SVD_OK;
}

int svd_set_trace_mode(svd_ctx* c, uint32_t mode) {
  if (!c || (mode != SVD_TRACE_TILED && mode != SVD_TRACE_PERSISTENT)) return SVD_ERR_ARG;
  c->traceMode = mode;
  return SVD_OK;
}

int svd_get_trace_mode(svd_ctx* c, uint32_t* mode) {
  if (!c || !mode) return SVD_ERR_ARG;
  *mode = c->traceMode;
  return SVD_OK;
}

namespace {
// mean kernel time of `reps` frames (after one warm-up) with the context's current shape
double time_frames(svd_ctx* c, svd_scene* s, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad,
                   uint32_t reps) {
  trace_frame(s, p, prim, shad);
  SVD_HIP(hipEventRecord(c->ev[0], c->stream));
  for (uint32_t r = 0; r < reps; ++r) trace_frame(s, p, prim, shad);
  SVD_HIP(hipEventRecord(c->ev[1], c->stream));
  SVD_HIP(hipEventSynchronize(c->ev[1]));
  float e = 0.f;
  SVD_HIP(hipEventElapsedTime(&e, c->ev[0], c->ev[1]));
  return (double)e / reps;
}
}  // namespace

int svd_autotune_trace(svd_scene* s, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad,
                       uint32_t reps, uint32_t* best, double* ms) {
  if (!s || !p || !prim || reps < 1) return SVD_ERR_ARG;
  svd_ctx* c = s->ctx;
  return guarded(c, [&] {
    const uint32_t shapes[4] = {8, 16, 32, 64};
    double t[4];
    for (int k = 0; k < 4; ++k) {
      c->traceTileW = shapes[k];
      t[k] = time_frames(c, s, p, prim, shad, reps);
    }
    int b = 0;
    for (int k = 1; k < 4; ++k) if (t[k] < t[b]) b = k;
    c->traceTileW = shapes[b];
    if (best) *best = shapes[b];
    if (ms) for (int k = 0; k < 4; ++k) ms[k] = t[k];
  });
}

int svd_set_build_limits(svd_ctx* c, const uint32_t* limits) {
  if (!c || !limits) return SVD_ERR_ARG;
  c->smallSvo = limits[0];
  c->smallDag = limits[1];
  c->smallSdag = limits[2];
  c->smallLayout = limits[3];
  return SVD_OK;
}

int svd_get_build_limits(svd_ctx* c, uint32_t* limits) {
  if (!c || !limits) return SVD_ERR_ARG;
  limits[0] = c->smallSvo;
  limits[1] = c->smallDag;
  limits[2] = c->smallSdag;
  limits[3] = c->smallLayout;
  return SVD_OK;
}

// The lossless build of a device voxel list under each preset of the
// one-workgroup level limits: 1 warm-up + reps builds each (host wall time
// to a drained stream, median), the fastest preset kept.
int svd_autotune_build(svd_ctx* c, const uint64_t* dcodes, uint64_t n, uint32_t levels, uint32_t reps,
                       uint32_t* best, double* ms) {
  if (!c || !dcodes || n == 0 || reps < 1 || levels < 2 || levels > kMaxLevels) return SVD_ERR_ARG;
  static const uint32_t presets[SVD_BUILD_PRESETS][4] = {
      {2048, 1024, 512, 2048}, {8192, 4096, 1024, 4096}, {4096, 2048, 1024, 4096},
      {1024, 512, 256, 1024}, {2048, 1024, 1024, 2048}};
  uint32_t keep[4];
  svd_get_build_limits(c, keep);
  auto build_once = [&]() -> int {
    svd_octree* o = nullptr;
    int rc = svd_octree_from_morton(c, dcodes, n, levels, 1, &o);
    if (rc) return rc;
    rc = svd_reduce_dag(o);
    if (!rc) rc = svd_reduce_sym(o);
    const int f[4] = {SVD_FMT_SVDAG, SVD_FMT_USSVDAG, SVD_FMT_ESVDAG, SVD_FMT_SSVDAG};
    if (!rc) rc = svd_encode_formats(o, f, levels >= 3 ? 4u : 2u, nullptr);
    if (!rc) rc = svd_ctx_synchronize(c);
    svd_octree_free(o);
    return rc;
  };
  double t[SVD_BUILD_PRESETS];
  for (int k = 0; k < SVD_BUILD_PRESETS; ++k) {
    svd_set_build_limits(c, presets[k]);
    std::vector<double> v;
    for (uint32_t r = 0; r <= reps; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = build_once();
      if (rc) {
        svd_set_build_limits(c, keep);
        return rc;
      }
      const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (r) v.push_back(dt);
    }
    std::sort(v.begin(), v.end());
    t[k] = v[v.size() / 2];
  }
  int b = 0;
  for (int k = 1; k < SVD_BUILD_PRESETS; ++k) if (t[k] < t[b]) b = k;
  svd_set_build_limits(c, presets[b]);
  if (best) for (int i = 0; i < 4; ++i) best[i] = presets[b][i];
  if (ms) for (int k = 0; k < SVD_BUILD_PRESETS; ++k) ms[k] = t[k];
  return SVD_OK;
}

int svd_autotune_trace_modes(svd_scene* s, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad,
                             uint32_t reps, uint32_t* best_tile, uint32_t* best_mode, double* ms) {
  if (!s || !p || !prim || reps < 1) return SVD_ERR_ARG;
  svd_ctx* c = s->ctx;
  return guarded(c, [&] {
    const uint32_t shapes[4] = {8, 16, 32, 64};
    double t[8];
    for (int m = 0; m < 2; ++m)
      for (int k = 0; k < 4; ++k) {
        c->traceMode = (uint32_t)m;
        c->traceTileW = shapes[k];
        t[4 * m + k] = time_frames(c, s, p, prim, shad, reps);
      }
    int b = 0;
    for (int k = 1; k < 8; ++k) if (t[k] < t[b]) b = k;
    c->traceMode = (uint32_t)(b / 4);
    c->traceTileW = shapes[b % 4];
    if (best_tile) *best_tile = c->traceTileW;
    if (best_mode) *best_mode = c->traceMode;
    if (ms) for (int k = 0; k < 8; ++k) ms[k] = t[k];
  });
}

int svd_workload_materials(svd_ctx* c, const uint64_t* codes, uint64_t n, uint32_t levels, uint64_t seed,
                           uint32_t nValues, double baseCells, uint8_t* out) {
  if (!c || (n && (!codes || !out)) || nValues < 1 || nValues > 256 || levels < 1 || levels > 21) return SVD_ERR_ARG;
  return guarded(c, [&] { gen_materials(c, codes, n, levels, seed, nValues, baseCells, out); });
}

int svd_codes_union(svd_ctx* c, const uint64_t* a, uint64_t na, const uint64_t* b, uint64_t nb, uint64_t** out,
                    uint64_t* n) {
  if (!c || !out || !n || (na && !a) || (nb && !b)) return SVD_ERR_ARG;
  return guarded(c, [&] { *n = codes_union(c, a, na, b, nb, out); });
}
