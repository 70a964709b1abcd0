// symvox_amd.cpp -- implementation of the reference-compatible C++ API on top
// of the C ABI of libsvdag_amd.so (GPU reductions, encoders, traversal).
#include "symvox_amd.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <stack>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace svd {

template <typename T>
void aabox<T>::to_empty() {
  for (int i = 0; i < 3; ++i) { b[0][i] = (T)1e30; b[1][i] = (T)-1e30; }
}
template <typename T>
void aabox<T>::merge(const vec3<T>& p) {
  for (int i = 0; i < 3; ++i) { b[0][i] = std::min(b[0][i], p[i]); b[1][i] = std::max(b[1][i], p[i]); }
}
template <typename T>
vec3<T> aabox<T>::center() const {
  return vec3<T>((b[0][0] + b[1][0]) * (T)0.5, (b[0][1] + b[1][1]) * (T)0.5, (b[0][2] + b[1][2]) * (T)0.5);
}
template <typename T>
vec3<T> aabox<T>::half_side_lengths() const {
  return vec3<T>((b[1][0] - b[0][0]) * (T)0.5, (b[1][1] - b[0][1]) * (T)0.5, (b[1][2] - b[0][2]) * (T)0.5);
}
template struct aabox<float>;
template struct aabox<double>;

svd_ctx* default_context() {
  static svd_ctx* ctx = nullptr;
  if (!ctx) {
    int rc = svd_ctx_create(0, &ctx);
    if (rc != SVD_OK) {
      fprintf(stderr, "ERROR: cannot create the GPU context (status %d)\n", rc);
      exit(1);
    }
  }
  return ctx;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void check(int rc, const char* what) {
  if (rc != SVD_OK) {
    fprintf(stderr, "ERROR: %s failed (%d): %s\n", what, rc, svd_last_error(default_context()));
    exit(1);
  }
}

}  // namespace svd

using namespace svd;

// ============================================================================ Scene
Scene::Scene() { _bbox.to_empty(); }

void Scene::loadObj(std::string fileName, bool, bool, bool, bool, bool doBuildTriVector) {
  printf("* Loading '%s'...\n", fileName.c_str());
  FILE* file = fopen(fileName.c_str(), "r");
  if (!file) {
    printf("ERROR: Scene:loadOBJ: Can't open file %s\n", fileName.c_str());
    exit(1);
  }
  _bbox.to_empty();
  _vertices.clear();
  _indexed.clear();
  const int len = 10000;
  std::vector<char> buf(len);
  char* str = buf.data();
  size_t nTris = 0;
  std::vector<int64_t> tri(3);
  while (fgets(str, len, file) != NULL) {
    if (str[0] == 'v') {
      if (str[1] == 'n' || str[1] == 't') continue;
      point3f p;
      if (sscanf(str + 1, "%f %f %f", &p[0], &p[1], &p[2]) != 3) continue;
      _bbox.merge(p);
      _vertices.push_back(p);
    } else if (str[0] == 'f') {
      // scene.cpp:138-216: tokens split on ' ', the last one loses its final char
      std::vector<std::string> triples;
      char* save = nullptr;
      for (char* pch = strtok_r(str + 1, " ", &save); pch; pch = strtok_r(nullptr, " ", &save)) triples.push_back(pch);
      if (triples.empty()) continue;
      if (!triples.back().empty()) triples.back().resize(triples.back().size() - 1);
      auto parse = [&](const std::string& s, int64_t& out) {
        int idxV = 0, idxN = 0;
        int n = sscanf(s.c_str(), "%d//%d", &idxV, &idxN);
        if (idxV < 0) idxV = int(_vertices.size()) + idxV + 1;
        if (n >= 1) out = (int64_t)idxV - 1;
      };
      for (int i = 0; i < 3 && i < (int)triples.size(); ++i) parse(triples[i], tri[i]);
      for (int i = 0; i < 3; ++i) _indexed.push_back((uint32_t)tri[i]);
      nTris++;
      if (triples.size() > 4) {
        for (int i = 0; i < 3; ++i) parse(triples[(i + 2) % 4], tri[i]);
        for (int i = 0; i < 3; ++i) _indexed.push_back((uint32_t)tri[i]);
        nTris++;
      }
    }
  }
  fclose(file);
  printf("\t- Loaded %zu triangles\n", nTris);
  printf("\t- Bbox: [%.3f %.3f %.3f]  [%.3f %.3f %.3f]\n", _bbox[0][0], _bbox[0][1], _bbox[0][2], _bbox[1][0],
         _bbox[1][1], _bbox[1][2]);
  if (doBuildTriVector) buildTriVector(true);
}

bool Scene::loadPly(std::string fileName) {
  printf("* Loading '%s'...\n", fileName.c_str());
  std::ifstream in(fileName);
  if (!in.is_open()) { printf("ERROR: can't open %s\n", fileName.c_str()); return false; }
  std::string line;
  size_t nv = 0, nf = 0;
  while (std::getline(in, line)) {
    if (line.rfind("element vertex", 0) == 0) nv = std::stoull(line.substr(15));
    else if (line.rfind("element face", 0) == 0) nf = std::stoull(line.substr(13));
    else if (line == "end_header") break;
  }
  _bbox.to_empty();
  _vertices.resize(nv);
  for (size_t i = 0; i < nv; ++i) {
    in >> _vertices[i][0] >> _vertices[i][1] >> _vertices[i][2];
    _bbox.merge(_vertices[i]);
  }
  _indexed.clear();
  for (size_t f = 0; f < nf; ++f) {
    size_t k;
    in >> k;
    std::vector<uint32_t> idx(k);
    for (size_t j = 0; j < k; ++j) in >> idx[j];
    for (size_t j = 1; j + 1 < k; ++j) { _indexed.push_back(idx[0]); _indexed.push_back(idx[j]); _indexed.push_back(idx[j + 1]); }
  }
  buildTriVector(true);
  printf("\t- Loaded %zu triangles\n", getNRawTriangles());
  return true;
}

// scene.cpp:394-416: only triangles whose three indices are valid
void Scene::buildTriVector(bool clearOtherData) {
  _triangles.clear();
  for (size_t t = 0; t + 2 < _indexed.size(); t += 3) {
    const uint32_t a = _indexed[t], b = _indexed[t + 1], c = _indexed[t + 2];
    if (a < _vertices.size() && b < _vertices.size() && c < _vertices.size()) {
      _triangles.push_back(_vertices[a]);
      _triangles.push_back(_vertices[b]);
      _triangles.push_back(_vertices[c]);
    }
  }
  if (clearOtherData) { _vertices.clear(); _vertices.shrink_to_fit(); _indexed.clear(); _indexed.shrink_to_fit(); }
}

// ============================================================================ GeomOctree
GeomOctree::GeomOctree(Scene* scene) : _scene(scene) { _bbox.to_empty(); }

GeomOctree::~GeomOctree() {
  if (_oct) svd_octree_free(_oct);
}

void GeomOctree::resizeSceneBbox(aabox3f bbox) {
  _bbox = bbox;
  const point3f s = _bbox.half_side_lengths();
  _rootSide = std::max(std::max(s[0] * 2.0f, s[1] * 2.0f), s[2] * 2.0f);
}

static_assert(sizeof(point3f) == 12, "triangles are handed over as packed float triples");

// geom_octree.cpp:171-280 on the device (svd_octree_from_mesh): the
// reference's creation order and FP64 triangle/box tests.
void GeomOctree::buildSVO(unsigned int levels, aabox3d bbox, bool internalCall) {
  if (!internalCall) printf("* Building SVO... ");
  fflush(stdout);
  const double t0 = now_s();
  if (_oct) { svd_octree_free(_oct); _oct = nullptr; }
  const double bb[6] = {bbox[0][0], bbox[0][1], bbox[0][2], bbox[1][0], bbox[1][1], bbox[1][2]};
  const size_t nt = _scene->getNRawTriangles();
  const float* tris = nt ? &_scene->getTrianglePtr(0)[0][0] : nullptr;
  check(svd_octree_from_mesh(default_context(), tris, nt, levels, bb, 0, &_oct), "buildSVO");
  _bbox = aabox3f(point3f((float)bb[0], (float)bb[1], (float)bb[2]), point3f((float)bb[3], (float)bb[4], (float)bb[5]));
  const point3f sides = _bbox.half_side_lengths();
  _rootSide = std::max(std::max(sides[0] * 2.0f, sides[1] * 2.0f), sides[2] * 2.0f);
  _levels = levels;
  _stats.buildSVOTime = now_s() - t0;
  uint64_t nn = 0, nv = 0, nl = 0;
  svd_octree_counts(_oct, &nn, &nv);
  svd_level_size(_oct, SVD_VIEW_SVO, levels - 1, &nl);
  _stats.nNodesSVO = nn;
  _stats.nNodesLastLevSVO = nl;
  _stats.simulatedEncodedSVOSize = (_stats.nNodesSVO - _stats.nNodesLastLevSVO) * 4;
  _stats.nTotalVoxels = nv;
  _state = S_SVO;
  _cacheState = -1;
  if (!internalCall) printf("OK! [%.3f s]\n", _stats.buildSVOTime);
}

// geom_octree.cpp:289-435 on the device (svd_octree_from_mesh_subtrees): the
// top SVO, one subtree per set voxel of its finest level voxelized in its own
// box, joined in the reference's order; one GPU reduction of the joined
// subtrees gives the same DAG as the reference's per-subtree toDAG followed by
// the final toDAG, because both keep the first occurrence of every key in join
// order.
void GeomOctree::buildDAG(unsigned int levels, unsigned int stepLevel, aabox3d bbox, bool verbose) {
  printf("* Building DAG [stepLevel: %i]\n", stepLevel);
  fflush(stdout);
  const double t0 = now_s();
  // no subtrees (a top of one level, or of every level): build the SVO and reduce it
  if (stepLevel == 0 || stepLevel + 1 >= levels) { buildSVO(levels, bbox, true); toDAG(true); return; }
  if (_oct) { svd_octree_free(_oct); _oct = nullptr; }
  const double bb[6] = {bbox[0][0], bbox[0][1], bbox[0][2], bbox[1][0], bbox[1][1], bbox[1][2]};
  const size_t nt = _scene->getNRawTriangles();
  const float* tris = nt ? &_scene->getTrianglePtr(0)[0][0] : nullptr;
  check(svd_octree_from_mesh_subtrees(default_context(), tris, nt, levels, stepLevel, bb, 0, &_oct), "buildDAG");
  _bbox = aabox3f(point3f((float)bb[0], (float)bb[1], (float)bb[2]), point3f((float)bb[3], (float)bb[4], (float)bb[5]));
  const point3f sides = _bbox.half_side_lengths();
  _rootSide = std::max(std::max(sides[0] * 2.0f, sides[1] * 2.0f), sides[2] * 2.0f);
  _levels = levels;
  uint64_t nn = 0, nv = 0, nl = 0;
  svd_octree_counts(_oct, &nn, &nv);
  svd_level_size(_oct, SVD_VIEW_SVO, levels - 1, &nl);
  _stats.nNodesSVO = nn;
  _stats.nNodesLastLevSVO = nl;
  _stats.simulatedEncodedSVOSize = (nn - nl) * 4;
  _stats.nTotalVoxels = nv;
  _state = S_SVO;
  _cacheState = -1;
  toDAG(true);
  _stats.buildDAGTime = now_s() - t0;
  if (verbose) printf("\t- Finished! Total time [%.3f s]\n", _stats.buildDAGTime);
}

bool GeomOctree::buildFromMorton(const uint64_t* codes, size_t n, unsigned int levels, bool onDevice) {
  if (_oct) { svd_octree_free(_oct); _oct = nullptr; }
  int rc = svd_octree_from_morton(default_context(), codes, n, levels, onDevice ? 1 : 0, &_oct);
  if (rc != SVD_OK) { printf("ERROR: %s\n", svd_last_error(default_context())); return false; }
  _levels = levels;
  _bbox = aabox3f(point3f(0, 0, 0), point3f(1, 1, 1));
  _rootSide = 1.0;
  uint64_t nn = 0, nv = 0;
  svd_octree_counts(_oct, &nn, &nv);
  _stats.nTotalVoxels = nv;
  _stats.nNodesSVO = nn;
  _state = S_SVO;
  _cacheState = -1;
  return true;
}

bool GeomOctree::fromSVDAG(const std::vector<uint8_t>& file) {
  if (_oct) { svd_octree_free(_oct); _oct = nullptr; }
  printf("Decoding nodes...\n");
  int rc = svd_octree_from_svdag(default_context(), file.data(), file.size(), 0, &_oct);
  if (rc != SVD_OK) { printf("ERROR: %s\n", svd_last_error(default_context())); _state = S_EMPTY; return false; }
  printf("Updating pointers...\n");
  float bb[6], rs = 0;
  uint32_t lv = 0;
  memcpy(bb, file.data(), 24);
  memcpy(&rs, file.data() + 24, 4);
  memcpy(&lv, file.data() + 28, 4);
  _bbox = aabox3f(point3f(bb[0], bb[1], bb[2]), point3f(bb[3], bb[4], bb[5]));
  _rootSide = rs;
  _levels = lv;
  uint64_t nn = 0, nv = 0;
  svd_octree_counts(_oct, &nn, &nv);
  _stats = Stats();
  _stats.nTotalVoxels = nv;     // 0: EncodedSVDAG::load reads no voxel count
  _stats.nNodesDAG = nn;        // the header's node count
  _state = S_DAG;
  _cacheState = -1;
  return true;
}

bool EncodedSVDAG::decode(GeomOctree& octree) const {
  if (_file.size() < 44) { printf("ERROR: no .svdag loaded\n"); return false; }
  return octree.fromSVDAG(_file);
}

void GeomOctree::toDAG(bool internalCall) {
  if (_state == S_SVO || _state == S_DAG) {
    if (!internalCall) printf("* Transforming SVO -> DAG ... ");
    fflush(stdout);
  } else {
    printf("ERROR! This is not a SVO!\n");
    return;
  }
  if (_state == S_DAG) return;   // the device DAG is already reduced
  const double t0 = now_s();
  check(svd_reduce_dag(_oct), "toDAG");
  _stats.toDAGTime = now_s() - t0;
  uint64_t nn = 0;
  svd_octree_counts(_oct, &nn, nullptr);
  _state = S_DAG;
  _stats.nNodesDAG = nn;
  svd_level_size(_oct, SVD_VIEW_DAG, _levels - 1, &nn);
  _stats.nNodesLastLevDAG = nn;
  _cacheState = -1;
  if (!internalCall) {
    printf("\n");
    for (unsigned lev = _levels - 1; lev > 0; --lev) {
      uint64_t a = 0, b = 0;
      svd_level_size(_oct, SVD_VIEW_SVO, lev, &a);
      svd_level_size(_oct, SVD_VIEW_DAG, lev, &b);
      printf("Reduced level %u from %lu to %lu nodes\n", lev, (unsigned long)a, (unsigned long)b);
    }
    printf("OK! [%.3f s]\n", _stats.toDAGTime);
  }
}

void GeomOctree::toSDAG(bool internalCall, bool skipSymmetry) {
  if (!internalCall) {
    if (_state == S_SVO) printf("* Transforming SVO -> SDAG (not best way)... ");
    else if (_state == S_DAG) printf("* Transforming DAG -> SDAG (normal)... ");
    else { printf("ERROR! This is not a DAG or SDAG!\n"); return; }
    fflush(stdout);
  }
  if (skipSymmetry) { printf("ERROR! pointer-only SDAG (skipSymmetry) is not supported\n"); return; }
  if (_state == S_SVO) toDAG(true);
  const double t0 = now_s();
  check(svd_reduce_sym(_oct), "toSDAG");
  _stats.toSDAGTime = now_s() - t0;
  uint64_t nn = 0;
  svd_octree_counts(_oct, &nn, nullptr);
  _stats.nNodesSDAG = nn;
  _state = S_SDAG;
  _cacheState = -1;
  if (!internalCall) printf("OK! [%.3f s]\n", _stats.toSDAGTime);
}

unsigned int GeomOctree::mergeAcrossAllLevels() {
  if (_state != S_DAG) { printf("ERROR! This is not a DAG!\n"); return 0; }
  const double t0 = now_s();
  uint64_t n = 0;
  int rc = svd_cross_level_merge(_oct, &n);
  _stats.crossMergeTime = now_s() - t0;
  if (rc != SVD_OK) printf("\t\t- %s\n", svd_last_error(default_context()));
  uint64_t nn = 0;
  svd_octree_counts(_oct, &nn, nullptr);
  _stats.nNodesDAG = nn;
  _stats.nCrossLevelMerged = n;
  _cacheState = -1;
  printf("Total number of nodes that was removed: %lu\n", (unsigned long)n);
  return (unsigned)n;
}

void GeomOctree::toLossyDag(float lossyInflation, float allowedLossyDiffFactor, int includedNodeRefCount) {
  (void)lossyInflation;
  if (_state != S_DAG) { printf("ERROR! This is not a DAG!\n"); return; }
  if (_levels < 3) { printf("ERROR! A lossy DAG needs at least 3 levels\n"); return; }
  printf("* Transforming DAG -> Lossy DAG3 ... \n");
  svd_lossy_params p = {};
  p.lev_lo = 1;
  p.lev_hi = _levels - 2;
  p.max_diff = 0;
  p.diff_factor = allowedLossyDiffFactor;
  p.max_refs = (uint32_t)std::max(1, includedNodeRefCount);
  p.list_cap = 32;
  p.proposer = SVD_PROPOSER_EXHAUSTIVE;
  std::vector<svd_lossy_level_stats> st(_levels);
  const double t0 = now_s();
  const int rc = svd_lossy_merge(_oct, &p, st.data(), (uint32_t)st.size());
  _stats.toLSVDAGTime = now_s() - t0;
  if (rc != SVD_OK) { printf("\t\t- %s\n", svd_last_error(default_context())); return; }
  size_t err = 0, merged = 0;
  for (unsigned i = 0; i + 2 < _levels; ++i) {
    printf("Level %u: max diff %u, %zu -> %zu nodes (%zu merged)\n", (unsigned)st[i].level, (unsigned)st[i].bound,
           (size_t)st[i].nodes_before, (size_t)st[i].nodes_after, (size_t)st[i].merged);
    err += st[i].error_bound;
    merged += st[i].merged;
  }
  const size_t before = _stats.nNodesDAG;
  uint64_t nn = 0;
  svd_octree_counts(_oct, &nn, nullptr);
  _stats.nNodesDAG = nn;
  _stats.totalLossyVoxelDifference = err;
  _stats.nClusteredNodes = merged;
  _cacheState = -1;
  printf("Lossy compressed from %zu to %zu nodes.\n", before, (size_t)nn);
}

size_t GeomOctree::getNNodes() const {
  uint64_t nn = 0;
  if (_oct) svd_octree_counts(_oct, &nn, nullptr);
  return nn;
}

const GeomOctree::NodeData& GeomOctree::getNodeData() const {
  if (_cacheState == (int)_state) return _cache;
  _cache.assign(_levels, {});
  const int view = _state == S_SDAG ? SVD_VIEW_SDAG : (_state == S_DAG ? SVD_VIEW_DAG : SVD_VIEW_SVO);
  for (unsigned l = 0; l < _levels && _oct; ++l) {
    uint64_t n = 0;
    svd_level_size(_oct, view, l, &n);
    std::vector<uint32_t> w(9 * n + 9);
    std::vector<uint8_t> inv(n + 1);
    std::vector<uint32_t> cl(8 * n + 8);
    if (n) svd_get_level(_oct, view, l, w.data(), inv.data(), cl.data());
    _cache[l].resize(n);
    for (uint64_t i = 0; i < n; ++i) {
      Node& nd = _cache[l][i];
      nd.childrenBitmask = (uint8_t)(w[9 * i] & 0xFF);
      nd.childrenMirroredBitmask[0] = (uint8_t)(w[9 * i] >> 8);
      nd.childrenMirroredBitmask[1] = (uint8_t)(w[9 * i] >> 16);
      nd.childrenMirroredBitmask[2] = (uint8_t)(w[9 * i] >> 24);
      nd.invariantBitmask = inv[i];
      for (int c = 0; c < 8; ++c) { nd.children[c] = w[9 * i + 1 + c]; nd.childLevels[c] = cl[8 * i + c]; }
    }
  }
  _cacheState = (int)_state;
  return _cache;
}

// ============================================================================ encoders
bool EncodedOctree::encodeWith(const GeomOctree& octree, int fmt, int needState, const char* what) {
  if (needState >= 0 && octree.getState() != needState) return false;
  (void)what;
  uint64_t bytes = 0;
  check(svd_encode(octree.handle(), fmt, &bytes), "encode");
  _file.resize(bytes);
  check(svd_encoded_copy(octree.handle(), fmt, _file.data(), bytes), "encoded copy");
  parse();
  return true;
}

bool EncodedOctree::save(const std::string filename) const {
  printf("* Saving '%s'... ", filename.c_str());
  fflush(stdout);
  std::ofstream ofs(filename, std::ios::out | std::ios::binary);
  if (!ofs.is_open()) { printf("FAILED!!!\n"); return false; }
  ofs.write((const char*)_file.data(), (std::streamsize)_file.size());
  printf("OK!\n");
  return true;
}

bool EncodedOctree::load(const std::string filename) {
  printf("* Loading '%s'... ", filename.c_str());
  fflush(stdout);
  std::ifstream ifs(filename, std::ios::in | std::ios::binary);
  if (!ifs.is_open()) { printf("FAILED!!!\n"); return false; }
  _file.assign(std::istreambuf_iterator<char>(ifs), std::istreambuf_iterator<char>());
  parse();
  printf("OK!\n");
  return true;
}

static void parse_header(const std::vector<uint8_t>& f, aabox3f& bb, float& rs, unsigned& lv, size_t& nn) {
  if (f.size() < 36) return;
  memcpy(bb[0].to_pointer(), f.data(), 12);
  memcpy(bb[1].to_pointer(), f.data() + 12, 12);
  memcpy(&rs, f.data() + 24, 4);
  memcpy(&lv, f.data() + 28, 4);
  uint32_t n32 = 0;
  memcpy(&n32, f.data() + 32, 4);
  nn = n32;
}

void EncodedSVDAG::parse() {
  parse_header(_file, _sceneBBox, _rootSide, _levels, _nNodes);
  if (_file.size() >= 44) {
    memcpy(&_firstLeafPtr, _file.data() + 36, 4);
    uint32_t cnt = 0;
    memcpy(&cnt, _file.data() + 40, 4);
    _nWords = cnt;
  }
}

void EncodedSVDAG::encode(const GeomOctree& octree) {
  printf("* Encoding to SVDAG... ");
  if (!encodeWith(octree, SVD_FMT_SVDAG, GeomOctree::S_DAG, "SVDAG")) {
    printf("FAILED! Octree is not in DAG state\n");
    return;
  }
  printf("OK!\n");
}

void EncodedUSSVDAG::encode(const GeomOctree& octree) {
  printf("* Encoding to USSVDAG... ");
  if (!encodeWith(octree, SVD_FMT_USSVDAG, GeomOctree::S_SDAG, "USSVDAG")) {
    printf("FAILED! Octree is not in SDAG state\n");
    return;
  }
  printf("OK!\n");
}

void EncodedSSVDAG::parse() {
  parse_header(_file, _sceneBBox, _rootSide, _levels, _nNodes);
  if (_file.size() < 40) return;
  uint32_t nInner = 0, nLeaf = 0, nOff = 0;
  memcpy(&nInner, _file.data() + 36, 4);
  const size_t lc = 40 + 2ull * nInner;
  if (lc + 4 > _file.size()) return;
  memcpy(&nLeaf, _file.data() + lc, 4);
  const size_t oc = lc + 4 + nLeaf;
  if (oc + 4 > _file.size()) return;
  memcpy(&nOff, _file.data() + oc, 4);
  _nInner = nInner;
  _nLeafBytes = nLeaf;
  _levelOffsets.resize(nOff);
  if (oc + 4 + 4ull * nOff <= _file.size()) memcpy(_levelOffsets.data(), _file.data() + oc + 4, 4ull * nOff);
}

void EncodedSSVDAG::encode(const GeomOctree& octree) {
  printf("* Encoding to SSVDAG... ");
  if (octree.getState() != GeomOctree::S_SDAG && octree.getState() != GeomOctree::S_DAG) {
    printf("FAILED! Octree is not in SDAG state\n");
    return;
  }
  _esvdag = octree.getState() == GeomOctree::S_DAG;
  encodeWith(octree, _esvdag ? SVD_FMT_ESVDAG : SVD_FMT_SSVDAG, -1, "SSVDAG");
  printf("OK!\n");
}

// ============================================================================ traversal
bool traceHeadless(const EncodedOctree& eo, const svd_trace_params& p, HeadlessTraceResult& out,
                   std::vector<svd_ray_result>* primary, float beamProjection) {
  svd_ctx* c = default_context();
  const std::vector<uint8_t>& f = eo.fileImage();
  int fmt = eo.format();
  if (fmt == SVD_FMT_ESVDAG) fmt = SVD_FMT_SSVDAG;
  svd_scene* sc = nullptr;
  if (svd_scene_create(c, fmt, f.data(), f.size(), 0, &sc) != SVD_OK) {
    printf("ERROR: %s\n", svd_last_error(c));
    return false;
  }
  const size_t n = (size_t)p.width * p.height;
  svd_ray_result *dp = nullptr, *ds = nullptr;
  if (hipMalloc((void**)&dp, n * sizeof(svd_ray_result)) != hipSuccess ||
      hipMalloc((void**)&ds, n * sizeof(svd_ray_result)) != hipSuccess) {
    printf("ERROR: device allocation failed\n");
    svd_scene_free(sc);
    return false;
  }
  int rc = beamProjection > 0.f ? svd_trace_beam(sc, &p, beamProjection, dp, ds, nullptr) : svd_trace(sc, &p, dp, ds);
  double ms = 0;
  svd_last_kernel_ms(c, "trace", &ms);
  std::vector<svd_ray_result> hp(n), hs(n);
  if (rc == SVD_OK) {
    (void)hipMemcpy(hp.data(), dp, n * sizeof(svd_ray_result), hipMemcpyDeviceToHost);
    (void)hipMemcpy(hs.data(), ds, n * sizeof(svd_ray_result), hipMemcpyDeviceToHost);
  }
  (void)hipFree(dp);
  (void)hipFree(ds);
  svd_scene_free(sc);
  if (rc != SVD_OK) { printf("ERROR: %s\n", svd_last_error(c)); return false; }
  out.kernelMs = ms;
  out.primaryRays = n;
  out.hits = 0;
  out.shadowRays = 0;
  for (size_t i = 0; i < n; ++i) {
    if (hp[i].t >= 0) out.hits++;
    if (hs[i].t != -5.f) out.shadowRays++;
  }
  out.mraysPerSec = (double)(out.primaryRays + out.shadowRays) / (ms * 1e-3) / 1e6;
  if (primary) *primary = hp;
  return true;
}

namespace {

// Scene of an encoded file on the default context (ESVDAG traces as SSVDAG).
svd_scene* headless_scene(const EncodedOctree& eo) {
  svd_ctx* c = default_context();
  const std::vector<uint8_t>& f = eo.fileImage();
  int fmt = eo.format();
  if (fmt == SVD_FMT_ESVDAG) fmt = SVD_FMT_SSVDAG;
  svd_scene* sc = nullptr;
  if (svd_scene_create(c, fmt, f.data(), f.size(), 0, &sc) != SVD_OK) {
    printf("ERROR: %s\n", svd_last_error(c));
    return nullptr;
  }
  return sc;
}

}  // namespace

bool renderHeadless(const EncodedOctree& eo, const svd_trace_params& p, const svd_view_params& v,
                    float beamProjection, std::vector<float>& rgb, double* kernelMs) {
  svd_ctx* c = default_context();
  svd_scene* sc = headless_scene(eo);
  if (!sc) return false;
  const size_t n = (size_t)p.width * p.height;
  float* d = nullptr;
  if (hipMalloc((void**)&d, 3 * n * sizeof(float) + 4) != hipSuccess) {
    printf("ERROR: device allocation failed\n");
    svd_scene_free(sc);
    return false;
  }
  const int rc = svd_render(sc, &p, &v, beamProjection, d, nullptr);
  if (rc == SVD_OK) {
    rgb.resize(3 * n);
    (void)hipMemcpy(rgb.data(), d, 3 * n * sizeof(float), hipMemcpyDeviceToHost);
    if (kernelMs) svd_last_kernel_ms(c, "trace", kernelMs);
  }
  (void)hipFree(d);
  svd_scene_free(sc);
  if (rc != SVD_OK) { printf("ERROR: %s\n", svd_last_error(c)); return false; }
  return true;
}

bool aoHeadless(const EncodedOctree& eo, const svd_trace_params& p, uint32_t numRays, std::vector<float>& vis,
                double* kernelMs) {
  svd_ctx* c = default_context();
  svd_scene* sc = headless_scene(eo);
  if (!sc) return false;
  const size_t n = (size_t)p.width * p.height;
  float* d = nullptr;
  if (hipMalloc((void**)&d, n * sizeof(float) + 4) != hipSuccess) {
    printf("ERROR: device allocation failed\n");
    svd_scene_free(sc);
    return false;
  }
  const int rc = svd_trace_ao(sc, &p, numRays, d, nullptr);
  if (rc == SVD_OK) {
    vis.resize(n);
    (void)hipMemcpy(vis.data(), d, n * sizeof(float), hipMemcpyDeviceToHost);
    if (kernelMs) svd_last_kernel_ms(c, "trace", kernelMs);
  }
  (void)hipFree(d);
  svd_scene_free(sc);
  if (rc != SVD_OK) { printf("ERROR: %s\n", svd_last_error(c)); return false; }
  return true;
}
