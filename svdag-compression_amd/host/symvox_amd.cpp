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

void Scene::loadObj(std::string fileName, bool tryLoadBinCache, bool, bool loadNormals, bool recomputeNormals,
                    bool doBuildTriVector) {
  printf("* Loading '%s'...\n", fileName.c_str());
  const std::string cache = fileName + ".bincache";
  if (tryLoadBinCache) {
    std::ifstream probe(cache);
    if (probe.good()) {
      probe.close();
      printf("\t- Binary cache found! Loading '%s'... ", cache.c_str());
      loadBinObj(cache);
      printf("OK!\n");
      if (doBuildTriVector) buildTriVector(true);
      printf("\t- Loaded %zu triangles\n", getNRawTriangles());
      printf("\t- Bbox: [%.3f %.3f %.3f]  [%.3f %.3f %.3f]\n", _bbox[0][0], _bbox[0][1], _bbox[0][2], _bbox[1][0],
             _bbox[1][1], _bbox[1][2]);
      return;
    }
  }
  FILE* file = fopen(fileName.c_str(), "r");
  if (!file) {
    printf("ERROR: Scene:loadOBJ: Can't open file %s\n", fileName.c_str());
    exit(1);
  }
  _bbox.to_empty();
  _vertices.clear();
  _normals.clear();
  _indexedTris.clear();
  const int len = 10000;
  std::vector<char> buf(len);
  char* str = buf.data();
  size_t nLine = 0;
  // one face's indices; a token that does not parse leaves the previous value
  TIndexedTri itri{};
  while (fgets(str, len, file) != NULL) {
    nLine++;
    if (str[0] == 'v') {
      if (str[1] == 'n') {
        if (loadNormals) {
          point3f nv;
          sscanf(str + 2, "%f %f %f", &nv[0], &nv[1], &nv[2]);
          _normals.push_back(nv);
        }
        continue;
      }
      if (str[1] == 't') continue;
      point3f p;
      if (sscanf(str + 1, "%f %f %f", &p[0], &p[1], &p[2]) != 3) {
        printf("Read in line %zu a bad vertex format\n", nLine);
        continue;
      }
      _bbox.merge(p);
      _vertices.push_back(p);
    } else if (str[0] == 'f') {
      // scene.cpp:138-216: tokens split on ' ', the last one loses its final char
      std::vector<std::string> triples;
      char* save = nullptr;
      for (char* pch = strtok_r(str + 1, " ", &save); pch; pch = strtok_r(nullptr, " ", &save)) triples.push_back(pch);
      if (triples.empty()) continue;
      if (!triples.back().empty()) triples.back().resize(triples.back().size() - 1);
      bool noNormal = false;
      auto parse = [&](const std::string& t, int i) {
        int idxV = 0, idxN = 0;
        const int n = sscanf(t.c_str(), "%d//%d", &idxV, &idxN);
        if (idxV < 0) idxV = int(_vertices.size()) + idxV + 1;
        if (idxN < 0) idxN = int(_normals.size()) + idxN + 1;
        if (n == 0) printf("ERROR:Scene:loadObj: problem parsing face instruction\n");
        else if (n == 1) { itri.verticesIdx[i] = (std::size_t)(int64_t)(idxV - 1); noNormal = true; }
        else if (n == 2) {
          itri.verticesIdx[i] = (std::size_t)(int64_t)(idxV - 1);
          if (!recomputeNormals) itri.normalsIdx[i] = (std::size_t)(int64_t)(idxN - 1);
        }
      };
      auto faceNormal = [&]() {
        if (!(loadNormals && (noNormal || recomputeNormals))) return;
        const size_t nv = _vertices.size();
        point3f nrm;
        if (itri.verticesIdx[0] < nv && itri.verticesIdx[1] < nv && itri.verticesIdx[2] < nv) {
          const point3f& a = _vertices[itri.verticesIdx[0]];
          const point3f& b = _vertices[itri.verticesIdx[1]];
          const point3f& c = _vertices[itri.verticesIdx[2]];
          const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
          const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
          nrm = point3f(e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]);
        }
        itri.normalsIdx[0] = itri.normalsIdx[1] = itri.normalsIdx[2] = _normals.size();
        _normals.push_back(nrm);
      };
      itri.material = 0;
      for (int i = 0; i < 3 && i < (int)triples.size(); ++i) parse(triples[i], i);
      faceNormal();
      _indexedTris.push_back(itri);
      if (triples.size() > 4) {          // a quad, when a trailing token follows its 4 indices
        for (int i = 0; i < 3; ++i) parse(triples[(i + 2) % 4], i);
        faceNormal();
        _indexedTris.push_back(itri);
      }
    }
  }
  fclose(file);
  printf("\t- Loaded %zu triangles in %.2fM obj lines\n", _indexedTris.size(), nLine / 1e6);
  printf("\t- Bbox: [%.3f %.3f %.3f]  [%.3f %.3f %.3f]\n", _bbox[0][0], _bbox[0][1], _bbox[0][2], _bbox[1][0],
         _bbox[1][1], _bbox[1][2]);
  printf("\t- Saving '%s' binary cache... ", cache.c_str());
  saveBinObj(cache);
  printf("OK!\n");
  if (doBuildTriVector) buildTriVector(true);
}

namespace {
// scene.hpp:109-119 and the 300-byte TMaterial record (char name[256], three
// colours, specular coefficient, transparency)
struct BinObjHeader {
  uint64_t nVertices, nNormals, nMaterials, nIndexedTris;
  float bboxMin[3], bboxMax[3];
};
static_assert(sizeof(BinObjHeader) == 56, "bincache header layout");
static_assert(sizeof(Scene::TIndexedTri) == 56, "bincache triangle layout");
constexpr size_t kMaterialBytes = 300;
}  // namespace

void Scene::saveBinObj(std::string fileName) const {
  BinObjHeader h{};
  h.nVertices = _vertices.size();
  h.nNormals = _normals.size();
  h.nMaterials = 0;
  h.nIndexedTris = _indexedTris.size();
  for (int i = 0; i < 3; ++i) { h.bboxMin[i] = _bbox[0][i]; h.bboxMax[i] = _bbox[1][i]; }
  std::ofstream f(fileName, std::ios::out | std::ios::binary);
  if (!f.is_open()) return;
  f.write((const char*)&h, sizeof(h));
  if (!_vertices.empty()) f.write((const char*)_vertices.data(), _vertices.size() * sizeof(point3f));
  if (!_normals.empty()) f.write((const char*)_normals.data(), _normals.size() * sizeof(point3f));
  if (!_indexedTris.empty()) f.write((const char*)_indexedTris.data(), _indexedTris.size() * sizeof(TIndexedTri));
}

void Scene::loadBinObj(std::string fileName) {
  std::ifstream f(fileName, std::ios::in | std::ios::binary);
  BinObjHeader h{};
  f.read((char*)&h, sizeof(h));
  for (int i = 0; i < 3; ++i) { _bbox[0][i] = h.bboxMin[i]; _bbox[1][i] = h.bboxMax[i]; }
  _vertices.resize(h.nVertices);
  if (h.nVertices) f.read((char*)_vertices.data(), h.nVertices * sizeof(point3f));
  _normals.resize(h.nNormals);
  if (h.nNormals) f.read((char*)_normals.data(), h.nNormals * sizeof(point3f));
  f.seekg((std::streamoff)(h.nMaterials * kMaterialBytes), std::ios::cur);   // materials are not kept
  _indexedTris.resize(h.nIndexedTris);
  if (h.nIndexedTris) f.read((char*)_indexedTris.data(), h.nIndexedTris * sizeof(TIndexedTri));
  if (!f) {
    printf("ERROR: Scene:loadBinObj: '%s' is truncated\n", fileName.c_str());
    exit(1);
  }
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
  _indexedTris.clear();
  for (size_t f = 0; f < nf; ++f) {
    size_t k;
    in >> k;
    std::vector<size_t> idx(k);
    for (size_t j = 0; j < k; ++j) in >> idx[j];
    for (size_t j = 1; j + 1 < k; ++j) {
      TIndexedTri t{};
      t.verticesIdx[0] = idx[0];
      t.verticesIdx[1] = idx[j];
      t.verticesIdx[2] = idx[j + 1];
      _indexedTris.push_back(t);
    }
  }
  buildTriVector(true);
  printf("\t- Loaded %zu triangles\n", getNRawTriangles());
  return true;
}

// scene.cpp:394-416: only triangles whose three indices are valid
void Scene::buildTriVector(bool clearOtherData) {
  _triangles.clear();
  for (const TIndexedTri& t : _indexedTris) {
    const size_t a = t.verticesIdx[0], b = t.verticesIdx[1], c = t.verticesIdx[2];
    if (a < _vertices.size() && b < _vertices.size() && c < _vertices.size()) {
      _triangles.push_back(_vertices[a]);
      _triangles.push_back(_vertices[b]);
      _triangles.push_back(_vertices[c]);
    }
  }
  if (clearOtherData) {
    _vertices.clear(); _vertices.shrink_to_fit();
    _normals.clear(); _normals.shrink_to_fit();
    _indexedTris.clear(); _indexedTris.shrink_to_fit();
  }
}

// ============================================================================ GeomOctree
GeomOctree::GeomOctree(Scene* scene) : _scene(scene) { _bbox.to_empty(); }

GeomOctree::GeomOctree(const GeomOctree& other)
    : _scene(other._scene), _bbox(other._bbox), _rootSide(other._rootSide), _levels(other._levels),
      _state(other._state), _stats(other._stats) {
  if (other._oct) check(svd_octree_copy(other._oct, &_oct), "octree copy");
}

GeomOctree& GeomOctree::operator=(const GeomOctree& other) {
  if (this == &other) return *this;
  GeomOctree tmp(other);
  *this = std::move(tmp);
  return *this;
}

GeomOctree::GeomOctree(GeomOctree&& other) noexcept
    : _scene(other._scene), _oct(other._oct), _bbox(other._bbox), _rootSide(other._rootSide),
      _levels(other._levels), _state(other._state), _stats(other._stats), _cache(std::move(other._cache)),
      _cacheState(other._cacheState) {
  other._oct = nullptr;
  other._state = S_EMPTY;
  other._cacheState = -1;
}

GeomOctree& GeomOctree::operator=(GeomOctree&& other) noexcept {
  if (this == &other) return *this;
  if (_oct) svd_octree_free(_oct);
  _scene = other._scene;
  _oct = other._oct;
  _bbox = other._bbox;
  _rootSide = other._rootSide;
  _levels = other._levels;
  _state = other._state;
  _stats = other._stats;
  _cache = std::move(other._cache);
  _cacheState = other._cacheState;
  other._oct = nullptr;
  other._state = S_EMPTY;
  other._cacheState = -1;
  return *this;
}

GeomOctree::GeomOctree(const NodeData& data, aabox3f bbox, double rootSide, unsigned int levels, Stats stats)
    : _scene(nullptr), _bbox(bbox), _rootSide(rootSide), _levels(levels), _state(S_DAG), _stats(stats) {
  std::vector<std::vector<uint32_t>> words(levels);
  std::vector<uint64_t> sizes(levels, 0);
  std::vector<const uint32_t*> ptrs(levels, nullptr);
  for (unsigned l = 0; l < levels && l < data.size(); ++l) {
    sizes[l] = data[l].size();
    words[l].resize(9 * data[l].size() + 9);
    for (size_t i = 0; i < data[l].size(); ++i) {
      const Node& n = data[l][i];
      words[l][9 * i] = n.childrenBitmask;
      for (int c = 0; c < 8; ++c) words[l][9 * i + 1 + c] = n.children[c];
    }
    ptrs[l] = words[l].data();
  }
  const float bb[6] = {bbox[0][0], bbox[0][1], bbox[0][2], bbox[1][0], bbox[1][1], bbox[1][2]};
  check(svd_octree_from_levels(default_context(), levels, sizes.data(), ptrs.data(), bb, rootSide, stats.nNodesDAG,
                               stats.nTotalVoxels, &_oct),
        "GeomOctree(NodeData)");
}

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

void GeomOctree::buildSVOFromPoints(const double* xyz, size_t n, unsigned int levels, aabox3d bbox,
                                    bool internalCall) {
  if (!internalCall) printf("* Building SVO... ");
  fflush(stdout);
  const double t0 = now_s();
  if (_oct) { svd_octree_free(_oct); _oct = nullptr; }
  const double bb[6] = {bbox[0][0], bbox[0][1], bbox[0][2], bbox[1][0], bbox[1][1], bbox[1][2]};
  check(svd_octree_from_points(default_context(), xyz, n, levels, bb, 0, &_oct), "buildSVOFromPoints");
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

void GeomOctree::buildSVOFromPoints(std::string fileName, unsigned int levels, aabox3d bbox, bool internalCall) {
  std::vector<double> xyz;
  if (!readLasPoints(fileName, xyz)) throw std::runtime_error("buildSVOFromPoints: cannot read LAS file " + fileName);
  buildSVOFromPoints(xyz.data(), xyz.size() / 3, levels, bbox, internalCall);
}

bool readLasPoints(const std::string& fileName, std::vector<double>& xyz) {
  std::ifstream f(fileName, std::ios::binary);
  if (!f) return false;
  uint8_t h[375] = {0};
  f.read((char*)h, sizeof(h));
  const std::streamsize got = f.gcount();
  if (got < 227 || memcmp(h, "LASF", 4) != 0) return false;
  auto u16 = [&](int o) { uint16_t v; memcpy(&v, h + o, 2); return v; };
  auto u32 = [&](int o) { uint32_t v; memcpy(&v, h + o, 4); return v; };
  auto u64 = [&](int o) { uint64_t v; memcpy(&v, h + o, 8); return v; };
  auto f64 = [&](int o) { double v; memcpy(&v, h + o, 8); return v; };
  const uint8_t verMinor = h[25];
  const uint32_t dataOffset = u32(96);
  const uint16_t recLen = u16(105);
  uint64_t count = u32(107);
  if (count == 0 && verMinor >= 4 && got >= 255) count = u64(247);   // LAS 1.4: 64-bit count
  const double scale[3] = {f64(131), f64(139), f64(147)}, offset[3] = {f64(155), f64(163), f64(171)};
  if (recLen < 12) return false;
  f.clear();
  f.seekg(dataOffset, std::ios::beg);
  xyz.assign(3 * count, 0.0);
  std::vector<uint8_t> rec(recLen);
  for (uint64_t i = 0; i < count; ++i) {
    f.read((char*)rec.data(), recLen);
    if (f.gcount() != recLen) return false;
    for (int a = 0; a < 3; ++a) {
      int32_t raw;
      memcpy(&raw, rec.data() + 4 * a, 4);
      xyz[3 * i + a] = (raw * scale[a]) + offset[a];
    }
  }
  return true;
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

GeomOctree EncodedSVDAG::decode(Scene& scene) const {
  GeomOctree octree(&scene);
  decode(octree);
  return octree;
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
  if (_state != S_SVO && _state != S_DAG) return;
  if (_state == S_SVO && !skipSymmetry) toDAG(true);
  const double t0 = now_s();
  // skipSymmetry: pointer-only compression, every node its own representative (geom_octree.cpp:591-596)
  check(skipSymmetry ? svd_reduce_sym_pointers(_oct) : svd_reduce_sym(_oct), "toSDAG");
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
  // the reference's statistics (geom_octree_extension.cpp:930, 1050, 1061, 1073-1089) in this rule's terms:
  // hashing = candidate lists (and the learned proposer's embedding and kNN), similar nodes = exact
  // verification, clustering = the representative resolution and merge; an edge is a verified pair, a
  // cluster a representative with its merged members
  size_t err = 0, merged = 0, clusters = 0, edges = 0;
  double msHash = 0, msSim = 0, msClust = 0;
  for (unsigned i = 0; i + 2 < _levels; ++i) {
    printf("Level %u: max diff %u, %zu -> %zu nodes (%zu merged)\n", (unsigned)st[i].level, (unsigned)st[i].bound,
           (size_t)st[i].nodes_before, (size_t)st[i].nodes_after, (size_t)st[i].merged);
    err += st[i].error_bound;
    merged += st[i].merged;
    clusters += st[i].clusters;
    edges += st[i].verified;
    msHash += st[i].ms_embed + st[i].ms_knn + st[i].ms_lists;
    msSim += st[i].ms_verify;
    msClust += st[i].ms_merge;
  }
  const size_t before = _stats.nNodesDAG;
  uint64_t nn = 0;
  svd_octree_counts(_oct, &nn, nullptr);
  _stats.nNodesDAG = nn;
  _stats.totalLossyVoxelDifference = err;
  _stats.nClusters = clusters;
  _stats.nClusteredNodes = merged + clusters;
  _stats.nEdges = edges;
  _stats.lHashing = msHash * 1e-3;
  _stats.lSimNodes = msSim * 1e-3;
  _stats.lClustering = msClust * 1e-3;
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
  _nVoxels = octree.getNVoxels();       // encoded_svdag.cpp:118, encoded_ssvdag.cpp:226
  return true;
}

EncodedOctree::TravNode EncodedOctree::getRootTravNode() const { return TravNode(); }

// ---- TravNode API over the host file image (encoded_svdag.cpp:318-354,
// encoded_ussvdag.cpp:172-210, encoded_ssvdag.cpp:468-524).  Indices outside
// the arrays (a malformed image) read as absent children instead of memory.
bool EncodedSVDAG::hasChild(const TravNode& node, const int c) const {
  const uint32_t* d = getDataPtr();
  if (!d || node.idx >= _nWords) return false;
  return (d[node.idx] & (1u << c)) != 0;
}

EncodedOctree::TravNode EncodedSVDAG::getChild(const TravNode& node, const int c, bool& mX, bool& mY,
                                               bool& mZ) const {
  TravNode tn;
  if (node.level >= _levels - 1) {
    printf("EncodedSVDAG::getChild: WARNING! Asking for a node child, but node is in its max level\n");
    return tn;
  }
  const uint32_t* d = getDataPtr();
  if (!d || node.idx >= _nWords) return tn;
  const uint8_t mask = (uint8_t)(d[node.idx] & 0xFFu);
  const size_t at = node.idx + (size_t)__builtin_popcount((unsigned)(mask >> c));
  tn.idx = at < _nWords ? d[at] : 0;
  tn.level = node.level + 1;
  mX = mY = mZ = false;
  return tn;
}

bool EncodedSVDAG::hasVoxel(const TravNode& leaf, const int i, const int j, const int k) const {
  return hasChild(leaf, 4 * i + 2 * j + k);
}

EncodedOctree::TravNode EncodedUSSVDAG::getChild(const TravNode& node, const int c, bool& mX, bool& mY,
                                                 bool& mZ) const {
  TravNode tn;
  if (node.level >= _levels - 1) {
    printf("EncodedUSSVDAG::getChild: WARNING! Asking for a node child, but node is in its max level\n");
    return tn;
  }
  const uint32_t* d = getDataPtr();
  if (!d || node.idx >= _nWords) return tn;
  const uint32_t h = d[node.idx];
  const size_t at = node.idx + (size_t)__builtin_popcount((unsigned)((h & 0xFFu) >> c));
  tn.idx = at < _nWords ? d[at] : 0;
  tn.level = node.level + 1;
  mX = (h & (1u << (c + 8))) != 0;
  mY = (h & (1u << (c + 16))) != 0;
  mZ = (h & (1u << (c + 24))) != 0;
  return tn;
}

bool EncodedSSVDAG::hasChild(const TravNode& node, const int c) const {
  const uint16_t* in = innerPtr();
  if (!in || node.idx >= _nInner) return false;
  return ((in[node.idx] >> (2u * c)) & 3u) != 0;
}

EncodedOctree::TravNode EncodedSSVDAG::getChild(const TravNode& node, const int c, bool& mX, bool& mY,
                                                bool& mZ) const {
  TravNode tn;
  if (node.level >= _levels - 2) {
    printf("EncodedSSVDAG::getChild: WARNING! Asking for a node child, but node is in its max level\n");
    return tn;
  }
  const uint16_t* in = innerPtr();
  if (!in || node.idx >= _nInner) return tn;
  const uint32_t hdr = in[node.idx];
  // 16-bit pointers take one u16, 32-bit ones two (size codes 1 / 2, 3)
  size_t off = 1;
  for (int i = 7; i > c; --i) off += std::min<uint32_t>((hdr >> (2u * i)) & 3u, 2u);
  if (node.idx + off >= _nInner) return tn;
  uint32_t idx = in[node.idx + off];
  const uint32_t code = (hdr >> (2u * c)) & 3u;
  mX = ((idx >> 13u) & 1u) == 1;
  mY = ((idx >> 14u) & 1u) == 1;
  mZ = ((idx >> 15u) & 1u) == 1;
  idx &= ~(7u << 13u);
  if (code > 1) {
    if (node.idx + off + 1 >= _nInner) return tn;
    idx = (idx << 16) | in[node.idx + off + 1];
    idx |= (code & 1u) << 29u;
  }
  tn.level = node.level + 1;
  if (tn.level < _levels - 2) tn.idx = idx + (tn.level < _levelOffsets.size() ? _levelOffsets[tn.level] : 0u);
  else tn.idx = idx * 8;
  return tn;
}

bool EncodedSSVDAG::hasVoxel(const TravNode& leaf, const int i, const int j, const int k) const {
  const uint8_t* lv = leafPtr();
  const int c = i + 4 * j + 16 * k;
  const size_t byte = (size_t)leaf.idx + (size_t)(c / 8);
  if (!lv || byte >= _nLeafBytes) return false;
  return (lv[byte] & (1u << (c % 8))) != 0;
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
  _nVoxels = 0;                         // the files hold no voxel count
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
