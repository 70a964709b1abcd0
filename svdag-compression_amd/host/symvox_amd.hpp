// symvox_amd.hpp -- C++ API of the MI355X sparse-voxel-DAG engine, mirroring
// the reference's GeomOctree / EncodedOctree / Scene classes
// (src/symvox/geom_octree.hpp:32-190, encoded_octree.hpp:28-98,
// encoded_svdag.hpp, encoded_ussvdag.hpp, encoded_ssvdag.hpp, scene.hpp).
//
// Same class names, method names, argument meaning, state machine and
// error behaviour (printf + return false / silent return on a wrong state).
// The only source-level difference: the SpaceLand geometry types of the
// reference signatures are replaced by the small svd::vec3 / svd::aabox types
// below.  The mesh voxelizer (FP64 triangle/box tests in the reference's
// triangle-major order), the reductions and the encoders all run on the GPU
// through the C ABI (include/svdag_amd.h); the host parses the input files.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/svdag_amd.h"

namespace svd {

template <typename T>
struct vec3 {
  T v[3];
  vec3() : v{0, 0, 0} {}
  vec3(T x, T y, T z) : v{x, y, z} {}
  T& operator[](int i) { return v[i]; }
  const T& operator[](int i) const { return v[i]; }
  T* to_pointer() { return v; }
  const T* to_pointer() const { return v; }
};
typedef vec3<float> point3f;
typedef vec3<double> point3d;

template <typename T>
struct aabox {
  vec3<T> b[2];
  aabox() { to_empty(); }
  aabox(const vec3<T>& lo, const vec3<T>& hi) { b[0] = lo; b[1] = hi; }
  void to_empty();
  bool is_empty() const { return b[0][0] > b[1][0]; }
  void merge(const vec3<T>& p);
  vec3<T> center() const;
  vec3<T> half_side_lengths() const;
  vec3<T>& operator[](int i) { return b[i]; }
  const vec3<T>& operator[](int i) const { return b[i]; }
};
typedef aabox<float> aabox3f;
typedef aabox<double> aabox3d;

// One context per process by default (device 0); created lazily.
svd_ctx* default_context();

}  // namespace svd

// ----------------------------------------------------------------------------
// Scene (src/symvox/scene.hpp:55-91): ASCII OBJ (and PLY) triangle soup.
// ----------------------------------------------------------------------------
class Scene {
 public:
  // scene.hpp:40-44: vertex and normal indices and the material of one triangle
  struct TIndexedTri {
    std::size_t verticesIdx[3];
    std::size_t normalsIdx[3];
    std::size_t material;
  };
  Scene();
  // Scene::loadObj (scene.cpp:51-268): same line parser and quirks (face
  // tokens split on ' ', the last token loses its final character, a face
  // is split into two triangles only when it has more than 4 tokens,
  // negative indices count from the end, "v//n" gives a normal index, the
  // box merges all vertices).  With tryLoadBinCache, `<fileName>.bincache`
  // is loaded instead of the text when it exists; after parsing the text the
  // cache is always written (scene.cpp:55-71, 257-258, 324-374), in the
  // reference's layout: a header of four size_t counts and the box, then the
  // vertices, normals, materials (300-byte records) and indexed triangles.
  // loadNormals keeps "vn" normals (recomputeNormals or a face without them
  // adds the face's cross product, :166-174); materials (mtllib / usemtl) are
  // not parsed, so every triangle has material 0 and the cache holds none.
  void loadObj(std::string fileName, bool tryLoadBinCache = true, bool loadMaterials = false,
               bool loadNormals = false, bool recomputeNormals = false, bool doBuildTriVector = true);
  // scene.cpp:324-374
  void saveBinObj(std::string fileName) const;
  void loadBinObj(std::string fileName);
  // ASCII PLY with "vertex" (x y z floats) and "face" (vertex_indices lists).
  bool loadPly(std::string fileName);
  void buildTriVector(bool clearOtherData = false);
  inline std::size_t getNRawTriangles() const { return _triangles.size() / 3; }
  inline const svd::point3f* getTrianglePtr(std::size_t id) const {
    return (id * 3 < _triangles.size()) ? &_triangles[id * 3] : nullptr;
  }
  inline svd::aabox3f getAABB() const { return _bbox; }
  inline void setAABB(svd::aabox3f bbox) { _bbox = bbox; }
  inline std::size_t getNTriangles() const { return _indexedTris.size() > 0 ? _indexedTris.size() : _triangles.size() / 3; }
  inline std::size_t getNVertices() const { return _vertices.size(); }
  inline std::vector<TIndexedTri>* getIndexedGeom() { return &_indexedTris; }
  inline const std::vector<svd::point3f>& getNormals() const { return _normals; }

 private:
  svd::aabox3f _bbox;
  std::vector<svd::point3f> _vertices;
  std::vector<svd::point3f> _normals;
  std::vector<TIndexedTri> _indexedTris;
  std::vector<svd::point3f> _triangles; // 3 points per triangle
};

// ----------------------------------------------------------------------------
// GeomOctree (src/symvox/geom_octree.hpp:32-190)
// ----------------------------------------------------------------------------
class GeomOctree {
 public:
  typedef uint32_t id_t;
  static const id_t nullNode = 0xFFFFFFFEu;
  enum State { S_EMPTY = 0, S_SVO = 1, S_DAG = 2, S_SDAG = 3 };

  struct Stats {
    size_t nTotalVoxels = 0;
    size_t nNodesSVO = 0, nNodesDAG = 0, nNodesSDAG = 0;
    size_t nNodesLastLevSVO = 0, nNodesLastLevDAG = 0, nNodesLastLevSDAG = 0;
    size_t simulatedEncodedSVOSize = 0, memFootprint = 0;
    double buildSVOTime = 0, buildDAGTime = 0, toDAGTime = 0, toLSVDAGTime = 0, toSDAGTime = 0;
    double lHashing = 0, lSimNodes = 0, lClustering = 0, crossMergeTime = 0;   // seconds
    size_t totalLossyVoxelDifference = 0, nClusteredNodes = 0, nClusters = 0, nEdges = 0, nCrossLevelMerged = 0;
  };

  struct Node {
    id_t children[8];
    uint8_t childrenBitmask = 0;
    unsigned int childLevels[8];
    uint8_t childrenMirroredBitmask[3] = {0, 0, 0};
    uint8_t invariantBitmask = 0;
    Node() { for (int i = 0; i < 8; ++i) { children[i] = nullNode; childLevels[i] = 0xFFFFFFFFu; } }
    bool existsChild(int c) const { return (childrenBitmask & (1u << c)) != 0; }
    bool existsChildPointer(int c) const { return children[c] != nullNode; }
    bool hasChildren() const { return childrenBitmask != 0; }
    unsigned int getNChildren() const { return (unsigned)__builtin_popcount(childrenBitmask); }
    bool getChildMirrrorBit(int axis, int c) const { return (childrenMirroredBitmask[axis] >> c) & 1u; }
    bool getInvariantBit(int axis) const { return (invariantBitmask >> axis) & 1u; }
  };
  typedef std::vector<std::vector<Node>> NodeData;

  explicit GeomOctree(Scene* scene);
  // geom_octree.cpp:41-46: a deep copy (its device levels and encoded images too)
  GeomOctree(const GeomOctree& other);
  GeomOctree& operator=(const GeomOctree& other);
  GeomOctree(GeomOctree&& other) noexcept;
  GeomOctree& operator=(GeomOctree&& other) noexcept;
  // geom_octree.cpp:49-60: a DAG-state octree from host levels; it reports
  // stats.nNodesDAG nodes and stats.nTotalVoxels voxels
  GeomOctree(const NodeData& data, svd::aabox3f bbox, double rootSide, unsigned int levels, Stats stats);
  ~GeomOctree();

  // geom_octree.cpp:171-280: voxelization on the device in the reference's node order
  void buildSVO(unsigned int levels, svd::aabox3d bbox, bool internalCall = false);
  // geom_octree.cpp:62-169 on the device (svd_octree_from_points): the points
  // of a LAS file, read by readLasPoints in place of libLAS.  leavesCenters
  // (the reference's buildDAG hook, unused by its point path) is not offered.
  void buildSVOFromPoints(std::string fileName, unsigned int levels, svd::aabox3d bbox, bool internalCall = false);
  // the same from points in memory: xyz = n points, x y z doubles each
  void buildSVOFromPoints(const double* xyz, size_t n, unsigned int levels, svd::aabox3d bbox,
                          bool internalCall = false);
  // geom_octree.cpp:289-435: the top SVO and every subtree (each in its own
  // box) voxelized on the device, joined in the reference's order and reduced.
  void buildDAG(unsigned int levels, unsigned int stepLevel, svd::aabox3d bbox, bool verbose = false);
  // Voxel list entry (Morton codes, svd_octree_from_morton).
  bool buildFromMorton(const uint64_t* codes, size_t n, unsigned int levels, bool codesOnDevice = false);
  // An .svdag file image decoded on the device (svd_octree_from_svdag); used by EncodedSVDAG::decode.
  bool fromSVDAG(const std::vector<uint8_t>& file);
  void toDAG(bool internalCall = false);
  // geom_octree.cpp:551-697; skipSymmetry: every node its own representative
  // (pointer-only compression, :591-596)
  void toSDAG(bool internalCall = false, bool skipSymmetry = false);
  unsigned int mergeAcrossAllLevels();
  // geom_octree_extension.cpp:797-1179 on the GPU: levels L-2..1, bound
  // round((L-lev)^2 * allowedLossyDiffFactor), nodes with at most
  // includedNodeRefCount effective references.  lossyInflation steered the
  // reference's MCL run, which this build replaces by its deterministic
  // first-representative rule; it is accepted and ignored.
  void toLossyDag(float lossyInflation, float allowedLossyDiffFactor, int includedNodeRefCount);
  void initChildLevels() {}   // child levels default to lev+1 on the device

  // getters (geom_octree.hpp:111-117, octree.hpp:101-114)
  const NodeData& getNodeData() const;   // copied from the device on demand
  Stats& getStats() { return _stats; }
  State getState() const { return _state; }
  size_t getNNodes() const;
  size_t getNVoxels() const { return _stats.nTotalVoxels; }
  unsigned int getLevels() const { return _levels; }
  svd::aabox3f getSceneBBox() const { return _bbox; }
  float getRootSide() const { return float(_rootSide); }
  double getRootSideD() const { return _rootSide; }
  void resizeSceneBbox(svd::aabox3f bbox);
  svd_octree* handle() const { return _oct; }

 private:

  Scene* _scene;
  svd_octree* _oct = nullptr;
  svd::aabox3f _bbox;
  double _rootSide = 0;
  unsigned int _levels = 0;
  State _state = S_EMPTY;
  Stats _stats;
  mutable NodeData _cache;
  mutable int _cacheState = -1;
};

// ----------------------------------------------------------------------------
// Encoded octrees (src/symvox/encoded_octree.hpp:28-98 and the three encoders)
// ----------------------------------------------------------------------------
// The point records of a LAS file (public header block of LAS 1.0-1.4, any
// point data format; the first 12 bytes of a record are the raw X Y Z int32s)
// as libLAS's Point::GetX/GetY/GetZ return them: raw * scale + offset, in
// file order.  false when the file is not a readable LAS file.
bool readLasPoints(const std::string& fileName, std::vector<double>& xyz);

class EncodedOctree {
 public:
  virtual ~EncodedOctree() {}
  virtual void encode(const GeomOctree& octree) = 0;
  virtual bool load(const std::string filename);
  virtual bool save(const std::string filename) const;
  virtual std::string getDescription() const = 0;
  virtual size_t getDataSize() const = 0;

  // Traversal interface (encoded_octree.hpp:60-76), read from the host file
  // image.  A TravNode's idx is the node's position in the format's array:
  // a word index (SVDAG, USSVDAG), a u16 index of an inner node or the byte
  // offset of a 4^3 leaf brick (SSVDAG).  Like the reference, getChild
  // leaves the child's center at (0,0,0) and prints a warning when asked for
  // a child of a node at the maximum level.
  struct TravNode {
    TravNode() : idx(0), level(0), center(0.f, 0.f, 0.f), mX(false), mY(false), mZ(false) {}
    uint32_t idx;
    unsigned int level;
    svd::point3f center;
    bool mX, mY, mZ;
  };
  virtual TravNode getRootTravNode() const;
  virtual bool hasChild(const TravNode& node, const int c) const = 0;
  virtual TravNode getChild(const TravNode& node, const int c, bool& mX, bool& mY, bool& mZ) const = 0;
  virtual bool isLeaf(const TravNode& node) const = 0;
  virtual int getLeafSize() const = 0;
  virtual bool hasVoxel(const TravNode& leaf, const int i, const int j, const int k) const = 0;

  inline svd::aabox3f getSceneBBox() const { return _sceneBBox; }
  inline float getRootSide() const { return _rootSide; }
  inline svd::point3f getCenter() const { return _sceneBBox.center(); }
  inline unsigned int getNLevels() const { return _levels; }
  inline size_t getNNodes() const { return _nNodes; }
  // encode() takes the octree's voxel count; a loaded file has none (0)
  inline size_t getNVoxels() const { return _nVoxels; }
  inline size_t getNLeaves() const { return _nLeaves; }
  inline float getHalfSide(const unsigned int level = 0) const { return _rootSide / float(1u << (level + 1)); }
  // the complete file image (what save() writes)
  const std::vector<uint8_t>& fileImage() const { return _file; }
  virtual int format() const = 0;

 protected:
  bool encodeWith(const GeomOctree& octree, int fmt, int needState, const char* what);
  virtual void parse() = 0;
  svd::aabox3f _sceneBBox;
  float _rootSide = 0;
  unsigned int _levels = 0;
  size_t _nNodes = 0, _nVoxels = 0, _nLeaves = 0;
  std::vector<uint8_t> _file;
};

class EncodedSVDAG : public EncodedOctree {
 public:
  void encode(const GeomOctree& octree) override;
  std::string getDescription() const override { return "SVDAG. Encoded just like [Kampe et al.] paper."; }
  size_t getDataSize() const override { return _nWords * sizeof(uint32_t); }
  uint32_t getFirstLeafPtr() const { return _firstLeafPtr; }
  const uint32_t* getDataPtr() const { return _file.empty() ? nullptr : (const uint32_t*)(_file.data() + 44); }
  int format() const override { return SVD_FMT_SVDAG; }
  // encoded_svdag.cpp:216-281: the DAG-state octree of this file, by value
  // (on a malformed file: printed, and an empty octree comes back)
  GeomOctree decode(Scene& scene) const;
  // the same into an existing octree; false on a malformed file
  bool decode(GeomOctree& octree) const;
  bool hasChild(const TravNode& node, const int c) const override;
  TravNode getChild(const TravNode& node, const int c, bool& mX, bool& mY, bool& mZ) const override;
  bool isLeaf(const TravNode& node) const override { return node.level == (_levels - 1); }
  int getLeafSize() const override { return 2; }
  bool hasVoxel(const TravNode& leaf, const int i, const int j, const int k) const override;

 protected:
  void parse() override;
  uint32_t _firstLeafPtr = 0;
  size_t _nWords = 0;
};

class EncodedUSSVDAG : public EncodedSVDAG {
 public:
  void encode(const GeomOctree& octree) override;
  std::string getDescription() const override { return "USSVDAG. Symmetry-aware DAG, uncompressed encoding."; }
  int format() const override { return SVD_FMT_USSVDAG; }
  // encoded_ussvdag.cpp:185-199: the child's mirror bits come from the header word
  TravNode getChild(const TravNode& node, const int c, bool& mX, bool& mY, bool& mZ) const override;
};

class EncodedSSVDAG : public EncodedOctree {
 public:
  void encode(const GeomOctree& octree) override;
  std::string getDescription() const override { return "SSVDAG. Symmetry-aware DAG, with compressed encoding."; }
  size_t getDataSize() const override {
    return _nInner * sizeof(uint16_t) + _nLeafBytes + _levelOffsets.size() * sizeof(uint32_t);
  }
  unsigned int getInnerNodesSize() const { return (unsigned)(_nInner * sizeof(uint16_t)); }
  unsigned int getNInnerNodes() const { return (unsigned)_nInner; }
  unsigned int getLeafNodesSize() const { return (unsigned)_nLeafBytes; }
  unsigned int getNLevelOffsets() const { return (unsigned)_levelOffsets.size(); }
  const uint32_t* getLevelOffsetsPtr() const { return _levelOffsets.data(); }
  // encoded_ssvdag.hpp:58-61: the u16 inner-node array and the leaf bytes
  void* getInnerNodesDataPtr() const { return (void*)innerPtr(); }
  void* getLeafNodesDataPtr() const { return (void*)leafPtr(); }
  int format() const override { return _esvdag ? SVD_FMT_ESVDAG : SVD_FMT_SSVDAG; }
  bool hasChild(const TravNode& node, const int c) const override;
  TravNode getChild(const TravNode& node, const int c, bool& mX, bool& mY, bool& mZ) const override;
  bool isLeaf(const TravNode& node) const override { return node.level == (_levels - 2); }
  int getLeafSize() const override { return 4; }
  bool hasVoxel(const TravNode& leaf, const int i, const int j, const int k) const override;

 protected:
  void parse() override;
  const uint16_t* innerPtr() const { return _file.size() >= 40 ? (const uint16_t*)(_file.data() + 40) : nullptr; }
  const uint8_t* leafPtr() const {
    return _file.size() >= 44 + 2 * _nInner ? _file.data() + 44 + 2 * _nInner : nullptr;
  }
  size_t _nInner = 0, _nLeafBytes = 0;
  std::vector<uint32_t> _levelOffsets;
  bool _esvdag = false;
};

// Headless traversal of an encoded file (svviewer's OctreeDDARenderer).
struct HeadlessTraceResult {
  double kernelMs = 0, mraysPerSec = 0;
  uint64_t primaryRays = 0, shadowRays = 0, hits = 0;
};
// beamProjection > 0: the viewer's minimum-depth pre-pass with that projection
// factor (svd_trace_beam); 0: every primary ray starts at the camera.
bool traceHeadless(const EncodedOctree& eo, const svd_trace_params& params, HeadlessTraceResult& out,
                   std::vector<svd_ray_result>* primary = nullptr, float beamProjection = 0.f);
// The viewer's colour per pixel (3 floats, VIEWER_MODE main) with its render
// mode, normal samples and colour switches in `view` (svd_render).
bool renderHeadless(const EncodedOctree& eo, const svd_trace_params& params, const svd_view_params& view,
                    float beamProjection, std::vector<float>& rgb, double* kernelMs = nullptr);
// AO_MODE visibility per pixel over the primary rays' hits (svd_trace_ao).
bool aoHeadless(const EncodedOctree& eo, const svd_trace_params& params, uint32_t numRays, std::vector<float>& vis,
                double* kernelMs = nullptr);
