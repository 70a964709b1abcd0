// ============================================================================
// svdag_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A plain CPU restatement of the reference (SymVox / SVDAG-Compression)
// lossless builder, its on-disk encoders and its GLSL octree-DDA traversal.
// It exists to CHECK the MI355X product path: only tests/, the smoke() entry
// and bench.py's cpu_baseline leg may load it.  It is never linked into, or
// called by, the product library.
//
// Every function cites the reference file:line it restates.  The data layout
// (vector<vector<Node>>, std::map dedup, libstdc++ std::sort) is kept on
// purpose: the oracle must reproduce the reference's node order and the
// reference's CPU cost profile (it is also the CPU baseline).
//
// Pinning status of each part is recorded in oracle/README.md.
// ============================================================================
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <map>
#include <set>
#include <stack>
#include <vector>
#include <string>
#include <algorithm>
#include <array>
#include <utility>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef uint32_t id_t;
static const id_t kNull = 0xFFFFFFFEu;  // octree.cpp:22 (numeric_limits<uint32>::max()-1)

// --------------------------------------------------------------------------
// Node record: octree.hpp:57-85 + geom_octree.hpp:59-96.  Key fields for
// operator==/operator< are (mask, children[0..7], mirror[0..2]);
// octree_node.cpp:39-86.  invariant / childLevels are payload only.
// --------------------------------------------------------------------------
struct Node {
  id_t children[8];
  uint8_t mask;
  uint32_t childLevels[8];
  uint8_t mirror[3];
  uint8_t invariant;
  Node() : mask(0), invariant(0) {
    for (int i = 0; i < 8; ++i) { children[i] = kNull; childLevels[i] = 0xFFFFFFFFu; }
    mirror[0] = mirror[1] = mirror[2] = 0;
  }
  bool existsChild(int c) const { return (mask >> c) & 1u; }
  bool existsChildPointer(int c) const { return children[c] != kNull; }
  bool hasChildren() const { return mask != 0; }
  unsigned nChildren() const { return (unsigned)__builtin_popcount(mask); }
  bool getMirrorBit(int a, int c) const { return (mirror[a] >> c) & 1u; }
  void setMirrorBit(int a, int c) { mirror[a] |= (uint8_t)(1u << c); }
  void unsetMirrorBit(int a, int c) { mirror[a] &= (uint8_t)~(1u << c); }
  void invMirrorBit(int a, int c) { mirror[a] ^= (uint8_t)(1u << c); }
  bool getInv(int a) const { return (invariant >> a) & 1u; }

  // octree_node.cpp:39-54
  bool operator==(const Node& o) const {
    if (mask != o.mask) return false;
    for (int i = 0; i < 8; ++i) if (children[i] != o.children[i]) return false;
    for (int a = 0; a < 3; ++a) if (mirror[a] != o.mirror[a]) return false;
    return true;
  }
  // octree_node.cpp:56-86 (lexicographic on the same fields)
  bool operator<(const Node& o) const {
    if (mask != o.mask) return mask < o.mask;
    for (int i = 0; i < 8; ++i) if (children[i] != o.children[i]) return children[i] < o.children[i];
    for (int a = 0; a < 3; ++a) if (mirror[a] != o.mirror[a]) return mirror[a] < o.mirror[a];
    return false;
  }
};

// Reflect the child-slot order of one axis: slot c moves to c ^ flip.
// X flip = 4, Y flip = 2, Z flip = 1 (child index 4x+2y+z, octree.hpp:33-42).
static inline uint8_t permute_bits(uint8_t m, int flip) {
  uint8_t r = 0;
  for (int c = 0; c < 8; ++c) if ((m >> c) & 1u) r |= (uint8_t)(1u << (c ^ flip));
  return r;
}

// Node::mirror (octree_node.cpp:88-195).  The returned node is a fresh
// Node (invariant bits cleared, childLevels reset) exactly like the
// reference, which builds it from a default-constructed n2.
static Node mirror_node(const Node& in, bool x, bool y, bool z, bool applyToChildren = true) {
  Node n = in;
  const bool ax[3] = {x, y, z};
  const int flips[3] = {4, 2, 1};
  for (int a = 0; a < 3; ++a) {
    if (!ax[a]) continue;
    Node n2;
    n2.mask = permute_bits(n.mask, flips[a]);
    for (int i = 0; i < 3; ++i) n2.mirror[i] = permute_bits(n.mirror[i], flips[a]);
    for (int c = 0; c < 8; ++c) n2.children[c ^ flips[a]] = n.children[c];
    n = n2;
  }
  if (applyToChildren) {
    for (int c = 0; c < 8; ++c) {
      if (n.existsChildPointer(c)) {
        if (x) n.invMirrorBit(0, c);
        if (y) n.invMirrorBit(1, c);
        if (z) n.invMirrorBit(2, c);
      }
    }
  }
  // The reference's single-axis path copies n2 over n only when an axis is
  // applied; with no axis it returns *this unchanged (payload included).
  return n;
}

enum State { S_EMPTY = 0, S_SVO = 1, S_DAG = 2, S_SDAG = 3 };

struct Octree {
  std::vector<std::vector<Node>> data;
  unsigned levels = 0;
  State state = S_EMPTY;
  size_t nNodes = 0, nVoxels = 0;
  float bboxMin[3] = {0, 0, 0}, bboxMax[3] = {1, 1, 1};
  double rootSide = 1.0;
  size_t nNodesSVO = 0, nNodesDAG = 0, nNodesSDAG = 0;
  size_t nCrossLevelMerged = 0;
  std::vector<size_t> dagCounts;   // "Reduced level l from N to U" (geom_octree.cpp:508-509)
  // toSDAG's correspondence of every DAG node, per level (the MirroredNode
  // list of geom_octree.cpp:566-572): sdag id | mX << 29 | mY << 30 | mZ << 31
  std::vector<std::vector<uint32_t>> sdagMap;
  int lastError = 0;
};

// --------------------------------------------------------------------------
// Akenine-Moller triangle/box overlap in double (test_triangle_box.cpp:38-184)
// --------------------------------------------------------------------------
static bool planeBoxOverlap(const double normal[3], const double vert[3], double maxbox) {
  double vmin[3], vmax[3];
  for (int q = 0; q < 3; ++q) {
    double v = vert[q];
    if (normal[q] > 0.0f) { vmin[q] = -maxbox - v; vmax[q] = maxbox - v; }
    else { vmin[q] = maxbox - v; vmax[q] = -maxbox - v; }
  }
  double d1 = normal[0] * vmin[0] + normal[1] * vmin[1] + normal[2] * vmin[2];
  if (d1 > 0.0) return false;
  double d2 = normal[0] * vmax[0] + normal[1] * vmax[1] + normal[2] * vmax[2];
  if (d2 >= 0.0) return true;
  return false;
}

static bool testTriBox(const double c[3], double h, const float* tri) {
  double v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
  for (int i = 0; i < 3; ++i) {
    v0[i] = (double)tri[0 * 3 + i] - c[i];
    v1[i] = (double)tri[1 * 3 + i] - c[i];
    v2[i] = (double)tri[2 * 3 + i] - c[i];
  }
  for (int i = 0; i < 3; ++i) { e0[i] = v1[i] - v0[i]; e1[i] = v2[i] - v1[i]; e2[i] = v0[i] - v2[i]; }
  double mn, mx, p0, p1, p2, rad, fex, fey, fez;
  const int X = 0, Y = 1, Z = 2;
#define AXT_X01(a, b, fa, fb) p0 = a * v0[Y] - b * v0[Z]; p2 = a * v2[Y] - b * v2[Z]; \
  if (p0 < p2) { mn = p0; mx = p2; } else { mn = p2; mx = p0; } rad = (fa + fb) * h; if (mn > rad || mx < -rad) return false;
#define AXT_X2(a, b, fa, fb) p0 = a * v0[Y] - b * v0[Z]; p1 = a * v1[Y] - b * v1[Z]; \
  if (p0 < p1) { mn = p0; mx = p1; } else { mn = p1; mx = p0; } rad = (fa + fb) * h; if (mn > rad || mx < -rad) return false;
#define AXT_Y02(a, b, fa, fb) p0 = -a * v0[X] + b * v0[Z]; p2 = -a * v2[X] + b * v2[Z]; \
  if (p0 < p2) { mn = p0; mx = p2; } else { mn = p2; mx = p0; } rad = (fa + fb) * h; if (mn > rad || mx < -rad) return false;
#define AXT_Y1(a, b, fa, fb) p0 = -a * v0[X] + b * v0[Z]; p1 = -a * v1[X] + b * v1[Z]; \
  if (p0 < p1) { mn = p0; mx = p1; } else { mn = p1; mx = p0; } rad = (fa + fb) * h; if (mn > rad || mx < -rad) return false;
#define AXT_Z12(a, b, fa, fb) p1 = a * v1[X] - b * v1[Y]; p2 = a * v2[X] - b * v2[Y]; \
  if (p2 < p1) { mn = p2; mx = p1; } else { mn = p1; mx = p2; } rad = (fa + fb) * h; if (mn > rad || mx < -rad) return false;
#define AXT_Z0(a, b, fa, fb) p0 = a * v0[X] - b * v0[Y]; p1 = a * v1[X] - b * v1[Y]; \
  if (p0 < p1) { mn = p0; mx = p1; } else { mn = p1; mx = p0; } rad = (fa + fb) * h; if (mn > rad || mx < -rad) return false;
  fex = fabs(e0[X]); fey = fabs(e0[Y]); fez = fabs(e0[Z]);
  AXT_X01(e0[Z], e0[Y], fez, fey); AXT_Y02(e0[Z], e0[X], fez, fex); AXT_Z12(e0[Y], e0[X], fey, fex);
  fex = fabs(e1[X]); fey = fabs(e1[Y]); fez = fabs(e1[Z]);
  AXT_X01(e1[Z], e1[Y], fez, fey); AXT_Y02(e1[Z], e1[X], fez, fex); AXT_Z0(e1[Y], e1[X], fey, fex);
  fex = fabs(e2[X]); fey = fabs(e2[Y]); fez = fabs(e2[Z]);
  AXT_X2(e2[Z], e2[Y], fez, fey); AXT_Y1(e2[Z], e2[X], fez, fex); AXT_Z12(e2[Y], e2[X], fey, fex);
#undef AXT_X01
#undef AXT_X2
#undef AXT_Y02
#undef AXT_Y1
#undef AXT_Z12
#undef AXT_Z0
  auto minmax3 = [](double a, double b, double c, double& lo, double& hi) {
    lo = hi = a; if (b < lo) lo = b; if (b > hi) hi = b; if (c < lo) lo = c; if (c > hi) hi = c; };
  minmax3(v0[X], v1[X], v2[X], mn, mx); if (mn > h || mx < -h) return false;
  minmax3(v0[Y], v1[Y], v2[Y], mn, mx); if (mn > h || mx < -h) return false;
  minmax3(v0[Z], v1[Z], v2[Z], mn, mx); if (mn > h || mx < -h) return false;
  double normal[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
  if (!planeBoxOverlap(normal, v0, h)) return false;
  return true;
}

// cleanEmptyNodes (geom_octree.cpp:437-456): unset parent bits of empty
// children (the empty nodes themselves stay in the level arrays).
static void cleanEmptyNodes(Octree& o) {
  for (int lev = (int)o.levels - 1; lev > 0; --lev) {
    std::set<id_t> empty;
    for (id_t i = 0; i < o.data[lev].size(); ++i) if (!o.data[lev][i].hasChildren()) empty.insert(i);
    if (empty.empty()) continue;
    for (id_t i = 0; i < o.data[lev - 1].size(); ++i)
      for (int j = 0; j < 8; ++j)
        if (empty.count(o.data[lev - 1][i].children[j])) {
          o.data[lev - 1][i].mask &= (uint8_t)~(1u << j);
          o.data[lev - 1][i].children[j] = kNull;
        }
  }
}

static void finishSVO(Octree& o) {
  o.nVoxels = 0;
  for (const Node& n : o.data[o.levels - 1]) o.nVoxels += n.nChildren();
  o.state = S_SVO;
  o.nNodesSVO = o.nNodes;
}

// buildSVO (geom_octree.cpp:171-280): triangle-major DFS with an explicit
// stack; children visited 7..0; a missing child is appended to the next
// level, which fixes the reference SVO node order.
static void buildSVOMesh(Octree& o, const float* tris, size_t ntris, unsigned levels, const double bmin[3], const double bmax[3]) {
  for (int i = 0; i < 3; ++i) { o.bboxMin[i] = (float)bmin[i]; o.bboxMax[i] = (float)bmax[i]; }
  o.levels = levels;
  float sides[3];
  for (int i = 0; i < 3; ++i) sides[i] = ((o.bboxMax[i] - o.bboxMin[i]) * 0.5f) * 2.0f;   // half_side_lengths()*2
  float rs = std::max(std::max(sides[0], sides[1]), sides[2]);
  o.rootSide = rs;
  o.data.assign(levels, {});
  o.data[0].push_back(Node());
  o.nNodes = 0;
  double center[3];
  for (int i = 0; i < 3; ++i) center[i] = (bmin[i] + bmax[i]) * 0.5;
  struct QI { id_t id; uint8_t level; double c[3]; };
  std::stack<QI> st;
  for (size_t t = 0; t < ntris; ++t) {
    const float* tri = tris + 9 * t;
    st.push(QI{0, 0, {center[0], center[1], center[2]}});
    while (!st.empty()) {
      QI qi = st.top(); st.pop();
      double k = o.rootSide / double(1u << (qi.level + 2));  // getHalfSideD(level+1)
      double cc[8][3];
      for (int c = 0; c < 8; ++c) {
        cc[c][0] = qi.c[0] + ((c & 4) ? +k : -k);
        cc[c][1] = qi.c[1] + ((c & 2) ? +k : -k);
        cc[c][2] = qi.c[2] + ((c & 1) ? +k : -k);
      }
      for (int i = 7; i >= 0; --i) {
        if (testTriBox(cc[i], k, tri)) {
          Node& node = o.data[qi.level][qi.id];
          node.mask |= (uint8_t)(1u << i);
          if (!node.existsChildPointer(i) && (qi.level < (levels - 1))) {
            node.children[i] = (id_t)o.data[qi.level + 1].size();
            o.data[qi.level + 1].emplace_back();
            o.nNodes++;
          }
          if ((qi.level + 1u) < levels) {
            const Node& nn = o.data[qi.level][qi.id];
            st.push(QI{nn.children[i], (uint8_t)(qi.level + 1), {cc[i][0], cc[i][1], cc[i][2]}});
          }
        }
      }
    }
  }
  cleanEmptyNodes(o);
  finishSVO(o);
}

// buildSVOFromPoints (geom_octree.cpp:62-169) with the points given in
// memory (the reference reads them with libLAS): point-major DFS with an
// explicit stack, the child centres of buildSVO, a point in child i when
// c - k <= p < c + k on every axis (:131-133), children created in the
// 7..0 loop, then cleanEmptyNodes (:155).
static void buildSVOPoints(Octree& o, const double* xyz, size_t npts, unsigned levels, const double bmin[3],
                           const double bmax[3]) {
  for (int i = 0; i < 3; ++i) { o.bboxMin[i] = (float)bmin[i]; o.bboxMax[i] = (float)bmax[i]; }
  o.levels = levels;
  float sides[3];
  for (int i = 0; i < 3; ++i) sides[i] = ((o.bboxMax[i] - o.bboxMin[i]) * 0.5f) * 2.0f;   // half_side_lengths()*2
  o.rootSide = std::max(std::max(sides[0], sides[1]), sides[2]);
  o.data.assign(levels, {});
  o.data[0].push_back(Node());
  o.nNodes = 0;
  double center[3];
  for (int i = 0; i < 3; ++i) center[i] = (bmin[i] + bmax[i]) * 0.5;
  struct QI { id_t id; uint8_t level; double c[3]; };
  std::stack<QI> st;
  for (size_t t = 0; t < npts; ++t) {
    const double* p = xyz + 3 * t;
    st.push(QI{0, 0, {center[0], center[1], center[2]}});
    while (!st.empty()) {
      QI qi = st.top(); st.pop();
      const double k = o.rootSide / double(1u << (qi.level + 2));   // getHalfSideD(level+1)
      double cc[8][3];
      for (int c = 0; c < 8; ++c) {
        cc[c][0] = qi.c[0] + ((c & 4) ? +k : -k);
        cc[c][1] = qi.c[1] + ((c & 2) ? +k : -k);
        cc[c][2] = qi.c[2] + ((c & 1) ? +k : -k);
      }
      for (int i = 7; i >= 0; --i) {
        if ((p[0] >= cc[i][0] - k && p[0] < cc[i][0] + k) && (p[1] >= cc[i][1] - k && p[1] < cc[i][1] + k) &&
            (p[2] >= cc[i][2] - k && p[2] < cc[i][2] + k)) {
          Node& node = o.data[qi.level][qi.id];
          node.mask |= (uint8_t)(1u << i);
          if (!node.existsChildPointer(i) && (qi.level < (levels - 1))) {
            node.children[i] = (id_t)o.data[qi.level + 1].size();
            o.data[qi.level + 1].emplace_back();
            o.nNodes++;
          }
          if ((qi.level + 1u) < levels) {
            const Node& nn = o.data[qi.level][qi.id];
            st.push(QI{nn.children[i], (uint8_t)(qi.level + 1), {cc[i][0], cc[i][1], cc[i][2]}});
          }
        }
      }
    }
  }
  cleanEmptyNodes(o);
  finishSVO(o);
}

static int toDAG(Octree& o);

// GeomOctree::buildDAG (geom_octree.cpp:289-435): a top SVO of stepLevel+1
// levels; every set child of its finest nodes is re-voxelized as its own
// octree (box = that cell) and reduced with toDAG; the subtrees are joined in
// (i ascending, j = 7..0) order with pointer offsets; a final toDAG.
static void buildDAGMesh(Octree& o, const float* tris, size_t ntris, unsigned levels, unsigned stepLevel,
                         const double bmin[3], const double bmax[3]) {
  const unsigned stepLevels = stepLevel + 1;
  std::vector<std::array<double, 3>> leaves;
  // top SVO with leaf centres (buildSVO(stepLevels, bbox, true, &leavesCenters))
  {
    for (int i = 0; i < 3; ++i) { o.bboxMin[i] = (float)bmin[i]; o.bboxMax[i] = (float)bmax[i]; }
    o.levels = stepLevels;
    float sides[3];
    for (int i = 0; i < 3; ++i) sides[i] = ((o.bboxMax[i] - o.bboxMin[i]) * 0.5f) * 2.0f;
    o.rootSide = std::max(std::max(sides[0], sides[1]), sides[2]);
    o.data.assign(stepLevels, {});
    o.data[0].push_back(Node());
    o.nNodes = 0;
    double center[3];
    for (int i = 0; i < 3; ++i) center[i] = (bmin[i] + bmax[i]) * 0.5;
    struct QI { id_t id; uint8_t level; double c[3]; };
    std::stack<QI> st;
    for (size_t t = 0; t < ntris; ++t) {
      const float* tri = tris + 9 * t;
      st.push(QI{0, 0, {center[0], center[1], center[2]}});
      while (!st.empty()) {
        QI qi = st.top(); st.pop();
        double k = o.rootSide / double(1u << (qi.level + 2));
        double cc[8][3];
        for (int c = 0; c < 8; ++c) {
          cc[c][0] = qi.c[0] + ((c & 4) ? +k : -k);
          cc[c][1] = qi.c[1] + ((c & 2) ? +k : -k);
          cc[c][2] = qi.c[2] + ((c & 1) ? +k : -k);
        }
        for (int i = 7; i >= 0; --i) {
          if (!testTriBox(cc[i], k, tri)) continue;
          Node& node = o.data[qi.level][qi.id];
          node.mask |= (uint8_t)(1u << i);
          if (!node.existsChildPointer(i) && qi.level < stepLevels - 1) {
            node.children[i] = (id_t)o.data[qi.level + 1].size();
            o.data[qi.level + 1].emplace_back();
            o.nNodes++;
            if (qi.level == stepLevels - 2) leaves.push_back({cc[i][0], cc[i][1], cc[i][2]});
          }
          if (qi.level + 1u < stepLevels) {
            const Node& nn = o.data[qi.level][qi.id];
            st.push(QI{nn.children[i], (uint8_t)(qi.level + 1), {cc[i][0], cc[i][1], cc[i][2]}});
          }
        }
      }
    }
    cleanEmptyNodes(o);
    finishSVO(o);
  }
  const double lhs = o.rootSide / double(1u << stepLevels);   // getHalfSideD(stepLevels - 1)
  size_t nVox = 0;
  std::map<unsigned, Octree*> subs;
  for (size_t i = 0; i < o.data[stepLevels - 1].size(); ++i)
    for (int j = 7; j >= 0; --j) {
      if (!o.data[stepLevels - 1][i].existsChild(j)) continue;
      const double* c = leaves[i].data();
      double lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        const int bit = a == 0 ? 4 : (a == 1 ? 2 : 1);
        const double corner = c[a] + ((j & bit) ? lhs : -lhs);
        lo[a] = std::min(c[a], corner);
        hi[a] = std::max(c[a], corner);
      }
      Octree* s = new Octree();
      buildSVOMesh(*s, tris, ntris, levels - stepLevels, lo, hi);
      nVox += s->nVoxels;
      toDAG(*s);
      subs[(unsigned)(i * 8 + j)] = s;
    }
  o.data.resize(levels);
  o.levels = levels;
  for (size_t i = 0; i < o.data[stepLevels - 1].size(); ++i)
    for (int j = 7; j >= 0; --j) {
      if (!o.data[stepLevels - 1][i].existsChild(j)) continue;
      Octree* s = subs[(unsigned)(i * 8 + j)];
      o.data[stepLevels - 1][i].children[j] = (id_t)o.data[stepLevels].size();
      for (unsigned k = stepLevels; k < levels; ++k) {
        if (k < levels - 1) {
          const id_t off = (id_t)o.data[k + 1].size();
          for (Node& m : s->data[k - stepLevels])
            for (int n = 0; n < 8; ++n)
              if (m.existsChildPointer(n)) m.children[n] += off;
        }
        o.data[k].insert(o.data[k].end(), s->data[k - stepLevels].begin(), s->data[k - stepLevels].end());
      }
      delete s;
    }
  o.nVoxels = nVox;
  o.state = S_SVO;
  toDAG(o);
}

// SVO from an integer voxel set given as 3-bit interleaved Morton codes
// (x most significant within each triple, so code & 7 = child slot 4x+2y+z).
// Restates the point-major DFS of buildSVOFromPoints (geom_octree.cpp:62-169)
// with voxel centres as the points, fed in the caller's order: every level's
// node order is the order of first arrival.  Scene box = unit cube.
static void buildSVOMorton(Octree& o, const uint64_t* codes, size_t n, unsigned levels) {
  o.levels = levels;
  for (int i = 0; i < 3; ++i) { o.bboxMin[i] = 0.f; o.bboxMax[i] = 1.f; }
  o.rootSide = 1.0;
  o.data.assign(levels, {});
  o.data[0].push_back(Node());
  o.nNodes = 0;
  for (size_t v = 0; v < n; ++v) {
    const uint64_t code = codes[v];
    id_t cur = 0;
    for (unsigned lev = 0; lev < levels; ++lev) {
      const int c = (int)((code >> (3 * (levels - 1 - lev))) & 7u);
      Node& node = o.data[lev][cur];
      node.mask |= (uint8_t)(1u << c);
      if (lev + 1 < levels) {
        if (!node.existsChildPointer(c)) {
          node.children[c] = (id_t)o.data[lev + 1].size();
          o.data[lev + 1].emplace_back();
          o.nNodes++;
        }
        cur = o.data[lev][cur].children[c];
      }
    }
  }
  finishSVO(o);
}

// GeomOctree::buildDAG (geom_octree.cpp:289-435) for a voxel list, the
// reference's OpenMP CPU builder: the top SVO of stepLevel + 1 levels (its
// finest nodes in first-arrival order = ascending Morton prefix); every set
// child j of a finest node i is a subtree of levels - stepLevel - 1 levels
// whose voxels are the codes below that cell, built and reduced with toDAG by
// an OpenMP thread (parallel for schedule(dynamic,2) over i, :331-378); the
// subtrees are joined in (i ascending, j = 7..0) order with pointer offsets
// (:397-417) and the whole octree is reduced once more (:425).
static void buildDAGMorton(Octree& o, const uint64_t* codes, size_t n, unsigned levels, unsigned stepLevel) {
  const unsigned stepLevels = stepLevel + 1;
  const unsigned subLevels = levels - stepLevels;
  const unsigned cellShift = 3 * subLevels;                  // code bits below a subtree root cell
  std::vector<uint64_t> top;
  top.reserve(n ? 1 : 0);
  for (size_t v = 0; v < n; ++v) {
    const uint64_t c = codes[v] >> cellShift;
    if (top.empty() || top.back() != c) top.push_back(c);
  }
  buildSVOMorton(o, top.data(), top.size(), stepLevels);
  const auto& fin = o.data[stepLevels - 1];
  std::vector<uint64_t> finPrefix;                           // prefix of finest node i
  for (uint64_t c : top) {
    const uint64_t p = c >> 3;
    if (finPrefix.empty() || finPrefix.back() != p) finPrefix.push_back(p);
  }
  const int nTop = (int)fin.size();
  std::vector<Octree*> subs((size_t)nTop * 8, nullptr);
  const uint64_t lowMask = cellShift >= 64 ? ~0ull : ((1ull << cellShift) - 1);
#pragma omp parallel for schedule(dynamic, 2)
  for (int i = 0; i < nTop; ++i) {
    for (int j = 7; j >= 0; --j) {
      if (!fin[i].existsChild(j)) continue;
      const uint64_t cell = (finPrefix[i] << 3) | (uint64_t)j;
      const uint64_t* lo = std::lower_bound(codes, codes + n, cell << cellShift);
      const uint64_t* hi = std::lower_bound(lo, codes + n, (cell + 1) << cellShift);
      std::vector<uint64_t> local((size_t)(hi - lo));
      for (size_t k = 0; k < local.size(); ++k) local[k] = lo[k] & lowMask;
      Octree* s = new Octree();
      buildSVOMorton(*s, local.data(), local.size(), subLevels);
      toDAG(*s);
      subs[(size_t)i * 8 + j] = s;
    }
  }
  size_t nVox = 0;
  o.data.resize(levels);
  o.levels = levels;
  for (int i = 0; i < nTop; ++i)
    for (int j = 7; j >= 0; --j) {
      Octree* s = subs[(size_t)i * 8 + j];
      if (!s) continue;
      nVox += s->nVoxels;
      o.data[stepLevels - 1][i].children[j] = (id_t)o.data[stepLevels].size();
      for (unsigned k = stepLevels; k < levels; ++k) {
        if (k < levels - 1) {
          const id_t off = (id_t)o.data[k + 1].size();
          for (Node& m : s->data[k - stepLevels])
            for (int c = 0; c < 8; ++c)
              if (m.existsChildPointer(c)) m.children[c] += off;
        }
        o.data[k].insert(o.data[k].end(), s->data[k - stepLevels].begin(), s->data[k - stepLevels].end());
      }
      delete s;
    }
  o.nVoxels = nVox;
  o.state = S_SVO;
  toDAG(o);
}

// toDAG (geom_octree.cpp:462-548): per level L-1..1, dedup by exact key in a
// std::map, first occurrence wins, remap parent pointers.
static int toDAG(Octree& o) {
  if (!(o.state == S_SVO || o.state == S_DAG)) return -3;
  o.nNodes = 1;
  o.dagCounts.assign(o.levels, 0);
  o.dagCounts[0] = o.data[0].size();
  std::vector<id_t> corr;
  std::map<Node, id_t> checker;
  std::vector<Node> uniq;
  for (unsigned lev = o.levels - 1; lev > 0; --lev) {
    uniq.clear(); checker.clear();
    corr.assign(o.data[lev].size(), 0);
    for (id_t i = 0; i < o.data[lev].size(); ++i) {
      const Node& n = o.data[lev][i];
      if (!n.hasChildren()) continue;
      auto k = checker.find(n);
      if (k != checker.end()) corr[i] = k->second;
      else { checker[n] = (id_t)uniq.size(); corr[i] = (id_t)uniq.size(); uniq.push_back(n); }
    }
    o.dagCounts[lev] = uniq.size();
    o.data[lev] = uniq;
    o.nNodes += o.data[lev].size();
    for (Node& bn : o.data[lev - 1])
      for (int j = 0; j < 8; ++j)
        if (bn.existsChild(j)) bn.children[j] = corr[bn.children[j]];
  }
  o.state = S_DAG;
  o.nNodesDAG = o.nNodes;
  return 0;
}

// invertInvs (geom_octree.cpp:822-831)
static void invertInvs(const Octree& o, Node& n, int lev, bool inX, bool inY, bool inZ) {
  for (int i = 0; i < 8; ++i) {
    if (n.existsChildPointer(i)) {
      const Node& nc = o.data[lev + 1][n.children[i]];
      if (inX && nc.getInv(0)) n.unsetMirrorBit(0, i);
      if (inY && nc.getInv(1)) n.unsetMirrorBit(1, i);
      if (inZ && nc.getInv(2)) n.unsetMirrorBit(2, i);
    }
  }
}

// toSDAG (geom_octree.cpp:551-697)
static int toSDAG(Octree& o) {
  if (!(o.state == S_SVO || o.state == S_DAG)) return -3;
  struct MN { bool x, y, z; id_t id; };
  std::vector<MN> corr;
  std::map<Node, id_t> checker;
  std::vector<Node> uniq;
  o.nNodes = 0;
  o.sdagMap.assign(o.levels, {});
  for (unsigned lev = o.levels - 1; lev > 0; --lev) {
    uniq.clear(); checker.clear();
    corr.resize(o.data[lev].size());   // not cleared, like the reference
    for (id_t i = 0; i < o.data[lev].size(); ++i) {
      Node n = o.data[lev][i];
      if (!n.hasChildren()) continue;
      if (n == mirror_node(n, true, false, false)) n.invariant |= 1;
      if (n == mirror_node(n, false, true, false)) n.invariant |= 2;
      if (n == mirror_node(n, false, false, true)) n.invariant |= 4;
      Node niX = mirror_node(n, true, false, false);
      Node niY = mirror_node(n, false, true, false);
      Node niZ = mirror_node(n, false, false, true);
      Node niXY = mirror_node(niX, false, true, false);
      Node niXZ = mirror_node(niX, false, false, true);
      Node niYZ = mirror_node(niY, false, false, true);
      Node niXYZ = mirror_node(niXY, false, false, true);
      if (lev < o.levels - 1) {
        invertInvs(o, niX, lev, true, false, false);
        invertInvs(o, niY, lev, false, true, false);
        invertInvs(o, niZ, lev, false, false, true);
        invertInvs(o, niXY, lev, true, true, false);
        invertInvs(o, niXZ, lev, true, false, true);
        invertInvs(o, niYZ, lev, false, true, true);
        invertInvs(o, niXYZ, lev, true, true, true);
      }
      const Node* vars[8] = {&n, &niX, &niY, &niZ, &niXY, &niXZ, &niYZ, &niXYZ};
      static const bool tx[8] = {0, 1, 0, 0, 1, 1, 0, 1};
      static const bool ty[8] = {0, 0, 1, 0, 1, 0, 1, 1};
      static const bool tz[8] = {0, 0, 0, 1, 0, 1, 1, 1};
      bool found = false;
      for (int t = 0; t < 8 && !found; ++t) {
        auto it = checker.find(*vars[t]);
        if (it != checker.end()) { corr[i] = MN{tx[t], ty[t], tz[t], it->second}; found = true; }
      }
      if (!found) {
        checker[n] = (id_t)uniq.size();
        corr[i] = MN{false, false, false, (id_t)uniq.size()};
        uniq.push_back(n);
      }
    }
    auto& sm = o.sdagMap[lev];
    sm.resize(o.data[lev].size());
    for (size_t i = 0; i < sm.size(); ++i)
      sm[i] = corr[i].id | ((uint32_t)corr[i].x << 29) | ((uint32_t)corr[i].y << 30) | ((uint32_t)corr[i].z << 31);
    o.data[lev] = uniq;
    o.nNodes += o.data[lev].size();
    for (Node& n : o.data[lev - 1])
      for (int j = 0; j < 8; ++j)
        if (n.existsChildPointer(j)) {
          MN mn = corr[n.children[j]];
          n.children[j] = mn.id;
          if (mn.x) n.setMirrorBit(0, j);
          if (mn.y) n.setMirrorBit(1, j);
          if (mn.z) n.setMirrorBit(2, j);
        }
  }
  o.nNodesSDAG = o.nNodes;
  o.state = S_SDAG;
  return 0;
}

// initChildLevels (geom_octree_extension.cpp:21-30)
static void initChildLevels(Octree& o) {
  for (unsigned lev = 0; lev < o.levels; ++lev)
    for (Node& n : o.data[lev])
      for (int c = 0; c < 8; ++c) n.childLevels[c] = lev + 1;
}

// --------------------------------------------------------------------------
// Encoders.  Each returns the exact file bytes written by save().
// --------------------------------------------------------------------------
static void put_bytes(std::vector<uint8_t>& b, const void* p, size_t n) {
  const uint8_t* s = (const uint8_t*)p; b.insert(b.end(), s, s + n);
}
static void put_header(std::vector<uint8_t>& b, const Octree& o) {
  put_bytes(b, o.bboxMin, 12); put_bytes(b, o.bboxMax, 12);
  float rs = (float)o.rootSide; put_bytes(b, &rs, 4);
  uint32_t lv = o.levels; put_bytes(b, &lv, 4);
  uint32_t nn = (uint32_t)o.nNodes; put_bytes(b, &nn, 4);   // size_t truncated to 4 B (encoded_svdag.cpp:92)
}

// EncodedSVDAG::encode + save (encoded_svdag.cpp:76-199)
static int encodeSVDAG(const Octree& o, std::vector<uint8_t>& file) {
  if (o.state != S_DAG) return -3;
  std::vector<uint32_t> truePtrs, data;
  uint32_t counter = 0;
  for (unsigned lev = 0; lev < o.levels; ++lev)
    for (const Node& n : o.data[lev]) {
      truePtrs.push_back(counter);
      counter += (lev < o.levels - 1) ? n.nChildren() + 1 : 1;
    }
  const uint32_t firstLeafPtr = counter;
  std::vector<uint32_t> levAcum(o.levels);
  uint32_t acc = 0;
  for (unsigned lev = 0; lev < o.levels; ++lev) { acc += (uint32_t)o.data[lev].size(); levAcum[lev] = acc; }
  for (unsigned lev = 0; lev < o.levels; ++lev)
    for (const Node& n : o.data[lev]) {
      data.push_back((uint32_t)n.mask);
      if (data.size() < firstLeafPtr) {
        for (int k = 7; k >= 0; --k)
          if (n.children[k] != kNull) {
            size_t off = n.childLevels[k] == 0 ? 0 : levAcum[n.childLevels[k] - 1];
            size_t at = (size_t)n.children[k] + off;
            if (at >= truePtrs.size()) return -5;   // the reference would read out of bounds here
            data.push_back(truePtrs[at]);
          }
      }
    }
  file.clear();
  put_header(file, o);
  put_bytes(file, &firstLeafPtr, 4);
  uint32_t count = (uint32_t)data.size(); put_bytes(file, &count, 4);
  put_bytes(file, data.data(), 4 * data.size());
  return 0;
}

// EncodedUSSVDAG::encode + save (encoded_ussvdag.cpp:60-170)
static int encodeUSSVDAG(const Octree& o, std::vector<uint8_t>& file) {
  if (o.state != S_SDAG) return -3;
  std::vector<uint32_t> truePtrs, data;
  uint32_t counter = 0;
  for (unsigned lev = 0; lev < o.levels; ++lev)
    for (const Node& n : o.data[lev]) {
      truePtrs.push_back(counter);
      counter += (lev < o.levels - 1) ? n.nChildren() + 1 : 1;
    }
  const uint32_t firstLeafPtr = counter;
  uint32_t levSizeAcum = 0;
  for (unsigned lev = 0; lev < o.levels; ++lev) {
    levSizeAcum += (uint32_t)o.data[lev].size();
    for (const Node& n : o.data[lev]) {
      uint32_t hdr = ((uint32_t)n.mirror[2] << 24) | ((uint32_t)n.mirror[1] << 16) | ((uint32_t)n.mirror[0] << 8) | n.mask;
      data.push_back(hdr);
      if (data.size() < firstLeafPtr)
        for (int k = 7; k >= 0; --k)
          if (n.children[k] != kNull) data.push_back(truePtrs[n.children[k] + levSizeAcum]);
    }
  }
  file.clear();
  put_header(file, o);
  put_bytes(file, &firstLeafPtr, 4);
  uint32_t count = (uint32_t)data.size(); put_bytes(file, &count, 4);
  put_bytes(file, data.data(), 4 * data.size());
  return 0;
}

typedef std::pair<id_t, id_t> HistPair;
static void refSort(std::vector<HistPair>& hist) {
  // encoded_ssvdag.cpp:273-275: libstdc++ introsort, comparator on the
  // reference count only (unstable).  The oracle uses the same library call.
  std::sort(hist.begin(), hist.end(), [](HistPair a, HistPair b) { return a.second > b.second; });
}

// EncodedSSVDAG::encode + save (encoded_ssvdag.cpp:84-466), matrix-order leaves.
static int encodeSSVDAG(const Octree& o, std::vector<uint8_t>& file, std::vector<uint16_t>* innerOut = nullptr,
                        std::vector<uint8_t>* leavesOut = nullptr, std::vector<uint32_t>* offOut = nullptr) {
  if (o.state != S_SDAG && o.state != S_DAG) return -3;
  const int L = (int)o.levels;
  if (L < 3) return -2;
  std::vector<std::vector<uint16_t>> encInners(L - 2);
  std::vector<uint8_t> leaves;
  std::vector<HistPair> hist;
  std::vector<id_t> indirection;
  for (int lev = L - 2; lev >= 0; --lev) {
    const auto& lvl = o.data[lev];
    if (lvl.size() > (1u << 30)) return -6;
    hist.resize(lvl.size());
    for (id_t i = 0; i < lvl.size(); ++i) hist[i] = HistPair(i, 0);
    if (lev > 0) {
      for (const Node& n : o.data[lev - 1])
        for (int c = 0; c < 8; ++c)
          if (n.existsChildPointer(c)) hist[n.children[c]].second++;
      refSort(hist);
    }
    std::vector<id_t> newInd(lvl.size());
    if (lev == L - 2) {
      leaves.assign(hist.size() * 8, 0);
      for (id_t i = 0; i < hist.size(); ++i) {
        newInd[hist[i].first] = i;
        const Node& n = lvl[hist[i].first];
        uint8_t buf[8];
        for (int c = 7; c >= 0; --c) {
          if (n.existsChildPointer(c)) {
            const Node& cn = o.data[lev + 1][n.children[c]];
            Node cnm = mirror_node(cn, n.getMirrorBit(0, c), n.getMirrorBit(1, c), n.getMirrorBit(2, c));
            buf[c] = cnm.mask;
          } else buf[c] = 0;
        }
        uint8_t nb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int off = 0;
        for (int z = 0; z < 4; ++z)
          for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
              int byte = (x / 2) + (y / 2) * 2 + (z / 2) * 4;   // xyz_to_byte_voxel (encoded_ssvdag.cpp:119-135)
              int bit = (x % 2) + (y % 2) * 2 + (z % 2) * 4;
              uint8_t v = (buf[byte] >> bit) & 1;
              nb[off / 8] |= (uint8_t)(v << (off % 8));
              ++off;
            }
        for (int c = 0; c < 8; ++c) leaves[i * 8 + c] = nb[c];
      }
    } else {
      for (id_t i = 0; i < hist.size(); ++i) {
        const Node& n = lvl[hist[i].first];
        uint16_t tmp[17];
        tmp[0] = 0;
        int sz = 1;
        for (int c = 7; c >= 0; --c) {
          if (!n.existsChild(c)) continue;
          id_t addr = indirection[n.children[c]];
          if (addr < (1u << 13)) {
            uint16_t p = (uint16_t)addr;
            tmp[0] |= (uint16_t)(1u << (2 * c));
            if (n.getMirrorBit(0, c)) p |= 1u << 13;
            if (n.getMirrorBit(1, c)) p |= 1u << 14;
            if (n.getMirrorBit(2, c)) p |= 1u << 15;
            tmp[sz++] = p;
          } else if (addr < (1u << 30)) {
            uint32_t p = addr;
            if (p & (1u << 29)) { tmp[0] |= (uint16_t)(3u << (2 * c)); p &= ~(1u << 29); }
            else tmp[0] |= (uint16_t)(2u << (2 * c));
            if (n.getMirrorBit(0, c)) p |= 1u << 29;
            if (n.getMirrorBit(1, c)) p |= 1u << 30;
            if (n.getMirrorBit(2, c)) p |= 1u << 31;
            tmp[sz++] = (uint16_t)(p >> 16);
            tmp[sz++] = (uint16_t)(p & 0xFFFF);
          } else return -6;
        }
        newInd[hist[i].first] = (id_t)encInners[lev].size();
        encInners[lev].insert(encInners[lev].end(), tmp, tmp + sz);
      }
    }
    indirection.swap(newInd);
  }
  std::vector<uint32_t> levelOffsets(L - 2);
  levelOffsets[0] = 0;
  for (size_t i = 1; i < levelOffsets.size(); ++i) levelOffsets[i] = (uint32_t)encInners[i - 1].size() + levelOffsets[i - 1];
  std::vector<uint16_t> inner;
  for (auto& v : encInners) inner.insert(inner.end(), v.begin(), v.end());
  file.clear();
  put_header(file, o);
  uint32_t cnt = (uint32_t)inner.size(); put_bytes(file, &cnt, 4); put_bytes(file, inner.data(), 2 * inner.size());
  cnt = (uint32_t)leaves.size(); put_bytes(file, &cnt, 4); put_bytes(file, leaves.data(), leaves.size());
  cnt = (uint32_t)levelOffsets.size(); put_bytes(file, &cnt, 4); put_bytes(file, levelOffsets.data(), 4 * levelOffsets.size());
  if (innerOut) *innerOut = inner;
  if (leavesOut) *leavesOut = leaves;
  if (offOut) *offOut = levelOffsets;
  return 0;
}

// --------------------------------------------------------------------------
// CSVDAG cross-level merge (geom_octree_extension.cpp:34-74, 105-142,
// 238-290, 1192-1544).  std::hash<uint64_t> is the identity in libstdc++.
// --------------------------------------------------------------------------
static uint64_t nodeHash(const Octree& o, const Node& node, unsigned depth) {
  uint64_t key = 0;
  if (depth == 0) return (uint64_t)node.mask;
  if (depth == 1) {
    for (unsigned c = 0; c < 8; ++c) {
      key = key << 8u;
      if (node.existsChildPointer(c)) key += o.data[node.childLevels[c]][node.children[c]].mask;
    }
    return key;
  }
  for (unsigned c = 0; c < 8; ++c) {
    uint64_t orig = key;
    uint64_t ch = orig << (c + 1u);
    if (node.existsChildPointer(c)) ch = nodeHash(o, o.data[node.childLevels[c]][node.children[c]], depth - 1);
    key = (orig << 1u) ^ ch;
    key = key ^ orig;
  }
  return key;
}

static void buildMultiMap(const Octree& o, unsigned depth, std::vector<std::multimap<uint64_t, id_t>>& maps,
                          std::vector<std::vector<uint64_t>>& hashes) {
  for (int lev = (int)o.levels - 1; lev >= 1; --lev) {
    maps[lev].clear();
    hashes[lev].assign(o.data[lev].size(), 0);
    for (id_t i = 0; i < o.data[lev].size(); ++i) {
      uint64_t k = nodeHash(o, o.data[lev][i], depth);
      maps[lev].insert(std::make_pair(k, i));
      hashes[lev][i] = k;
    }
  }
}

static bool compareSubtrees(const Octree& o, unsigned levA, unsigned levB, const Node& nA, const Node& nB,
                            std::vector<std::map<id_t, std::pair<unsigned, id_t>>>& inSub) {
  if (nA.hasChildren() && !nB.hasChildren()) return false;
  else if (levA == o.levels - 1) return nA.mask == nB.mask;
  for (int i = 0; i < 8; ++i) {
    if (nA.existsChild(i) != nB.existsChild(i)) return false;
    else if (!nA.existsChild(i) && !nB.existsChild(i)) continue;
    const Node& cA = o.data[levA + 1][nA.children[i]];
    const Node& cB = o.data[levB + 1][nB.children[i]];
    inSub[levA + 1][nA.children[i]] = std::make_pair(levB + 1, nB.children[i]);
    if (!compareSubtrees(o, levA + 1, levB + 1, cA, cB, inSub)) return false;
  }
  return true;
}

static int mergeAcrossAllLevels(Octree& o) {
  if (o.state != S_DAG) return -3;
  const unsigned L = o.levels;
  std::vector<std::multimap<uint64_t, id_t>> maps(L);
  std::vector<std::vector<uint64_t>> hashes(L);
  unsigned curDepth = L / 2;
  buildMultiMap(o, curDepth, maps, hashes);
  std::set<id_t> cur, next;
  unsigned levA = 1;
  for (id_t i = 0; i < o.data[levA].size(); ++i) cur.insert(i);
  o.nNodes = 1;
  std::vector<std::map<id_t, std::pair<unsigned, id_t>>> mlc(L), inSub(L);
  int missing = 0;
  for (; levA < L; ++levA) {
    unsigned maxDepth = L - levA - 1;
    if (maxDepth < curDepth) { curDepth = maxDepth; buildMultiMap(o, curDepth, maps, hashes); }
    for (id_t idA : cur) {
      const Node& nA = o.data[levA][idA];
      uint64_t key = hashes[levA][idA];
      bool found = false;
      for (unsigned levB = 0; levB < levA && !found; ++levB) {
        auto rng = maps[levB].equal_range(key);
        for (auto it = rng.first; it != rng.second; ++it) {
          id_t j = it->second;
          for (unsigned i = 0; i < L; ++i) inSub[i].clear();
          if (compareSubtrees(o, levA, levB, nA, o.data[levB][j], inSub)) {
            found = true;
            mlc[levA][idA] = std::make_pair(levB, j);
            for (unsigned lr = 0; lr < L; ++lr) mlc[lr].insert(inSub[lr].begin(), inSub[lr].end());
            break;
          }
        }
      }
      if (!found)
        for (int i = 0; i < 8; ++i)
          if (nA.existsChild(i)) next.insert(nA.children[i]);
    }
    cur.clear(); cur.swap(next);
  }
  for (unsigned lev = L - 1; lev > 0; --lev) {
    std::vector<Node> uniq;
    std::vector<id_t> corr(o.data[lev].size(), (id_t)-1);
    for (id_t i = 0; i < o.data[lev].size(); ++i)
      if (mlc[lev].count(i) == 0) { corr[i] = (id_t)uniq.size(); uniq.push_back(o.data[lev][i]); }
    o.data[lev] = uniq;
    o.nNodes += o.data[lev].size();
    for (Node& nd : o.data[lev - 1])
      for (int j = 0; j < 8; ++j)
        if (nd.existsChild(j)) {
          auto it = mlc[lev].find(nd.children[j]);
          if (it != mlc[lev].end()) { nd.childLevels[j] = it->second.first; nd.children[j] = it->second.second; }
          else nd.children[j] = corr[nd.children[j]];
        }
    for (unsigned levLow = L - 2; levLow >= lev; --levLow) {
      for (Node& nd : o.data[levLow])
        for (int j = 0; j < 8; ++j)
          if (nd.existsChild(j) && nd.childLevels[j] == lev) {
            if (corr[nd.children[j]] == (id_t)-1) missing++;
            nd.children[j] = corr[nd.children[j]];
          }
      if (levLow == 0) break;
    }
  }
  o.nNodesDAG = o.nNodes;
  size_t tot = 0;
  for (unsigned i = 0; i < L; ++i) tot += mlc[i].size();
  o.nCrossLevelMerged = tot;
  return missing ? -7 : 0;
}

// --------------------------------------------------------------------------
// GLSL octree_dda traversal, restated on the CPU in FP32 with no FMA
// contraction (shaders/octree_dda.frag.glsl:99-581, 583-603, 803-820).
// min/max follow the GLSL definitions (min(x,y) = y<x ? y : x).
// --------------------------------------------------------------------------
struct TraceParams {           // identical layout to svd_trace_params (include/svdag_amd.h)
  float viewInv[16];           // column-major 4x4 (GL convention)
  float projInv[16];
  uint32_t width, height;
  uint32_t maxIters;           // octree_dda_renderer.cpp:66 default 300
  uint32_t drawLevel;          // renderer init sets it to the level count
  float projectionFactor;      // octree_dda_renderer.cpp:503-507
  float lightPos[3];           // default: bbox max (octree_dda_renderer.cpp:47)
  uint32_t shadows;
  uint32_t pad;
};

struct Scene {
  int fmt;                     // 0 SVDAG, 1 USSVDAG, 2 SSVDAG
  unsigned levels;
  int innerLevels, leafSize;
  float bmin[3], bmax[3], center[3], rootHalfSide;
  const uint32_t* nodes = nullptr; size_t nNodesWords = 0;
  const uint16_t* inner = nullptr; size_t nInner = 0;
  const uint64_t* leaves = nullptr; size_t nLeaves = 0;
  std::vector<uint32_t> levelOffsets;
  std::vector<uint8_t> storage;
};

static inline float gmin(float x, float y) { return y < x ? y : x; }
static inline float gmax(float x, float y) { return x < y ? y : x; }
static inline float gsign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
static inline int isign(float x) { return x > 0.f ? 1 : (x < 0.f ? -1 : 0); }

struct V3 { float x, y, z; };
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline float dot3(V3 a, V3 b) { float s = a.x * b.x; s = s + a.y * b.y; s = s + a.z * b.z; return s; }
static inline float length3(V3 a) { return sqrtf(dot3(a, a)); }
static inline V3 normalize3(V3 a) { float l = length3(a); return v3(a.x / l, a.y / l, a.z / l); }

struct TS {
  float t_current;
  int node_index;
  uint32_t hdr;
  int mm[3];
  uint32_t leaf_lo, leaf_hi;
  int idx[3];
  int local_idx[3];
  uint32_t child_linear_index;
  float t_next[3];
  float inv_d[3];
  int delta_idx[3];
  int current_node_size;
  float cell_size;
  uint32_t level;
};

struct StackEnt { int node; uint32_t hdr; int mask; };

struct Tracer {
  const Scene& s;
  StackEnt stack[40];
  uint32_t stack_size = 0;
  uint32_t path = 2166136261u;
  uint32_t bytes = 0;          // algorithmic bytes fetched by this ray
  explicit Tracer(const Scene& sc) : s(sc) {}

  void push(int node, uint32_t hdr, const int mm[3], uint32_t level) {
    int mask = mm[0] | (mm[1] << 1) | (mm[2] << 2) | (int)(level << 3);
    stack[stack_size++] = StackEnt{node, hdr, mask};
  }
  void pop(int& node, uint32_t& hdr, int mm[3], uint32_t& level) {
    --stack_size;
    StackEnt e = stack[stack_size];
    node = e.node; hdr = e.hdr;
    mm[0] = e.mask & 1; mm[1] = (e.mask >> 1) & 1; mm[2] = (e.mask >> 2) & 1;
    level = (uint32_t)(e.mask >> 3) & 255u;
  }
  static uint32_t lin_idx(const int mm[3], const int idx[3], int sz) {
    int i0 = (1 - 2 * mm[0]) * idx[0] + mm[0] * (sz - 1);
    int i1 = (1 - 2 * mm[1]) * idx[1] + mm[1] * (sz - 1);
    int i2 = (1 - 2 * mm[2]) * idx[2] + mm[2] * (sz - 1);
    return (uint32_t)(i2 + sz * (i1 + sz * i0));
  }
  bool fetch_voxel_bit(const TS& ts) const {
    if (s.fmt != 2) return (ts.hdr & (1u << ts.child_linear_index)) != 0;
    if ((int)ts.level < s.innerLevels) return ((ts.hdr >> (ts.child_linear_index << 1)) & 3u) != 0;
    uint32_t word = ts.child_linear_index & 32u;
    uint32_t bm = 1u << (ts.child_linear_index & 31u);
    return ((word == 0 ? ts.leaf_lo : ts.leaf_hi) & bm) != 0;
  }
  void fetch_data(TS& ts) {
    if (s.fmt != 2) { ts.hdr = s.nodes[(uint32_t)ts.node_index]; bytes += 4; return; }
    if ((int)ts.level < s.innerLevels) {
      ts.hdr = s.inner[(uint32_t)(ts.node_index + (int)s.levelOffsets[ts.level])]; bytes += 2;
    } else {
      uint64_t v = s.leaves[(uint32_t)ts.node_index]; bytes += 8;
      ts.leaf_lo = (uint32_t)v; ts.leaf_hi = (uint32_t)(v >> 32);
    }
  }
  void fetch_child_index_in(TS& ts) {
    if (s.fmt != 2) {
      int pos = __builtin_popcount((ts.hdr & 0xFFu) >> ts.child_linear_index);
      ts.node_index = (int)s.nodes[(uint32_t)(ts.node_index + pos)]; bytes += 4;
      if (s.fmt == 1) {
        int m0 = (ts.hdr & (1u << (ts.child_linear_index + 8))) != 0;
        int m1 = (ts.hdr & (1u << (ts.child_linear_index + 16))) != 0;
        int m2 = (ts.hdr & (1u << (ts.child_linear_index + 24))) != 0;
        ts.mm[0] ^= m0; ts.mm[1] ^= m1; ts.mm[2] ^= m2;
      }
      return;
    }
    int off = 1 + (int)s.levelOffsets[ts.level] + ts.node_index;
    for (int i = 7; i > (int)ts.child_linear_index; --i) { int m = (int)((ts.hdr >> (2 * i)) & 3u); off += m < 2 ? m : 2; }
    const int cm = (int)((ts.hdr >> (ts.child_linear_index << 1)) & 3u);
    int cx = (int)s.inner[(uint32_t)off]; bytes += 2;
    int m0 = (cx >> 13) & 1, m1 = (cx >> 14) & 1, m2 = (cx >> 15) & 1;
    cx &= -57345;
    if (cm > 1) {
      int cy = (int)s.inner[(uint32_t)off + 1]; bytes += 2;
      cx = (cx << 16) | cy;
      cx |= (cm & 1) << 29;
    }
    ts.mm[0] ^= m0; ts.mm[1] ^= m1; ts.mm[2] ^= m2;
    ts.node_index = cx;
  }
  static void set_next_crossing(TS& ts, const V3& o) {
    const float oo[3] = {o.x, o.y, o.z};
    for (int a = 0; a < 3; ++a) {
      int dic = ts.delta_idx[a] > 0 ? ts.delta_idx[a] : 0;
      float pn = (float)(ts.idx[a] + dic) * ts.cell_size;
      ts.t_next[a] = (pn - oo[a]) * ts.inv_d[a];
    }
  }
  void dda_init(const V3& o, const V3& d, TS& ts) {
    const float voxel_eps = 1.0f / (256.f * 1024.f);
    const float tt = ts.t_current + voxel_eps;
    const float pa[3] = {o.x + tt * d.x, o.y + tt * d.y, o.z + tt * d.z};
    for (int a = 0; a < 3; ++a) ts.idx[a] = (int)(pa[a] / ts.cell_size);
    set_next_crossing(ts, o);
    for (int a = 0; a < 3; ++a) ts.local_idx[a] = ts.idx[a] % 2;
  }
  static int next_mask(const TS& ts, int m[3]) {
    const float* t = ts.t_next;
    m[0] = (t[0] < t[1]) & (t[0] <= t[2]);
    m[1] = (t[1] < t[2]) & (t[1] <= t[0]);
    m[2] = (t[2] < t[0]) & (t[2] <= t[1]);
    return 0;
  }
  void dda_next(TS& ts, int delta[3]) {
    int m[3]; next_mask(ts, m);
    for (int a = 0; a < 3; ++a) { delta[a] = m[a] * ts.delta_idx[a]; ts.idx[a] += delta[a]; ts.local_idx[a] += delta[a]; }
    float tc = (float)m[0] * ts.t_next[0];
    tc = tc + (float)m[1] * ts.t_next[1];
    tc = tc + (float)m[2] * ts.t_next[2];
    ts.t_current = tc;
    for (int a = 0; a < 3; ++a) ts.t_next[a] = ts.t_next[a] + ((float)m[a] * ts.cell_size) * fabsf(ts.inv_d[a]);
  }
  static bool in_bounds(const int li[3], int sz) {
    return li[0] < sz && li[1] < sz && li[2] < sz && 0 <= li[0] && 0 <= li[1] && 0 <= li[2];
  }
  void up_in(const V3& o, TS& ts) {
    uint32_t dl = ts.level;
    pop(ts.node_index, ts.hdr, ts.mm, ts.level);
    dl -= ts.level;
    for (int a = 0; a < 3; ++a) ts.idx[a] >>= dl;
    ts.cell_size *= (float)(1 << dl);
    ts.current_node_size = (int)ts.level < s.innerLevels ? 2 : s.leafSize;
    for (int a = 0; a < 3; ++a) ts.local_idx[a] = ts.idx[a] & 1;
    set_next_crossing(ts, o);
  }
  void go_down_one_level(const V3& o, const V3& d, TS& ts) {
    ++ts.level;
    ts.cell_size *= 0.5f;
    const float pa[3] = {o.x + ts.t_current * d.x, o.y + ts.t_current * d.y, o.z + ts.t_current * d.z};
    for (int a = 0; a < 3; ++a) {
      float pc = (float)(ts.idx[a] * 2 + 1) * ts.cell_size;
      int dd = pc < pa[a] ? 1 : 0;
      ts.idx[a] = ts.idx[a] * 2 + dd;
    }
    set_next_crossing(ts, o);
    for (int a = 0; a < 3; ++a) ts.local_idx[a] = ts.idx[a] & (ts.current_node_size - 1);
  }
  void down_in(const V3& o, const V3& d, TS& ts) {
    int m[3]; next_mask(ts, m);
    int nl[3];
    for (int a = 0; a < 3; ++a) nl[a] = ts.local_idx[a] + m[a] * ts.delta_idx[a];
    if (in_bounds(nl, 2)) push(ts.node_index, ts.hdr, ts.mm, ts.level);
    fetch_child_index_in(ts);
    go_down_one_level(o, d, ts);
    if ((int)ts.level == s.innerLevels) {
      ts.current_node_size = s.leafSize;
      int vc = s.leafSize / 2;
      while (vc > 1) { go_down_one_level(o, d, ts); vc >>= 1; }
    }
    path = (path ^ (uint32_t)ts.node_index) * 16777619u;
    path = (path ^ ts.level) * 16777619u;
  }

  // trace_ray (octree_dda.frag.glsl:539-581).  Returns vec4 in out[0..3].
  void trace(V3 o, V3 d, float tmin, float tmax, float projf, uint32_t maxIters, uint32_t drawLevel,
             float out[4], int stepDir[3]) {
    stack_size = 0;
    stepDir[0] = stepDir[1] = stepDir[2] = 0;
    // transform_ray (:480-506)
    const float eps = 1e-4f;
    const float sgn[3] = {gsign(d.x), gsign(d.y), gsign(d.z)};
    const float scale = 1.0f / (2.0f * s.rootHalfSide);
    const float omin[3] = {s.center[0] - s.rootHalfSide, s.center[1] - s.rootHalfSide, s.center[2] - s.rootHalfSide};
    o = v3((o.x - omin[0]) * scale, (o.y - omin[1]) * scale, (o.z - omin[2]) * scale);
    tmin = tmin * scale; tmax = tmax * scale;
    float dd[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; ++a) if (dd[a] * sgn[a] < eps) dd[a] = sgn[a] * eps;
    d = v3(dd[0], dd[1], dd[2]);
    float cmin[3], cmax[3];
    for (int a = 0; a < 3; ++a) { cmin[a] = (s.bmin[a] - omin[a]) * scale; cmax[a] = (s.bmax[a] - omin[a]) * scale; }
    const float oo[3] = {o.x, o.y, o.z};
    float tMin[3], tMax[3];
    for (int a = 0; a < 3; ++a) {
      float t1 = (cmin[a] - oo[a]) / dd[a];
      float t2 = (cmax[a] - oo[a]) / dd[a];
      tMin[a] = gmin(t1, t2); tMax[a] = gmax(t1, t2);
    }
    const float tix = gmax(gmax(tMin[0], 0.0f), gmax(tMin[1], tMin[2]));
    const float tiy = gmin(tMax[0], gmin(tMax[1], tMax[2]));
    tmin = gmax(tix, tmin + 1e-10f);
    tmax = gmin(tiy, tmax);
    if (!(tix < tiy)) { out[0] = -4.f; out[1] = 0; out[2] = 0; out[3] = -1.f; return; }
    const float oscale = 2.0f * s.rootHalfSide;
    TS ts = {};
    ts.t_current = tmin;
    // init (:509-525)
    for (int a = 0; a < 3; ++a) { ts.inv_d[a] = 1.0f / dd[a]; ts.delta_idx[a] = isign(dd[a]); }
    ts.mm[0] = ts.mm[1] = ts.mm[2] = 0;
    ts.level = 0;
    ts.cell_size = 0.5f;
    ts.leaf_lo = ts.leaf_hi = 0;
    dda_init(o, d, ts);
    ts.current_node_size = 2;
    ts.node_index = 0;
    fetch_data(ts);
    ts.child_linear_index = lin_idx(ts.mm, ts.local_idx, ts.current_node_size);
    int it = 0;
    const uint32_t maxLevel = std::min((uint32_t)s.innerLevels, drawLevel - 1);
    do {
      bool full = fetch_voxel_bit(ts);
      if (!full) {
        dda_next(ts, stepDir);
        if (!in_bounds(ts.local_idx, ts.current_node_size)) {
          if (stack_size == 0) { out[0] = -1.f; out[1] = 0; out[2] = (float)it; out[3] = -1.f; return; }
          up_in(o, ts);
        }
      } else {
        bool hit = (ts.level >= maxLevel) || ((ts.cell_size * projf) < ts.t_current);
        if (hit) {
          out[0] = ts.t_current * oscale; out[1] = (float)ts.level; out[2] = (float)it; out[3] = (float)ts.node_index;
          return;
        }
        down_in(o, d, ts);
        fetch_data(ts);
      }
      ts.child_linear_index = lin_idx(ts.mm, ts.local_idx, ts.current_node_size);
      ++it;
    } while ((ts.t_current < tmax) && ((uint32_t)it < maxIters));
    if ((uint32_t)it >= maxIters) { out[0] = -3.f; out[1] = 0; out[2] = (float)it; out[3] = -1.f; return; }
    out[0] = -2.f; out[1] = 0; out[2] = (float)it; out[3] = -1.f;
  }
};

static inline void mat4_mul(const float m[16], const float v[4], float r[4]) {
  for (int i = 0; i < 4; ++i) {
    float s = m[0 * 4 + i] * v[0];
    s = s + m[1 * 4 + i] * v[1];
    s = s + m[2 * 4 + i] * v[2];
    s = s + m[3 * 4 + i] * v[3];
    r[i] = s;
  }
}

// computeCameraRay (octree_dda.frag.glsl:587-603) for pixel (px, py);
// (sw, sh) = the screenRes uniform.
// camera_ray_sc: the same from the shader's screenCoords (sx, sy).
static void camera_ray_sc(const TraceParams& p, float sx, float sy, V3& o, V3& d) {
  const float s0[4] = {sx, sy, 0.f, 1.f}, s1[4] = {sx, sy, 1.f, 1.f};
  float w0[4], w1[4];
  mat4_mul(p.projInv, s0, w0); mat4_mul(p.projInv, s1, w1);
  V3 a = v3(w0[0] / w0[3], w0[1] / w0[3], w0[2] / w0[3]);
  V3 b = v3(w1[0] / w1[3], w1[1] / w1[3], w1[2] / w1[3]);
  V3 rd = normalize3(v3(b.x - a.x, b.y - a.y, b.z - a.z));
  const float z4[4] = {0.f, 0.f, 0.f, 1.f}, d4[4] = {rd.x, rd.y, rd.z, 1.f};
  float op[4], ep[4];
  mat4_mul(p.viewInv, z4, op); mat4_mul(p.viewInv, d4, ep);
  o = v3(op[0], op[1], op[2]);
  d = normalize3(v3(ep[0] - op[0], ep[1] - op[1], ep[2] - op[2]));
}

static float screen_coord(uint32_t px, float sw) { return (((float)px + 0.5f) / sw) * 2.0f - 1.0f; }

static void camera_ray(const TraceParams& p, uint32_t px, uint32_t py, float sw, float sh, V3& o, V3& d) {
  camera_ray_sc(p, screen_coord(px, sw), screen_coord(py, sh), o, d);
}

// --------------------------------------------------------------------------
// Gray-code attribute bit-tree.  The reference names the feature only
// (readme.md:8); the definition restated here is the one documented for
// this repository: attributes in voxel (Morton / DFS) order, code =
// Gray (b ^ b>>1) or binary, per bit plane a pruned binary tree over the
// sequence padded to a power of two with its last value, emitted in
// pre-order: single element -> its bit; uniform node -> 0, bit; mixed -> 1,
// left, right.  Bits are packed LSB-first in u32 words, each plane starting
// on a word boundary.  PARITY UNPINNED against the reference (no code).
// --------------------------------------------------------------------------
static void bittree_plane(const std::vector<uint8_t>& bits, std::vector<uint8_t>& out) {
  const size_t V = bits.size();
  size_t P = 1;
  while (P < V) P <<= 1;
  auto bit = [&](size_t i) -> uint8_t { return bits[i < V ? i : V - 1]; };
  // change-point prefix for O(1) uniformity queries
  std::vector<uint32_t> C(P + 1, 0);
  for (size_t i = 1; i < P; ++i) C[i] = C[i - 1] + (bit(i) != bit(i - 1) ? 1u : 0u);
  struct F { size_t lo, hi; };
  std::vector<F> st{{0, P}};
  while (!st.empty()) {
    F f = st.back(); st.pop_back();
    if (f.hi - f.lo == 1) { out.push_back(bit(f.lo)); continue; }
    const bool uni = C[f.hi - 1] == C[f.lo];
    if (uni) { out.push_back(0); out.push_back(bit(f.lo)); continue; }
    out.push_back(1);
    const size_t mid = (f.lo + f.hi) / 2;
    st.push_back({mid, f.hi});   // right after left (stack)
    st.push_back({f.lo, mid});
  }
}

static std::vector<uint8_t> bittree_encode(const uint8_t* a, size_t V, int gray) {
  std::vector<uint32_t> planeWords[8];
  uint32_t planeBits[8] = {0};
  for (int k = 0; k < 8; ++k) {
    std::vector<uint8_t> bits(V), out;
    for (size_t i = 0; i < V; ++i) {
      const uint8_t c = gray ? (uint8_t)(a[i] ^ (a[i] >> 1)) : a[i];
      bits[i] = (c >> k) & 1u;
    }
    if (V) bittree_plane(bits, out);
    planeBits[k] = (uint32_t)out.size();
    planeWords[k].assign((out.size() + 31) / 32, 0u);
    for (size_t i = 0; i < out.size(); ++i) planeWords[k][i / 32] |= (uint32_t)out[i] << (i % 32);
  }
  std::vector<uint8_t> file;
  const uint32_t magic = 0x42415653u;   // "SVAB"
  const uint32_t hdr[4] = {magic, (uint32_t)V, (uint32_t)gray, 8u};
  put_bytes(file, hdr, 16);
  put_bytes(file, planeBits, 32);
  for (int k = 0; k < 8; ++k) put_bytes(file, planeWords[k].data(), 4 * planeWords[k].size());
  return file;
}

static int bittree_decode(const uint8_t* f, size_t len, std::vector<uint8_t>& a) {
  if (len < 48) return -4;
  uint32_t hdr[4], planeBits[8];
  memcpy(hdr, f, 16);
  memcpy(planeBits, f + 16, 32);
  if (hdr[0] != 0x42415653u || hdr[3] != 8) return -4;
  const size_t V = hdr[1];
  const int gray = (int)hdr[2];
  size_t P = 1;
  while (P < V) P <<= 1;
  std::vector<uint8_t> code(V, 0);
  size_t off = 48;
  for (int k = 0; k < 8; ++k) {
    const size_t nw = (planeBits[k] + 31) / 32;
    if (off + 4 * nw > len) return -4;
    const uint32_t* w = (const uint32_t*)(f + off);
    size_t pos = 0;
    auto next = [&]() -> uint32_t { uint32_t b = (w[pos / 32] >> (pos % 32)) & 1u; ++pos; return b; };
    struct Fr { size_t lo, hi; };
    std::vector<Fr> st;
    if (V) st.push_back({0, P});
    while (!st.empty()) {
      Fr fr = st.back(); st.pop_back();
      if (pos >= planeBits[k]) return -4;
      if (fr.hi - fr.lo == 1) {
        const uint32_t b = next();
        if (fr.lo < V) code[fr.lo] |= (uint8_t)(b << k);
        continue;
      }
      if (next() == 0) {
        const uint32_t b = next();
        for (size_t i = fr.lo; i < fr.hi && i < V; ++i) code[i] |= (uint8_t)(b << k);
        continue;
      }
      const size_t mid = (fr.lo + fr.hi) / 2;
      st.push_back({mid, fr.hi});
      st.push_back({fr.lo, mid});
    }
    off += 4 * nw;
  }
  a.resize(V);
  for (size_t i = 0; i < V; ++i) {
    uint8_t c = code[i];
    if (gray) { uint8_t b = c; for (int s = 1; s < 8; s <<= 1) b ^= (uint8_t)(b >> s); c = b; }
    a[i] = c;
  }
  return 0;
}

}  // namespace orc

// ============================================================================
// C interface (ctypes) -- test infrastructure only.
// ============================================================================
using namespace orc;

struct OrcTraceOut {           // per ray; identical layout to svd_ray_result (include/svdag_amd.h)
  float t; float level; float iters; float node;
  uint32_t path; uint32_t bytes;
};

extern "C" {

void* orc_new() { return new Octree(); }
void orc_free(void* h) { delete (Octree*)h; }

int orc_build_morton(void* h, const uint64_t* codes, uint64_t n, uint32_t levels) {
  if (levels < 1 || levels > 21) return -2;
  buildSVOMorton(*(Octree*)h, codes, (size_t)n, levels);
  return 0;
}

int orc_build_mesh(void* h, const float* tris, uint64_t ntris, uint32_t levels, const double bbox[6]) {
  if (levels < 1 || levels > 21) return -2;
  buildSVOMesh(*(Octree*)h, tris, (size_t)ntris, levels, bbox, bbox + 3);
  return 0;
}

int orc_build_points(void* h, const double* xyz, uint64_t npts, uint32_t levels, const double bbox[6]) {
  if (levels < 1 || levels > 21) return -2;
  buildSVOPoints(*(Octree*)h, xyz, (size_t)npts, levels, bbox, bbox + 3);
  return 0;
}

int orc_build_dag_mesh(void* h, const float* tris, uint64_t ntris, uint32_t levels, uint32_t stepLevel,
                       const double bbox[6]) {
  // stepLevel 0 leaves a one-level top with no finest-node centres (the
  // reference's buildDAG then indexes an empty leavesCenters): rejected here
  if (levels < 2 || levels > 21 || stepLevel == 0 || stepLevel + 1 >= levels) return -2;
  buildDAGMesh(*(Octree*)h, tris, (size_t)ntris, levels, stepLevel, bbox, bbox + 3);
  return 0;
}

// testTriBox (test_triangle_box.cpp:105-184) on one triangle (9 floats) and
// one cube (centre, half side), for the exact-arithmetic check.
int orc_tri_box(const double c[3], double h, const float* tri) { return testTriBox(c, h, tri) ? 1 : 0; }

// toSDAG's correspondence of the DAG nodes of level lev (before the
// reduction): sdag id | mX << 29 | mY << 30 | mZ << 31.  Returns the count.
int64_t orc_sdag_map(void* h, uint32_t lev, uint32_t* out, uint64_t cap) {
  Octree& o = *(Octree*)h;
  if (lev >= o.sdagMap.size()) return -2;
  const auto& m = o.sdagMap[lev];
  if (out && cap >= m.size()) memcpy(out, m.data(), 4 * m.size());
  return (int64_t)m.size();
}

// The OpenMP subtree builder on a sorted voxel list; ends in the DAG state.
int orc_build_dag_morton(void* h, const uint64_t* codes, uint64_t n, uint32_t levels, uint32_t stepLevel) {
  if (levels < 2 || levels > 21 || stepLevel == 0 || stepLevel + 1 >= levels) return -2;
  buildDAGMorton(*(Octree*)h, codes, (size_t)n, levels, stepLevel);
  return 0;
}

// Threads the OpenMP loops use (omp_get_max_threads), 1 without OpenMP.
int orc_omp_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_to_dag(void* h) { return toDAG(*(Octree*)h); }
int orc_to_sdag(void* h) { return toSDAG(*(Octree*)h); }
void orc_init_child_levels(void* h) { initChildLevels(*(Octree*)h); }
int orc_merge_cross_levels(void* h) { return mergeAcrossAllLevels(*(Octree*)h); }

uint32_t orc_levels(void* h) { return ((Octree*)h)->levels; }
int orc_state(void* h) { return (int)((Octree*)h)->state; }
uint64_t orc_level_size(void* h, uint32_t lev) { return ((Octree*)h)->data[lev].size(); }
uint64_t orc_nnodes(void* h) { return ((Octree*)h)->nNodes; }
uint64_t orc_nvoxels(void* h) { return ((Octree*)h)->nVoxels; }
uint64_t orc_nnodes_svo(void* h) { return ((Octree*)h)->nNodesSVO; }
uint64_t orc_cross_merged(void* h) { return ((Octree*)h)->nCrossLevelMerged; }
double orc_root_side(void* h) { return ((Octree*)h)->rootSide; }

// Level dump: words[9*i] = mask | mX<<8 | mY<<16 | mZ<<24, words[9*i+1+c] = children[c];
// inv[i] = invariant bits; childLevels optional (8 per node).
int orc_get_level(void* h, uint32_t lev, uint32_t* words, uint8_t* inv, uint32_t* childLevels) {
  Octree& o = *(Octree*)h;
  if (lev >= o.levels) return -2;
  const auto& v = o.data[lev];
  for (size_t i = 0; i < v.size(); ++i) {
    const Node& n = v[i];
    words[9 * i] = (uint32_t)n.mask | ((uint32_t)n.mirror[0] << 8) | ((uint32_t)n.mirror[1] << 16) | ((uint32_t)n.mirror[2] << 24);
    for (int c = 0; c < 8; ++c) words[9 * i + 1 + c] = n.children[c];
    if (inv) inv[i] = n.invariant;
    if (childLevels) for (int c = 0; c < 8; ++c) childLevels[8 * i + c] = n.childLevels[c];
  }
  return 0;
}

// Replace a level wholesale (lets tests feed hand-made or fuzzed DAGs).
int orc_set_level(void* h, uint32_t lev, const uint32_t* words, uint64_t n) {
  Octree& o = *(Octree*)h;
  if (lev >= o.levels) return -2;
  auto& v = o.data[lev];
  v.assign(n, Node());
  for (size_t i = 0; i < n; ++i) {
    Node& nd = v[i];
    nd.mask = (uint8_t)(words[9 * i] & 0xFF);
    nd.mirror[0] = (uint8_t)(words[9 * i] >> 8);
    nd.mirror[1] = (uint8_t)(words[9 * i] >> 16);
    nd.mirror[2] = (uint8_t)(words[9 * i] >> 24);
    for (int c = 0; c < 8; ++c) nd.children[c] = words[9 * i + 1 + c];
  }
  return 0;
}

int orc_init_levels(void* h, uint32_t levels, int state) {
  Octree& o = *(Octree*)h;
  o.levels = levels; o.data.assign(levels, {}); o.state = (State)state;
  for (int i = 0; i < 3; ++i) { o.bboxMin[i] = 0.f; o.bboxMax[i] = 1.f; }
  o.rootSide = 1.0;
  return 0;
}

// fmt: 0 = .svdag, 1 = .ussvdag, 2 = .ssvdag/.esvdag (by state).  Returns the
// file size; copies when cap suffices.  Negative on error.
int64_t orc_encode(void* h, int fmt, uint8_t* out, uint64_t cap) {
  Octree& o = *(Octree*)h;
  std::vector<uint8_t> file;
  int rc = fmt == 0 ? encodeSVDAG(o, file) : fmt == 1 ? encodeUSSVDAG(o, file) : encodeSSVDAG(o, file);
  if (rc) return rc;
  if (out && cap >= file.size()) memcpy(out, file.data(), file.size());
  return (int64_t)file.size();
}

// libstdc++ std::sort of (idx, val) pairs by val descending -- the exact
// call the reference encoder makes (encoded_ssvdag.cpp:273-275).
void orc_std_sort_refcount(uint32_t* idx, uint32_t* val, uint64_t n) {
  std::vector<HistPair> hist(n);
  for (uint64_t i = 0; i < n; ++i) hist[i] = HistPair(idx[i], val[i]);
  refSort(hist);
  for (uint64_t i = 0; i < n; ++i) { idx[i] = hist[i].first; val[i] = hist[i].second; }
}

// Trace one frame over an encoded file (bytes of .svdag/.ussvdag/.ssvdag).
// primary[w*h] and shadow[w*h] results; shadow entries are filled only for
// primary hits when params->shadows != 0 (others get t = -5).
// The viewer's minimum-depth pre-pass (octree_dda_renderer.cpp:389-408,
// DEPTH_MODE octree_dda.frag.glsl:862-872) and getMinT(8) (:607-617); texels
// outside the texture read as 0 (GL leaves them undefined).
struct MinDepthTex { const float* t = nullptr; uint32_t w = 0, h = 0; };
static float depth_texel(const MinDepthTex& m, int x, int y) {
  return (x >= 0 && y >= 0 && (uint32_t)x < m.w && (uint32_t)y < m.h) ? m.t[(size_t)y * m.w + x] : 0.f;
}
static float min_t(const MinDepthTex& m, uint32_t px, uint32_t py) {
  const int x = (int)((((float)px + 0.5f) - 4.0f) / 8.0f);   // ivec2((gl_FragCoord.xy - 8/2) / 8)
  const int y = (int)((((float)py + 0.5f) - 4.0f) / 8.0f);
  const float tl = depth_texel(m, x, y), tr = depth_texel(m, x + 1, y);
  const float bl = depth_texel(m, x, y + 1), br = depth_texel(m, x + 1, y + 1);
  return gmin(gmin(tl, tr), gmin(bl, br));
}

static int trace_frame_impl(const uint8_t* file, uint64_t len, int fmt, const TraceParams* p,
                            OrcTraceOut* primary, OrcTraceOut* shadow, uint32_t row0, uint32_t nrows,
                            const float* depthProjf, float* minDepthOut);

int orc_trace_frame(const uint8_t* file, uint64_t len, int fmt, const TraceParams* p,
                    OrcTraceOut* primary, OrcTraceOut* shadow, uint32_t row0, uint32_t nrows) {
  return trace_frame_impl(file, len, fmt, p, primary, shadow, row0, nrows, nullptr, nullptr);
}

// Beam-optimised frame: minDepthOut receives the (width/8) x (height/8) texture.
int orc_trace_frame_beam(const uint8_t* file, uint64_t len, int fmt, const TraceParams* p, float depthProjf,
                         OrcTraceOut* primary, OrcTraceOut* shadow, uint32_t row0, uint32_t nrows,
                         float* minDepthOut) {
  return trace_frame_impl(file, len, fmt, p, primary, shadow, row0, nrows, &depthProjf, minDepthOut);
}

// An encoded file image (.svdag / .ussvdag / .ssvdag bytes) as a traversal scene.
struct LoadedScene {
  Scene s;
  std::vector<uint32_t> nodes;
  std::vector<uint16_t> inner;
  std::vector<uint64_t> leaves;
};

static int load_scene(const uint8_t* file, uint64_t len, int fmt, LoadedScene& ls) {
  Scene& s = ls.s;
  s.fmt = fmt;
  size_t off = 0;
  auto rd = [&](void* dst, size_t n) { if (off + n > len) return false; memcpy(dst, file + off, n); off += n; return true; };
  float rs;
  uint32_t lv, nn;
  if (!rd(s.bmin, 12) || !rd(s.bmax, 12) || !rd(&rs, 4) || !rd(&lv, 4) || !rd(&nn, 4)) return -4;
  s.levels = lv;
  if (fmt == 0 || fmt == 1) {
    uint32_t flp, cnt;
    if (!rd(&flp, 4) || !rd(&cnt, 4)) return -4;
    ls.nodes.resize(cnt);
    if (!rd(ls.nodes.data(), 4ull * cnt)) return -4;
    s.innerLevels = (int)lv - 1; s.leafSize = 2;
    s.nodes = ls.nodes.data(); s.nNodesWords = cnt;
  } else {
    uint32_t cnt;
    if (!rd(&cnt, 4)) return -4;
    ls.inner.resize(cnt);
    if (!rd(ls.inner.data(), 2ull * cnt)) return -4;
    if (!rd(&cnt, 4)) return -4;
    ls.leaves.resize(cnt / 8);
    if (!rd(ls.leaves.data(), cnt)) return -4;
    if (!rd(&cnt, 4)) return -4;
    s.levelOffsets.resize(cnt);
    if (!rd(s.levelOffsets.data(), 4ull * cnt)) return -4;
    s.innerLevels = (int)lv - 2; s.leafSize = 4;
    s.inner = ls.inner.data(); s.nInner = ls.inner.size();
    s.leaves = ls.leaves.data(); s.nLeaves = ls.leaves.size();
  }
  s.rootHalfSide = rs / 2.0f;   // EncodedOctree::getHalfSide(0) (encoded_octree.hpp:68-70)
  for (int a = 0; a < 3; ++a) s.center[a] = (s.bmin[a] + s.bmax[a]) * 0.5f;
  return 0;
}

static int trace_frame_impl(const uint8_t* file, uint64_t len, int fmt, const TraceParams* p,
                            OrcTraceOut* primary, OrcTraceOut* shadow, uint32_t row0, uint32_t nrows,
                            const float* depthProjf, float* minDepthOut) {
  LoadedScene ls;
  if (int rc = load_scene(file, len, fmt, ls)) return rc;
  const Scene& s = ls.s;
  Tracer tr(s);
  MinDepthTex md;
  std::vector<float> tex;
  if (depthProjf) {
    md.w = p->width / 8;
    md.h = p->height / 8;
    tex.assign((size_t)md.w * md.h, 0.f);
    const float sw = (float)p->width / 8.0f, sh = (float)p->height / 8.0f;
    for (uint32_t y = 0; y < md.h; ++y)
      for (uint32_t x = 0; x < md.w; ++x) {
        V3 o, d;
        camera_ray(*p, x, y, sw, sh, o, d);
        float res[4]; int step[3];
        tr.trace(o, d, 0.0f, 1e30f, *depthProjf, p->maxIters, p->drawLevel, res, step);
        tex[(size_t)y * md.w + x] = res[0] > 0.f ? res[0] : 0.f;
      }
    md.t = tex.data();
    if (minDepthOut && !tex.empty()) memcpy(minDepthOut, tex.data(), 4 * tex.size());
  }
  const uint32_t rowEnd = std::min(p->height, row0 + nrows);
  for (uint32_t py = row0; py < rowEnd; ++py)
    for (uint32_t px = 0; px < p->width; ++px) {
      const size_t k = (size_t)py * p->width + px;
      V3 o, d;
      camera_ray(*p, px, py, (float)p->width, (float)p->height, o, d);
      float res[4]; int step[3];
      tr.path = 2166136261u; tr.bytes = 0;
      const float tmin = md.t ? min_t(md, px, py) - 1e-3f : 0.0f;   // octree_dda.frag.glsl:681
      tr.trace(o, d, tmin, 1e30f, p->projectionFactor, p->maxIters, p->drawLevel, res, step);
      primary[k] = OrcTraceOut{res[0], res[1], res[2], res[3], tr.path, tr.bytes};
      if (!shadow) continue;
      shadow[k] = OrcTraceOut{-5.f, 0, 0, -1.f, 0, 0};
      if (p->shadows && res[0] >= 0.f) {
        // octree_dda.frag.glsl:686-688, 753-754, 803-815 (normSamples = 0)
        V3 hitPos = v3(o.x + res[0] * d.x, o.y + res[0] * d.y, o.z + res[0] * d.z);
        const float cellSize = (2.0f * s.rootHalfSide) / ldexpf(1.0f, (int)res[1]);
        V3 hn = normalize3(v3(-(float)step[0], -(float)step[1], -(float)step[2]));
        V3 pp = v3(hitPos.x + (hn.x * cellSize) * 0.5f, hitPos.y + (hn.y * cellSize) * 0.5f, hitPos.z + (hn.z * cellSize) * 0.5f);
        V3 l2p = v3(pp.x - p->lightPos[0], pp.y - p->lightPos[1], pp.z - p->lightPos[2]);
        const float l2pLen = length3(l2p);
        const float sProj = l2pLen / cellSize;
        V3 nd = normalize3(l2p);
        V3 sd = v3(-nd.x, -nd.y, -nd.z);
        float sres[4]; int sstep[3];
        tr.path = 2166136261u; tr.bytes = 0;
        tr.trace(pp, sd, 1e-3f * 0.5f, l2pLen * 0.5f, sProj, p->maxIters, p->drawLevel, sres, sstep);
        shadow[k] = OrcTraceOut{sres[0], sres[1], sres[2], sres[3], tr.path, tr.bytes};
      }
    }
  return 0;
}


// ---- viewer shading: VIEWER_MODE main() (octree_dda.frag.glsl:671-858) and
// AO_MODE main() (:884-919) --------------------------------------------------
// GLSL leaves sin, cos, log2 and pow to the implementation.  The definitions
// restated here are fixed sequences of float operations:
//  - sin/cos: k = floor(x * 2/pi + 0.5); r = x - k*P1 - k*P2 - k*P3 (pi/2 in
//    two 12-bit parts and a remainder); Taylor polynomials in r^2 on
//    [-pi/4, pi/4] (sin to r^11, cos to r^12); quadrant k mod 4.
//  - log2: x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh((m-1)/(m+1))
//    by its series to the 13th power; log2 x = e + ln m / ln 2.
//  - mod(x, y) = x - y floor(x / y) (GLSL); pow(x, 2) = x * x.
// They are pinned against float64 sin / cos / log2 (tests/test_oracle_viewer.py).
static void vm_sincos(float x, float& sn, float& cs) {
  const float k = floorf(x * 0.636619747f + 0.5f);
  float r = x - k * 1.5703125f;
  r = r - k * 4.8375129699707031e-4f;
  r = r - k * 7.5497901264043321e-8f;
  const float r2 = r * r;
  float ps = -2.50521084e-8f;
  ps = ps * r2 + 2.75573192e-6f;
  ps = ps * r2 + -1.98412698e-4f;
  ps = ps * r2 + 8.33333333e-3f;
  ps = ps * r2 + -1.66666667e-1f;
  const float sv = r + (r * r2) * ps;
  float pc = 2.08767570e-9f;
  pc = pc * r2 + -2.75573192e-7f;
  pc = pc * r2 + 2.48015873e-5f;
  pc = pc * r2 + -1.38888889e-3f;
  pc = pc * r2 + 4.16666667e-2f;
  const float cv = (1.0f - 0.5f * r2) + (r2 * r2) * pc;
  const int q = (int)k & 3;
  sn = q == 0 ? sv : q == 1 ? cv : q == 2 ? -sv : -cv;
  cs = q == 0 ? cv : q == 1 ? -sv : q == 2 ? -cv : sv;
}

static float vm_log2(float x) {
  if (x != x || x < 0.f) return NAN;
  if (x == 0.f) return -INFINITY;
  if (x == INFINITY) return x;
  int e;
  float m = frexpf(x, &e);
  if (m < 0.707106769f) { m = m * 2.0f; e = e - 1; }
  const float f = (m - 1.0f) / (m + 1.0f);
  const float q = f * f;
  float p = 7.69230769e-2f;
  p = p * q + 9.09090909e-2f;
  p = p * q + 1.11111111e-1f;
  p = p * q + 1.42857143e-1f;
  p = p * q + 2.00000000e-1f;
  p = p * q + 3.33333333e-1f;
  const float f2 = 2.0f * f;
  const float ln = f2 + (f2 * q) * p;
  return (float)e + ln * 1.44269502f;
}

static float vm_mod(float x, float y) { return x - y * floorf(x / y); }

struct ViewParams {            // identical layout to svd_view_params (include/svdag_amd.h)
  int32_t renderMode, normSamples;
  uint32_t selectedVoxel, randomColors, freqColors, pad;
};

// The DEPTH_MODE pre-pass texture (octree_dda_renderer.cpp:389-408).
static MinDepthTex min_depth_texture(Tracer& tr, const TraceParams& p, float projf, std::vector<float>& tex) {
  MinDepthTex md;
  md.w = p.width / 8;
  md.h = p.height / 8;
  tex.assign((size_t)md.w * md.h, 0.f);
  const float sw = (float)p.width / 8.0f, sh = (float)p.height / 8.0f;
  for (uint32_t y = 0; y < md.h; ++y)
    for (uint32_t x = 0; x < md.w; ++x) {
      V3 o, d;
      camera_ray(p, x, y, sw, sh, o, d);
      float res[4]; int step[3];
      tr.trace(o, d, 0.0f, 1e30f, projf, p.maxIters, p.drawLevel, res, step);
      tex[(size_t)y * md.w + x] = res[0] > 0.f ? res[0] : 0.f;
    }
  md.t = tex.data();
  return md;
}

// One fragment of VIEWER_MODE main().
static void render_pixel(Tracer& tr, const Scene& s, const TraceParams& p, const ViewParams& v,
                         const MinDepthTex& md, uint32_t px, uint32_t py, float col[3], OrcTraceOut* prim) {
  const float sw = (float)p.width, sh = (float)p.height;
  const float scx = screen_coord(px, sw), scy = screen_coord(py, sh);          // :672
  V3 o, d;
  camera_ray_sc(p, scx, scy, o, d);                                             // :673
  const float tmin = md.t ? min_t(md, px, py) - 1e-3f : 0.0f;                   // :680-681
  float res[4]; int step[3];
  tr.path = 2166136261u; tr.bytes = 0;
  tr.trace(o, d, tmin, 1e30f, p.projectionFactor, p.maxIters, p.drawLevel, res, step);
  if (prim) *prim = OrcTraceOut{res[0], res[1], res[2], res[3], tr.path, tr.bytes};
  if (res[0] >= 0.f) {                                                          // :694
    const V3 hitPos = v3(o.x + res[0] * d.x, o.y + res[0] * d.y, o.z + res[0] * d.z);   // :686
    const float cellSize = (2.0f * s.rootHalfSide) / ldexpf(1.0f, (int)res[1]);         // :687
    const uint32_t nodeIndex = (uint32_t)res[3];                                          // :688
    V3 hitNorm = v3(-(float)step[0], -(float)step[1], -(float)step[2]);                  // trace_ray :566
    const float diag = length3(v3(s.bmax[0] - s.bmin[0], s.bmax[1] - s.bmin[1], s.bmax[2] - s.bmin[2]));
    float t = 0.f;
    if (v.renderMode == 0) {
      t = 1.0f - (res[2] / (float)p.maxIters);                                  // :697-698
    } else if (v.renderMode == 1) {
      t = res[0] / diag;                                                        // :699-700
    } else if (v.renderMode == 2) {
      t = vm_log2((2.0f * s.rootHalfSide) / res[1]) / (float)s.levels;           // :701-702
    } else if (v.renderMode == 3) {                                             // :703-764
      t = 1.0f;
      const float normalRayLength = 64.0f * cellSize;
      V3 smooth = hitNorm;
      const float normSampleRot = (3.14159f * 2.0f) / (float)v.normSamples;
      const float ni = (float)nodeIndex;
      const float initAngleOffset = vm_mod(((res[0] * scx) * ni) * 0.01f + (ni * 0.001f) * scy, 1.0f);
      for (int i = 0; i < v.normSamples; ++i) {
        const float rad = 0.1f + vm_mod((((float)i / (float)v.normSamples) * ni) * 0.001f, 0.9f);
        const float angle = (float)i * normSampleRot + initAngleOffset;
        float sn, cs;
        vm_sincos(angle, sn, cs);
        const float offx = (rad * cs) / sw, offy = (rad * sn) / sh;
        V3 so, sd;
        camera_ray_sc(p, scx + offx, scy + offy, so, sd);
        float sres[4]; int sstep[3];
        tr.trace(so, sd, res[0] - normalRayLength * 0.5f, res[0] + normalRayLength * 0.5f, p.projectionFactor,
                 p.maxIters, p.drawLevel, sres, sstep);
        if (sres[0] >= 0.f && fabsf(sres[0] - res[0]) <= normalRayLength)
          smooth = v3(smooth.x + -(float)sstep[0], smooth.y + -(float)sstep[1], smooth.z + -(float)sstep[2]);
      }
      hitNorm = normalize3(smooth);
      const V3 lightDir = normalize3(v3(p.lightPos[0] - hitPos.x, p.lightPos[1] - hitPos.y, p.lightPos[2] - hitPos.z));
      t = t * (0.5f + 0.5f * gmax(dot3(hitNorm, lightDir), 0.0f));
      const float dist = length3(v3(hitPos.x - p.lightPos[0], hitPos.y - p.lightPos[1], hitPos.z - p.lightPos[2]));
      t = t - (0.2f * dist) / diag;
    }
    col[0] = col[1] = col[2] = t;                                               // :767
    if (nodeIndex == v.selectedVoxel) {                                         // :771-776
      col[0] = col[0] * 0.5f; col[1] = col[1] * 0.5f; col[2] = 1.0f;
    } else if (s.fmt == 2 && v.freqColors) {                                    // :777-792 (SSVDAG)
      const int hitLev = (int)res[1];
      const int L = (int)s.levels;
      const size_t nOff = s.levelOffsets.size();
      float levSize = 0.f;
      if (hitLev > L - 3) levSize = (float)(uint32_t)s.nLeaves;                 // textureSize(leafNodes)
      else if (hitLev == L - 3) {
        if ((size_t)hitLev < nOff) levSize = (float)((uint32_t)(s.nInner / 4) * 4u - s.levelOffsets[hitLev]);
      } else if ((size_t)hitLev + 1 < nOff) {
        levSize = (float)(s.levelOffsets[hitLev + 1] - s.levelOffsets[hitLev]);
      }
      const float indexInLevel = (float)nodeIndex / levSize;
      const float u = 1.0f - indexInLevel;
      const float a = 1.0f - u * u;
      // mix(vec3(1), vec3(0, 0.2, 0.8), a) = x (1 - a) + y a
      col[0] = col[0] * (1.0f * (1.0f - a) + 0.0f * a);
      col[1] = col[1] * (1.0f * (1.0f - a) + 0.2f * a);
      col[2] = col[2] * (1.0f * (1.0f - a) + 0.8f * a);
    } else if (v.randomColors) {                                                // :795-802
      const V3 rc = normalize3(v3((float)(nodeIndex % 100u) / 100.0f, (float)((3u * nodeIndex) % 200u) / 200.0f,
                                  (float)((2u * nodeIndex) % 300u) / 300.0f));
      col[0] = col[0] * rc.x; col[1] = col[1] * rc.y; col[2] = col[2] * rc.z;
    }
    if (p.shadows) {                                                            // :805-820
      const V3 sp = v3(hitPos.x + (hitNorm.x * cellSize) * 0.5f, hitPos.y + (hitNorm.y * cellSize) * 0.5f,
                       hitPos.z + (hitNorm.z * cellSize) * 0.5f);
      const V3 l2p = v3(sp.x - p.lightPos[0], sp.y - p.lightPos[1], sp.z - p.lightPos[2]);
      const float l2pLen = length3(l2p);
      const float sProj = l2pLen / cellSize;
      const V3 nd = normalize3(l2p);
      float sres[4]; int sstep[3];
      tr.trace(sp, v3(-nd.x, -nd.y, -nd.z), 1e-3f * 0.5f, l2pLen * 0.5f, sProj, p.maxIters, p.drawLevel, sres, sstep);
      const float f = sres[0] > 0.f ? 0.8f : 1.0f;
      col[0] = col[0] * f; col[1] = col[1] * f; col[2] = col[2] * f;
    }
  } else if (res[0] == -1.f || res[0] == -2.f || res[0] == -4.f) {             // :823-835
    col[0] = col[1] = col[2] = 1.0f;
  } else if (res[0] == -3.f) {                                                  // :837-840
    col[0] = 1.0f; col[1] = 0.0f; col[2] = 0.0f;
  } else {                                                                      // discard
    col[0] = col[1] = col[2] = 0.0f;
  }
}

// rows [row0, row0 + nrows) of the viewer's frame; rgb holds 3 floats per
// pixel of those rows (row-major from row0); primary (optional) likewise.
int orc_render_frame(const uint8_t* file, uint64_t len, int fmt, const TraceParams* p, const ViewParams* v,
                     float beamProjf, float* rgb, OrcTraceOut* primary, uint32_t row0, uint32_t nrows) {
  LoadedScene ls;
  if (int rc = load_scene(file, len, fmt, ls)) return rc;
  if (v->normSamples < 0 || v->normSamples > 64 || p->drawLevel == 0) return -2;
  Tracer tr(ls.s);
  std::vector<float> tex;
  MinDepthTex md;
  if (beamProjf > 0.f) md = min_depth_texture(tr, *p, beamProjf, tex);
  const uint32_t rowEnd = std::min(p->height, row0 + nrows);
  for (uint32_t py = row0; py < rowEnd; ++py)
    for (uint32_t px = 0; px < p->width; ++px) {
      const size_t k = (size_t)(py - row0) * p->width + px;
      render_pixel(tr, ls.s, *p, *v, md, px, py, rgb + 3 * k, primary ? primary + k : nullptr);
    }
  return 0;
}

// OctreeDDARenderer::generateHSSamples(256) (octree_dda_renderer.cpp:510-531)
// over HaltonSequence(2) (halton.cpp:24-40, 110-122: bases 2 and 3, index
// from 1); cos / sin evaluated in double and rounded to float.
static std::vector<float> hemisphere_samples() {
  std::vector<float> out;
  for (int i = 1; i <= 256; ++i) {
    float h[2];
    const int base[2] = {2, 3};
    for (int b = 0; b < 2; ++b) {
      float result = 0.f, fraction = 1.0f / (float)base[b];
      int idx = i;
      while (idx > 0) {
        int digit = idx % base[b];
        result += fraction * (float)digit;
        idx = (idx - digit) / base[b];
        fraction /= (float)base[b];
      }
      h[b] = result;
    }
    const float s1 = sqrtf(h[0]);
    const float l = sqrtf(1.0f - s1 * s1);
    const float ang = 2.0f * h[1] * 3.141592654f;
    out.push_back(l * (float)cos((double)ang));
    out.push_back(s1);
    out.push_back(l * (float)sin((double)ang));
  }
  return out;
}

void orc_hemisphere_samples(float* out) {
  const std::vector<float> hs = hemisphere_samples();
  memcpy(out, hs.data(), hs.size() * sizeof(float));
}

// AO_MODE over the G-buffer of the primary rays (hit position, -stepDir,
// cell size 2 rootHalfSide / 2^level; a pixel without a hit has position 0
// and is discarded: -1).  Sample rays whose direction has a NaN component
// count as misses (the bounding-box test of transform_ray rejects them).
int orc_trace_ao(const uint8_t* file, uint64_t len, int fmt, const TraceParams* p, uint32_t numRays, float* vis,
                 OrcTraceOut* primary, uint32_t row0, uint32_t nrows) {
  LoadedScene ls;
  if (int rc = load_scene(file, len, fmt, ls)) return rc;
  if (numRays > 256 || p->drawLevel == 0) return -2;
  const Scene& s = ls.s;
  const std::vector<float> hs = hemisphere_samples();
  Tracer tr(s);
  const uint32_t rowEnd = std::min(p->height, row0 + nrows);
  for (uint32_t py = row0; py < rowEnd; ++py)
    for (uint32_t px = 0; px < p->width; ++px) {
      const size_t k = (size_t)(py - row0) * p->width + px;
      V3 o, d;
      camera_ray(*p, px, py, (float)p->width, (float)p->height, o, d);
      float res[4]; int step[3];
      tr.path = 2166136261u; tr.bytes = 0;
      tr.trace(o, d, 0.0f, 1e30f, p->projectionFactor, p->maxIters, p->drawLevel, res, step);
      if (primary) primary[k] = OrcTraceOut{res[0], res[1], res[2], res[3], tr.path, tr.bytes};
      V3 hitPos = v3(0.f, 0.f, 0.f);
      if (res[0] >= 0.f) hitPos = v3(o.x + res[0] * d.x, o.y + res[0] * d.y, o.z + res[0] * d.z);
      if (hitPos.x == 0.f && hitPos.y == 0.f && hitPos.z == 0.f) { vis[k] = -1.0f; continue; }   // :888
      const V3 n = v3(-(float)step[0], -(float)step[1], -(float)step[2]);
      const float cellSize = (2.0f * s.rootHalfSide) / ldexpf(1.0f, (int)res[1]);
      hitPos = v3(hitPos.x + n.x * 1e-3f, hitPos.y + n.y * 1e-3f, hitPos.z + n.z * 1e-3f);            // :891
      const V3 rvec = v3(0.f, 1.f, 0.f);                                                            // :897
      const float dn = dot3(rvec, n);
      const V3 tangent = normalize3(v3(rvec.x - n.x * dn, rvec.y - n.y * dn, rvec.z - n.z * dn));  // :900
      // cross(n, tangent) (GLSL: x = a.y b.z - b.y a.z, ...)
      const V3 bitangent = v3(n.y * tangent.z - tangent.y * n.z, n.z * tangent.x - tangent.z * n.x,
                              n.x * tangent.y - tangent.x * n.y);
      const float largo = 0.5f;
      const float projFactor = largo / cellSize;
      const int base = (int)(((float)px + 0.5f) * ((float)py + 0.5f));
      float hits = 0.f;
      for (uint32_t i = 0; i < numRays; ++i) {                                                     // :911-916
        const float* sm = &hs[3 * (((int)i + base) % 256)];
        // mat3(tangent, hitNorm, bitangent) * sample
        const V3 dir = normalize3(v3((tangent.x * sm[0] + n.x * sm[1]) + bitangent.x * sm[2],
                                     (tangent.y * sm[0] + n.y * sm[1]) + bitangent.y * sm[2],
                                     (tangent.z * sm[0] + n.z * sm[1]) + bitangent.z * sm[2]));
        if (dir.x != dir.x || dir.y != dir.y || dir.z != dir.z) continue;
        float ares[4]; int astep[3];
        tr.trace(hitPos, dir, 0.0f, largo, projFactor, p->maxIters, p->drawLevel, ares, astep);
        if (ares[0] > 0.f) hits = hits + 1.0f;
      }
      const float visibility = numRays > 0 ? 1.0f - (hits / (float)numRays) : 1.0f;
      vis[k] = visibility * visibility;
    }
  return 0;
}

// Pinning hooks for the math definitions above.
void orc_vm_math(const float* x, uint64_t n, float* sn, float* cs, float* lg) {
  for (uint64_t i = 0; i < n; ++i) {
    vm_sincos(x[i], sn[i], cs[i]);
    lg[i] = vm_log2(x[i]);
  }
}

}  // extern "C"

extern "C" {

// Attribute bit-tree blob (two-call: returns the size; copies when cap suffices).
int64_t orc_attr_bittree_encode(const uint8_t* a, uint64_t n, int gray, uint8_t* out, uint64_t cap) {
  std::vector<uint8_t> f = bittree_encode(a, (size_t)n, gray);
  if (out && cap >= f.size()) memcpy(out, f.data(), f.size());
  return (int64_t)f.size();
}

int orc_attr_bittree_decode(const uint8_t* f, uint64_t len, uint8_t* a, uint64_t cap) {
  std::vector<uint8_t> v;
  int rc = bittree_decode(f, (size_t)len, v);
  if (rc) return rc;
  if (cap < v.size()) return -1;
  memcpy(a, v.data(), v.size());
  return 0;
}

}  // extern "C"

// ============================================================================
// Lossy DAG (LSVDAG) -- the build's deterministic definition of
// GeomOctree::toLossyDag (geom_octree_extension.cpp:797-1179).  The reference
// clusters its similarity graph with the external MCL binary
// (cluster.hpp:134-192), which is absent, so the clustering rule below is this
// project's own; the parts kept from the reference are the effective
// reference count (sortByEffectiveRefCount, :521-560), the bottom-up level
// order (:882), the per-level bound round((L-lev)^2 * factor) (:901), the
// includedNodeRefCount filter (:945) and the final toDAG (:1176).
//
//   * distance d(a,b) = number of voxels in which the two subtrees differ
//     (exact; the reference's diffSubtrees :294-368 approximates it);
//   * signature = 512-bit occupancy of the node's cells three levels down
//     (voxels for level L-3); Hamming(sig) <= d, so it is an exact pre-filter;
//   * node j (effective refcount <= max_refs) lists the first list_cap
//     earlier nodes i < j that pass the pre-filter (exhaustive), the
//     list_cap nearest of them by (Hamming, i) (ranked), or the proposed
//     ones (learned), keeps those with d <= bound, and walks them in order:
//     a representative becomes its target; an entry that merged earlier
//     offers its own target, taken when within the bound of j.
// ============================================================================
namespace orc {

struct LossyParams {             // identical layout to svd_lossy_params' leading fields
  uint32_t lev_lo, lev_hi, max_diff, max_refs, list_cap;
  float diff_factor;
  uint32_t proposer;             // 2: ranked (used when no proposals are given)
};

struct LossyLevelStats {         // per processed level, hi..lo
  uint64_t level, bound, nodes_before, nodes_after, candidates, listed, verified, merged, error_bound;
};

static std::vector<std::vector<uint64_t>> effectiveRefCounts(const Octree& o) {
  std::vector<std::vector<uint64_t>> r(o.levels);
  r[0].assign(o.data[0].size(), 1);
  for (unsigned lev = 1; lev < o.levels; ++lev) {
    r[lev].assign(o.data[lev].size(), 0);
    for (size_t p = 0; p < o.data[lev - 1].size(); ++p) {
      const Node& n = o.data[lev - 1][p];
      for (int c = 0; c < 8; ++c)
        if (n.existsChild(c) && n.existsChildPointer(c)) r[lev][n.children[c]] += r[lev - 1][p];
    }
  }
  return r;
}

static void voxelCounts(const Octree& o, unsigned lev, std::vector<std::vector<uint64_t>>& vox) {
  const auto& v = o.data[lev];
  vox[lev].assign(v.size(), 0);
  for (size_t i = 0; i < v.size(); ++i) {
    if (lev == o.levels - 1) { vox[lev][i] = (uint64_t)__builtin_popcount(v[i].mask); continue; }
    uint64_t s = 0;
    for (int c = 0; c < 8; ++c) if (v[i].existsChild(c)) s += vox[lev + 1][v[i].children[c]];
    vox[lev][i] = s;
  }
}

static void lossySignature(const Octree& o, unsigned lev, id_t n, uint32_t sig[16]) {
  for (int w = 0; w < 16; ++w) sig[w] = 0;
  const unsigned L = o.levels;
  auto set = [&](unsigned b) { sig[b >> 5] |= 1u << (b & 31); };
  const Node& a = o.data[lev][n];
  for (int c1 = 0; c1 < 8; ++c1) {
    if (!a.existsChild(c1)) continue;
    if (lev + 1 >= L) { set(c1); continue; }
    const Node& b = o.data[lev + 1][a.children[c1]];
    for (int c2 = 0; c2 < 8; ++c2) {
      if (!b.existsChild(c2)) continue;
      if (lev + 2 >= L) { set(c1 * 8 + c2); continue; }
      const Node& d = o.data[lev + 2][b.children[c2]];
      for (int c3 = 0; c3 < 8; ++c3) if (d.existsChild(c3)) set(c1 * 64 + c2 * 8 + c3);
    }
  }
}

static unsigned hamming512(const uint32_t* a, const uint32_t* b) {
  unsigned h = 0;
  for (int w = 0; w < 16; ++w) h += (unsigned)__builtin_popcount(a[w] ^ b[w]);
  return h;
}

// The ranked proposer's list of node j: every earlier node within the bound,
// the cap nearest by (Hamming distance, index).
static void rankedList(const uint32_t* sig, size_t j, uint64_t bound, uint32_t cap, std::vector<id_t>& pre) {
  std::vector<std::pair<unsigned, id_t>> all;
  for (size_t i = 0; i < j; ++i) {
    const unsigned h = hamming512(&sig[16 * i], &sig[16 * j]);
    if (h <= bound) all.push_back({h, (id_t)i});
  }
  std::sort(all.begin(), all.end());
  pre.clear();
  for (size_t q = 0; q < all.size() && q < cap; ++q) pre.push_back(all[q].second);
}

// acc + d(a, b), returning as soon as the running total exceeds cap.
static uint64_t subtreeDiff(const Octree& o, const std::vector<std::vector<uint64_t>>& vox, unsigned lev, id_t a,
                            id_t b, uint64_t cap, uint64_t acc) {
  if (a == b) return acc;
  const Node& A = o.data[lev][a];
  const Node& B = o.data[lev][b];
  if (lev == o.levels - 1) return acc + (uint64_t)__builtin_popcount((unsigned)(A.mask ^ B.mask));
  for (int c = 0; c < 8; ++c) {
    const bool ea = A.existsChild(c), eb = B.existsChild(c);
    if (ea && eb) acc = subtreeDiff(o, vox, lev + 1, A.children[c], B.children[c], cap, acc);
    else if (ea) acc += vox[lev + 1][A.children[c]];
    else if (eb) acc += vox[lev + 1][B.children[c]];
    if (acc > cap) return acc;
  }
  return acc;
}

static uint64_t levelBound(const LossyParams& p, unsigned L, unsigned lev) {
  if (p.max_diff) return p.max_diff;
  const double d = (double)(L - lev);
  return (uint64_t)std::llround(d * d * (double)p.diff_factor);
}

// proposals: per processed level (hi..lo order), n_l x k neighbour indices
// (kNone-padded) from the learned proposer, or null for the exhaustive one.
static int lossyMerge(Octree& o, const LossyParams& p, const uint32_t* proposals, uint32_t k,
                      std::vector<LossyLevelStats>& stats) {
  if (o.state != S_DAG) return -3;
  const unsigned L = o.levels;
  if (L < 3 || p.lev_lo < 1 || p.lev_lo > p.lev_hi || p.lev_hi > L - 2 || p.list_cap < 1 || p.list_cap > 64)
    return -1;
  const auto refs = effectiveRefCounts(o);
  std::vector<std::vector<uint64_t>> vox(L);
  for (int lev = (int)L - 1; lev > (int)p.lev_hi; --lev) voxelCounts(o, (unsigned)lev, vox);
  size_t propOff = 0;
  for (int lev = (int)p.lev_hi; lev >= (int)p.lev_lo; --lev) {
    voxelCounts(o, (unsigned)lev, vox);
    const size_t N = o.data[lev].size();
    const uint64_t bound = levelBound(p, L, (unsigned)lev);
    LossyLevelStats st = {};
    st.level = (uint64_t)lev; st.bound = bound; st.nodes_before = N;
    std::vector<uint32_t> sig(16 * N);
    for (size_t i = 0; i < N; ++i) lossySignature(o, (unsigned)lev, (id_t)i, &sig[16 * i]);
    std::vector<char> cand(N);
    for (size_t j = 0; j < N; ++j) { cand[j] = refs[lev][j] <= p.max_refs; st.candidates += cand[j]; }
    // earlier-neighbour lists
    std::vector<std::vector<std::pair<id_t, uint64_t>>> lists(N);
    std::vector<std::vector<id_t>> prop;
    if (proposals) {
      prop.assign(N, {});
      for (size_t x = 0; x < N; ++x)
        for (uint32_t t = 0; t < k; ++t) {
          const uint32_t y = proposals[propOff + x * k + t];
          if (y == 0xFFFFFFFFu || y == x) continue;
          if (y >= N) return -1;
          const size_t lo = std::min<size_t>(x, y), hi = std::max<size_t>(x, y);
          prop[hi].push_back((id_t)lo);
        }
      propOff += N * k;
    }
    for (size_t j = 0; j < N; ++j) {
      if (!cand[j]) continue;
      std::vector<id_t> pre;
      if (proposals) {
        std::sort(prop[j].begin(), prop[j].end());
        prop[j].erase(std::unique(prop[j].begin(), prop[j].end()), prop[j].end());
        for (id_t i : prop[j]) {
          if (hamming512(&sig[16 * i], &sig[16 * j]) <= bound) pre.push_back(i);
          if (pre.size() == p.list_cap) break;
        }
      } else if (p.proposer == 2) {
        rankedList(sig.data(), j, bound, p.list_cap, pre);
      } else {
        for (size_t i = 0; i < j && pre.size() < p.list_cap; ++i)
          if (hamming512(&sig[16 * i], &sig[16 * j]) <= bound) pre.push_back((id_t)i);
      }
      st.listed += pre.size();
      for (id_t i : pre) {
        const uint64_t d = subtreeDiff(o, vox, (unsigned)lev, i, (id_t)j, bound, 0);
        if (d <= bound) lists[j].push_back({i, d});
      }
      st.verified += lists[j].size();
    }
    // sequential resolution in index order: the first verified entry that is
    // a representative, or whose representative (it merged earlier) is within
    // the bound of j, becomes j's target
    std::vector<char> rep(N, 1);
    std::vector<id_t> target(N);
    for (size_t j = 0; j < N; ++j) {
      target[j] = (id_t)j;
      for (const auto& e : lists[j]) {
        id_t t = e.first;
        uint64_t d = e.second;
        if (!rep[t]) {
          t = target[t];
          d = subtreeDiff(o, vox, (unsigned)lev, t, (id_t)j, bound, 0);
          if (d > bound) continue;
        }
        rep[j] = 0; target[j] = t;
        st.merged += 1; st.error_bound += d * refs[lev][j];
        break;
      }
    }
    std::vector<id_t> newId(N, 0);
    std::vector<Node> kept;
    std::vector<uint64_t> keptVox;
    for (size_t j = 0; j < N; ++j)
      if (rep[j]) { newId[j] = (id_t)kept.size(); kept.push_back(o.data[lev][j]); keptVox.push_back(vox[lev][j]); }
    std::vector<id_t> corr(N);
    for (size_t j = 0; j < N; ++j) corr[j] = newId[target[j]];
    o.data[lev] = kept;
    vox[lev] = keptVox;
    for (Node& bn : o.data[lev - 1])
      for (int c = 0; c < 8; ++c)
        if (bn.existsChild(c)) bn.children[c] = corr[bn.children[c]];
    st.nodes_after = kept.size();
    stats.push_back(st);
  }
  return toDAG(o);
}

// ---- LSH kNN over embeddings (the learned proposer's neighbour search) ----
// Hyperplanes: SplitMix64(seed) stream, component = ((z >> 40) * 2^-24) * 2 - 1.
// Key of table t = bits b (b = 0..bits-1) set where dot(e, h[t][b]) > 0, with
// dot = sequential float sum e0*h0 + e1*h1 + ... (no contraction).  Per table
// the nodes are stably sorted by key; node x at sorted position q takes the
// positions q-W..q+W (not q) as candidates.  Distance = sequential float sum
// of squared differences.  Result: the k smallest (distance, index) pairs over
// all tables, distinct indices, ascending; kNone-padded.
static void lshHyperplanes(uint64_t seed, uint32_t tables, uint32_t bits, uint32_t dim, std::vector<float>& h) {
  h.resize((size_t)tables * bits * dim);
  uint64_t s = seed;
  for (size_t i = 0; i < h.size(); ++i) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    h[i] = (float)(z >> 40) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
  }
}

// sig (optional): candidates are scored by the Hamming distance of the
// 512-bit signatures (16 u32 per point) instead of the embedding distance.
static void knnLSH(const float* e, size_t n, uint32_t dim, uint32_t k, uint32_t tables, uint32_t bits, uint32_t W,
                   uint64_t seed, uint32_t* out, const uint32_t* sig = nullptr) {
  std::vector<float> h;
  lshHyperplanes(seed, tables, bits, dim, h);
  std::vector<std::vector<std::pair<float, uint32_t>>> best(n);
  std::vector<uint32_t> key(n), order(n), pos(n);
  auto dist = [&](size_t a, size_t b) {
    if (sig) return (float)hamming512(sig + 16 * a, sig + 16 * b);
    float s = 0.f;
    for (uint32_t d = 0; d < dim; ++d) { const float t = e[a * dim + d] - e[b * dim + d]; const float q = t * t; s = s + q; }
    return s;
  };
  for (uint32_t t = 0; t < tables; ++t) {
    for (size_t x = 0; x < n; ++x) {
      uint32_t kk = 0;
      for (uint32_t b = 0; b < bits; ++b) {
        const float* hp = &h[((size_t)t * bits + b) * dim];
        float s = 0.f;
        for (uint32_t d = 0; d < dim; ++d) { const float q = e[x * dim + d] * hp[d]; s = s + q; }
        if (s > 0.f) kk |= 1u << b;
      }
      key[x] = kk;
      order[x] = (uint32_t)x;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    for (size_t q = 0; q < n; ++q) pos[order[q]] = (uint32_t)q;
    for (size_t x = 0; x < n; ++x) {
      const int64_t q = pos[x];
      for (int64_t r = q - (int64_t)W; r <= q + (int64_t)W; ++r) {
        if (r < 0 || r >= (int64_t)n || r == q) continue;
        const uint32_t y = order[(size_t)r];
        best[x].push_back({dist(x, y), y});
      }
    }
  }
  for (size_t x = 0; x < n; ++x) {
    auto& b = best[x];
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end(), [](const std::pair<float, uint32_t>& u, const std::pair<float, uint32_t>& v) {
              return u.second == v.second; }), b.end());
    for (uint32_t i = 0; i < k; ++i) out[x * k + i] = i < b.size() ? b[i].second : 0xFFFFFFFFu;
  }
}

// One neighbour-of-neighbour round on the signature Hamming metric: x keeps
// the k smallest distinct (distance, index) among its list and its
// neighbours' lists (kNone-padded).
static void knnRefine(const uint32_t* sig, size_t n, uint32_t k, const uint32_t* in, uint32_t* out) {
  for (size_t x = 0; x < n; ++x) {
    std::vector<std::pair<uint32_t, uint32_t>> c;
    auto add = [&](uint32_t z) {
      if (z == 0xFFFFFFFFu || z == x || z >= n) return;
      c.push_back({hamming512(sig + 16 * x, sig + 16 * (size_t)z), z});
    };
    for (uint32_t t = 0; t < k; ++t) {
      const uint32_t y = in[x * k + t];
      add(y);
      if (y == 0xFFFFFFFFu || y >= n) continue;
      for (uint32_t u = 0; u < k; ++u) add(in[(size_t)y * k + u]);
    }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    for (uint32_t t = 0; t < k; ++t) out[x * k + t] = t < c.size() ? c[t].second : 0xFFFFFFFFu;
  }
}

}  // namespace orc

extern "C" {

void orc_knn_refine(const uint32_t* sig, uint64_t n, uint32_t k, const uint32_t* in, uint32_t* out) {
  orc::knnRefine(sig, (size_t)n, k, in, out);
}

int orc_lossy_merge(void* h, const orc::LossyParams* p, const uint32_t* proposals, uint32_t k,
                    orc::LossyLevelStats* stats, uint32_t stats_cap) {
  std::vector<orc::LossyLevelStats> st;
  const int rc = orc::lossyMerge(*(orc::Octree*)h, *p, proposals, k, st);
  for (size_t i = 0; i < st.size() && i < stats_cap; ++i) stats[i] = st[i];
  return rc;
}

// rankedList of every node j < n of a signature array: lists[j * cap + q], counts[j]
int orc_ranked_lists(const uint32_t* sig, uint64_t n, uint32_t bound, uint32_t cap, uint32_t* lists,
                     uint32_t* counts) {
  std::vector<orc::id_t> pre;
  for (uint64_t j = 0; j < n; ++j) {
    orc::rankedList(sig, (size_t)j, bound, cap, pre);
    counts[j] = (uint32_t)pre.size();
    for (size_t q = 0; q < pre.size(); ++q) lists[j * cap + q] = pre[q];
  }
  return 0;
}

int orc_lossy_signatures(void* h, uint32_t lev, uint32_t* sig) {
  orc::Octree& o = *(orc::Octree*)h;
  if (lev >= o.levels) return -2;
  for (size_t i = 0; i < o.data[lev].size(); ++i) orc::lossySignature(o, lev, (orc::id_t)i, sig + 16 * i);
  return 0;
}

int orc_effective_refcounts(void* h, uint32_t lev, uint64_t* out) {
  orc::Octree& o = *(orc::Octree*)h;
  if (lev >= o.levels) return -2;
  const auto r = orc::effectiveRefCounts(o);
  for (size_t i = 0; i < r[lev].size(); ++i) out[i] = r[lev][i];
  return 0;
}

// d(a, b) at level lev, capped: min(d, cap + 1).
uint64_t orc_subtree_diff(void* h, uint32_t lev, uint32_t a, uint32_t b, uint64_t cap) {
  orc::Octree& o = *(orc::Octree*)h;
  std::vector<std::vector<uint64_t>> vox(o.levels);
  for (int l = (int)o.levels - 1; l >= (int)lev; --l) orc::voxelCounts(o, (unsigned)l, vox);
  const uint64_t d = orc::subtreeDiff(o, vox, lev, a, b, cap, 0);
  return d > cap ? cap + 1 : d;
}

void orc_knn_lsh(const float* e, uint64_t n, uint32_t dim, uint32_t k, uint32_t tables, uint32_t bits, uint32_t window,
                 uint64_t seed, uint32_t* out, const uint32_t* sig) {
  orc::knnLSH(e, (size_t)n, dim, k, tables, bits, window, seed, out, sig);
}

}  // extern "C"
