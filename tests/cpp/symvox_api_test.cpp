// symvox_api_test -- drives the C++ drop-in API (host/symvox_amd.hpp) the way
// code written against the reference's headers does, and writes what it sees
// for tests/test_gpu_cli.py to check against the oracle.
//
//   symvox_api_test <codes.bin> <levels> <outdir>
//
// codes.bin: sorted unique uint64 Morton codes.  Writes into outdir:
//   <fmt>.trav   voxels enumerated through the TravNode API of each encoded
//                format (svdag, esvdag, ussvdag, ssvdag), sorted uint64 codes
//   dag.svdag, copy.svdag, nodedata.svdag, decoded.svdag
//                the DAG's .svdag from the octree, its copy, an octree made
//                from its NodeData, and from decode() of the saved file
//   ptr.ussvdag, ptr.ssvdag   the pointer-only SDAG (toSDAG skipSymmetry)
// and prints one "key value" line per scalar it checks.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../svdag-compression_amd/host/symvox_amd.hpp"

namespace {

typedef EncodedOctree::TravNode TravNode;

// Every voxel below node n through the TravNode API.  acc: the mirror flags
// accumulated on the way down (4 x | 2 y | z); geometric slot g is stored
// slot g ^ acc.  SSVDAG bricks store bit z + 4y + 16x of their stored
// coordinates, which hasVoxel(leaf, i, j, k) reads as i + 4j + 16k.
void walk(const EncodedOctree& e, const TravNode& n, uint64_t prefix, unsigned acc, std::vector<uint64_t>& out) {
  if (e.isLeaf(n)) {
    if (e.getLeafSize() == 2) {
      for (unsigned g = 0; g < 8; ++g) {
        const unsigned c = g ^ acc;
        if (e.hasVoxel(n, (c >> 2) & 1, (c >> 1) & 1, c & 1)) out.push_back((prefix << 3) | g);
      }
      return;
    }
    for (unsigned gx = 0; gx < 4; ++gx)
      for (unsigned gy = 0; gy < 4; ++gy)
        for (unsigned gz = 0; gz < 4; ++gz) {
          const unsigned sx = (acc & 4) ? 3 - gx : gx, sy = (acc & 2) ? 3 - gy : gy, sz = (acc & 1) ? 3 - gz : gz;
          if (!e.hasVoxel(n, (int)sz, (int)sy, (int)sx)) continue;
          const uint64_t hi = 4 * (gx >> 1) + 2 * (gy >> 1) + (gz >> 1);
          const uint64_t lo = 4 * (gx & 1) + 2 * (gy & 1) + (gz & 1);
          out.push_back((prefix << 6) | (hi << 3) | lo);
        }
    return;
  }
  for (unsigned g = 0; g < 8; ++g) {
    const unsigned c = g ^ acc;
    if (!e.hasChild(n, (int)c)) continue;
    bool mx = false, my = false, mz = false;
    const TravNode ch = e.getChild(n, (int)c, mx, my, mz);
    walk(e, ch, (prefix << 3) | g, acc ^ ((mx ? 4u : 0u) | (my ? 2u : 0u) | (mz ? 1u : 0u)), out);
  }
}

void write_codes(const std::string& path, std::vector<uint64_t> v) {
  std::sort(v.begin(), v.end());
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)v.data(), (std::streamsize)(v.size() * 8));
}

void traverse(const EncodedOctree& e, const std::string& path, const char* name) {
  std::vector<uint64_t> v;
  walk(e, e.getRootTravNode(), 0, 0, v);
  write_codes(path, v);
  printf("%s_voxels %zu\n", name, v.size());
}

}  // namespace

// symvox_api_test --las <file.las> <levels> <outdir> x0 y0 z0 x1 y1 z1:
// buildSVOFromPoints from the LAS file in the box, toDAG, outdir/points.svdag.
int las_mode(int argc, char** argv) {
  if (argc < 11) {
    fprintf(stderr, "usage: symvox_api_test --las file.las levels outdir x0 y0 z0 x1 y1 z1\n");
    return 2;
  }
  const unsigned L = (unsigned)atoi(argv[3]);
  const std::string out = argv[4];
  double b[6];
  for (int i = 0; i < 6; ++i) b[i] = atof(argv[5 + i]);
  const svd::aabox3d box(svd::point3d(b[0], b[1], b[2]), svd::point3d(b[3], b[4], b[5]));
  Scene scene;
  GeomOctree g(&scene);
  g.buildSVOFromPoints(std::string(argv[2]), L, box);
  printf("points_voxels %zu\n", (size_t)g.getStats().nTotalVoxels);
  g.toDAG(true);
  EncodedSVDAG e;
  e.encode(g);
  e.save(out + "/points.svdag");
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "--las") return las_mode(argc, argv);
  if (argc < 4) {
    fprintf(stderr, "usage: symvox_api_test codes.bin levels outdir\n");
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  in.seekg(0, std::ios::end);
  const size_t n = (size_t)in.tellg() / 8;
  in.seekg(0);
  std::vector<uint64_t> c(n);
  in.read((char*)c.data(), (std::streamsize)(8 * n));
  const unsigned L = (unsigned)atoi(argv[2]);
  const std::string out = argv[3];

  Scene scene;
  GeomOctree g(&scene);
  if (!g.buildFromMorton(c.data(), c.size(), L)) return 1;
  g.toDAG(true);
  g.initChildLevels();
  EncodedSVDAG svdag;
  svdag.encode(g);
  svdag.save(out + "/dag.svdag");
  printf("svdag_nvoxels %zu\n", svdag.getNVoxels());
  const svd::point3f ctr = svdag.getCenter();
  printf("svdag_center %.3f %.3f %.3f\n", ctr[0], ctr[1], ctr[2]);
  traverse(svdag, out + "/svdag.trav", "svdag");
  EncodedSSVDAG esvdag;
  esvdag.encode(g);
  traverse(esvdag, out + "/esvdag.trav", "esvdag");

  // the copy constructor, then the NodeData constructor
  {
    GeomOctree copy(g);
    EncodedSVDAG e;
    e.encode(copy);
    e.save(out + "/copy.svdag");
    GeomOctree assigned(&scene);
    assigned = copy;
    printf("assigned_nodes %zu\n", assigned.getNNodes());
  }
  {
    GeomOctree nd(g.getNodeData(), g.getSceneBBox(), g.getRootSideD(), L, g.getStats());
    printf("nodedata_state %d\n", (int)nd.getState());
    printf("nodedata_nodes %zu\n", nd.getNNodes());
    EncodedSVDAG e;
    e.encode(nd);
    e.save(out + "/nodedata.svdag");
  }
  // decode() by value, from the file just saved
  {
    EncodedSVDAG f;
    if (!f.load(out + "/dag.svdag")) return 1;
    printf("loaded_nvoxels %zu\n", f.getNVoxels());
    GeomOctree d = f.decode(scene);
    printf("decoded_state %d\n", (int)d.getState());
    EncodedSVDAG e;
    e.encode(d);
    e.save(out + "/decoded.svdag");
  }
  // pointer-only SDAG
  {
    GeomOctree p(g);
    p.toSDAG(true, true);
    printf("pointer_sdag_nodes %zu\n", p.getNNodes());
    EncodedUSSVDAG u;
    u.encode(p);
    u.save(out + "/ptr.ussvdag");
    EncodedSSVDAG s;
    s.encode(p);
    s.save(out + "/ptr.ssvdag");
  }
  // the symmetric DAG and its two encodings
  g.toSDAG(true, false);
  EncodedUSSVDAG ussvdag;
  ussvdag.encode(g);
  traverse(ussvdag, out + "/ussvdag.trav", "ussvdag");
  EncodedSSVDAG ssvdag;
  ssvdag.encode(g);
  ssvdag.save(out + "/sym.ssvdag");
  traverse(ssvdag, out + "/ssvdag.trav", "ssvdag");
  // the data pointers are the arrays of the file image
  const uint8_t* img = ssvdag.fileImage().data();
  printf("inner_ptr_ok %d\n", memcmp(ssvdag.getInnerNodesDataPtr(), img + 40, ssvdag.getInnerNodesSize()) == 0);
  printf("leaf_ptr_ok %d\n",
         memcmp(ssvdag.getLeafNodesDataPtr(), img + 44 + ssvdag.getInnerNodesSize(), ssvdag.getLeafNodesSize()) == 0);
  // asking for a child of a leaf prints the reference's warning and returns a default node
  bool mx, my, mz;
  TravNode leaf = ssvdag.getRootTravNode();
  while (!ssvdag.isLeaf(leaf)) {
    int cc = 0;
    while (cc < 8 && !ssvdag.hasChild(leaf, cc)) ++cc;
    leaf = ssvdag.getChild(leaf, cc, mx, my, mz);
  }
  const TravNode none = ssvdag.getChild(leaf, 0, mx, my, mz);
  printf("leaf_child_default %d\n", none.idx == 0 && none.level == 0);
  return 0;
}
