// Mesh voxelization into the reference's SVO: GeomOctree::buildSVO
// (geom_octree.cpp:171-280) + cleanEmptyNodes (:437-456), with the FP64
// triangle/box overlap test of test_triangle_box.cpp:105-184; and the point
// form, buildSVOFromPoints (:62-169), whose DFS is the same with a point
// inside a half-open child box in place of the overlap test.
//
// The reference walks the triangles in order; for each it runs a DFS from the
// root with an explicit stack.  Popping node p tests p's 8 child boxes, slot 7
// first: a hit sets p's mask bit and, if the child does not exist yet, appends
// it to the next level; every hit child is pushed, so the DFS visits children
// in ascending slot order.  A level's node order is therefore the order of
// creation events: by the first triangle that reaches the child, then within
// that triangle's DFS by the parent's pre-order position (its path from the
// root, ascending) and by descending slot.
//
// Device form, one level at a time: the (triangle, node) pairs of a level, in
// triangle order, test their 8 child boxes (one lane per pair); the hits are
// the next level's pairs.  A stable radix sort of the hits by (parent, slot)
// groups the visits of each child; the first of a group is its first triangle.
// A second pair of stable sorts orders the children by (first triangle,
// parent path * 8 + 7 - slot), which is the creation order.  Child centres are
// the parent's centre plus or minus k, the same FP64 operations as the
// reference, so every box test sees the reference's numbers.
#include <rocprim/device/device_radix_sort.hpp>
#include "svd_internal.hpp"

namespace svd {

namespace {

constexpr int kB = 256;

// ---- test_triangle_box.cpp:38-184 (Akenine-Moller), in double ---------------
__device__ __forceinline__ bool plane_box_overlap(const double n[3], const double v[3], double h) {
  double vmin[3], vmax[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (n[q] > 0.0f) { vmin[q] = -h - v[q]; vmax[q] = h - v[q]; }
    else { vmin[q] = h - v[q]; vmax[q] = -h - v[q]; }
  }
  if (n[0] * vmin[0] + n[1] * vmin[1] + n[2] * vmin[2] > 0.0) return false;
  if (n[0] * vmax[0] + n[1] * vmax[1] + n[2] * vmax[2] >= 0.0) return true;
  return false;
}

// one separating-axis family test: p_a = s0*a*p[i0] + s1*b*p[i1] for two vertices
__device__ __forceinline__ bool sat_axis(double a, double b, double fa, double fb, const double* p, const double* q,
                                         int i0, int i1, double s0, double s1, double h) {
  const double pa = s0 * a * p[i0] + s1 * b * p[i1];
  const double pb = s0 * a * q[i0] + s1 * b * q[i1];
  double mn, mx;
  if (pa < pb) { mn = pa; mx = pb; } else { mn = pb; mx = pa; }
  const double rad = (fa + fb) * h;
  return !(mn > rad || mx < -rad);
}

__device__ __forceinline__ bool sat_z(const double* e, double fex, double fey, const double* v1, const double* v2,
                                      double h) {
  const double p1 = e[1] * v1[0] - e[0] * v1[1], p2 = e[1] * v2[0] - e[0] * v2[1];
  double mn, mx;
  if (p2 < p1) { mn = p2; mx = p1; } else { mn = p1; mx = p2; }
  const double rad = (fey + fex) * h;
  return !(mn > rad || mx < -rad);
}

__device__ bool tri_box(const double c[3], double h, const float* tri) {
  double v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v0[i] = (double)tri[i] - c[i];
    v1[i] = (double)tri[3 + i] - c[i];
    v2[i] = (double)tri[6 + i] - c[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) { e0[i] = v1[i] - v0[i]; e1[i] = v2[i] - v1[i]; e2[i] = v0[i] - v2[i]; }
  double fex = fabs(e0[0]), fey = fabs(e0[1]), fez = fabs(e0[2]);
  if (!sat_axis(e0[2], e0[1], fez, fey, v0, v2, 1, 2, 1, -1, h)) return false;   // AXISTEST_X01
  if (!sat_axis(e0[2], e0[0], fez, fex, v0, v2, 0, 2, -1, 1, h)) return false;   // AXISTEST_Y02
  if (!sat_z(e0, fex, fey, v1, v2, h)) return false;                             // AXISTEST_Z12
  fex = fabs(e1[0]); fey = fabs(e1[1]); fez = fabs(e1[2]);
  if (!sat_axis(e1[2], e1[1], fez, fey, v0, v2, 1, 2, 1, -1, h)) return false;   // AXISTEST_X01
  if (!sat_axis(e1[2], e1[0], fez, fex, v0, v2, 0, 2, -1, 1, h)) return false;   // AXISTEST_Y02
  if (!sat_axis(e1[1], e1[0], fey, fex, v0, v1, 0, 1, 1, -1, h)) return false;   // AXISTEST_Z0
  fex = fabs(e2[0]); fey = fabs(e2[1]); fez = fabs(e2[2]);
  if (!sat_axis(e2[2], e2[1], fez, fey, v0, v1, 1, 2, 1, -1, h)) return false;   // AXISTEST_X2
  if (!sat_axis(e2[2], e2[0], fez, fex, v0, v1, 0, 2, -1, 1, h)) return false;   // AXISTEST_Y1
  if (!sat_z(e2, fex, fey, v1, v2, h)) return false;                             // AXISTEST_Z12
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double mn = v0[a], mx = v0[a];
    if (v1[a] < mn) mn = v1[a];
    if (v1[a] > mx) mx = v1[a];
    if (v2[a] < mn) mn = v2[a];
    if (v2[a] > mx) mx = v2[a];
    if (mn > h || mx < -h) return false;
  }
  const double nrm[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
  return plane_box_overlap(nrm, v0, h);
}

// per node: centre, path from its root (3 bits per level) and root index
struct NodeGeo { double* cx; double* cy; double* cz; uint64_t* path; uint32_t* root; };

// 1. every (triangle, node) pair tests the node's 8 child boxes; hit bits go
//    to the node's mask; the number of hits is the pair's visit count
// k = getHalfSideD(level + 1) of the node's root: rootSide / 2^(level + 2)
// The primitives: a triangle overlaps the child box (test_triangle_box.cpp,
// buildSVO), or a point lies in it, half-open on every axis
// (childrenCenters[i] -/+ k, buildSVOFromPoints, geom_octree.cpp:131-133).
struct TriPrim {
  const float* __restrict__ tris;
  __device__ bool hit(uint32_t t, const double cc[3], double k) const { return tri_box(cc, k, tris + 9ull * t); }
};
struct PointPrim {
  const double* __restrict__ xyz;
  __device__ bool hit(uint32_t t, const double cc[3], double k) const {
    const double* q = xyz + 3ull * t;
    return (q[0] >= cc[0] - k && q[0] < cc[0] + k) && (q[1] >= cc[1] - k && q[1] < cc[1] + k) &&
           (q[2] >= cc[2] - k && q[2] < cc[2] + k);
  }
};

template <typename Prim>
__global__ void k_vox_test(Prim prim, const uint32_t* __restrict__ pt, const uint32_t* __restrict__ pn, uint64_t np,
                           NodeGeo g, const double* __restrict__ side, double scale, uint32_t* __restrict__ mask,
                           uint32_t* __restrict__ cnt, uint8_t* __restrict__ hits) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const uint32_t t = pt[j];
  const uint32_t p = pn[j];
  const double k = side[g.root[p]] / scale;
  const double c0 = g.cx[p], c1 = g.cy[p], c2 = g.cz[p];
  uint32_t m = 0;
  for (int i = 7; i >= 0; --i) {
    const double cc[3] = {c0 + ((i & 4) ? +k : -k), c1 + ((i & 2) ? +k : -k), c2 + ((i & 1) ? +k : -k)};
    if (prim.hit(t, cc, k)) m |= 1u << i;
  }
  if (m) atomicOr(&mask[p], m);
  if (cnt) { cnt[j] = (uint32_t)__popc(m); hits[j] = (uint8_t)m; }
}

// 2. the hits of each pair, in pair (triangle) order: key = parent * 8 + slot
__global__ void k_vox_visits(const uint32_t* __restrict__ pt, const uint32_t* __restrict__ pn, uint64_t np,
                             const uint8_t* __restrict__ hits, const uint32_t* __restrict__ off,
                             uint64_t* __restrict__ vkey, uint32_t* __restrict__ vt, uint32_t* __restrict__ vpos) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const uint32_t p = pn[j], m = hits[j];
  uint32_t o = off[j];
  for (int i = 0; i < 8; ++i) {
    if (!((m >> i) & 1u)) continue;
    vkey[o] = (uint64_t)p * 8 + (uint64_t)i;
    vt[o] = pt[j];
    vpos[o] = o;
    ++o;
  }
}

// 3. group heads of the (parent, slot)-sorted visits
__global__ void k_vox_heads(const uint64_t* __restrict__ sk, uint64_t nv, uint32_t* __restrict__ head) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nv) head[j] = (j == 0 || sk[j] != sk[j - 1]) ? 1u : 0u;
}

// 4. per child (group): first triangle and creation key; per visit: its group
//    (gid = exclusive count of heads, so a visit's group is gid + head - 1)
__global__ void k_vox_groups(const uint64_t* __restrict__ sk, const uint32_t* __restrict__ sv,
                             const uint32_t* __restrict__ vt, const uint32_t* __restrict__ head,
                             const uint32_t* __restrict__ gid, uint64_t nv, const uint64_t* __restrict__ ppath,
                             uint32_t* __restrict__ gKey, uint64_t* __restrict__ gPath, uint32_t* __restrict__ gVal,
                             uint64_t* __restrict__ gChild, uint32_t* __restrict__ visitGroup) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const uint32_t gi = gid[j] + head[j] - 1u;
  visitGroup[sv[j]] = gi;
  if (!head[j]) return;
  const uint64_t key = sk[j];
  const uint32_t p = (uint32_t)(key >> 3), slot = (uint32_t)(key & 7u);
  gKey[gi] = vt[sv[j]];                        // first triangle (stable sort of a triangle-ordered list)
  gPath[gi] = ppath[p] * 8 + (7u - slot);      // DFS position of the creation inside that triangle
  gVal[gi] = gi;
  gChild[gi] = key;
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                             uint32_t* __restrict__ dst) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[idx[r]];
}

__global__ void k_inverse(const uint32_t* __restrict__ order, uint64_t n, uint32_t* __restrict__ rank) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) rank[order[r]] = (uint32_t)r;
}

// 5. node records of the parent level and geometry of the new level
__global__ void k_vox_children(const uint64_t* __restrict__ gChild, const uint32_t* __restrict__ rank, uint64_t ng,
                               uint32_t* __restrict__ rec, NodeGeo pg, NodeGeo cg, const double* __restrict__ side,
                               double scale) {
  const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ng) return;
  const uint64_t key = gChild[gi];
  const uint32_t p = (uint32_t)(key >> 3), i = (uint32_t)(key & 7u);
  const uint32_t x = rank[gi];
  const uint32_t r = pg.root[p];
  const double k = side[r] / scale;
  rec[(uint64_t)p * kRecWords + 1 + i] = x;
  cg.cx[x] = pg.cx[p] + ((i & 4) ? +k : -k);
  cg.cy[x] = pg.cy[p] + ((i & 2) ? +k : -k);
  cg.cz[x] = pg.cz[p] + ((i & 1) ? +k : -k);
  cg.path[x] = pg.path[p] * 8 + i;
  cg.root[x] = r;
}

// roots' key for the last (most significant) creation-order sort
__global__ void k_gather_root(const uint64_t* __restrict__ gChild, const uint32_t* __restrict__ order, uint64_t ng,
                              const uint32_t* __restrict__ proot, uint32_t* __restrict__ key) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < ng) key[r] = proot[(uint32_t)(gChild[order[r]] >> 3)];
}

// level-0 pairs of R roots, triangle-major: pair t*R + r = (t, r)
__global__ void k_vox_root_pairs(uint64_t ntris, uint32_t R, uint32_t* __restrict__ pt, uint32_t* __restrict__ pn) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ntris * R) return;
  pt[j] = (uint32_t)(j / R);
  pn[j] = (uint32_t)(j % R);
}

// The top SVO's finest level becomes an inner level whose children are the
// subtree roots, numbered i ascending, slot 7..0 (geom_octree.cpp:397-417).
__global__ void k_vox_join_top(const uint32_t* __restrict__ mask, const uint32_t* __restrict__ base, uint64_t n,
                               uint32_t* __restrict__ rec) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = mask[i] & 0xFFu;
  uint32_t* r = rec + i * kRecWords;
  r[0] = m;
  for (int j = 0; j < 8; ++j)
    r[1 + j] = ((m >> j) & 1u) ? base[i] + (uint32_t)__popc(m >> (j + 1)) : kNull;
}

__global__ void k_popc_mask(const uint32_t* __restrict__ mask, uint64_t n, uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)__popc(mask[i] & 0xFFu);
}

__global__ void k_vox_next_pairs(const uint32_t* __restrict__ vt, const uint32_t* __restrict__ visitGroup,
                                 const uint32_t* __restrict__ rank, uint64_t nv, uint32_t* __restrict__ pt,
                                 uint32_t* __restrict__ pn) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  pt[v] = vt[v];
  pn[v] = rank[visitGroup[v]];
}

__global__ void k_vox_init_rec(uint32_t* __restrict__ rec, const uint32_t* __restrict__ mask, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* r = rec + i * kRecWords;
  r[0] = mask[i];
#pragma unroll
  for (int c = 0; c < 8; ++c) r[1 + c] = kNull;
}

// cleanEmptyNodes: parents drop children without children
__global__ void k_vox_clean(uint32_t* __restrict__ rec, uint64_t n, const uint32_t* __restrict__ child, int childLeaf) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* r = rec + i * kRecWords;
  for (int c = 0; c < 8; ++c) {
    const uint32_t ch = r[1 + c];
    if (ch == kNull) continue;
    const uint32_t m = childLeaf ? child[ch] : child[(uint64_t)ch * kRecWords];
    if ((m & 0xFFu) == 0) { r[0] &= ~(1u << c); r[1 + c] = kNull; }
  }
}

template <typename K>
void stable_sort_pairs(svd_ctx* c, K* keys, uint32_t* vals, uint64_t n, int bits) {
  if (n < 2 || bits <= 0) return;
  K* k2 = dalloc<K>(c, n);
  uint32_t* v2 = dalloc<uint32_t>(c, n);
  size_t bytes = 0;
  SVD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, k2, vals, v2, (size_t)n, 0, bits, c->stream));
  void* tmp = c->arena->get(bytes + 16);
  SVD_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys, k2, vals, v2, (size_t)n, 0, bits, c->stream));
  SVD_HIP(hipMemcpyAsync(keys, k2, sizeof(K) * n, hipMemcpyDeviceToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(vals, v2, 4 * n, hipMemcpyDeviceToDevice, c->stream));
  c->arena->put(tmp);
  dfree(c, k2);
  dfree(c, v2);
}

int bits_for(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b)) ++b;
  return b;
}

NodeGeo alloc_geo(svd_ctx* c, uint64_t n) {
  return NodeGeo{dalloc<double>(c, n + 1), dalloc<double>(c, n + 1), dalloc<double>(c, n + 1),
                 dalloc<uint64_t>(c, n + 1), dalloc<uint32_t>(c, n + 1)};
}
void free_geo(svd_ctx* c, NodeGeo& g) {
  dfree(c, g.cx); dfree(c, g.cy); dfree(c, g.cz); dfree(c, g.path); dfree(c, g.root);
}

// One root box of buildSVO (geom_octree.cpp:176-184): the float box of the
// double one, rootSide = its largest side in float, centre of the double box.
struct VoxRoot { double c[3]; double side; float bb[6]; };
VoxRoot root_of_box(const double bmin[3], const double bmax[3]) {
  VoxRoot r;
  for (int i = 0; i < 3; ++i) { r.bb[i] = (float)bmin[i]; r.bb[3 + i] = (float)bmax[i]; }
  float side = 0.f;
  for (int i = 0; i < 3; ++i) {
    const float s = ((r.bb[3 + i] - r.bb[i]) * 0.5f) * 2.0f;
    side = i == 0 ? s : std::max(side, s);
  }
  r.side = side;
  for (int i = 0; i < 3; ++i) r.c[i] = (bmin[i] + bmax[i]) * 0.5;
  return r;
}

// R independent SVOs of L levels (one per root box) over the same triangles,
// joined level by level in root order (each root's nodes in its creation
// order, child pointers into the joined next level), cleanEmptyNodes applied
// inside every SVO.  Returns the nodes created below the roots; lastCenters
// (optional) receives the centres of the finest level's nodes.
template <typename Prim>
uint64_t voxelize_roots(svd_ctx* c, Prim prim, uint64_t ntris, uint32_t L, const std::vector<VoxRoot>& roots,
                        std::vector<Level>& out, std::vector<double>* lastCenters) {
  const uint64_t R = roots.size();
  if (ntris >= 0xFFFFFFFFull || R == 0 || ntris * R >= 0xFFFFFFF0ull)
    throw Error{SVD_ERR_CAPACITY, "too many triangles (x roots)"};
  out.assign(L, Level());
  uint64_t np = ntris * R, n = R, nNodes = 0;
  uint32_t* pt = dalloc<uint32_t>(c, np + 1);
  uint32_t* pn = dalloc<uint32_t>(c, np + 1);
  if (np) k_vox_root_pairs<<<div_up(np, kB), kB, 0, c->stream>>>(ntris, (uint32_t)R, pt, pn);
  NodeGeo geo = alloc_geo(c, R);
  double* side = dalloc<double>(c, R + 1);
  {
    std::vector<double> cx(R), cy(R), cz(R), sd(R);
    std::vector<uint64_t> path(R, 0);
    std::vector<uint32_t> rid(R);
    for (uint64_t r = 0; r < R; ++r) {
      cx[r] = roots[r].c[0]; cy[r] = roots[r].c[1]; cz[r] = roots[r].c[2]; sd[r] = roots[r].side;
      rid[r] = (uint32_t)r;
    }
    SVD_HIP(hipMemcpyAsync(geo.cx, cx.data(), 8 * R, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(geo.cy, cy.data(), 8 * R, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(geo.cz, cz.data(), 8 * R, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(geo.path, path.data(), 8 * R, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(geo.root, rid.data(), 4 * R, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipMemcpyAsync(side, sd.data(), 8 * R, hipMemcpyHostToDevice, c->stream));
    SVD_HIP(hipStreamSynchronize(c->stream));   // host staging goes out of scope
  }
  uint32_t* total = dalloc<uint32_t>(c, 1);
  for (uint32_t l = 0; l < L; ++l) {
    const double scale = double(1ull << (l + 2));   // k = rootSide / 2^(l+2) = getHalfSideD(level + 1)
    uint32_t* mask = dalloc<uint32_t>(c, n + 1);
    fill_u32(c, mask, 0u, n);
    Level& lv = out[l];
    lv.n = n;
    lv.leaf = (l == L - 1);
    if (l > 0) nNodes += n;
    if (lv.leaf) {   // voxel bits of the leaf-level nodes
      if (np) k_vox_test<<<div_up(np, kB), kB, 0, c->stream>>>(prim, pt, pn, np, geo, side, scale, mask, nullptr,
                                                                 nullptr);
      check_launch("k_vox_test");
      lv.rec = mask;
      break;
    }
    uint32_t* cnt = dalloc<uint32_t>(c, np + 1);
    uint8_t* hits = dalloc<uint8_t>(c, np + 1);
    if (np) k_vox_test<<<div_up(np, kB), kB, 0, c->stream>>>(prim, pt, pn, np, geo, side, scale, mask, cnt, hits);
    check_launch("k_vox_test");
    scan_exclusive(c, cnt, cnt, np, total);
    const uint64_t nv = read_u32(c, total);
    if (nv >= 0xFFFFFFF0ull) throw Error{SVD_ERR_CAPACITY, "too many triangle/node visits in one level"};
    lv.rec = dalloc<uint32_t>(c, n * kRecWords + 1);
    if (n) k_vox_init_rec<<<div_up(n, kB), kB, 0, c->stream>>>(lv.rec, mask, n);
    dfree(c, mask);
    uint64_t* vkey = dalloc<uint64_t>(c, nv + 1);
    uint32_t* vt = dalloc<uint32_t>(c, nv + 1);
    uint32_t* vpos = dalloc<uint32_t>(c, nv + 1);
    if (np) k_vox_visits<<<div_up(np, kB), kB, 0, c->stream>>>(pt, pn, np, hits, cnt, vkey, vt, vpos);
    check_launch("k_vox_visits");
    dfree(c, cnt);
    dfree(c, hits);
    // group the visits by child (parent, slot); the sort is stable, so the
    // first visit of a group comes from the child's first triangle
    stable_sort_pairs<uint64_t>(c, vkey, vpos, nv, 3 + bits_for(n));
    uint32_t* head = dalloc<uint32_t>(c, nv + 1);
    uint32_t* gid = dalloc<uint32_t>(c, nv + 1);
    if (nv) k_vox_heads<<<div_up(nv, kB), kB, 0, c->stream>>>(vkey, nv, head);
    scan_exclusive(c, head, gid, nv, total);
    const uint64_t ng = read_u32(c, total);   // children of this level = nodes of the next
    if (ng >= 0xFFFFFFF0ull) throw Error{SVD_ERR_CAPACITY, "level too large"};
    uint32_t* gKey = dalloc<uint32_t>(c, ng + 1);
    uint64_t* gPath = dalloc<uint64_t>(c, ng + 1);
    uint32_t* gVal = dalloc<uint32_t>(c, ng + 1);
    uint64_t* gChild = dalloc<uint64_t>(c, ng + 1);
    uint32_t* visitGroup = dalloc<uint32_t>(c, nv + 1);
    if (nv)
      k_vox_groups<<<div_up(nv, kB), kB, 0, c->stream>>>(vkey, vpos, vt, head, gid, nv, geo.path, gKey, gPath, gVal,
                                                          gChild, visitGroup);
    check_launch("k_vox_groups");
    // creation order: by root, then first triangle, then DFS position inside it
    stable_sort_pairs<uint64_t>(c, gPath, gVal, ng, 3 * (int)(l + 1));
    uint32_t* gKey2 = dalloc<uint32_t>(c, ng + 1);
    uint32_t* rank = dalloc<uint32_t>(c, ng + 1);
    if (ng) k_gather_u32<<<div_up(ng, kB), kB, 0, c->stream>>>(gKey, gVal, ng, gKey2);
    stable_sort_pairs<uint32_t>(c, gKey2, gVal, ng, bits_for(ntris));
    if (R > 1) {
      if (ng) k_gather_root<<<div_up(ng, kB), kB, 0, c->stream>>>(gChild, gVal, ng, geo.root, gKey2);
      stable_sort_pairs<uint32_t>(c, gKey2, gVal, ng, bits_for(R));
    }
    if (ng) k_inverse<<<div_up(ng, kB), kB, 0, c->stream>>>(gVal, ng, rank);
    NodeGeo next = alloc_geo(c, ng);
    if (ng) k_vox_children<<<div_up(ng, kB), kB, 0, c->stream>>>(gChild, rank, ng, lv.rec, geo, next, side, scale);
    dfree(c, pt);
    dfree(c, pn);
    pt = dalloc<uint32_t>(c, nv + 1);
    pn = dalloc<uint32_t>(c, nv + 1);
    if (nv) k_vox_next_pairs<<<div_up(nv, kB), kB, 0, c->stream>>>(vt, visitGroup, rank, nv, pt, pn);
    check_launch("k_vox_children");
    free_geo(c, geo);
    geo = next;
    dfree(c, vkey); dfree(c, vt); dfree(c, vpos); dfree(c, head); dfree(c, gid); dfree(c, gKey); dfree(c, gKey2);
    dfree(c, gPath); dfree(c, gVal); dfree(c, gChild); dfree(c, visitGroup); dfree(c, rank);
    np = nv;
    n = ng;
  }
  if (lastCenters) {
    const uint64_t nl = out[L - 1].n;
    lastCenters->assign(3 * nl, 0.0);
    std::vector<double> tmp(nl);
    const double* srcs[3] = {geo.cx, geo.cy, geo.cz};
    for (int a = 0; a < 3; ++a) {
      if (nl) SVD_HIP(hipMemcpyAsync(tmp.data(), srcs[a], 8 * nl, hipMemcpyDeviceToHost, c->stream));
      SVD_HIP(hipStreamSynchronize(c->stream));
      for (uint64_t i = 0; i < nl; ++i) (*lastCenters)[3 * i + a] = tmp[i];
    }
  }
  dfree(c, pt);
  dfree(c, pn);
  dfree(c, side);
  dfree(c, total);
  free_geo(c, geo);
  // cleanEmptyNodes inside every SVO, bottom-up (parents never leave their SVO)
  for (uint32_t l = L - 1; l > 0; --l) {
    const Level& child = out[l];
    const Level& par = out[l - 1];
    if (par.n && child.n)
      k_vox_clean<<<div_up(par.n, kB), kB, 0, c->stream>>>(par.rec, par.n, child.rec, child.leaf ? 1 : 0);
  }
  check_launch("k_vox_clean");
  SVD_HIP(hipStreamSynchronize(c->stream));
  return nNodes;
}

}  // namespace

void build_svo_from_mesh(svd_octree* o, const float* dtris, uint64_t ntris, const double bmin[3],
                         const double bmax[3]) {
  const VoxRoot root = root_of_box(bmin, bmax);
  for (int i = 0; i < 6; ++i) o->bbox[i] = root.bb[i];
  o->rootSide = root.side;
  o->nNodesSVO = voxelize_roots(o->ctx, TriPrim{dtris}, ntris, o->levels, {root}, o->svo, nullptr);
}

// GeomOctree::buildSVOFromPoints (geom_octree.cpp:62-169) for points given in
// memory (x y z doubles, the file's order) instead of read from a LAS file:
// the same DFS per point, child boxes half-open, the root box and the child
// centres as buildSVO's, creation order and cleanEmptyNodes as there.
void build_svo_from_points(svd_octree* o, const double* dxyz, uint64_t npts, const double bmin[3],
                           const double bmax[3]) {
  const VoxRoot root = root_of_box(bmin, bmax);
  for (int i = 0; i < 6; ++i) o->bbox[i] = root.bb[i];
  o->rootSide = root.side;
  o->nNodesSVO = voxelize_roots(o->ctx, PointPrim{dxyz}, npts, o->levels, {root}, o->svo, nullptr);
}

// GeomOctree::buildDAG's SVO (geom_octree.cpp:289-435): a top SVO of
// stepLevel + 1 levels; one subtree per set voxel of its finest nodes,
// voxelized from all triangles inside the voxel's own box (:331-378; the box
// spans the parent centre and the voxel's far corner); the subtrees joined
// under the top in order (finest node ascending, slot 7..0, :397-417).  The
// node count is the reference's: the top's created nodes plus each subtree's,
// without the subtree roots.
void build_svo_from_mesh_subtrees(svd_octree* o, const float* dtris, uint64_t ntris, const double bmin[3],
                                  const double bmax[3], uint32_t stepLevel) {
  svd_ctx* c = o->ctx;
  const uint32_t L = o->levels, stepLevels = stepLevel + 1;
  if (stepLevels >= L) throw Error{SVD_ERR_ARG, "buildDAG: step level must be below levels - 1"};
  const VoxRoot top = root_of_box(bmin, bmax);
  for (int i = 0; i < 6; ++i) o->bbox[i] = top.bb[i];
  o->rootSide = top.side;
  std::vector<Level> T;
  std::vector<double> centers;
  const uint64_t topNodes = voxelize_roots(c, TriPrim{dtris}, ntris, stepLevels, {top}, T, &centers);
  Level& fin = T[stepLevels - 1];
  std::vector<uint32_t> masks(fin.n);
  if (fin.n) SVD_HIP(hipMemcpyAsync(masks.data(), fin.rec, 4 * fin.n, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
  const double lhs = top.side / double(1u << stepLevels);   // getHalfSideD(stepLevels - 1)
  std::vector<VoxRoot> roots;
  for (uint64_t i = 0; i < fin.n; ++i)
    for (int j = 7; j >= 0; --j) {
      if (!((masks[i] >> j) & 1u)) continue;
      const double* cc = &centers[3 * i];
      double lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        const double corner = cc[a] + ((j & (4 >> a)) ? lhs : -lhs);
        lo[a] = std::min(cc[a], corner);
        hi[a] = std::max(cc[a], corner);
      }
      roots.push_back(root_of_box(lo, hi));
    }
  std::vector<Level> sub;
  uint64_t subNodes = 0;
  if (!roots.empty()) subNodes = voxelize_roots(c, TriPrim{dtris}, ntris, L - stepLevels, roots, sub, nullptr);
  // join: the top's finest level points at the subtree roots
  uint32_t* joined = dalloc<uint32_t>(c, fin.n * kRecWords + 1);
  uint32_t* base = dalloc<uint32_t>(c, fin.n + 1);
  if (fin.n) {
    k_popc_mask<<<div_up(fin.n, kB), kB, 0, c->stream>>>(fin.rec, fin.n, base);
    scan_exclusive(c, base, base, fin.n, nullptr);
    k_vox_join_top<<<div_up(fin.n, kB), kB, 0, c->stream>>>(fin.rec, base, fin.n, joined);
  }
  check_launch("k_vox_join_top");
  SVD_HIP(hipStreamSynchronize(c->stream));
  dfree(c, base);
  dfree(c, fin.rec);
  fin.rec = joined;
  fin.leaf = false;
  o->svo.assign(L, Level());
  for (uint32_t l = 0; l < stepLevels; ++l) o->svo[l] = T[l];
  for (uint32_t l = stepLevels; l < L; ++l) {
    if (roots.empty()) { o->svo[l].n = 0; o->svo[l].leaf = (l == L - 1); o->svo[l].rec = dalloc<uint32_t>(c, 1); }
    else o->svo[l] = sub[l - stepLevels];
  }
  o->nNodesSVO = topNodes + subNodes;
}

}  // namespace svd
