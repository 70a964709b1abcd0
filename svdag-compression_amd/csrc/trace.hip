// Per-ray octree DDA traversal (shaders/octree_dda.frag.glsl:99-581) as a
// compute kernel: one ray per lane, 16x16 pixel tiles per workgroup, the
// descent stack (INNER_LEVELS+1 entries) in LDS laid out [depth][lane].
// FP32 throughout; this file is compiled with -ffp-contract=off so every
// product / sum rounds exactly like the shader's individual GLSL operations.
// min/max follow the GLSL definitions: min(x,y) = y<x ? y : x.
#include <algorithm>
#include <mutex>
#include "svd_internal.hpp"

namespace svd {

namespace {

constexpr int kBT = 256;                    // threads per workgroup (any tile shape of 256)
constexpr int kAOSamples = 256;             // NUM_MAX_HS_SAMPLES (octree_dda_renderer.cpp:32)

struct Dev {
  const uint32_t* nodes;
  const uint16_t* inner;
  const uint2* leaves;
  const uint32_t* levelOffsets;
  uint32_t nWords, nInner, nLeaves;   // element counts, for bounds checks
  int innerLevels, leafSize;
  float bmin[3], bmax[3], center[3], rootHalfSide;
};

__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float gsign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ int isign(float x) { return x > 0.f ? 1 : (x < 0.f ? -1 : 0); }

struct TS {
  float t_current;
  int node_index;
  uint32_t hdr;       // node header; an SSVDAG leaf's low mask word
  uint32_t mm;        // mirror flags, bit a for axis a
  uint32_t leaf_hi;   // an SSVDAG leaf's high mask word
  int idx[3];
  int li[3];
  uint32_t cli;
  float tn[3];
  float inv[3];
  int di[3];
  int cns;
  float cell;
  uint32_t level;
};

// Descent stack in LDS, [depth][lane]: the node, and (mirror bits, level,
// header) packed in one word when the header fits 16 bits (SSVDAG inner
// headers, SVDAG masks); USSVDAG headers carry mirror masks and keep a word of
// their own.
template <int FMT>
constexpr int kStackWords = FMT == SVD_FMT_USSVDAG ? 3 : 2;

struct Stack {
  uint32_t* node;   // [depth * kBT + lane]
  uint32_t* aux;    // mirror bits | level << 3 | header << 16 (header word apart for USSVDAG)
  uint32_t* hdr;    // USSVDAG only
  int lane;
  uint32_t size;
};

template <int FMT>
__device__ __forceinline__ Stack make_stack(uint32_t* mem, int depth, int lane) {
  return Stack{mem, mem + depth * kBT, kStackWords<FMT> == 3 ? mem + 2 * depth * kBT : nullptr, lane, 0};
}

// INSTR: keep the per-ray descent-path hash and the fetched-byte count (the
// test and byte-count variant); without it the two stay at their initial
// values and cost no instructions.
template <int FMT, bool INSTR = true>
struct Tracer {
  const Dev& s;
  Stack st;
  uint32_t path;
  uint32_t bytes;
  bool bad;   // a fetch fell outside the scene arrays (malformed file)

  __device__ Tracer(const Dev& d, Stack k) : s(d), st(k), path(2166136261u), bytes(0), bad(false) {}

  __device__ __forceinline__ uint32_t word(uint32_t i) {
    if (i < s.nWords) return s.nodes[i];
    bad = true;
    return 0u;
  }
  __device__ __forceinline__ uint32_t half(uint32_t i) {
    if (i < s.nInner) return s.inner[i];
    bad = true;
    return 0u;
  }

  // SSVDAG inner headers are 16 bits; of an SVDAG header word only the mask
  // (bits 0-7) is ever read, so both fit beside the mirror bits and level.
  __device__ __forceinline__ void push(int node, uint32_t hdr, uint32_t mm, uint32_t level) {
    const uint32_t mask = mm | (level << 3);
    const uint32_t o = st.size * kBT + st.lane;
    st.node[o] = (uint32_t)node;
    if (kStackWords<FMT> == 3) { st.aux[o] = mask; st.hdr[o] = hdr; }
    else st.aux[o] = mask | ((FMT == SVD_FMT_SSVDAG ? hdr : (hdr & 0xFFu)) << 16);
    ++st.size;
  }
  __device__ __forceinline__ void pop(TS& ts) {
    --st.size;
    const uint32_t o = st.size * kBT + st.lane;
    ts.node_index = (int)st.node[o];
    const uint32_t a = st.aux[o];
    ts.hdr = kStackWords<FMT> == 3 ? st.hdr[o] : (a >> 16);
    ts.mm = a & 7u;
    ts.level = (a >> 3) & 255u;
  }
  // voxel_to_linear_idx: (1 - 2m) li + m (sz - 1) per axis, then i2 + sz (i1 + sz i0).
  // sz is 2 or 4 and li is always in [0, sz) here, so the mirrored index is
  // li ^ (sz - 1) and the products are shifts.
  __device__ __forceinline__ static uint32_t lin_idx(const TS& ts, int sz) {
    const int hm = sz - 1, sh = sz == 4 ? 2 : 1;
    const int i0 = ts.li[0] ^ ((ts.mm & 1u) ? hm : 0);
    const int i1 = ts.li[1] ^ ((ts.mm & 2u) ? hm : 0);
    const int i2 = ts.li[2] ^ ((ts.mm & 4u) ? hm : 0);
    return (uint32_t)((((i0 << sh) | i1) << sh) | i2);
  }
  __device__ __forceinline__ bool fetch_voxel_bit(const TS& ts) const {
    if (FMT != SVD_FMT_SSVDAG) return (ts.hdr & (1u << ts.cli)) != 0;
    if ((int)ts.level < s.innerLevels) return ((ts.hdr >> (ts.cli << 1)) & 3u) != 0;
    // the leaf's 64 bits as one value: a select between the two words can
    // become a select between their addresses, which keeps them in memory
    const uint64_t bits = ((uint64_t)ts.leaf_hi << 32) | ts.hdr;
    return ((bits >> ts.cli) & 1u) != 0;
  }
  __device__ __forceinline__ void fetch_data(TS& ts) {
    if (FMT != SVD_FMT_SSVDAG) { ts.hdr = word((uint32_t)ts.node_index); if (INSTR) bytes += 4; return; }
    // both words assigned on both paths: stores of one field or the other
    // could be merged into one store through a selected address, which keeps
    // the two in scratch memory
    uint32_t lo, hi = ts.leaf_hi;
    if ((int)ts.level < s.innerLevels) {
      lo = half((uint32_t)(ts.node_index + (int)s.levelOffsets[ts.level]));
      if (INSTR) bytes += 2;
    } else {
      uint2 v = make_uint2(0u, 0u);
      if ((uint32_t)ts.node_index < s.nLeaves) v = s.leaves[(uint32_t)ts.node_index];
      else bad = true;
      if (INSTR) bytes += 8;
      lo = v.x; hi = v.y;
    }
    ts.hdr = lo;
    ts.leaf_hi = hi;
  }
  __device__ __forceinline__ void fetch_child_index_in(TS& ts) {
    if (FMT != SVD_FMT_SSVDAG) {
      const int pos = __popc((ts.hdr & 0xFFu) >> ts.cli);
      ts.node_index = (int)word((uint32_t)(ts.node_index + pos));
      if (INSTR) bytes += 4;
      if (FMT == SVD_FMT_USSVDAG) {
        const uint32_t b = ts.hdr >> ts.cli;
        ts.mm ^= ((b >> 8) & 1u) | ((b >> 15) & 2u) | ((b >> 22) & 4u);
      }
      return;
    }
    // childIndir table (octree_dda_renderer.cpp:533-561), computed in registers:
    // children i > cli take min(code_i, 2) words, i.e. one per non-empty code
    // plus one more per code with its high bit set
    const uint32_t lo = ts.hdr & 0x5555u, hi = (ts.hdr >> 1) & 0x5555u;
    const uint32_t above = 0xFFFFu & ~((2u << (2 * ts.cli + 1)) - 1u);   // fields cli+1 .. 7
    int off = 1 + (int)s.levelOffsets[ts.level] + ts.node_index +
              __popc((lo | hi) & above) + __popc(hi & above);
    const int cm = (int)((ts.hdr >> (ts.cli << 1)) & 3u);
    int cx = (int)half((uint32_t)off);
    if (INSTR) bytes += 2;
    const uint32_t mf = ((uint32_t)cx >> 13) & 7u;
    cx &= -57345;
    if (cm > 1) {
      const int cy = (int)half((uint32_t)off + 1);
      if (INSTR) bytes += 2;
      cx = (cx << 16) | cy;
      cx |= (cm & 1) << 29;
    }
    ts.mm ^= mf;
    ts.node_index = cx;
  }
  __device__ __forceinline__ static void next_crossing(TS& ts, const float o[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int dic = ts.di[a] > 0 ? ts.di[a] : 0;
      const float pn = (float)(ts.idx[a] + dic) * ts.cell;
      ts.tn[a] = (pn - o[a]) * ts.inv[a];
    }
  }
  __device__ __forceinline__ static void next_mask(const TS& ts, int m[3]) {
    m[0] = (ts.tn[0] < ts.tn[1]) & (ts.tn[0] <= ts.tn[2]);
    m[1] = (ts.tn[1] < ts.tn[2]) & (ts.tn[1] <= ts.tn[0]);
    m[2] = (ts.tn[2] < ts.tn[0]) & (ts.tn[2] <= ts.tn[1]);
  }
  // The DDA step of the current cell: m, the axis whose plane is crossed
  // first, and delta = m * di.  An empty cell takes the step (dda_next); a
  // full one only tests whether li + delta stays in its node (down_in).
  __device__ __forceinline__ void dda_next(TS& ts, const int m[3], const int delta[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) { ts.idx[a] += delta[a]; ts.li[a] += delta[a]; }
    float tc = (float)m[0] * ts.tn[0];
    tc = tc + (float)m[1] * ts.tn[1];
    tc = tc + (float)m[2] * ts.tn[2];
    ts.t_current = tc;
#pragma unroll
    for (int a = 0; a < 3; ++a) ts.tn[a] = ts.tn[a] + ((float)m[a] * ts.cell) * fabsf(ts.inv[a]);
  }
  __device__ __forceinline__ static bool in_bounds(const int li[3], int sz) {
    return (uint32_t)li[0] < (uint32_t)sz && (uint32_t)li[1] < (uint32_t)sz && (uint32_t)li[2] < (uint32_t)sz;
  }
  __device__ __forceinline__ void up_in(const float o[3], TS& ts) {
    uint32_t dl = ts.level;
    pop(ts);
    dl -= ts.level;
#pragma unroll
    for (int a = 0; a < 3; ++a) ts.idx[a] >>= dl;
    ts.cell *= (float)(1 << dl);
    ts.cns = (int)ts.level < s.innerLevels ? 2 : s.leafSize;
#pragma unroll
    for (int a = 0; a < 3; ++a) ts.li[a] = ts.idx[a] & 1;
    next_crossing(ts, o);
  }
  // The position step of go_down_one_level (octree_dda.frag.glsl:434-451).
  // Its next-crossing and local-index updates read nothing this step reads,
  // so a descent that steps several levels (into a leaf brick,
  // octree_dda.frag.glsl:466-476) computes them once, at its end.
  __device__ __forceinline__ void step_down(const float o[3], const float d[3], TS& ts) {
    ++ts.level;
    ts.cell *= 0.5f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float pa = o[a] + ts.t_current * d[a];
      const float pc = (float)(ts.idx[a] * 2 + 1) * ts.cell;
      ts.idx[a] = ts.idx[a] * 2 + (pc < pa ? 1 : 0);
    }
  }
  // next_in_node: li + delta is inside the current node (a full cell below the
  // leaf level is in a node of side 2).
  __device__ __forceinline__ void down_in(const float o[3], const float d[3], TS& ts, bool next_in_node) {
    if (next_in_node) push(ts.node_index, ts.hdr, ts.mm, ts.level);
    fetch_child_index_in(ts);
    step_down(o, d, ts);
    if ((int)ts.level == s.innerLevels) {
      ts.cns = s.leafSize;
      int vc = s.leafSize / 2;
      while (vc > 1) { step_down(o, d, ts); vc >>= 1; }
    }
    next_crossing(ts, o);
#pragma unroll
    for (int a = 0; a < 3; ++a) ts.li[a] = ts.idx[a] & (ts.cns - 1);
    if (INSTR) {
      path = (path ^ (uint32_t)ts.node_index) * 16777619u;
      path = (path ^ ts.level) * 16777619u;
    }
  }

  // Per-ray state between iterations of the traversal loop.
  struct Ray {
    TS ts;
    float o[3], d[3];
    float tmax, projf;
    int it;
    uint32_t maxLevel, maxIters;
    uint32_t stepm;   // axis mask of the last DDA step; the step is mask * di
    __device__ __forceinline__ void steps(int st[3]) const {
#pragma unroll
      for (int a = 0; a < 3; ++a) st[a] = (stepm >> a) & 1u ? ts.di[a] : 0;
    }
  };

  // Setup of traceRay; returns true when the ray is already finished (out set).
  __device__ __forceinline__ bool begin(Ray& r, const float oin[3], const float din[3], float tmin, float tmax,
                                        float projf, uint32_t maxIters, uint32_t drawLevel, float out[4]) {
    st.size = 0;
    r.stepm = 0;
    const float eps = 1e-4f;
    const float scale = 1.0f / (2.0f * s.rootHalfSide);
    float omin[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      omin[a] = s.center[a] - s.rootHalfSide;
      r.o[a] = (oin[a] - omin[a]) * scale;
      const float sg = gsign(din[a]);
      r.d[a] = din[a];
      if (r.d[a] * sg < eps) r.d[a] = sg * eps;
    }
    tmin = tmin * scale;
    tmax = tmax * scale;
    float tMin[3], tMax[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float cmn = (s.bmin[a] - omin[a]) * scale;
      const float cmx = (s.bmax[a] - omin[a]) * scale;
      const float t1 = (cmn - r.o[a]) / r.d[a];
      const float t2 = (cmx - r.o[a]) / r.d[a];
      tMin[a] = gmin(t1, t2);
      tMax[a] = gmax(t1, t2);
    }
    const float tix = gmax(gmax(tMin[0], 0.0f), gmax(tMin[1], tMin[2]));
    const float tiy = gmin(tMax[0], gmin(tMax[1], tMax[2]));
    tmin = gmax(tix, tmin + 1e-10f);
    r.tmax = gmin(tiy, tmax);
    if (!(tix < tiy)) { out[0] = -4.f; out[1] = 0.f; out[2] = 0.f; out[3] = -1.f; return true; }
    TS& ts = r.ts;
    ts = TS{};
    ts.t_current = tmin;
#pragma unroll
    for (int a = 0; a < 3; ++a) { ts.inv[a] = 1.0f / r.d[a]; ts.di[a] = isign(r.d[a]); }
    ts.mm = 0;
    ts.level = 0;
    ts.cell = 0.5f;
    ts.leaf_hi = 0;
    {
      const float voxel_eps = 1.0f / (256.f * 1024.f);
      const float tt = ts.t_current + voxel_eps;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float pa = r.o[a] + tt * r.d[a];
        ts.idx[a] = (int)(pa / ts.cell);
      }
      next_crossing(ts, r.o);
#pragma unroll
      for (int a = 0; a < 3; ++a) ts.li[a] = ts.idx[a] % 2;
    }
    ts.cns = 2;
    ts.node_index = 0;
    fetch_data(ts);
    ts.cli = lin_idx(ts, ts.cns);
    r.it = 0;
    r.projf = projf;
    r.maxIters = maxIters;
    r.maxLevel = min((uint32_t)s.innerLevels, drawLevel - 1);
    return false;
  }

  // One iteration of traceRay's loop; returns true when the ray is finished.
  __device__ __forceinline__ bool step(Ray& r, float out[4]) {
    TS& ts = r.ts;
    const bool full = fetch_voxel_bit(ts);
    int m[3], delta[3], nl[3];
    next_mask(ts, m);
#pragma unroll
    for (int a = 0; a < 3; ++a) { delta[a] = m[a] ? ts.di[a] : 0; nl[a] = ts.li[a] + delta[a]; }
    const bool next_in_node = in_bounds(nl, ts.cns);
    if (!full) {
      r.stepm = (uint32_t)(m[0] | (m[1] << 1) | (m[2] << 2));
      dda_next(ts, m, delta);
      if (!next_in_node) {
        if (st.size == 0) { out[0] = -1.f; out[1] = 0.f; out[2] = (float)r.it; out[3] = -1.f; return true; }
        up_in(r.o, ts);
      }
    } else {
      const bool hit = (ts.level >= r.maxLevel) || ((ts.cell * r.projf) < ts.t_current);
      if (hit) {
        out[0] = ts.t_current * (2.0f * s.rootHalfSide); out[1] = (float)ts.level; out[2] = (float)r.it;
        out[3] = (float)ts.node_index;
        return true;
      }
      down_in(r.o, r.d, ts, next_in_node);
      fetch_data(ts);
    }
    ts.cli = lin_idx(ts, ts.cns);
    ++r.it;
    if ((ts.t_current < r.tmax) && ((uint32_t)r.it < r.maxIters)) return false;
    if ((uint32_t)r.it >= r.maxIters) { out[0] = -3.f; out[1] = 0.f; out[2] = (float)r.it; out[3] = -1.f; return true; }
    out[0] = -2.f; out[1] = 0.f; out[2] = (float)r.it; out[3] = -1.f;
    return true;
  }

  __device__ void trace(const float oin[3], const float din[3], float tmin, float tmax, float projf,
                        uint32_t maxIters, uint32_t drawLevel, float out[4], int step_out[3]) {
    Ray r;
    if (!begin(r, oin, din, tmin, tmax, projf, maxIters, drawLevel, out)) {
      while (!step(r, out)) {}
    }
    r.steps(step_out);
  }
};

__device__ __forceinline__ void mat4_mul(const float* m, const float v[4], float r[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float s = m[0 * 4 + i] * v[0];
    s = s + m[1 * 4 + i] * v[1];
    s = s + m[2 * 4 + i] * v[2];
    s = s + m[3 * 4 + i] * v[3];
    r[i] = s;
  }
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

__device__ __forceinline__ void normalize3(float v[3]) {
  const float l = sqrtf(dot3(v, v));
  v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
}

// computeCameraRay (octree_dda.frag.glsl:587-603)
// (sw, sh) = the shader's screenRes uniform: the frame size, or a fraction of
// it for the depth pre-pass
// camera_ray_sc takes the shader's screenCoords (sx, sy) in [-1, 1].
__device__ __forceinline__ void camera_ray_sc(const svd_trace_params& p, float sx, float sy, float o[3], float d[3]) {
  const float s0[4] = {sx, sy, 0.f, 1.f}, s1[4] = {sx, sy, 1.f, 1.f};
  float w0[4], w1[4];
  mat4_mul(p.proj_inv, s0, w0);
  mat4_mul(p.proj_inv, s1, w1);
  float rd[3] = {w1[0] / w1[3] - w0[0] / w0[3], w1[1] / w1[3] - w0[1] / w0[3], w1[2] / w1[3] - w0[2] / w0[3]};
  normalize3(rd);
  const float z4[4] = {0.f, 0.f, 0.f, 1.f}, d4[4] = {rd[0], rd[1], rd[2], 1.f};
  float op[4], ep[4];
  mat4_mul(p.view_inv, z4, op);
  mat4_mul(p.view_inv, d4, ep);
  o[0] = op[0]; o[1] = op[1]; o[2] = op[2];
  d[0] = ep[0] - op[0]; d[1] = ep[1] - op[1]; d[2] = ep[2] - op[2];
  normalize3(d);
}

__device__ __forceinline__ float screen_coord(uint32_t px, float sw) { return (((float)px + 0.5f) / sw) * 2.0f - 1.0f; }

__device__ __forceinline__ void camera_ray(const svd_trace_params& p, uint32_t px, uint32_t py, float sw, float sh,
                                           float o[3], float d[3]) {
  camera_ray_sc(p, screen_coord(px, sw), screen_coord(py, sh), o, d);
}

// Shadow ray from a hit position, its normal and cell size (octree_dda.frag.glsl:803-815).
struct ShadowRay { float o[3], d[3], tmin, tmax, projf; };
__device__ __forceinline__ ShadowRay shadow_from(const svd_trace_params& p, const float hit[3], const float hn[3],
                                                 float cellSize) {
  ShadowRay r;
  const float pp[3] = {hit[0] + (hn[0] * cellSize) * 0.5f, hit[1] + (hn[1] * cellSize) * 0.5f,
                       hit[2] + (hn[2] * cellSize) * 0.5f};
  const float l2p[3] = {pp[0] - p.light_pos[0], pp[1] - p.light_pos[1], pp[2] - p.light_pos[2]};
  const float len = sqrtf(dot3(l2p, l2p));
  r.projf = len / cellSize;
  r.d[0] = -(l2p[0] / len); r.d[1] = -(l2p[1] / len); r.d[2] = -(l2p[2] / len);
  r.o[0] = pp[0]; r.o[1] = pp[1]; r.o[2] = pp[2];
  r.tmin = 1e-3f * 0.5f;
  r.tmax = len * 0.5f;
  return r;
}

// Shadow ray of a primary hit as the pretty mode casts it with no normal
// samples: hitNorm = normalize(-stepDir) (octree_dda.frag.glsl:686-688, 753-754, 803-815).
__device__ __forceinline__ ShadowRay shadow_ray(const Dev& s, const svd_trace_params& p, const float o[3],
                                                const float d[3], const float res[4], const int step[3]) {
  const float hit[3] = {o[0] + res[0] * d[0], o[1] + res[0] * d[1], o[2] + res[0] * d[2]};
  const float cellSize = (2.0f * s.rootHalfSide) / ldexpf(1.0f, (int)res[1]);
  float hn[3] = {-(float)step[0], -(float)step[1], -(float)step[2]};
  normalize3(hn);
  return shadow_from(p, hit, hn, cellSize);
}

// The minimum-depth pre-pass ("beam optimisation", octree_dda_renderer.cpp:389-408):
// a texture of (width/8) x (height/8) texels holding the hit distance of a
// coarse ray per texel (0 where the ray does not hit).  getMinT(8)
// (octree_dda.frag.glsl:607-617) reads the 2x2 texels around the pixel; GL
// leaves reads outside the texture undefined, this build reads them as 0.
struct MinDepth { const float* t; uint32_t w, h; };

__device__ __forceinline__ float depth_texel(const MinDepth& md, int x, int y) {
  return (x >= 0 && y >= 0 && (uint32_t)x < md.w && (uint32_t)y < md.h) ? md.t[(uint32_t)y * md.w + (uint32_t)x] : 0.f;
}

__device__ __forceinline__ float min_t(const MinDepth& md, uint32_t px, uint32_t py) {
  // ivec2((gl_FragCoord.xy - delta / 2) / delta), delta = 8: truncation toward zero
  const int x = (int)((((float)px + 0.5f) - 4.0f) / 8.0f);
  const int y = (int)((((float)py + 0.5f) - 4.0f) / 8.0f);
  const float tl = depth_texel(md, x, y), tr = depth_texel(md, x + 1, y);
  const float bl = depth_texel(md, x, y + 1), br = depth_texel(md, x + 1, y + 1);
  return gmin(gmin(tl, tr), gmin(bl, br));
}

// One pixel of one camera: primary ray, then the shadow ray of a hit.  With a
// minimum-depth texture the primary ray starts at getMinT(8) - 1e-3
// (octree_dda.frag.glsl:681).  INSTR writes svd_ray_result records (prim,
// shad; a pixel without a shadow ray gets the -5 record); otherwise traceRay's
// vec4 (t, level, iters, node) goes to primHit, and to shadHit only where a
// shadow ray is cast.
template <int FMT, bool INSTR = true>
__device__ __forceinline__ void trace_pixel(const Dev& s, const svd_trace_params& p, uint32_t px, uint32_t py,
                                            uint32_t* stackMem, svd_ray_result* __restrict__ prim,
                                            svd_ray_result* __restrict__ shad, MinDepth md = MinDepth{nullptr, 0, 0},
                                            float4* __restrict__ primHit = nullptr,
                                            float4* __restrict__ shadHit = nullptr) {
  const int depth = s.innerLevels + 1;
  const int lane = threadIdx.y * blockDim.x + threadIdx.x;
  Stack stk = make_stack<FMT>(stackMem, depth, lane);
  Tracer<FMT, INSTR> tr(s, stk);
  float o[3], d[3];
  camera_ray(p, px, py, (float)p.width, (float)p.height, o, d);
  float res[4];
  int step[3];
  const float tmin = md.t ? min_t(md, px, py) - 1e-3f : 0.0f;
  tr.trace(o, d, tmin, 1e30f, p.projection_factor, p.max_iters, p.draw_level, res, step);
  if (tr.bad) { res[0] = -6.f; res[1] = 0.f; res[3] = -1.f; }
  const size_t k = (size_t)py * p.width + px;
  if (INSTR) prim[k] = svd_ray_result{res[0], res[1], res[2], res[3], tr.path, tr.bytes};
  else primHit[k] = make_float4(res[0], res[1], res[2], res[3]);
  if (INSTR ? !shad : !shadHit) return;
  svd_ray_result sr{-5.f, 0.f, 0.f, -1.f, 0u, 0u};
  if (p.shadows && res[0] >= 0.f) {
    const ShadowRay sh = shadow_ray(s, p, o, d, res, step);
    float sres[4];
    int sstep[3];
    tr.path = 2166136261u;
    tr.bytes = 0;
    tr.bad = false;
    tr.trace(sh.o, sh.d, sh.tmin, sh.tmax, sh.projf, p.max_iters, p.draw_level, sres, sstep);
    if (tr.bad) { sres[0] = -6.f; sres[1] = 0.f; sres[3] = -1.f; }
    if (!INSTR) {
      shadHit[k] = make_float4(sres[0], sres[1], sres[2], sres[3]);
      return;
    }
    sr = svd_ray_result{sres[0], sres[1], sres[2], sres[3], tr.path, tr.bytes};
  }
  if (INSTR) shad[k] = sr;
}

// ---- viewer shading ---------------------------------------------------------
// GLSL leaves sin, cos and log2 to the implementation.  This build defines
// them with plain float operations in a fixed order, so that a CPU
// restatement reproduces every bit:
//  - sin/cos: k = floor(x * 2/pi + 0.5); r = x - k*P1 - k*P2 - k*P3 with pi/2
//    split into two 12-bit parts and a remainder (k*P1 and k*P2 are exact for
//    |k| < 2^12); Taylor polynomials in r^2 on [-pi/4, pi/4] (sin to r^11,
//    cos to r^12); the quadrant k mod 4 picks the sign and function.
//  - log2: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh f with
//    f = (m - 1) / (m + 1), series to f^13; log2 x = e + ln m / ln 2.
//  - mod(x, y) = x - y * floor(x / y) (the GLSL definition).
__device__ __forceinline__ void vm_sincos(float x, float& sn, float& cs) {
  const float k = floorf(x * 0.636619747f + 0.5f);
  float r = x - k * 1.5703125f;
  r = r - k * 4.8375129699707031e-4f;
  r = r - k * 7.5497901264043321e-8f;
  const float r2 = r * r;
  float ps = -2.50521084e-8f;          // -1/11!
  ps = ps * r2 + 2.75573192e-6f;       //  1/9!
  ps = ps * r2 + -1.98412698e-4f;      // -1/7!
  ps = ps * r2 + 8.33333333e-3f;       //  1/5!
  ps = ps * r2 + -1.66666667e-1f;      // -1/3!
  const float s = r + (r * r2) * ps;
  float pc = 2.08767570e-9f;           //  1/12!
  pc = pc * r2 + -2.75573192e-7f;      // -1/10!
  pc = pc * r2 + 2.48015873e-5f;       //  1/8!
  pc = pc * r2 + -1.38888889e-3f;      // -1/6!
  pc = pc * r2 + 4.16666667e-2f;       //  1/4!
  const float c = (1.0f - 0.5f * r2) + (r2 * r2) * pc;
  const int q = (int)k & 3;
  sn = q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
  cs = q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}

__device__ __forceinline__ float vm_log2(float x) {
  if (x != x || x < 0.f) return __builtin_nanf("");
  if (x == 0.f) return -__builtin_inff();
  if (x == __builtin_inff()) return x;
  int e;
  float m = frexpf(x, &e);
  if (m < 0.707106769f) { m = m * 2.0f; e = e - 1; }
  const float f = (m - 1.0f) / (m + 1.0f);
  const float s = f * f;
  float p = 7.69230769e-2f;            // 1/13
  p = p * s + 9.09090909e-2f;          // 1/11
  p = p * s + 1.11111111e-1f;          // 1/9
  p = p * s + 1.42857143e-1f;          // 1/7
  p = p * s + 2.00000000e-1f;          // 1/5
  p = p * s + 3.33333333e-1f;          // 1/3
  const float f2 = 2.0f * f;
  const float ln = f2 + (f2 * s) * p;
  return (float)e + ln * 1.44269502f;
}

__device__ __forceinline__ float vm_mod(float x, float y) { return x - y * floorf(x / y); }

// Uniforms of the viewer program beyond the traversal's.
struct ViewArgs {
  int mode, normSamples;
  uint32_t selected, randomColors, freqColors;
  uint32_t levels;                  // LEVELS
  uint32_t leafTexels, innerTexels; // textureSize(leafNodes), textureSize(innerNodes) (SSVDAG)
  uint32_t nOffsets;                // entries of levelOffsets
};

// VIEWER_MODE main() for one pixel (octree_dda.frag.glsl:671-858): primary
// ray (from getMinT(8) - 1e-3 with a depth texture), the render mode's
// intensity, the normal samples of the pretty mode, the highlight / frequency /
// random colouring and the shadow ray.  rgb gets the colour the fragment
// writes; a fragment the shader discards gets 0.
template <int FMT>
__device__ __forceinline__ void render_pixel(const Dev& s, const svd_trace_params& p, const ViewArgs& va,
                                             uint32_t px, uint32_t py, uint32_t* stackMem, float* __restrict__ rgb,
                                             svd_ray_result* __restrict__ prim, MinDepth md) {
  const int lane = threadIdx.y * blockDim.x + threadIdx.x;
  Tracer<FMT> tr(s, make_stack<FMT>(stackMem, s.innerLevels + 1, lane));
  const float sw = (float)p.width, sh = (float)p.height;
  const float scx = screen_coord(px, sw), scy = screen_coord(py, sh);
  float o[3], d[3];
  camera_ray_sc(p, scx, scy, o, d);
  float res[4];
  int step[3];
  const float tmin = md.t ? min_t(md, px, py) - 1e-3f : 0.0f;
  tr.trace(o, d, tmin, 1e30f, p.projection_factor, p.max_iters, p.draw_level, res, step);
  if (tr.bad) { res[0] = -6.f; res[1] = 0.f; res[3] = -1.f; }
  const size_t k = (size_t)py * p.width + px;
  if (prim) prim[k] = svd_ray_result{res[0], res[1], res[2], res[3], tr.path, tr.bytes};
  float col[3];
  if (res[0] >= 0.f) {
    const float hit[3] = {o[0] + res[0] * d[0], o[1] + res[0] * d[1], o[2] + res[0] * d[2]};
    const float cellSize = (2.0f * s.rootHalfSide) / ldexpf(1.0f, (int)res[1]);
    const uint32_t nodeIndex = (uint32_t)res[3];
    float hn[3] = {-(float)step[0], -(float)step[1], -(float)step[2]};
    const float ext[3] = {s.bmax[0] - s.bmin[0], s.bmax[1] - s.bmin[1], s.bmax[2] - s.bmin[2]};
    const float diag = sqrtf(dot3(ext, ext));
    float t = 0.f;
    if (va.mode == 0) {
      t = 1.0f - (res[2] / (float)p.max_iters);
    } else if (va.mode == 1) {
      t = res[0] / diag;
    } else if (va.mode == 2) {
      t = vm_log2((2.0f * s.rootHalfSide) / res[1]) / (float)va.levels;
    } else if (va.mode == 3) {
      t = 1.0f;
      const float nrl = 64.0f * cellSize;
      const float fni = (float)nodeIndex;
      const float rot = (3.14159f * 2.0f) / (float)va.normSamples;
      const float off = vm_mod(((res[0] * scx) * fni) * 0.01f + (fni * 0.001f) * scy, 1.0f);
      for (int i = 0; i < va.normSamples; ++i) {
        const float rad = 0.1f + vm_mod((((float)i / (float)va.normSamples) * fni) * 0.001f, 0.9f);
        const float ang = (float)i * rot + off;
        float sn, cs;
        vm_sincos(ang, sn, cs);
        float so[3], sd[3];
        camera_ray_sc(p, scx + (rad * cs) / sw, scy + (rad * sn) / sh, so, sd);
        float sres[4];
        int sstep[3];
        tr.trace(so, sd, res[0] - nrl * 0.5f, res[0] + nrl * 0.5f, p.projection_factor, p.max_iters, p.draw_level,
                 sres, sstep);
        if (sres[0] >= 0.f && fabsf(sres[0] - res[0]) <= nrl) {
          hn[0] = hn[0] + -(float)sstep[0];
          hn[1] = hn[1] + -(float)sstep[1];
          hn[2] = hn[2] + -(float)sstep[2];
        }
      }
      normalize3(hn);
      float ld[3] = {p.light_pos[0] - hit[0], p.light_pos[1] - hit[1], p.light_pos[2] - hit[2]};
      normalize3(ld);
      t = t * (0.5f + 0.5f * gmax(dot3(hn, ld), 0.0f));
      const float dl[3] = {hit[0] - p.light_pos[0], hit[1] - p.light_pos[1], hit[2] - p.light_pos[2]};
      t = t - (0.2f * sqrtf(dot3(dl, dl))) / diag;
    }
    col[0] = col[1] = col[2] = t;
    if (nodeIndex == va.selected) {
      col[0] = col[0] * 0.5f;
      col[1] = col[1] * 0.5f;
      col[2] = 1.0f;
    } else if (FMT == SVD_FMT_SSVDAG && va.freqColors) {
      const int hitLev = (int)res[1];
      float levSize = 0.f;
      if (hitLev > (int)va.levels - 3) {
        levSize = (float)va.leafTexels;
      } else if (hitLev == (int)va.levels - 3) {
        if ((uint32_t)hitLev < va.nOffsets) levSize = (float)(va.innerTexels * 4u - s.levelOffsets[hitLev]);
      } else if ((uint32_t)hitLev + 1 < va.nOffsets) {
        levSize = (float)(s.levelOffsets[hitLev + 1] - s.levelOffsets[hitLev]);
      }
      const float u = 1.0f - (float)nodeIndex / levSize;
      const float a = 1.0f - u * u;
      col[0] = col[0] * (1.0f * (1.0f - a) + 0.0f * a);
      col[1] = col[1] * (1.0f * (1.0f - a) + 0.2f * a);
      col[2] = col[2] * (1.0f * (1.0f - a) + 0.8f * a);
    } else if (va.randomColors) {
      float rc[3] = {(float)(nodeIndex % 100u) / 100.0f, (float)((3u * nodeIndex) % 200u) / 200.0f,
                     (float)((2u * nodeIndex) % 300u) / 300.0f};
      normalize3(rc);
      col[0] = col[0] * rc[0];
      col[1] = col[1] * rc[1];
      col[2] = col[2] * rc[2];
    }
    if (p.shadows) {
      const ShadowRay sh = shadow_from(p, hit, hn, cellSize);
      float sres[4];
      int sstep[3];
      tr.trace(sh.o, sh.d, sh.tmin, sh.tmax, sh.projf, p.max_iters, p.draw_level, sres, sstep);
      const float f = sres[0] > 0.f ? 0.8f : 1.0f;
      col[0] = col[0] * f;
      col[1] = col[1] * f;
      col[2] = col[2] * f;
    }
  } else if (res[0] == -1.f || res[0] == -2.f || res[0] == -4.f) {
    col[0] = col[1] = col[2] = 1.0f;
  } else if (res[0] == -3.f) {
    col[0] = 1.0f; col[1] = 0.0f; col[2] = 0.0f;
  } else {
    col[0] = col[1] = col[2] = 0.0f;
  }
  rgb[3 * k + 0] = col[0];
  rgb[3 * k + 1] = col[1];
  rgb[3 * k + 2] = col[2];
}

// AO_MODE main() (octree_dda.frag.glsl:884-919), in three passes so that
// the hemisphere rays of the pixels with a hit fill every lane:
//  1. ao_gbuffer: per pixel, the primary ray and the G-buffer entry.  The
//     reference's AO program reads a G-buffer (hitPosTex, hitNormTex,
//     depthTex) that nothing in the reference writes; this build defines it
//     from the primary ray: hitPos = o + t d of a hit and (0, 0, 0)
//     otherwise, hitNorm = -stepDir as trace_ray returns it, cell size
//     2 rootHalfSide / 2^level.  A discarded fragment (hitPos = 0) gets -1;
//     the others are appended to a list.
//  2. k_ao_rays: one (entry, ray) pair per lane over the list; a ray that
//     hits adds 1 to its pixel's count.  A sample direction with a NaN
//     component counts as a miss (the box test rejects such rays).
//  3. k_ao_finish: visibility (1 - hits / numAORays)^2.  The count is an
//     integer, so the order of the additions does not matter.
struct AOEntry {
  float hit[3];       // hit position moved 1e-3 along the normal
  float projf;        // 0.5 / cell size
  int8_t step[4];     // stepDir: x, y, z (the normal is -stepDir)
  int base;           // int(gl_FragCoord.x * gl_FragCoord.y)
  uint32_t pixel;
};

template <int FMT>
__device__ __forceinline__ void ao_gbuffer(const Dev& s, const svd_trace_params& p, uint32_t px, uint32_t py,
                                           uint32_t* stackMem, float* __restrict__ vis, uint32_t* __restrict__ cnt,
                                           AOEntry* __restrict__ entries, uint32_t* __restrict__ nEntries,
                                           svd_ray_result* __restrict__ prim) {
  const int lane = threadIdx.y * blockDim.x + threadIdx.x;
  Tracer<FMT> tr(s, make_stack<FMT>(stackMem, s.innerLevels + 1, lane));
  float o[3], d[3];
  camera_ray(p, px, py, (float)p.width, (float)p.height, o, d);
  float res[4];
  int step[3];
  tr.trace(o, d, 0.0f, 1e30f, p.projection_factor, p.max_iters, p.draw_level, res, step);
  if (tr.bad) { res[0] = -6.f; res[1] = 0.f; res[3] = -1.f; }
  const size_t k = (size_t)py * p.width + px;
  if (prim) prim[k] = svd_ray_result{res[0], res[1], res[2], res[3], tr.path, tr.bytes};
  float hit[3] = {0.f, 0.f, 0.f};
  if (res[0] >= 0.f) {
    hit[0] = o[0] + res[0] * d[0];
    hit[1] = o[1] + res[0] * d[1];
    hit[2] = o[2] + res[0] * d[2];
  }
  if (hit[0] == 0.f && hit[1] == 0.f && hit[2] == 0.f) { vis[k] = -1.0f; return; }
  const float hn[3] = {-(float)step[0], -(float)step[1], -(float)step[2]};
  const float cellSize = (2.0f * s.rootHalfSide) / ldexpf(1.0f, (int)res[1]);
  AOEntry e;
  e.hit[0] = hit[0] + hn[0] * 1e-3f;
  e.hit[1] = hit[1] + hn[1] * 1e-3f;
  e.hit[2] = hit[2] + hn[2] * 1e-3f;
  e.projf = 0.5f / cellSize;
  e.step[0] = (int8_t)step[0];
  e.step[1] = (int8_t)step[1];
  e.step[2] = (int8_t)step[2];
  e.step[3] = 0;
  e.base = (int)((((float)px + 0.5f)) * (((float)py + 0.5f)));
  e.pixel = (uint32_t)k;
  vis[k] = 0.0f;
  cnt[k] = 0u;
  entries[atomicAdd(nEntries, 1u)] = e;
}

// Persistent form: lanes that finished their rays take new ones from a
// queue, so a wavefront is not held by its slowest ray.
//  - Refill only once kRefill lanes (or all) are idle: the setup of a new
//    ray is paid by the whole wavefront, so it is batched.
//  - Each wavefront owns a chunk of kChunk consecutive rays and takes a new
//    chunk with one atomic when it runs out (no hot device-wide counter).
//  - Ray r is pixel (tile-major tiles of tw x 256/tw, pixel-major inside a
//    tile), so a wavefront's rays stay spatially coherent.
//  - A lane that finishes a primary hit continues with its shadow ray.
// Every wavefront leaves once the queue is exhausted and its lanes are idle;
// each ray ends within max_iters iterations.
#ifndef SVD_TRACE_REFILL
#define SVD_TRACE_REFILL 32
#endif
constexpr uint32_t kRefill = SVD_TRACE_REFILL;
constexpr uint32_t kChunk = 64;

template <int FMT>
__global__ void __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(5)))
k_trace_persist(Dev s, svd_trace_params p, svd_ray_result* __restrict__ prim, svd_ray_result* __restrict__ shad,
                uint32_t* __restrict__ queue, uint32_t tw, uint32_t tilesX, uint32_t total) {
  enum : uint32_t { kIdle = 0, kWantShadow, kPrimary, kShadow };
  extern __shared__ uint32_t stackMem[];
  const int depth = s.innerLevels + 1;
  const uint32_t lane = threadIdx.x & 63u;
  Stack stk = make_stack<FMT>(stackMem, depth, (int)threadIdx.x);
  Tracer<FMT> tr(s, stk);
  typename Tracer<FMT>::Ray ray;
  const uint32_t th = kBT / tw;
  uint32_t state = kIdle, px = 0, py = 0;
  uint32_t qNext = 0, qEnd = 0;   // this wavefront's remaining rays [qNext, qEnd)
  bool exhausted = false;
  float out[4] = {0.f, 0.f, 0.f, 0.f};
  while (true) {
    // ---- refill: idle lanes take new pixels, finished primaries start shadows
    const uint64_t idle = __ballot(state == kIdle);
    const uint32_t nIdle = (uint32_t)__popcll(idle);
    bool start = state == kWantShadow;
    if (nIdle && !exhausted) {
      const uint32_t avail = qEnd - qNext;
      uint32_t newBase = 0;
      if (avail < nIdle) {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(queue, kChunk);
        newBase = __shfl(b, 0);
      }
      if (state == kIdle) {
        const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        const uint32_t r = rank < avail ? qNext + rank : (newBase < total ? newBase + (rank - avail) : total);
        if (r < total) {
          const uint32_t tile = r / kBT, in = r % kBT;
          px = (tile % tilesX) * tw + in % tw;
          py = (tile / tilesX) * th + in / tw;
          start = px < p.width && py < p.height;
        }
      }
      if (avail >= nIdle) {
        qNext += nIdle;
      } else if (newBase >= total) {
        qNext = qEnd;
        exhausted = true;
      } else {
        qNext = newBase + (nIdle - avail);
        qEnd = newBase + kChunk;
      }
    }
    bool done = false;
    if (start) {
      float o[3], d[3];
      camera_ray(p, px, py, (float)p.width, (float)p.height, o, d);
      float tmin = 0.0f, tmax = 1e30f, projf = p.projection_factor;
      if (state == kWantShadow) {   // from the primary's result and steps
        int st[3];
        ray.steps(st);
        const ShadowRay sh = shadow_ray(s, p, o, d, out, st);
#pragma unroll
        for (int a = 0; a < 3; ++a) { o[a] = sh.o[a]; d[a] = sh.d[a]; }
        tmin = sh.tmin;
        tmax = sh.tmax;
        projf = sh.projf;
      }
      state = state == kWantShadow ? kShadow : kPrimary;
      tr.path = 2166136261u;
      tr.bytes = 0;
      tr.bad = false;
      done = tr.begin(ray, o, d, tmin, tmax, projf, p.max_iters, p.draw_level, out);
    }
    if (!__ballot(state != kIdle)) {
      if (exhausted) break;
      continue;
    }
    // ---- traverse until kRefill lanes (or all) wait for work
    while (true) {
      if (state == kPrimary || state == kShadow) {
        if (!done) done = tr.step(ray, out);
        if (done) {
          done = false;
          if (tr.bad) { out[0] = -6.f; out[1] = 0.f; out[3] = -1.f; }
          const size_t k = (size_t)py * p.width + px;
          const svd_ray_result res{out[0], out[1], out[2], out[3], tr.path, tr.bytes};
          if (state == kShadow) {
            shad[k] = res;
            state = kIdle;
          } else {
            prim[k] = res;
            state = kIdle;
            if (shad) {
              if (p.shadows && out[0] >= 0.f) state = kWantShadow;
              else shad[k] = svd_ray_result{-5.f, 0.f, 0.f, -1.f, 0u, 0u};
            }
          }
        }
      }
      const uint32_t waiting = (uint32_t)__popcll(__ballot(state == kIdle || state == kWantShadow));
      if (waiting >= kRefill || waiting == 64u) break;
    }
  }
}

// Six wavefronts per SIMD (80 VGPRs, a 16-byte spill) measured 5% faster on
// the 1920x1080 and 3840x2160 terrain frames than the compiler's five (86
// VGPRs); seven and eight spill more and gain nothing at 1080p.
#ifndef SVD_TRACE_WAVES_PER_EU
#define SVD_TRACE_WAVES_PER_EU 6
#endif
#define SVD_TRACE_OCC __attribute__((amdgpu_waves_per_eu(SVD_TRACE_WAVES_PER_EU)))
template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_trace(Dev s, svd_trace_params p, svd_ray_result* __restrict__ prim,
                                               svd_ray_result* __restrict__ shad, MinDepth md) {
  extern __shared__ uint32_t stackMem[];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= p.width || py >= p.height) return;
  trace_pixel<FMT>(s, p, px, py, stackMem, prim, shad, md);
}

// The same frame without instrumentation: traceRay's vec4 per ray (16 bytes),
// shadow records only where a shadow ray is cast.
template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_trace_hits(Dev s, svd_trace_params p, float4* __restrict__ prim,
                                                    float4* __restrict__ shad, MinDepth md) {
  extern __shared__ uint32_t stackMem[];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= p.width || py >= p.height) return;
  trace_pixel<FMT, false>(s, p, px, py, stackMem, nullptr, nullptr, md, prim, shad);
}

// DEPTH_MODE pre-pass (octree_dda.frag.glsl:862-872) over the (w x h)
// texture: screenRes = (sw, sh), its own projection factor; a texel keeps the
// hit distance when it is > 0 and the cleared 0 otherwise.
template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_trace_depth(Dev s, svd_trace_params p, float sw, float sh,
                                                                     float projf, uint32_t w, uint32_t h,
                                                                     float* __restrict__ depth) {
  extern __shared__ uint32_t stackMem[];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= w || py >= h) return;
  const int lane = threadIdx.y * blockDim.x + threadIdx.x;
  Tracer<FMT, false> tr(s, make_stack<FMT>(stackMem, s.innerLevels + 1, lane));
  float o[3], d[3];
  camera_ray(p, px, py, sw, sh, o, d);
  float res[4];
  int step[3];
  tr.trace(o, d, 0.0f, 1e30f, projf, p.max_iters, p.draw_level, res, step);
  depth[(size_t)py * w + px] = res[0] > 0.f ? res[0] : 0.f;
}

// Many cameras over many scenes in one launch: blockIdx.z is the camera;
// every camera has the same resolution, results are [camera][row][column].
template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_trace_batch(const Dev* __restrict__ scenes,
                                                     const svd_trace_params* __restrict__ cams,
                                                     const uint32_t* __restrict__ camScene,
                                                     svd_ray_result* __restrict__ prim, svd_ray_result* __restrict__ shad) {
  extern __shared__ uint32_t stackMem[];
  const uint32_t cam = blockIdx.z;
  const svd_trace_params p = cams[cam];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= p.width || py >= p.height) return;
  const Dev s = scenes[camScene[cam]];
  const size_t off = (size_t)cam * p.width * p.height;
  trace_pixel<FMT>(s, p, px, py, stackMem, prim + off, shad ? shad + off : nullptr);
}

template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_render(Dev s, svd_trace_params p, ViewArgs va,
                                                                float* __restrict__ rgb,
                                                                svd_ray_result* __restrict__ prim, MinDepth md) {
  extern __shared__ uint32_t stackMem[];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= p.width || py >= p.height) return;
  render_pixel<FMT>(s, p, va, px, py, stackMem, rgb, prim, md);
}

template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_ao_gbuffer(Dev s, svd_trace_params p, float* __restrict__ vis,
                                                                    uint32_t* __restrict__ cnt,
                                                                    AOEntry* __restrict__ entries,
                                                                    uint32_t* __restrict__ nEntries,
                                                                    svd_ray_result* __restrict__ prim) {
  extern __shared__ uint32_t stackMem[];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= p.width || py >= p.height) return;
  ao_gbuffer<FMT>(s, p, px, py, stackMem, vis, cnt, entries, nEntries, prim);
}

// Grid-stride over the (entry, ray) pairs; consecutive lanes take the rays
// of one entry.  The pair count is read once, so every lane's loop ends.
template <int FMT>
__global__ void __launch_bounds__(kBT) SVD_TRACE_OCC k_ao_rays(Dev s, svd_trace_params p, const float* __restrict__ hs,
                                                                 uint32_t nrays, const AOEntry* __restrict__ entries,
                                                                 const uint32_t* __restrict__ nEntries,
                                                                 uint32_t* __restrict__ cnt) {
  extern __shared__ uint32_t stackMem[];
  Tracer<FMT> tr(s, make_stack<FMT>(stackMem, s.innerLevels + 1, threadIdx.x));
  const uint64_t total = (uint64_t)(*nEntries) * nrays;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += stride) {
    const uint32_t i = (uint32_t)(w % nrays);
    const AOEntry e = entries[w / nrays];
    const float hn[3] = {-(float)e.step[0], -(float)e.step[1], -(float)e.step[2]};
    const float rv[3] = {0.f, 1.f, 0.f};
    const float dn = dot3(rv, hn);
    float tg[3] = {rv[0] - hn[0] * dn, rv[1] - hn[1] * dn, rv[2] - hn[2] * dn};
    normalize3(tg);
    const float bt[3] = {hn[1] * tg[2] - tg[1] * hn[2], hn[2] * tg[0] - tg[2] * hn[0], hn[0] * tg[1] - tg[0] * hn[1]};
    const float* sm = hs + 3 * (((int)i + e.base) % kAOSamples);
    float rd[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) rd[a] = (tg[a] * sm[0] + hn[a] * sm[1]) + bt[a] * sm[2];
    normalize3(rd);
    if (rd[0] != rd[0] || rd[1] != rd[1] || rd[2] != rd[2]) continue;
    float ares[4];
    int astep[3];
    tr.trace(e.hit, rd, 0.0f, 0.5f, e.projf, p.max_iters, p.draw_level, ares, astep);
    if (ares[0] > 0.f) atomicAdd(&cnt[e.pixel], 1u);
  }
}

__global__ void k_ao_finish(float* __restrict__ vis, const uint32_t* __restrict__ cnt, uint64_t n, uint32_t nrays) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || vis[k] == -1.0f) return;
  const float v = nrays > 0 ? 1.0f - ((float)cnt[k] / (float)nrays) : 1.0f;
  vis[k] = v * v;
}

}  // namespace

void trace_frame_impl(svd_scene* sc, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad,
                      const float* depthProjf, float* minDepthOut);

void trace_frame(svd_scene* sc, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad) {
  trace_frame_impl(sc, p, prim, shad, nullptr, nullptr);
}

// Beam-optimised frame: the DEPTH_MODE pre-pass at (width/8) x (height/8)
// with its own projection factor, then the frame with every primary ray
// starting at getMinT(8) - 1e-3 (octree_dda_renderer.cpp:389-408).
void trace_frame_beam(svd_scene* sc, const svd_trace_params* p, float depthProjf, svd_ray_result* prim,
                      svd_ray_result* shad, float* minDepthOut) {
  trace_frame_impl(sc, p, prim, shad, &depthProjf, minDepthOut);
}

Dev scene_dev(const svd_scene* sc) {
  Dev d{};
  d.nodes = sc->nodes;
  d.inner = sc->inner;
  d.leaves = (const uint2*)sc->leaves;
  d.levelOffsets = sc->levelOffsets;
  d.nWords = sc->nWords;
  d.nInner = sc->nInner;
  d.nLeaves = sc->nLeaves;
  if (sc->fmt == SVD_FMT_SSVDAG || sc->fmt == SVD_FMT_ESVDAG) { d.innerLevels = (int)sc->levels - 2; d.leafSize = 4; }
  else { d.innerLevels = (int)sc->levels - 1; d.leafSize = 2; }
  for (int a = 0; a < 3; ++a) { d.bmin[a] = sc->bmin[a]; d.bmax[a] = sc->bmax[a]; d.center[a] = sc->center[a]; }
  d.rootHalfSide = sc->rootHalfSide;
  return d;
}

namespace {

// LDS of one workgroup: the descent stacks of its kBT lanes
size_t stack_bytes(const svd_scene* sc, const Dev& d) {
  return (size_t)(sc->fmt == SVD_FMT_USSVDAG ? kStackWords<SVD_FMT_USSVDAG> : 2) * (d.innerLevels + 1) * kBT *
         sizeof(uint32_t);
}

// The DEPTH_MODE pre-pass into `depth` ((width/8) x (height/8) floats).  The
// texture is (int)width / 8 x (int)height / 8; its screenRes uniform is the
// float width / 8, height / 8.
MinDepth depth_prepass(svd_scene* sc, const Dev& d, const svd_trace_params* p, float projf, float* depth) {
  svd_ctx* c = sc->ctx;
  MinDepth md{depth, p->width / 8, p->height / 8};
  if (!md.w || !md.h) return md;
  const uint32_t tw = c->traceTileW, th = kBT / tw;
  const size_t shmem = stack_bytes(sc, d);
  const float sw = (float)p->width / 8.0f, sh = (float)p->height / 8.0f;
  const dim3 dg(div_up(md.w, tw), div_up(md.h, th)), db(tw, th);
  switch (sc->fmt) {
    case SVD_FMT_SVDAG:
      k_trace_depth<SVD_FMT_SVDAG><<<dg, db, shmem, c->stream>>>(d, *p, sw, sh, projf, md.w, md.h, depth);
      break;
    case SVD_FMT_USSVDAG:
      k_trace_depth<SVD_FMT_USSVDAG><<<dg, db, shmem, c->stream>>>(d, *p, sw, sh, projf, md.w, md.h, depth);
      break;
    default:
      k_trace_depth<SVD_FMT_SSVDAG><<<dg, db, shmem, c->stream>>>(d, *p, sw, sh, projf, md.w, md.h, depth);
      break;
  }
  check_launch("k_trace_depth");
  return md;
}

}  // namespace

void trace_frame_impl(svd_scene* sc, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad,
                      const float* depthProjf, float* minDepthOut) {
  svd_ctx* c = sc->ctx;
  const Dev d = scene_dev(sc);
  if (p->width == 0 || p->height == 0) return;
  if (p->draw_level == 0) throw Error{SVD_ERR_ARG, "draw_level must be >= 1"};
  const uint32_t tw = c->traceTileW, th = kBT / tw;
  const dim3 grid(div_up(p->width, tw), div_up(p->height, th));
  const size_t shmem = stack_bytes(sc, d);
  MinDepth md{nullptr, 0, 0};
  float* depth = nullptr;
  if (depthProjf) {
    depth = minDepthOut ? minDepthOut : dalloc<float>(c, (uint64_t)(p->width / 8) * (p->height / 8) + 1);
    md = depth_prepass(sc, d, p, *depthProjf, depth);
  }
  if (c->traceMode == SVD_TRACE_PERSISTENT && !depthProjf) {
    const uint64_t total = (uint64_t)grid.x * grid.y * kBT;
    if (total > 0xF0000000ull) throw Error{SVD_ERR_ARG, "frame too large"};
    const void* fn = sc->fmt == SVD_FMT_SVDAG     ? (const void*)k_trace_persist<SVD_FMT_SVDAG>
                     : sc->fmt == SVD_FMT_USSVDAG ? (const void*)k_trace_persist<SVD_FMT_USSVDAG>
                                                  : (const void*)k_trace_persist<SVD_FMT_SSVDAG>;
    // the occupancy query costs milliseconds of host time: ask once per (format, LDS size)
    const uint64_t key = ((uint64_t)sc->fmt << 32) | (uint64_t)shmem;
    auto it = c->occupancy.find(key);
    if (it == c->occupancy.end()) {
      int q = 0;
      SVD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&q, fn, kBT, shmem));
      it = c->occupancy.emplace(key, q < 1 ? 1 : q).first;
    }
    const int perCU = it->second;
    const uint64_t want = (uint64_t)perCU * (uint64_t)(c->numCU > 0 ? c->numCU : 256);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, div_up(total, kBT));
    uint32_t* queue = dalloc<uint32_t>(c, 1);
    fill_u32(c, queue, 0u, 1);
    switch (sc->fmt) {
      case SVD_FMT_SVDAG:
        k_trace_persist<SVD_FMT_SVDAG><<<blocks, kBT, shmem, c->stream>>>(d, *p, prim, shad, queue, tw, grid.x,
                                                                          (uint32_t)total);
        break;
      case SVD_FMT_USSVDAG:
        k_trace_persist<SVD_FMT_USSVDAG><<<blocks, kBT, shmem, c->stream>>>(d, *p, prim, shad, queue, tw, grid.x,
                                                                            (uint32_t)total);
        break;
      default:
        k_trace_persist<SVD_FMT_SSVDAG><<<blocks, kBT, shmem, c->stream>>>(d, *p, prim, shad, queue, tw, grid.x,
                                                                           (uint32_t)total);
        break;
    }
    check_launch("k_trace_persist");
    dfree(c, queue);
    return;
  }
  const dim3 block(tw, th);
  switch (sc->fmt) {
    case SVD_FMT_SVDAG: k_trace<SVD_FMT_SVDAG><<<grid, block, shmem, c->stream>>>(d, *p, prim, shad, md); break;
    case SVD_FMT_USSVDAG: k_trace<SVD_FMT_USSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, prim, shad, md); break;
    default: k_trace<SVD_FMT_SSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, prim, shad, md); break;
  }
  check_launch("k_trace");
  if (depth && depth != minDepthOut) dfree(c, depth);
}

// The uninstrumented frame (k_trace_hits, always one workgroup per pixel
// tile): prim gets width*height vec4, shad (optional) the vec4 of every cast
// shadow ray.
void trace_frame_hits(svd_scene* sc, const svd_trace_params* p, float* prim, float* shad) {
  svd_ctx* c = sc->ctx;
  const Dev d = scene_dev(sc);
  if (p->width == 0 || p->height == 0) return;
  if (p->draw_level == 0) throw Error{SVD_ERR_ARG, "draw_level must be >= 1"};
  const uint32_t tw = c->traceTileW, th = kBT / tw;
  const dim3 grid(div_up(p->width, tw), div_up(p->height, th)), block(tw, th);
  const size_t shmem = stack_bytes(sc, d);
  const MinDepth md{nullptr, 0, 0};
  float4* ph = reinterpret_cast<float4*>(prim);
  float4* sh = reinterpret_cast<float4*>(shad);
  switch (sc->fmt) {
    case SVD_FMT_SVDAG: k_trace_hits<SVD_FMT_SVDAG><<<grid, block, shmem, c->stream>>>(d, *p, ph, sh, md); break;
    case SVD_FMT_USSVDAG: k_trace_hits<SVD_FMT_USSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, ph, sh, md); break;
    default: k_trace_hits<SVD_FMT_SSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, ph, sh, md); break;
  }
  check_launch("k_trace_hits");
}

namespace {

// Scratch of svd_render_noalloc: the minimum-depth pre-pass texture,
// (width / 8) x (height / 8) floats, in a static device array so that the
// entry point allocates nothing.  Every GPU the library runs on holds its own
// copy until the process ends.  k_trace_depth writes every texel of the
// current frame's texture before k_render reads it, so nothing needs clearing;
// calls that overlap on one GPU (two streams) share the array and race.
static __device__ __align__(256) float scratch[kRenderScratchTexels];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMu;
float* g_scratchPtr[kMaxScratchDevices] = {};
hipError_t g_scratchErr[kMaxScratchDevices] = {};
bool g_scratchDone[kMaxScratchDevices] = {};

}  // namespace

// Resolve the device address of `scratch` on the current device (once per
// device; later contexts of the same device reuse it).  The array's name is a
// host-side stand-in: host code only ever passes the resolved pointer.
void render_scratch_init(svd_ctx* c) {
  c->renderScratch = nullptr;
  c->renderScratchErr = hipErrorInvalidDevice;
  if (c->device < 0 || c->device >= kMaxScratchDevices) return;
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchDone[c->device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    g_scratchPtr[c->device] = e == hipSuccess ? (float*)ptr : nullptr;
    g_scratchErr[c->device] = e;
    g_scratchDone[c->device] = true;
  }
  c->renderScratch = g_scratchPtr[c->device];
  c->renderScratchErr = g_scratchErr[c->device];
}

// The viewer's frame (VIEWER_MODE, octree_dda.frag.glsl:671-858), optionally
// after the minimum-depth pre-pass (beamProjf > 0).  rgb: 3 floats per pixel.
// depthBuf: the pre-pass texture, or null to take it from the context arena.
void render_frame(svd_scene* sc, const svd_trace_params* p, const svd_view_params* v, float beamProjf, float* rgb,
                  svd_ray_result* prim, float* depthBuf) {
  svd_ctx* c = sc->ctx;
  const Dev d = scene_dev(sc);
  if (v->norm_samples < 0 || v->norm_samples > 64) throw Error{SVD_ERR_ARG, "norm_samples must be in [0, 64]"};
  if (p->width == 0 || p->height == 0) return;
  if (p->draw_level == 0) throw Error{SVD_ERR_ARG, "draw_level must be >= 1"};
  if (!rgb) throw Error{SVD_ERR_ARG, "rgb is NULL"};
  ViewArgs va{};
  va.mode = v->render_mode;
  va.normSamples = v->norm_samples;
  va.selected = v->selected_voxel;
  va.randomColors = v->random_colors;
  va.freqColors = v->freq_colors;
  va.levels = sc->levels;
  va.leafTexels = sc->nLeaves;          // RG32UI texels: one per 8-byte leaf
  va.innerTexels = sc->nInner / 4;      // RGBA16UI texels: four u16 each
  va.nOffsets = sc->nLevelOffsets;
  const uint32_t tw = c->traceTileW, th = kBT / tw;
  const dim3 grid(div_up(p->width, tw), div_up(p->height, th)), block(tw, th);
  const size_t shmem = stack_bytes(sc, d);
  MinDepth md{nullptr, 0, 0};
  float* depth = nullptr;
  if (beamProjf > 0.f) {
    depth = depthBuf ? depthBuf : dalloc<float>(c, (uint64_t)(p->width / 8) * (p->height / 8) + 1);
    md = depth_prepass(sc, d, p, beamProjf, depth);
  }
  switch (sc->fmt) {
    case SVD_FMT_SVDAG: k_render<SVD_FMT_SVDAG><<<grid, block, shmem, c->stream>>>(d, *p, va, rgb, prim, md); break;
    case SVD_FMT_USSVDAG:
      k_render<SVD_FMT_USSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, va, rgb, prim, md);
      break;
    default: k_render<SVD_FMT_SSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, va, rgb, prim, md); break;
  }
  check_launch("k_render");
  if (depth != depthBuf) dfree(c, depth);
}

// svd_render with the pre-pass texture in the static scratch array: no device
// allocation, no synchronisation.
void render_frame_noalloc(svd_scene* sc, const svd_trace_params* p, const svd_view_params* v, float beamProjf,
                          float* rgb, svd_ray_result* prim) {
  svd_ctx* c = sc->ctx;
  float* depth = nullptr;
  if (beamProjf > 0.f) {
    const uint64_t texels = (uint64_t)(p->width / 8) * (p->height / 8);
    if (texels > kRenderScratchTexels)
      throw Error{SVD_ERR_ARG, "frame too large for the no-alloc render: the depth pre-pass needs " +
                                   std::to_string(texels) + " texels, the scratch holds " +
                                   std::to_string(kRenderScratchTexels)};
    if (!c->renderScratch)
      throw Error{SVD_ERR_DEVICE, std::string("render scratch unavailable: hipGetSymbolAddress: ") +
                                      hipGetErrorString(c->renderScratchErr)};
    depth = c->renderScratch;
  }
  render_frame(sc, p, v, beamProjf, rgb, prim, depth);
}

namespace {
// The 256 hemisphere samples of OctreeDDARenderer::generateHSSamples
// (octree_dda_renderer.cpp:510-531): Halton points (bases 2 and 3,
// halton.cpp:24-40) mapped to cosine-weighted directions around +y.
const std::vector<float>& hemisphere_samples() {
  static const std::vector<float> hs = [] {
    std::vector<float> out;
    out.reserve(3 * kAOSamples);
    for (int i = 1; i <= kAOSamples; ++i) {
      float h[2];
      const int bases[2] = {2, 3};
      for (int b = 0; b < 2; ++b) {
        float r = 0.f, frac = 1.0f / (float)bases[b];
        int idx = i;
        while (idx > 0) {
          const int digit = idx % bases[b];
          r += frac * (float)digit;
          idx = (idx - digit) / bases[b];
          frac /= (float)bases[b];
        }
        h[b] = r;
      }
      const float y = std::sqrt(h[0]);
      const float l = std::sqrt(1.0f - y * y);
      const float ang = 2.0f * h[1] * 3.141592654f;
      // cos / sin in double, rounded once to float
      out.push_back(l * (float)std::cos((double)ang));
      out.push_back(y);
      out.push_back(l * (float)std::sin((double)ang));
    }
    return out;
  }();
  return hs;
}
}  // namespace

// AO_MODE (octree_dda.frag.glsl:884-919) over the G-buffer of the frame's
// primary rays: visibility^2 per pixel, -1 where the pixel has no hit.
void ao_frame(svd_scene* sc, const svd_trace_params* p, uint32_t nrays, float* vis, svd_ray_result* prim) {
  svd_ctx* c = sc->ctx;
  const Dev d = scene_dev(sc);
  if (nrays > (uint32_t)kAOSamples) throw Error{SVD_ERR_ARG, "num_rays must be <= 256"};
  if (p->width == 0 || p->height == 0) return;
  if (p->draw_level == 0) throw Error{SVD_ERR_ARG, "draw_level must be >= 1"};
  if (!vis) throw Error{SVD_ERR_ARG, "visibility is NULL"};
  const uint64_t npix = (uint64_t)p->width * p->height;
  const std::vector<float>& hs = hemisphere_samples();
  float* dhs = dalloc<float>(c, hs.size());
  SVD_HIP(hipMemcpyAsync(dhs, hs.data(), hs.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  uint32_t* cnt = dalloc<uint32_t>(c, npix);
  AOEntry* entries = dalloc<AOEntry>(c, npix);
  uint32_t* nEntries = dalloc<uint32_t>(c, 1);
  fill_u32(c, nEntries, 0u, 1);
  const uint32_t tw = c->traceTileW, th = kBT / tw;
  const dim3 grid(div_up(p->width, tw), div_up(p->height, th)), block(tw, th);
  const size_t shmem = stack_bytes(sc, d);
  // as many workgroups as run at once at six wavefronts per SIMD
  const uint32_t rayBlocks = (uint32_t)std::max(1, c->numCU > 0 ? c->numCU : 256) * 6u;
  switch (sc->fmt) {
    case SVD_FMT_SVDAG:
      k_ao_gbuffer<SVD_FMT_SVDAG><<<grid, block, shmem, c->stream>>>(d, *p, vis, cnt, entries, nEntries, prim);
      if (nrays) k_ao_rays<SVD_FMT_SVDAG><<<rayBlocks, kBT, shmem, c->stream>>>(d, *p, dhs, nrays, entries, nEntries, cnt);
      break;
    case SVD_FMT_USSVDAG:
      k_ao_gbuffer<SVD_FMT_USSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, vis, cnt, entries, nEntries, prim);
      if (nrays)
        k_ao_rays<SVD_FMT_USSVDAG><<<rayBlocks, kBT, shmem, c->stream>>>(d, *p, dhs, nrays, entries, nEntries, cnt);
      break;
    default:
      k_ao_gbuffer<SVD_FMT_SSVDAG><<<grid, block, shmem, c->stream>>>(d, *p, vis, cnt, entries, nEntries, prim);
      if (nrays)
        k_ao_rays<SVD_FMT_SSVDAG><<<rayBlocks, kBT, shmem, c->stream>>>(d, *p, dhs, nrays, entries, nEntries, cnt);
      break;
  }
  k_ao_finish<<<div_up(npix, 256), 256, 0, c->stream>>>(vis, cnt, npix, nrays);
  check_launch("ao passes");
  dfree(c, dhs); dfree(c, cnt); dfree(c, entries); dfree(c, nEntries);
}

// SVDAG scenes (nodes, nWords, levels, box) x cameras in one launch.
void trace_batch_svdag(svd_ctx* c, const std::vector<svd_scene>& scenes, const std::vector<svd_trace_params>& cams,
                       const std::vector<uint32_t>& camScene, svd_ray_result* prim, svd_ray_result* shad) {
  if (cams.empty()) return;
  const uint32_t W = cams[0].width, H = cams[0].height;
  int maxInner = 1;
  std::vector<Dev> devs(scenes.size());
  for (size_t i = 0; i < scenes.size(); ++i) {
    if (scenes[i].fmt != SVD_FMT_SVDAG) throw Error{SVD_ERR_ARG, "batched trace: SVDAG scenes only"};
    devs[i] = scene_dev(&scenes[i]);
    maxInner = devs[i].innerLevels > maxInner ? devs[i].innerLevels : maxInner;
  }
  for (size_t k = 0; k < cams.size(); ++k) {
    if (cams[k].width != W || cams[k].height != H) throw Error{SVD_ERR_ARG, "batched trace: one resolution for all cameras"};
    if (cams[k].draw_level == 0) throw Error{SVD_ERR_ARG, "draw_level must be >= 1"};
    if (camScene[k] >= scenes.size()) throw Error{SVD_ERR_ARG, "camera scene index out of range"};
  }
  if (W == 0 || H == 0) return;
  Dev* dDev = dalloc<Dev>(c, devs.size());
  svd_trace_params* dCam = dalloc<svd_trace_params>(c, cams.size());
  uint32_t* dMap = dalloc<uint32_t>(c, camScene.size());
  SVD_HIP(hipMemcpyAsync(dDev, devs.data(), devs.size() * sizeof(Dev), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(dCam, cams.data(), cams.size() * sizeof(svd_trace_params), hipMemcpyHostToDevice, c->stream));
  SVD_HIP(hipMemcpyAsync(dMap, camScene.data(), camScene.size() * 4, hipMemcpyHostToDevice, c->stream));
  const uint32_t tw = c->traceTileW, th = kBT / tw;
  const dim3 grid(div_up(W, tw), div_up(H, th), (uint32_t)cams.size());
  const dim3 block(tw, th);
  const size_t shmem = (size_t)kStackWords<SVD_FMT_SVDAG> * (maxInner + 1) * kBT * sizeof(uint32_t);
  k_trace_batch<SVD_FMT_SVDAG><<<grid, block, shmem, c->stream>>>(dDev, dCam, dMap, prim, shad);
  check_launch("k_trace_batch");
  SVD_HIP(hipStreamSynchronize(c->stream));
  dfree(c, dDev); dfree(c, dCam); dfree(c, dMap);
}

}  // namespace svd
