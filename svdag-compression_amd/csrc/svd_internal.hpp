// Internal declarations of libsvdag_amd (MI355X / gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <map>
#include "../../include/svdag_amd.h"

namespace svd {

constexpr uint32_t kNull = 0xFFFFFFFEu;    // octree.cpp:22 null child pointer
constexpr uint32_t kNone = 0xFFFFFFFFu;    // "no entry" marker in device tables
constexpr int kRecWords = 9;               // node record: mask|mX<<8|mY<<16|mZ<<24, children[8]
constexpr uint32_t kMaxLevels = 21;        // 3*21 = 63 Morton bits
// svd_render_noalloc: texels of its static depth pre-pass texture (512 KiB),
// enough for (width / 8) * (height / 8) of frames up to 4096 x 2048 or 3840 x 2160
constexpr uint64_t kRenderScratchTexels = 131072;

struct Error {
  int code;
  std::string msg;
};

// ---- device memory: a small caching allocator per context ----------------
class Arena {
 public:
  explicit Arena(int dev) : dev_(dev) { (void)dev_; }
  ~Arena();
  void* get(size_t bytes);                  // throws Error on failure
  void put(void* p);
  void trim();
 private:
  int dev_;
  std::multimap<size_t, void*> free_;        // size class -> block
  std::map<void*, size_t> live_;             // block -> size class
};

}  // namespace svd

struct svd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  svd::Arena* arena = nullptr;
  uint64_t* pinned = nullptr;                // small pinned staging for size read-backs
  hipEvent_t ev[2] = {nullptr, nullptr};
  // per timed family: events around its last call (the elapsed time is read
  // when asked for, so a call need not wait for its own kernels)
  std::map<std::string, std::pair<hipEvent_t, hipEvent_t>> timers;
  uint32_t traceTileW = 16;                  // traversal workgroup tile: traceTileW x (256 / traceTileW) pixels
  uint32_t traceMode = SVD_TRACE_TILED;      // SVD_TRACE_TILED or SVD_TRACE_PERSISTENT
  int numCU = 0;                             // compute units of the device
  bool sdagChecked = false;                  // toSDAG: check convergence after every round (SVD_TOSDAG_CHECKED=1)
  int refsortExtra = 4;                      // refcount order: rounds past the size bound before the first look
  // one-workgroup limits of the top levels (0: the compiled maximum; a limit
  // above it is cut to it): SVO child codes, toDAG / toSDAG nodes, SSVDAG
  // layout nodes (SVD_SMALL_LEVELS="svo,dag,sdag,layout").  The defaults were
  // the fastest of a sweep on the 2048^3 terrain build (1.654 ms against 1.679
  // ms for the compiled maxima 8192, 4096, 1024, 4096, interleaved runs).
  uint32_t smallSvo = 2048, smallDag = 1024, smallSdag = 512, smallLayout = 2048;
  std::map<uint64_t, int> occupancy;         // (kernel id, LDS bytes) -> resident workgroups per CU
  float* renderScratch = nullptr;            // device address of svd_render_noalloc's static texture
  hipError_t renderScratchErr = hipSuccess;  // why it is null, if it is
  unsigned long long* lbStatus = nullptr;    // single-pass scans: one published word per tile
  uint64_t lbCap = 0;                        // tiles lbStatus holds
  uint32_t* lbCounter = nullptr;             // tile counter, zero between passes
  uint32_t lbEpoch = 0;                      // epoch of the last pass (words of older passes are stale)
};

namespace svd {

// Level arrays on the device.  Inner levels hold 9-word records; the finest
// level (leaf 2^3 masks) holds one word per node (the mask) when leaf=true.
struct Level {
  uint32_t* rec = nullptr;        // n * (leaf ? 1 : 9) words
  uint8_t* inv = nullptr;         // invariant bits (SDAG view only)
  uint32_t* childLevels = nullptr;// 8 per node (cross-level view only)
  uint64_t n = 0;
  bool leaf = false;
};

struct Encoded {
  uint8_t* d = nullptr;           // full file image on the device
  uint64_t bytes = 0;
  bool valid = false;
};

}  // namespace svd

struct svd_octree {
  svd_ctx* ctx = nullptr;
  uint32_t levels = 0;
  int state = SVD_S_EMPTY;
  float bbox[6] = {0, 0, 0, 1, 1, 1};
  double rootSide = 1.0;
  uint64_t nVoxels = 0;
  uint64_t nNodesSVO = 0, nNodesDAG = 0, nNodesSDAG = 0, nNodesX = 0;
  uint64_t crossMerged = 0;
  bool hasX = false;              // cross-level view present
  std::vector<svd::Level> svo, dag, sdag, xdag;
  // toSDAG map of every DAG level: sdag id (bits 0..28) | mirror xyz << 29
  std::vector<uint32_t*> sdagMap;
  svd::Encoded enc[4];
};

// Learned candidate encoder: fp32 master parameters (flat, SVD_MLP_PARAMS),
// Adam moments, gradients and the bf16 weight copies the MFMA kernels read.
struct svd_mlp {
  svd_ctx* ctx = nullptr;
  float* params = nullptr;
  float* grads = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  uint16_t* wb = nullptr;          // bf16 W1 | W2 | W3 (same row-major layouts)
  uint64_t steps = 0;              // Adam steps taken
};

struct svd_scene {
  svd_ctx* ctx = nullptr;
  int fmt = 0;
  uint32_t levels = 0;
  float bmin[3], bmax[3], center[3], rootHalfSide;
  const uint32_t* nodes = nullptr;     // SVDAG / USSVDAG words
  const uint16_t* inner = nullptr;     // SSVDAG inner u16
  const uint8_t* leaves = nullptr;     // SSVDAG leaf bytes (8 per brick)
  uint32_t* levelOffsets = nullptr;    // device copy
  uint32_t nWords = 0, nInner = 0, nLeaves = 0;   // element counts (bounds checks)
  uint8_t* owned = nullptr;            // owned device image (when loaded from bytes)
  uint32_t nLevelOffsets = 0;
};

// Many scenes built together (batch.hip): the shared DAG levels and the
// SVDAG image of every scene in one device buffer.
struct svd_batch {
  svd_ctx* ctx = nullptr;
  uint32_t S = 0, Lmax = 0;
  std::vector<uint32_t> L;                     // depth of every scene
  std::vector<uint64_t> voxels, nodes;         // per scene
  std::vector<uint64_t> fileOff, fileBytes;    // SVDAG image of scene s at files + fileOff[s]
  std::vector<svd::Level> dag;                 // batched DAG levels (scene-major)
  uint8_t* files = nullptr;
  std::vector<svd_scene> scenes;               // traversal views of the images
};

namespace svd {

// ---- error helpers ----------------------------------------------------------
#define SVD_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::svd::Error{SVD_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error{SVD_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <typename T>
inline T* dalloc(svd_ctx* c, uint64_t count) {
  return (T*)c->arena->get(count * sizeof(T) + 16);
}
inline void dfree(svd_ctx* c, void* p) { if (p) c->arena->put(p); }

inline uint64_t small_limit(uint32_t want, uint32_t compiled) { return want && want < compiled ? want : compiled; }

// Read back a few u32/u64 device values into host memory (synchronises).
uint64_t read_u32(svd_ctx* c, const uint32_t* d);
void read_u32s(svd_ctx* c, const uint32_t* d, uint32_t* h, size_t n);
void read_u32s_2(svd_ctx* c, const uint32_t* d1, uint32_t* h1, size_t n1, const uint32_t* d2, uint32_t* h2,
                 size_t n2);

// ---- scan (scan.hip) ---------------------------------------------------------
// The look-back state of one single-pass scan over `tiles` tiles on the
// context's stream (lookback.hpp): a fresh epoch, the status array grown to
// fit.  Passes are ordered by the stream; a pass must not be replayed from a
// captured graph (its epoch would repeat).
struct LbPass;
LbPass lb_pass(svd_ctx* c, uint64_t tiles);
// Exclusive prefix sum of n u32 values; out may alias in.  *total (device) gets
// the sum (may be null).  Scratch comes from the context arena.
void scan_exclusive(svd_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total);
// in-place exclusive scans of two arrays of n elements (one launch when small)
void scan_exclusive2(svd_ctx* c, uint32_t* a, uint32_t* b, uint64_t n);
void fill_u32(svd_ctx* c, uint32_t* d, uint32_t v, uint64_t n);
void iota_u32(svd_ctx* c, uint32_t* d, uint64_t n);
void fill_u8(svd_ctx* c, uint8_t* d, uint8_t v, uint64_t n);

// ---- build / reduction (build.hip) -----------------------------------------
void build_svo_from_morton(svd_octree* o, const uint64_t* dcodes, uint64_t n);
// buildSVO of a triangle soup (voxelize.hip); dtris = 9 floats per triangle on the device
void build_svo_from_mesh(svd_octree* o, const float* dtris, uint64_t ntris, const double bmin[3], const double bmax[3]);
// buildSVOFromPoints of an in-memory point array (voxelize.hip); dxyz = 3 doubles per point on the device
void build_svo_from_points(svd_octree* o, const double* dxyz, uint64_t npts, const double bmin[3], const double bmax[3]);
// buildDAG's joined SVO: top of stepLevel+1 levels + one voxelized subtree per set voxel of its finest level
void build_svo_from_mesh_subtrees(svd_octree* o, const float* dtris, uint64_t ntris, const double bmin[3],
                                  const double bmax[3], uint32_t stepLevel);
// A batch's scenes in one code list (device arrays): scene s holds codes
// off[s] .. off[s+1]-1, each below 2^(3 * levels[s]); a build reads them
// tagged with s above bit kSceneTagShift.  tileScene is the build's own table.
constexpr int kSceneTagShift = 40;
struct SceneTags {
  const uint64_t* off;
  const uint32_t* levels;
  const uint32_t* tileScene;
  uint32_t S;
};
// SVO levels of n > 0 strictly ascending codes with no bit of hiMask set (or a
// batch's scenes, tagged as read, with tags); levelScenes (optional, with tags)
// receives every level's node scene ids.  Returns the non-root nodes.
uint64_t svo_levels_from_codes(svd_ctx* c, const uint64_t* dcodes, uint64_t n, uint32_t L, std::vector<Level>& svo,
                               std::vector<uint32_t*>* levelScenes, uint64_t hiMask,
                               const SceneTags* tags = nullptr);
// inner levels L-2 .. 0 from the finest level's parent codes (taken over and freed)
// and every level's node count (sizes, host)
uint64_t svo_inner_levels(svd_ctx* c, uint64_t* codeA, uint64_t nc, uint32_t L, std::vector<Level>& svo,
                          std::vector<uint32_t*>* levelScenes, const uint64_t* sizes);
// one inner level: exact-key dedup, first occurrence wins -> remap + compacted records
uint64_t dedup_level(svd_ctx* c, const uint32_t* rec, uint64_t n, uint32_t* remap, uint32_t** outRec);
void remap_children(svd_ctx* c, const uint32_t* in, uint64_t n, const uint32_t* remap, uint32_t* out);
void reduce_dag(svd_octree* o);
uint64_t dedup_all(svd_ctx* c, uint32_t L, const std::vector<Level>& src, std::vector<Level>& dst);
void reduce_sym(svd_octree* o);
// toSDAG(..., skipSymmetry = true): every non-empty node its own representative, no mirror bits
void reduce_sym_pointers(svd_octree* o);
void cross_level_merge(svd_octree* o);

// ---- refcount order: exact emulation of libstdc++ std::sort (refsort.hip) --
// Sorts (idx, val) pairs in place exactly like
// std::sort(first, last, [](a, b) { return a.second > b.second; })
// (encoded_ssvdag.cpp:273-275).
void refsort(svd_ctx* c, uint32_t* idx, uint32_t* val, uint64_t n);
// Several independent std::sort calls at once: range r is [rfirst[r], rlast[r])
// of the same arrays (n elements in all).  The ranges share every round.
// unfinished (optional device word): run the device-wide rounds the sizes
// call for without reading the list size back, and leave the count of
// segments still too big for a workgroup there (nonzero: the order is not
// finished and the caller must sort again without it).
void refsort_ranges(svd_ctx* c, uint32_t* idx, uint32_t* val, uint64_t n, const uint64_t* rfirst,
                    const uint64_t* rlast, size_t nr, uint32_t* unfinished = nullptr);

// ---- attribute bit-tree (attr.hip) ---------------------------------------------
uint64_t attr_bittree(svd_ctx* c, const uint8_t* dattrs, uint64_t V, int gray, uint8_t* dst, uint64_t cap);

// ---- encoders (encode.hip) -----------------------------------------------------
void encode(svd_octree* o, int fmt);
// several formats (each at most once); .svdag/.ussvdag share one size read,
// .ssvdag/.esvdag share every pass (one segmented refcount order for both)
void encode_formats(svd_octree* o, const int* fmts, int n);
// .svdag file image -> DAG-state octree (decode.hip)
void octree_from_svdag(svd_octree* o, const uint8_t* file, uint64_t bytes, bool onDevice);

// ---- traversal (trace.hip) -----------------------------------------------------
void trace_frame(svd_scene* s, const svd_trace_params* p, svd_ray_result* prim, svd_ray_result* shad);
void trace_frame_hits(svd_scene* s, const svd_trace_params* p, float* prim, float* shad);
// beam optimisation: depth pre-pass at 1/8 resolution, then the frame from getMinT(8)
void trace_frame_beam(svd_scene* s, const svd_trace_params* p, float depthProjf, svd_ray_result* prim,
                      svd_ray_result* shad, float* minDepthOut);
// viewer colour per pixel (beamProjf > 0: after the depth pre-pass) and the AO pass
// depthBuf: the pre-pass texture ((width/8)*(height/8) floats), or null to take it from the arena
void render_frame(svd_scene* s, const svd_trace_params* p, const svd_view_params* v, float beamProjf, float* rgb,
                  svd_ray_result* prim, float* depthBuf = nullptr);
// the same frame with the pre-pass texture in a static device array (no allocation)
void render_frame_noalloc(svd_scene* s, const svd_trace_params* p, const svd_view_params* v, float beamProjf,
                          float* rgb, svd_ray_result* prim);
// once per device, with it current: resolve the static texture's device address into the context
void render_scratch_init(svd_ctx* c);
void ao_frame(svd_scene* s, const svd_trace_params* p, uint32_t nrays, float* vis, svd_ray_result* prim);

// ---- lossy DAG (lossy.hip), learned encoder (mlp.hip), kNN (knn.hip) ----------
void sort_u64(svd_ctx* c, uint64_t* keys, uint64_t n);
void lossy_signatures(svd_octree* o, uint32_t lev, uint32_t* sig);
void lossy_merge(svd_octree* o, const svd_lossy_params* p, std::vector<svd_lossy_level_stats>& out);
void mlp_embed(svd_mlp* m, const uint32_t* sig, uint64_t n, float* emb);
void mlp_alloc(svd_mlp* m);
void mlp_release(svd_mlp* m);
void mlp_init(svd_mlp* m, uint64_t seed);
void mlp_set_params(svd_mlp* m, const float* host);
void mlp_get_params(svd_mlp* m, float* host);
void mlp_loss_grad(svd_mlp* m, const uint32_t* X, const uint32_t* labels, uint64_t P, float* lossHost, float* gradsHost);
void mlp_train(svd_mlp* m, svd_octree* o, const svd_mlp_train_params* p, svd_mlp_train_stats* st);
// sig != null: candidates scored by the Hamming distance of the signatures
void knn_lsh(svd_ctx* c, const float* emb, uint64_t n, uint32_t k, uint32_t tables, uint32_t bits, uint32_t window,
             uint64_t seed, uint32_t* out, const uint32_t* sig = nullptr);
// neighbour-of-neighbour rounds on the signature Hamming metric (in place)
void knn_refine(svd_ctx* c, const uint32_t* sig, uint64_t n, uint32_t k, uint32_t rounds, uint32_t* knn);

// ---- batches (batch.hip, trace.hip) ---------------------------------------------
void batch_build(svd_batch* b, const uint64_t* dcodes, uint64_t n, const std::vector<uint64_t>& off);
void trace_batch_svdag(svd_ctx* c, const std::vector<svd_scene>& scenes, const std::vector<svd_trace_params>& cams,
                       const std::vector<uint32_t>& camScene, svd_ray_result* prim, svd_ray_result* shad);

// ---- synthetic scenes (workload.hip) ---------------------------------------------
uint64_t gen_caves(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double threshold, double baseCells,
                   uint64_t** out);
uint64_t gen_balls(svd_ctx* c, uint32_t levels, const double* balls, uint32_t nBalls, uint64_t** out);
void gen_materials(svd_ctx* c, const uint64_t* codes, uint64_t n, uint32_t levels, uint64_t seed, uint32_t nValues,
                   double baseCells, uint8_t* out);
uint64_t codes_union(svd_ctx* c, const uint64_t* a, uint64_t na, const uint64_t* b, uint64_t nb, uint64_t** out);
uint64_t gen_menger(svd_ctx* c, uint32_t levels, uint32_t order, uint64_t** out);
uint64_t gen_terrain(svd_ctx* c, uint32_t levels, uint64_t seed, uint32_t octaves, double hmin, double hmax,
                     double baseCells, uint64_t** out);

// ---- timing helpers ------------------------------------------------------------
// Events around one API call on the context stream.  wait: return only once
// the call's kernels are done (false for calls whose results stay in the
// library until another call, itself on the stream, reads them).
struct Timer {
  svd_ctx* c;
  hipEvent_t end = nullptr;
  bool wait;
  Timer(svd_ctx* ctx, const char* n, bool w = true) : c(ctx), wait(w) {
    auto& ev = c->timers[n];
    if (!ev.first && (hipEventCreate(&ev.first) != hipSuccess || hipEventCreate(&ev.second) != hipSuccess)) {
      (void)hipGetLastError();
      return;
    }
    end = ev.second;
    (void)hipEventRecord(ev.first, c->stream);
  }
  ~Timer() {
    if (!end) return;
    (void)hipEventRecord(end, c->stream);
    if (wait) (void)hipEventSynchronize(end);
  }
};
// Elapsed ms of the family's last timed call (waits for it); < 0 if none.
double timer_ms(svd_ctx* c, const std::string& name);

// ---- device helpers --------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ull;
  z ^= z >> 27; z *= 0x94d049bb133111ebull;
  z ^= z >> 31;
  return z;
}

// Hash of a 9-word record for the device tables: two 32-bit lanes mixed with
// 32-bit multiplies (murmur3-style word steps and finalisers; a 64-bit
// multiply costs four 32-bit ones here), the low word indexing the table,
// the high word the fingerprint.  Only equality of records decides what a
// table returns, so the hash affects probe lengths, never results.
__device__ __forceinline__ uint64_t hash_rec(const uint32_t w[kRecWords]) {
  uint32_t a = 0x9E3779B9u, b = 0x7F4A7C15u;
#pragma unroll
  for (int i = 0; i < kRecWords; ++i) {
    a = (a ^ w[i]) * 0xCC9E2D51u;
    a = (a << 15) | (a >> 17);
    b = (b + w[i]) * 0x1B873593u;
    b ^= b >> 13;
  }
  a ^= a >> 16; a *= 0x85EBCA6Bu; a ^= a >> 13; a *= 0xC2B2AE35u; a ^= a >> 16;
  b ^= b >> 16; b *= 0x7FEB352Du; b ^= b >> 15; b *= 0x846CA68Bu; b ^= b >> 16;
  return ((uint64_t)b << 32) | a;
}

}  // namespace svd
