/*
 * svdag_amd.h -- C ABI of the MI355X sparse-voxel-DAG engine (libsvdag_amd.so).
 *
 * Plain C types only: opaque handles, pointers and sizes.  Every entry point
 * returns an int status (SVD_OK == 0, negative codes below) and records a
 * message retrievable with svd_last_error().  All work is enqueued on the
 * context's HIP stream; entry points that hand results back to the host
 * synchronise that stream before returning.
 *
 * Each declaration names the reference interface it replaces
 * (RvanderLaan/SVDAG-Compression, paths relative to the repository root).
 */
#ifndef SVDAG_AMD_H
#define SVDAG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SVD_OK            0
#define SVD_ERR_ARG      -1   /* bad argument / null handle                  */
#define SVD_ERR_LEVELS   -2   /* level count outside the supported range     */
#define SVD_ERR_STATE    -3   /* octree is not in the state the call needs   */
#define SVD_ERR_FORMAT   -4   /* malformed encoded buffer                    */
#define SVD_ERR_POINTER  -5   /* a child pointer has no target               */
#define SVD_ERR_CAPACITY -6   /* a level exceeds the encoding's pointer range*/
#define SVD_ERR_DEVICE   -8   /* HIP runtime error                           */
#define SVD_ERR_NOMEM    -9   /* device allocation failed                    */

/* ---- octree states (geom_octree.hpp:35-40) ------------------------------ */
#define SVD_S_EMPTY 0
#define SVD_S_SVO   1
#define SVD_S_DAG   2
#define SVD_S_SDAG  3

/* ---- encoded formats ---------------------------------------------------- */
#define SVD_FMT_SVDAG   0   /* .svdag   EncodedSVDAG   (encoded_svdag.cpp:76-199)    */
#define SVD_FMT_USSVDAG 1   /* .ussvdag EncodedUSSVDAG (encoded_ussvdag.cpp:60-170)  */
#define SVD_FMT_SSVDAG  2   /* .ssvdag  EncodedSSVDAG on the SDAG (encoded_ssvdag.cpp:84-466) */
#define SVD_FMT_ESVDAG  3   /* .esvdag  EncodedSSVDAG on the DAG  (svbuilder/main.cpp:243-244) */

/* ---- level views --------------------------------------------------------- */
#define SVD_VIEW_SVO  0     /* levels as built (before toDAG)                 */
#define SVD_VIEW_DAG  1     /* after toDAG                                    */
#define SVD_VIEW_SDAG 2     /* after toSDAG                                   */

typedef struct svd_ctx svd_ctx;
typedef struct svd_octree svd_octree;
typedef struct svd_scene svd_scene;

/* Context: one device, one stream.  Replaces nothing in the reference (the
 * reference is single-threaded host code); it owns device memory pools.   */
int svd_ctx_create(int device, svd_ctx** out);
int svd_ctx_destroy(svd_ctx* ctx);
const char* svd_last_error(const svd_ctx* ctx);
int svd_ctx_synchronize(svd_ctx* ctx);
/* The HIP stream all work of this context is enqueued on (hipStream_t). */
void* svd_ctx_stream(svd_ctx* ctx);

/* ---- construction --------------------------------------------------------
 * svd_octree_from_morton: SVO from a voxel set given as 3-bit interleaved
 * Morton codes (code & 7 == child slot 4x+2y+z of the finest level), sorted
 * ascending and unique.  Node order per level = order of first arrival of
 * the voxels, i.e. the point-major DFS of GeomOctree::buildSVOFromPoints
 * (geom_octree.cpp:62-169).  codes may live on the device
 * (codes_on_device = 1, e.g. a torch tensor) or on the host.  Scene box =
 * unit cube, rootSide = 1.                                                   */
int svd_octree_from_morton(svd_ctx* ctx, const uint64_t* codes, uint64_t n, uint32_t levels,
                           int codes_on_device, svd_octree** out);

/* svd_octree_from_svo: SVO given level by level on the host, as produced by
 * GeomOctree::buildSVO (geom_octree.cpp:171-280).  words[lev] holds
 * 9*level_sizes[lev] uint32: word 0 = childrenBitmask (bits 0-7), words
 * 1..8 = children[0..7] (0xFFFFFFFE = null).  bbox = min xyz, max xyz (f32),
 * root_side = GeomOctree::_rootSide.                                         */
int svd_octree_from_svo(svd_ctx* ctx, uint32_t levels, const uint64_t* level_sizes,
                        const uint32_t* const* words, const float bbox[6], double root_side,
                        svd_octree** out);
int svd_octree_free(svd_octree* oct);

/* ---- reductions -------------------------------------------------------- */
/* GeomOctree::toDAG (geom_octree.cpp:462-548). */
int svd_reduce_dag(svd_octree* oct);
/* GeomOctree::toSDAG(false,false) (geom_octree.cpp:551-697).  Needs S_DAG. */
int svd_reduce_sym(svd_octree* oct);
/* GeomOctree::mergeAcrossAllLevels (geom_octree_extension.cpp:1192-1544).
 * Needs S_DAG; writes the cross-level view used by SVD_FMT_SVDAG.          */
int svd_cross_level_merge(svd_octree* oct, uint64_t* nodes_eliminated);

/* ---- queries ------------------------------------------------------------- */
int svd_octree_state(svd_octree* oct, int* state);
int svd_octree_levels(svd_octree* oct, uint32_t* levels);
/* GeomOctree::getNNodes / getNVoxels (octree.hpp:101-102) for the current state. */
int svd_octree_counts(svd_octree* oct, uint64_t* n_nodes, uint64_t* n_voxels);
int svd_level_size(svd_octree* oct, int view, uint32_t lev, uint64_t* n);
/* GeomOctree::getNodeData (geom_octree.hpp:113) for one level, copied to the
 * host: 9 uint32 per node (word 0 = mask | mX<<8 | mY<<16 | mZ<<24, words
 * 1..8 = children), inv = invariant bits (may be NULL), child_levels = 8 per
 * node (may be NULL; only the cross-level view differs from lev+1).          */
int svd_get_level(svd_octree* oct, int view, uint32_t lev, uint32_t* words, uint8_t* inv,
                  uint32_t* child_levels);

/* ---- encoders -------------------------------------------------------------
 * EncodedSVDAG / EncodedUSSVDAG / EncodedSSVDAG ::encode + ::save.  The
 * encoded file is built in device memory; svd_encode returns its size in
 * bytes (exactly what save() writes), svd_encoded_copy copies it to the host
 * and svd_encoded_device exposes the device buffer.                          */
int svd_encode(svd_octree* oct, int fmt, uint64_t* file_bytes);
int svd_encoded_copy(svd_octree* oct, int fmt, void* dst, uint64_t cap);
int svd_encoded_device(svd_octree* oct, int fmt, const void** dptr, uint64_t* file_bytes);

/* ---- traversal (octree_dda.frag.glsl) --------------------------------------
 * A scene is an encoded file resident on the device (EncodedOctree::load +
 * OctreeDDARenderer::uploadData, octree_dda_renderer.cpp:81-214).           */
typedef struct svd_trace_params {
  float view_inv[16];          /* column-major (GL) inverse view matrix      */
  float proj_inv[16];          /* column-major inverse projection matrix     */
  uint32_t width, height;      /* screenRes                                  */
  uint32_t max_iters;          /* maxIters (octree_dda_renderer.cpp:66)      */
  uint32_t draw_level;         /* drawLevel                                  */
  float projection_factor;     /* octree_dda_renderer.cpp:503-507            */
  float light_pos[3];          /* lightPos (octree_dda_renderer.cpp:47)      */
  uint32_t shadows;            /* 1: one shadow ray per primary hit (:803-820) */
  uint32_t pad;
} svd_trace_params;

typedef struct svd_ray_result {  /* trace_ray's vec4 (octree_dda.frag.glsl:528-537) */
  float t;                     /* >= 0 hit distance; -1 miss, -2 t out, -3 iters, -4 box */
  float level;
  float iters;
  float node;
  uint32_t path;               /* FNV-1a hash of the (node, level) descent path */
  uint32_t bytes;              /* algorithmic bytes fetched by this ray       */
} svd_ray_result;

/* Load an encoded file image (host or device bytes) as a traversal scene.   */
int svd_scene_create(svd_ctx* ctx, int fmt, const void* file, uint64_t bytes, int file_on_device,
                     svd_scene** out);
/* Scene straight from an octree's encoded buffer (no host round trip).      */
int svd_scene_from_octree(svd_octree* oct, int fmt, svd_scene** out);
int svd_scene_free(svd_scene* sc);
/* One frame: width*height primary rays (+ shadow rays).  primary / shadow are
 * DEVICE arrays of width*height svd_ray_result; shadow may be NULL.          */
int svd_trace(svd_scene* sc, const svd_trace_params* p, svd_ray_result* primary, svd_ray_result* shadow);

/* ---- Gray-code attribute bit-tree --------------------------------------------
 * The reference lists the feature (readme.md:8) but has no code for it; the
 * encoding is the one documented with this library: per-voxel u8 attributes
 * in voxel Morton (DFS) order, coded Gray (gray=1, b^(b>>1)) or binary, one
 * pruned binary tree per bit plane emitted in pre-order.  Blob: u32 magic
 * 'SVAB', u32 n, u32 gray, u32 8, u32 bits[8], then the planes' u32 words.
 * attrs on the device (attrs_on_device=1) or the host; out is a HOST buffer.
 * Returns the blob size in *bytes; copies it when cap is large enough.      */
int svd_attr_bittree_encode(svd_ctx* ctx, const uint8_t* attrs, uint64_t n, int gray, int attrs_on_device,
                            void* out, uint64_t cap, uint64_t* bytes);

/* ---- reference-count order ------------------------------------------------
 * The permutation std::sort(hist, [](a,b){ return a.second > b.second; })
 * produces in EncodedSSVDAG::encode (encoded_ssvdag.cpp:273-275), computed on
 * the device.  idx / val are HOST arrays of n pairs, sorted in place.        */
int svd_refcount_order(svd_ctx* ctx, uint32_t* idx, uint32_t* val, uint64_t n);

/* ---- instrumentation ------------------------------------------------------ */
/* Average duration in milliseconds of the last launch of the named kernel
 * family, measured with HIP events on the context stream ("trace", "build",
 * "encode").  Returns SVD_ERR_ARG for an unknown name.                      */
int svd_last_kernel_ms(svd_ctx* ctx, const char* name, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* SVDAG_AMD_H */
