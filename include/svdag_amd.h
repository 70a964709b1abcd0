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
 * reference is single-threaded host code); it owns device memory pools.
 * Environment read at creation (tuning only; results never change):
 *   SVD_SMALL_LEVELS="svo,dag,sdag,layout"  the level sizes up to which the
 *       top levels of a pass run in one workgroup (SVO child codes, toDAG and
 *       toSDAG nodes, SSVDAG layout nodes; default 2048,1024,512,2048, each
 *       capped at its compiled maximum 8192,4096,1024,4096; 0 = the maximum);
 *   SVD_REFSORT_EXTRA_ROUNDS=k  device-wide refcount-order rounds run past the
 *       size bound before the list size is first read back (default 4, 0..16);
 *   SVD_TOSDAG_CHECKED=1  toSDAG reads its undecided count back after every
 *       resolution round instead of running two rounds speculatively.        */
int svd_ctx_create(int device, svd_ctx** out);
int svd_ctx_destroy(svd_ctx* ctx);
const char* svd_last_error(const svd_ctx* ctx);
int svd_ctx_synchronize(svd_ctx* ctx);
/* The HIP stream all work of this context is enqueued on (hipStream_t).
 * svd_octree_from_morton with device codes, svd_reduce_dag, svd_reduce_sym
 * and svd_encode_formats may return while their last kernels still run on
 * it (they have read back every size they need; the codes have been read).
 * Every call that returns data (counts, levels, sizes, copies, device
 * pointers) is ordered after them; work of the caller's own on other streams
 * is not: svd_ctx_synchronize first.                                        */
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

/* svd_octree_from_svdag: an .svdag file image (host or device bytes) decoded
 * into a DAG-state octree, as EncodedSVDAG::load + EncodedSVDAG::decode do
 * for svbuilder's .svdag input (encoded_svdag.cpp:43-66, 200-281;
 * svbuilder/main.cpp:97-101).  Node order per level = file order.  Like the
 * reference the octree reports the header's node count and 0 voxels (load
 * does not read a voxel count).  Files whose levels are not tiled exactly by
 * the referenced nodes (unreferenced or overlapping words, pointers into
 * other levels as in -multi files) fail with SVD_ERR_FORMAT.               */
int svd_octree_from_svdag(svd_ctx* ctx, const void* file, uint64_t bytes, int file_on_device, svd_octree** out);

/* svd_octree_from_mesh: GeomOctree::buildSVO (geom_octree.cpp:171-280) +
 * cleanEmptyNodes (:437-456) on the device, for a triangle soup of ntris
 * triangles (9 floats each: v0 xyz, v1 xyz, v2 xyz, Scene::getTrianglePtr
 * order) inside bbox = min xyz, max xyz (double, svbuilder's sceneBBoxD).
 * The triangle/box overlap test is test_triangle_box.cpp:105-184 in FP64 and
 * the node order per level is the reference's creation order, so the SVO
 * equals the reference's node for node.  tris on the host (on_device = 0) or
 * on the device.                                                             */
int svd_octree_from_mesh(svd_ctx* ctx, const float* tris, uint64_t ntris, uint32_t levels, const double bbox[6],
                         int on_device, svd_octree** out);
/* svd_octree_from_points: GeomOctree::buildSVOFromPoints
 * (geom_octree.cpp:62-169) on the device for npts points (x y z doubles each,
 * in the order the reference would read them from its LAS file) inside
 * bbox = min xyz, max xyz (double).  Every point walks the reference's DFS:
 * child centres are the parent's plus or minus getHalfSideD in FP64 and a
 * point is in a child box when min <= p < max on all three axes, so points on
 * rounding-overlapped boundaries reach every box that holds them, as there.
 * Node order per level is the order of creation; cleanEmptyNodes follows.
 * The reference reads the points with libLAS; the point array replaces the
 * file.  xyz on the host (on_device = 0) or on the device.                   */
int svd_octree_from_points(svd_ctx* ctx, const double* xyz, uint64_t npts, uint32_t levels, const double bbox[6],
                           int on_device, svd_octree** out);
/* svd_octree_from_mesh_subtrees: the SVO that GeomOctree::buildDAG
 * (geom_octree.cpp:289-435) reduces: a top SVO of step_level + 1 levels, then
 * one subtree per set voxel of its finest level, each voxelized from all
 * triangles in its own box, joined under the top in the reference's order.
 * The result is in the SVO state (svd_reduce_dag gives buildDAG's DAG); the
 * node count is the reference's (subtree roots not counted).  1 <= step_level
 * and step_level + 1 < levels (SVD_ERR_ARG otherwise: with step 0 the
 * reference's top has no finest-node centres, at levels - 1 no subtree level
 * is left; svbuilder only calls buildDAG for steps > 0).                    */
int svd_octree_from_mesh_subtrees(svd_ctx* ctx, const float* tris, uint64_t ntris, uint32_t levels,
                                  uint32_t step_level, const double bbox[6], int on_device, svd_octree** out);

/* svd_octree_from_svo: SVO given level by level on the host, as produced by
 * GeomOctree::buildSVO (geom_octree.cpp:171-280).  words[lev] holds
 * 9*level_sizes[lev] uint32: word 0 = childrenBitmask (bits 0-7), words
 * 1..8 = children[0..7] (0xFFFFFFFE = null).  bbox = min xyz, max xyz (f32),
 * root_side = GeomOctree::_rootSide.                                         */
int svd_octree_from_svo(svd_ctx* ctx, uint32_t levels, const uint64_t* level_sizes,
                        const uint32_t* const* words, const float bbox[6], double root_side,
                        svd_octree** out);
/* svd_octree_from_levels: GeomOctree(NodeData, bbox, rootSide, levels, Stats)
 * (geom_octree.cpp:49-60): host level arrays in the layout of
 * svd_octree_from_svo become a DAG-state octree (no reduction is run; the
 * levels are taken as they are).  Like the reference constructor, the octree
 * reports the node and voxel counts it is given (stats.nNodesDAG,
 * stats.nTotalVoxels).                                                      */
int svd_octree_from_levels(svd_ctx* ctx, uint32_t levels, const uint64_t* level_sizes,
                           const uint32_t* const* words, const float bbox[6], double root_side, uint64_t n_nodes,
                           uint64_t n_voxels, svd_octree** out);
/* svd_octree_copy: the GeomOctree copy constructor (geom_octree.cpp:41-46):
 * a deep copy on the same context of every level view, the toSDAG maps and
 * the encoded images.                                                        */
int svd_octree_copy(const svd_octree* src, svd_octree** out);
int svd_octree_free(svd_octree* oct);

/* ---- reductions -------------------------------------------------------- */
/* GeomOctree::toDAG (geom_octree.cpp:462-548). */
int svd_reduce_dag(svd_octree* oct);
/* GeomOctree::toSDAG(false,false) (geom_octree.cpp:551-697).  Needs S_DAG. */
int svd_reduce_sym(svd_octree* oct);
/* GeomOctree::toSDAG(internalCall, skipSymmetry = true) (geom_octree.cpp:
 * 591-596): pointer-only compression.  Every node with children becomes its
 * own representative with no mirror bits and no invariance bits, nodes
 * without children are dropped (parents remapped), the root is kept.  Needs
 * S_SVO or S_DAG; ends in S_SDAG with nNodes = sum of levels 1..L-1.       */
int svd_reduce_sym_pointers(svd_octree* oct);
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
/* svd_encode_formats: svd_encode of n distinct formats in one call (svbuilder
 * encodes all four, main.cpp:220-271), with the same images.  .ssvdag and
 * .esvdag share every pass: one reference-count launch, one segmented
 * refcount order over the levels of both (each level still its own std::sort
 * call), one layout read; .svdag and .ussvdag share one size read, the same
 * read when both kinds are asked for.
 * file_bytes (may be NULL) receives n sizes in the order of fmts.          */
int svd_encode_formats(svd_octree* oct, const int* fmts, uint32_t n, uint64_t* file_bytes);
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
/* svd_trace_hits: the same frame without the path/bytes instrumentation:
 * per ray trace_ray's vec4 alone (octree_dda.frag.glsl:528-537), 4 floats
 * (t, level, iters, node), the values svd_trace gives.  primary: DEVICE array
 * of width*height vec4.  shadow (DEVICE, optional, width*height vec4) is
 * written only for the pixels that cast a shadow ray (p->shadows set and the
 * primary t >= 0); its other entries are left as they are.  Always uses the
 * tiled launch.  No reference counterpart as an entry point: it is the
 * traversal of octree_dda.frag.glsl as the viewer runs it, whose results the
 * reference keeps on the GPU.                                                */
int svd_trace_hits(svd_scene* sc, const svd_trace_params* p, float* primary, float* shadow);
/* svd_trace_beam: the same frame with the viewer's minimum-depth optimisation
 * (OctreeDDARenderer::draw with _useMinDepthOptimization, its default,
 * octree_dda_renderer.cpp:389-408): a DEPTH_MODE pre-pass over a
 * (width/8) x (height/8) texture with its own projection factor
 * (getProjectionFactor(1, 8), :503-507; screenRes = width/8, height/8 as
 * floats) stores the hit distance of each coarse ray (0 on a miss); every
 * primary ray then starts at getMinT(8) - 1e-3 (octree_dda.frag.glsl:607-617,
 * 681).  Texels outside the texture read as 0 (GL leaves them undefined).
 * min_depth (optional DEVICE array of (width/8)*(height/8) floats) receives
 * the texture.  Always uses the tiled launch.                                */
int svd_trace_beam(svd_scene* sc, const svd_trace_params* p, float depth_projection_factor, svd_ray_result* primary,
                   svd_ray_result* shadow, float* min_depth);

/* ---- viewer shading ---------------------------------------------------------
 * The uniforms of the viewer program (octree_dda.frag.glsl:52-60, 81-83;
 * defaults octree_dda_renderer.cpp:41-45, 73).                               */
typedef struct svd_view_params {
  int32_t render_mode;         /* viewerRenderMode: 0 iterations (default), 1 depth, 2 voxel levels, 3 pretty */
  int32_t norm_samples;        /* normSamples: extra rays that smooth the normal in mode 3 (0..64; default 0) */
  uint32_t selected_voxel;     /* selectedVoxelIndex, drawn blue; 0xFFFFFFFF = none (the default -1) */
  uint32_t random_colors;      /* randomColors (default 1) */
  uint32_t freq_colors;        /* freqColors (default 1); SSVDAG / ESVDAG scenes only */
  uint32_t pad;
} svd_view_params;
/* svd_render: the colour VIEWER_MODE main() writes per pixel
 * (octree_dda.frag.glsl:671-858): the primary ray, the render mode's
 * intensity, the pretty mode's normal samples and diffuse / fall-off terms,
 * highlight, frequency or random colours, and the shadow ray when
 * p->shadows.  beam_projection_factor > 0 runs the minimum-depth pre-pass
 * first, as svd_trace_beam does (<= 0: off).  rgb: DEVICE array of
 * 3 * width * height floats (unclamped; a discarded fragment gets 0);
 * primary (optional, DEVICE) receives the primary rays.
 *  - GLSL leaves sin, cos, log2 and pow to the implementation; this build
 *    defines them (Cody-Waite reduction and Taylor polynomials, the atanh
 *    series, pow(x, 2) = x * x) so results are reproducible bit for bit.
 *  - freqColors reads the SSVDAG texture sizes as TBOs (textureSize(leafNodes)
 *    = leaf count, textureSize(innerNodes) = u16 count / 4).  The reference's
 *    USSVDAG viewer program does not compile with that branch (it names
 *    SSVDAG-only uniforms); here USSVDAG and SVDAG scenes fall through to
 *    randomColors.                                                          */
int svd_render(svd_scene* sc, const svd_trace_params* p, const svd_view_params* v, float beam_projection_factor,
               float* rgb, svd_ray_result* primary);
/* svd_render_noalloc: svd_render for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as svd_render, bit for bit; the minimum-depth pre-pass texture
 * lives in a static device array of 131072 floats (512 KiB) that every GPU
 * the library runs on holds until the process ends, so the call makes no
 * hipMalloc / hipFree and no allocating library call.  With
 * beam_projection_factor > 0, (width / 8) * (height / 8) must be at most
 * 131072 (3840 x 2160 fits), else SVD_ERR_ARG.  Work is enqueued on the
 * context stream and the call returns without synchronising (and without
 * recording the "trace" kernel time).  Every call on one GPU shares the
 * array: calls that overlap on one GPU (contexts on different streams) race. */
int svd_render_noalloc(svd_scene* sc, const svd_trace_params* p, const svd_view_params* v,
                       float beam_projection_factor, float* rgb, svd_ray_result* primary);
/* svd_trace_ao: AO_MODE main() (octree_dda.frag.glsl:884-919) with num_rays
 * (<= 256; the renderer's default is 16) hemisphere rays of length 0.5 per
 * pixel, over the 256 cosine-weighted Halton samples of
 * OctreeDDARenderer::generateHSSamples (octree_dda_renderer.cpp:510-531).
 * The reference's AO program reads a G-buffer that nothing in the reference
 * writes (and svviewer never selects the mode); here it is defined from each
 * pixel's primary ray: hit position, -stepDir as the normal, cell size
 * 2 rootHalfSide / 2^level.  visibility: DEVICE array of width * height
 * floats = (1 - hits / num_rays)^2, -1 where the primary ray has no hit.    */
int svd_trace_ao(svd_scene* sc, const svd_trace_params* p, uint32_t num_rays, float* visibility,
                 svd_ray_result* primary);

/* ---- autotuning (no reference counterpart) ----------------------------------
 * Launch shape of the traversal; results never depend on it.
 *  - SVD_TRACE_TILED (default): one workgroup per tile of tile_w x (256 / tile_w) pixels.
 *  - SVD_TRACE_PERSISTENT: as many workgroups as the device holds;
 *    a lane that finishes a ray takes the next one from a device-wide queue
 *    of rays ordered by the same tiles.
 * tile_w in {8, 16, 32, 64} (default 16).
 * svd_autotune_trace times every tile shape in the current mode on one frame
 * (1 warm-up + `reps` timed launches each, HIP events), keeps the fastest in
 * the context and reports it; ms[4] (may be NULL) receives the mean kernel
 * time per shape in the order 8, 16, 32, 64.
 * svd_autotune_trace_modes does the same over both modes: ms[8] (may be NULL)
 * = tiled 8..64, then persistent 8..64.                                       */
#define SVD_TRACE_TILED       0
#define SVD_TRACE_PERSISTENT  1
int svd_set_trace_tile(svd_ctx* ctx, uint32_t tile_w);
int svd_get_trace_tile(svd_ctx* ctx, uint32_t* tile_w);
int svd_set_trace_mode(svd_ctx* ctx, uint32_t mode);
int svd_get_trace_mode(svd_ctx* ctx, uint32_t* mode);
int svd_autotune_trace(svd_scene* sc, const svd_trace_params* p, svd_ray_result* primary, svd_ray_result* shadow,
                       uint32_t reps, uint32_t* best_tile_w, double* ms);
int svd_autotune_trace_modes(svd_scene* sc, const svd_trace_params* p, svd_ray_result* primary,
                             svd_ray_result* shadow, uint32_t reps, uint32_t* best_tile_w, uint32_t* best_mode,
                             double* ms);

/* One-workgroup level limits of the lossless build (results never depend on
 * them): limits[4] = SVO child codes, toDAG nodes, toSDAG nodes, SSVDAG layout
 * nodes up to which the top levels of a pass run in one workgroup (0 = the
 * compiled maximum 8192, 4096, 1024, 4096; larger values are cut to it).
 * svd_autotune_build times the lossless build of `levels`-level device codes
 * (svd_octree_from_morton, svd_reduce_dag, svd_reduce_sym, then the .svdag,
 * .ussvdag, .esvdag and .ssvdag images; the first two for levels < 3) under
 * SVD_BUILD_PRESETS presets of the limits (1 warm-up + `reps` builds each,
 * the median host time to a drained stream), keeps the fastest in the context
 * and reports it in best[4]; ms (may be NULL) receives the SVD_BUILD_PRESETS
 * medians.  The presets: {2048,1024,512,2048} (the default),
 * {8192,4096,1024,4096}, {4096,2048,1024,4096}, {1024,512,256,1024},
 * {2048,1024,1024,2048}.                                                     */
#define SVD_BUILD_PRESETS 5
int svd_set_build_limits(svd_ctx* ctx, const uint32_t* limits);
int svd_get_build_limits(svd_ctx* ctx, uint32_t* limits);
int svd_autotune_build(svd_ctx* ctx, const uint64_t* device_codes, uint64_t n, uint32_t levels, uint32_t reps,
                       uint32_t* best, double* ms);

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

/* ---- lossy DAG (LSVDAG) -----------------------------------------------------
 * GeomOctree::toLossyDag (geom_octree_extension.cpp:797-1179).  Needs S_DAG;
 * replaces the DAG view by the lossy DAG (then exactly re-deduplicated, :1176)
 * and stays in S_DAG, so the encoders and svd_reduce_sym run on it as in
 * svbuilder (main.cpp:234-257).  The reference clusters its similarity graph
 * with the external MCL binary (cluster.hpp:134-192); this library's rule is
 * deterministic instead: level by level from lev_hi down to lev_lo, a node j
 * whose effective reference count (sortByEffectiveRefCount, :521-560) is at
 * most max_refs lists the first list_cap earlier nodes i < j whose 512-bit
 * occupancy three levels down differs from its own in at most `bound` cells
 * (exhaustive proposer) or the proposed ones of those (learned proposer:
 * k-nearest neighbours of MLP embeddings), keeps those whose subtree differs
 * from its own in at most `bound` voxels (exact count), and merges into the
 * first of them that stays a representative.  bound = max_diff, or when
 * max_diff is 0, round((L-lev)^2 * diff_factor) (:901).                     */
typedef struct svd_mlp svd_mlp;

#define SVD_PROPOSER_EXHAUSTIVE 0
#define SVD_PROPOSER_LEARNED    1
/* every earlier node within the bound, ranked: the list_cap nearest by
 * (occupancy Hamming distance, index), list_cap <= 64, walked nearest first */
#define SVD_PROPOSER_RANKED     2

typedef struct svd_lossy_params {
  uint32_t lev_lo, lev_hi;     /* levels merged, processed lev_hi .. lev_lo (1 <= lo <= hi <= L-2) */
  uint32_t max_diff;           /* voxel bound per merge; 0 = use diff_factor   */
  uint32_t max_refs;           /* includedNodeRefCount (main.cpp:112)          */
  uint32_t list_cap;           /* earlier neighbours kept per node, 1..64      */
  float diff_factor;           /* allowedLossyDiffFactor (main.cpp:111)        */
  uint32_t proposer;           /* SVD_PROPOSER_*                               */
  uint32_t knn_k;              /* learned: neighbours per node (16)            */
  uint32_t knn_tables, knn_bits, knn_window;  /* learned: LSH search shape (8, 16, 64) */
  uint32_t knn_refine;         /* learned: neighbour-of-neighbour rounds (0..16) */
  uint64_t knn_seed;           /* learned: LSH hyperplane seed                 */
  svd_mlp* model;              /* learned: trained encoder                     */
  uint32_t* knn_record;        /* optional HOST buffer: the kNN lists used, level by level */
  uint64_t knn_record_cap;     /* its capacity in uint32 entries               */
} svd_lossy_params;

typedef struct svd_lossy_level_stats {
  uint64_t level, bound, nodes_before, nodes_after, candidates, listed, verified, merged;
  uint64_t error_bound;        /* sum over merges of voxel difference x effective refs */
  double ms_embed, ms_knn, ms_lists, ms_verify, ms_merge;
  double ms_total;             /* device time of the level, signatures to remap */
  uint64_t clusters;           /* representatives that took in at least one merged node */
} svd_lossy_level_stats;

/* stats: HOST array of stats_cap entries (levels hi..lo); may be NULL. */
int svd_lossy_merge(svd_octree* oct, const svd_lossy_params* p, svd_lossy_level_stats* stats, uint32_t stats_cap);
/* The 512-bit occupancy (16 uint32 per node, bit c1*64+c2*8+c3 = child slot
 * path) of every node of DAG level lev, written to the DEVICE array sig.     */
int svd_lossy_signatures(svd_octree* oct, uint32_t lev, uint32_t* sig_dev);

/* ---- learned candidate encoder (no reference counterpart: the reference's
 * lossy path has no learned part) ---------------------------------------------
 * MLP 512 -> 256 -> 128 -> 32, bf16 weights, fp32 accumulate on MFMA.
 * Parameters are fp32 masters in one flat array of SVD_MLP_PARAMS floats:
 * W1[512][256], b1[256], W2[256][128], b2[128], W3[128][32], b3[32].        */
#define SVD_MLP_PARAMS 168352u

typedef struct svd_mlp_train_params {
  uint32_t lev_lo, lev_hi;     /* DAG levels whose nodes are sampled           */
  uint32_t pairs_per_step;     /* P (multiple of 16)                           */
  uint32_t steps;
  float lr;                    /* Adam learning rate                           */
  uint32_t near_pct;           /* percent of pairs drawn as near pairs (0 = 50) */
  uint64_t seed;               /* pair sampling seed                           */
} svd_mlp_train_params;

typedef struct svd_mlp_train_stats {
  float first_loss, last_loss; /* mean loss of the first / last 16 steps       */
  double ms_train;             /* device time of all steps                     */
  uint64_t nodes, pairs;
} svd_mlp_train_stats;

int svd_mlp_create(svd_ctx* ctx, uint64_t seed, svd_mlp** out);
int svd_mlp_free(svd_mlp* m);
int svd_mlp_get_params(svd_mlp* m, float* host, uint64_t count);
int svd_mlp_set_params(svd_mlp* m, const float* host, uint64_t count);
/* Embeddings (n x 32 fp32, DEVICE) of n signatures (n x 16 uint32, DEVICE). */
int svd_mlp_embed(svd_mlp* m, const uint32_t* sig_dev, uint64_t n, float* emb_dev);
/* One forward + backward without update: rows (2p, 2p+1) of sig_dev are pair
 * p, labels_dev[p] its Hamming distance.  loss and grads (SVD_MLP_PARAMS
 * floats) are HOST outputs.                                                 */
int svd_mlp_loss_grad(svd_mlp* m, const uint32_t* sig_dev, const uint32_t* labels_dev, uint64_t npairs,
                      float* loss, float* grads);
/* Train on node pairs of oct's DAG levels with exact Hamming labels (Adam);
 * the step is captured once in a HIP graph and replayed.                     */
int svd_mlp_train(svd_mlp* m, svd_octree* oct, const svd_mlp_train_params* p, svd_mlp_train_stats* st);
/* k nearest neighbours of n points of dim 32 (DEVICE), searched in `tables`
 * LSH orderings of `bits` sign bits within +-window positions; distance =
 * squared L2 of the embeddings, or, when sig_dev (n x 16 uint32 signatures,
 * DEVICE) is given, the Hamming distance of the signatures (the embeddings
 * then only decide which points meet).  out_dev: n x k indices, ascending by
 * (distance, index), 0xFFFFFFFF-padded.                                      */
int svd_knn(svd_ctx* ctx, const float* emb_dev, uint64_t n, uint32_t k, uint32_t tables, uint32_t bits,
            uint32_t window, uint64_t seed, uint32_t* out_dev, const uint32_t* sig_dev);

/* ---- batches: many scenes per launch (no reference counterpart: the
 * reference builds one model per svbuilder process) ---------------------------
 * svd_batch_build: n_scenes voxel lists (each sorted, unique, at least one
 * voxel, below 2^(3*levels[s])) concatenated in `codes`; scene s is
 * codes[offsets[s] .. offsets[s+1]) (offsets: HOST array of n_scenes+1),
 * depth levels[s] (HOST array, 2..13, uneven depths allowed).  All scenes are
 * reduced to DAGs and encoded as .svdag images in shared launches; every
 * scene's image equals its own svd_octree_from_morton + svd_reduce_dag +
 * svd_encode(SVD_FMT_SVDAG) byte for byte.                                   */
typedef struct svd_batch svd_batch;
int svd_batch_build(svd_ctx* ctx, uint32_t n_scenes, const uint64_t* codes, const uint64_t* offsets,
                    const uint32_t* levels, int codes_on_device, svd_batch** out);
int svd_batch_free(svd_batch* b);
int svd_batch_scene_info(svd_batch* b, uint32_t scene, uint32_t* levels, uint64_t* n_voxels, uint64_t* n_nodes,
                         uint64_t* svdag_bytes);
/* The .svdag image of one scene: device pointer, or a copy into a HOST buffer. */
int svd_batch_svdag(svd_batch* b, uint32_t scene, const void** dptr, uint64_t* bytes);
int svd_batch_svdag_copy(svd_batch* b, uint32_t scene, void* dst, uint64_t cap);
/* n_cams cameras (HOST array, one resolution for all) over the batch's scenes
 * in one launch; camera k looks at scene cam_scene[k] (HOST array).  prim /
 * shadow: DEVICE arrays of n_cams * width * height results, camera-major.    */
int svd_batch_trace(svd_batch* b, uint32_t n_cams, const svd_trace_params* cams, const uint32_t* cam_scene,
                    svd_ray_result* primary, svd_ray_result* shadow);

/* ---- synthetic scenes (benchmark and test inputs; no reference counterpart)
 * Cave volume: surface voxels (solid with an empty 6-neighbour) of
 * {fbm3(p) - 0.5 > threshold}, p = (i + 0.5) * base_cells / 2^levels, value
 * noise with a SplitMix64 lattice hash, lacunarity 2, gain 0.5 -- the float64
 * definition of workloads.cave_codes, bit for bit.  *out_dev receives a
 * device array of *n sorted Morton codes; release it with svd_device_free.   */
int svd_workload_caves(svd_ctx* ctx, uint32_t levels, uint64_t seed, uint32_t octaves, double threshold,
                       double base_cells, uint64_t** out_dev, uint64_t* n);
/* Union of balls (cx, cy, cz, r per ball, HOST array of 4*n_balls doubles):
 * surface voxels of {|p - c| <= r}, p = (i + 0.5) / 2^levels.              */
int svd_workload_balls(svd_ctx* ctx, uint32_t levels, const double* balls, uint32_t n_balls, uint64_t** out_dev,
                       uint64_t* n);
/* Menger sponge of the given order filling the grid: its surface voxels.    */
int svd_workload_menger(svd_ctx* ctx, uint32_t levels, uint32_t order, uint64_t** out_dev, uint64_t* n);
/* fBm heightfield terrain (workloads.terrain_codes): every column filled from
 * the lowest of its 4-neighbour heights up to its own height.               */
int svd_workload_terrain(svd_ctx* ctx, uint32_t levels, uint64_t seed, uint32_t octaves, double hmin, double hmax,
                         double base_cells, uint64_t** out_dev, uint64_t* n);
/* Per-voxel u8 material of n Morton codes (DEVICE) into out_dev (DEVICE):
 * floor(value_noise3((voxel + 0.5) * base_cells / 2^levels) * n_values),
 * clamped to n_values - 1 (workloads.material_attrs).                       */
int svd_workload_materials(svd_ctx* ctx, const uint64_t* codes_dev, uint64_t n, uint32_t levels, uint64_t seed,
                           uint32_t n_values, double base_cells, uint8_t* out_dev);
/* Sorted union of two sorted unique code arrays (DEVICE); *out_dev receives a
 * device array of *n codes (release with svd_device_free).                  */
int svd_codes_union(svd_ctx* ctx, const uint64_t* a_dev, uint64_t na, const uint64_t* b_dev, uint64_t nb,
                    uint64_t** out_dev, uint64_t* n);
int svd_device_free(svd_ctx* ctx, void* dev_ptr);
/* Copy bytes between two device arrays (on the context stream, synchronous). */
int svd_device_copy(svd_ctx* ctx, void* dev_dst, const void* dev_src, uint64_t bytes);
/* Copy bytes from a device array to a HOST buffer (synchronous). */
int svd_device_read(svd_ctx* ctx, void* host_dst, const void* dev_src, uint64_t bytes);

/* ---- reference-count order ------------------------------------------------
 * The permutation std::sort(hist, [](a,b){ return a.second > b.second; })
 * produces in EncodedSSVDAG::encode (encoded_ssvdag.cpp:273-275), computed on
 * the device.  idx / val are HOST arrays of n pairs, sorted in place.        */
int svd_refcount_order(svd_ctx* ctx, uint32_t* idx, uint32_t* val, uint64_t n);
/* The same order applied to nr disjoint ranges [first[r], last[r]) of one pair
 * array at once, each range its own std::sort call (the per-level sorts of
 * encoded_ssvdag.cpp:257-275 run together).  Arrays on the HOST.            */
int svd_refcount_order_ranges(svd_ctx* ctx, uint32_t* idx, uint32_t* val, uint64_t n, const uint64_t* first,
                              const uint64_t* last, uint64_t nr);

/* ---- instrumentation ------------------------------------------------------ */
/* Average duration in milliseconds of the last launch of the named kernel
 * family, measured with HIP events on the context stream ("trace", "build",
 * "encode").  Returns SVD_ERR_ARG for an unknown name.                      */
int svd_last_kernel_ms(svd_ctx* ctx, const char* name, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* SVDAG_AMD_H */
