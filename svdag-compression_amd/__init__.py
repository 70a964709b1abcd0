"""MI355X sparse-voxel-DAG engine: ctypes binding of libsvdag_amd.so.

The product path is the HIP library built from ``csrc/`` (see ``Makefile``);
this module only marshals arguments across its C ABI (``include/svdag_amd.h``).
It never falls back to CPU code: if the library is missing or no GPU is
present, the calls raise.

Classes mirror the reference's C++ API:
    Octree   ~ GeomOctree      (src/symvox/geom_octree.hpp:32-190)
    encode() ~ Encoded*::encode + save (encoded_{svdag,ussvdag,ssvdag}.cpp)
    Scene    ~ EncodedOctree::load + OctreeDDARenderer (svviewer)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsvdag_amd.so")

# status codes / constants (include/svdag_amd.h)
OK = 0
S_EMPTY, S_SVO, S_DAG, S_SDAG = 0, 1, 2, 3
FMT_SVDAG, FMT_USSVDAG, FMT_SSVDAG, FMT_ESVDAG = 0, 1, 2, 3
VIEW_SVO, VIEW_DAG, VIEW_SDAG = 0, 1, 2
NULL_NODE = 0xFFFFFFFE
FMT_EXT = {FMT_SVDAG: ".svdag", FMT_USSVDAG: ".ussvdag", FMT_SSVDAG: ".ssvdag", FMT_ESVDAG: ".esvdag"}


TRACE_MODES = ("tiled", "persistent")   # SVD_TRACE_TILED, SVD_TRACE_PERSISTENT


class SvdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"svd error {code}: {msg}")
        self.code = code


class TraceParams(C.Structure):
    _fields_ = [("view_inv", C.c_float * 16), ("proj_inv", C.c_float * 16),
                ("width", C.c_uint32), ("height", C.c_uint32), ("max_iters", C.c_uint32),
                ("draw_level", C.c_uint32), ("projection_factor", C.c_float),
                ("light_pos", C.c_float * 3), ("shadows", C.c_uint32), ("pad", C.c_uint32)]


class ViewParams(C.Structure):
    """svd_view_params: the viewer program's uniforms (octree_dda_renderer.cpp:41-45, 73 defaults)."""
    _fields_ = [("render_mode", C.c_int32), ("norm_samples", C.c_int32), ("selected_voxel", C.c_uint32),
                ("random_colors", C.c_uint32), ("freq_colors", C.c_uint32), ("pad", C.c_uint32)]

    @classmethod
    def defaults(cls, **kw) -> "ViewParams":
        v = cls(0, 0, 0xFFFFFFFF, 1, 1, 0)
        for k, x in kw.items():
            setattr(v, k, x)
        return v


PROPOSER_EXHAUSTIVE, PROPOSER_LEARNED, PROPOSER_RANKED = 0, 1, 2
MLP_PARAMS = 168352


class LossyParams(C.Structure):
    _fields_ = [("lev_lo", C.c_uint32), ("lev_hi", C.c_uint32), ("max_diff", C.c_uint32), ("max_refs", C.c_uint32),
                ("list_cap", C.c_uint32), ("diff_factor", C.c_float), ("proposer", C.c_uint32),
                ("knn_k", C.c_uint32), ("knn_tables", C.c_uint32), ("knn_bits", C.c_uint32),
                ("knn_window", C.c_uint32), ("knn_refine", C.c_uint32), ("knn_seed", C.c_uint64), ("model", C.c_void_p),
                ("knn_record", C.c_void_p), ("knn_record_cap", C.c_uint64)]


LOSSY_STATS_DTYPE = np.dtype([(f, "<u8") for f in ("level", "bound", "nodes_before", "nodes_after", "candidates",
                                                   "listed", "verified", "merged", "error_bound")]
                             + [(f, "<f8") for f in ("ms_embed", "ms_knn", "ms_lists", "ms_verify", "ms_merge",
                                                     "ms_total")] + [("clusters", "<u8")])


class TrainParams(C.Structure):
    _fields_ = [("lev_lo", C.c_uint32), ("lev_hi", C.c_uint32), ("pairs_per_step", C.c_uint32),
                ("steps", C.c_uint32), ("lr", C.c_float), ("near_pct", C.c_uint32), ("seed", C.c_uint64)]


class TrainStats(C.Structure):
    _fields_ = [("first_loss", C.c_float), ("last_loss", C.c_float), ("ms_train", C.c_double),
                ("nodes", C.c_uint64), ("pairs", C.c_uint64)]


RAY_DTYPE = np.dtype([("t", "<f4"), ("level", "<f4"), ("iters", "<f4"), ("node", "<f4"),
                      ("path", "<u4"), ("bytes", "<u4")])

_lib = None


def build(verbose: bool = False) -> None:
    """Compile libsvdag_amd.so for gfx950 in-tree (always recompiles changed sources)."""
    out = subprocess.run(["make", "-C", _HERE, "-j8"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("libsvdag_amd build failed:\n" + out.stdout[-4000:] + out.stderr[-4000:])
    if verbose:
        print(out.stdout[-2000:])


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run build() first")
    # One HIP runtime per process: torch ships its own libamdhip64 (loaded by
    # file name), so let torch load it first; our library's NEEDED
    # libamdhip64.so.7 then binds to that copy instead of a second one.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    P = C.POINTER
    sig = {
        "svd_ctx_create": (i32, [i32, P(vp)]),
        "svd_ctx_destroy": (i32, [vp]),
        "svd_last_error": (C.c_char_p, [vp]),
        "svd_ctx_synchronize": (i32, [vp]),
        "svd_ctx_stream": (vp, [vp]),
        "svd_octree_from_morton": (i32, [vp, vp, u64, u32, i32, P(vp)]),
        "svd_octree_from_svo": (i32, [vp, u32, P(u64), P(vp), P(C.c_float), C.c_double, P(vp)]),
        "svd_octree_from_svdag": (i32, [vp, vp, u64, i32, P(vp)]),
        "svd_octree_from_mesh": (i32, [vp, vp, u64, u32, P(C.c_double), i32, P(vp)]),
        "svd_octree_from_mesh_subtrees": (i32, [vp, vp, u64, u32, u32, P(C.c_double), i32, P(vp)]),
        "svd_octree_from_points": (i32, [vp, vp, u64, u32, P(C.c_double), i32, P(vp)]),
        "svd_octree_free": (i32, [vp]),
        "svd_octree_from_levels": (i32, [vp, u32, P(u64), P(vp), P(C.c_float), C.c_double, u64, u64, P(vp)]),
        "svd_octree_copy": (i32, [vp, P(vp)]),
        "svd_reduce_sym_pointers": (i32, [vp]),
        "svd_reduce_dag": (i32, [vp]),
        "svd_reduce_sym": (i32, [vp]),
        "svd_cross_level_merge": (i32, [vp, P(u64)]),
        "svd_octree_state": (i32, [vp, P(i32)]),
        "svd_octree_levels": (i32, [vp, P(u32)]),
        "svd_octree_counts": (i32, [vp, P(u64), P(u64)]),
        "svd_level_size": (i32, [vp, i32, u32, P(u64)]),
        "svd_get_level": (i32, [vp, i32, u32, vp, vp, vp]),
        "svd_encode": (i32, [vp, i32, P(u64)]),
        "svd_encoded_copy": (i32, [vp, i32, vp, u64]),
        "svd_encode_formats": (i32, [vp, vp, u32, vp]),
        "svd_encoded_device": (i32, [vp, i32, P(vp), P(u64)]),
        "svd_scene_create": (i32, [vp, i32, vp, u64, i32, P(vp)]),
        "svd_scene_from_octree": (i32, [vp, i32, P(vp)]),
        "svd_scene_free": (i32, [vp]),
        "svd_trace": (i32, [vp, P(TraceParams), vp, vp]),
        "svd_trace_hits": (i32, [vp, P(TraceParams), vp, vp]),
        "svd_trace_beam": (i32, [vp, P(TraceParams), C.c_float, vp, vp, vp]),
        "svd_render": (i32, [vp, P(TraceParams), P(ViewParams), C.c_float, vp, vp]),
        "svd_trace_ao": (i32, [vp, P(TraceParams), C.c_uint32, vp, vp]),
        "svd_render_noalloc": (i32, [vp, P(TraceParams), P(ViewParams), C.c_float, vp, vp]),
        "svd_last_kernel_ms": (i32, [vp, C.c_char_p, P(C.c_double)]),
        "svd_refcount_order": (i32, [vp, vp, vp, u64]),
        "svd_refcount_order_ranges": (i32, [vp, vp, vp, u64, vp, vp, u64]),
        "svd_attr_bittree_encode": (i32, [vp, vp, u64, i32, i32, vp, u64, P(u64)]),
        "svd_lossy_merge": (i32, [vp, P(LossyParams), vp, u32]),
        "svd_lossy_signatures": (i32, [vp, u32, vp]),
        "svd_mlp_create": (i32, [vp, u64, P(vp)]),
        "svd_mlp_free": (i32, [vp]),
        "svd_mlp_get_params": (i32, [vp, vp, u64]),
        "svd_mlp_set_params": (i32, [vp, vp, u64]),
        "svd_mlp_embed": (i32, [vp, vp, u64, vp]),
        "svd_mlp_loss_grad": (i32, [vp, vp, vp, u64, P(C.c_float), vp]),
        "svd_mlp_train": (i32, [vp, vp, P(TrainParams), P(TrainStats)]),
        "svd_knn": (i32, [vp, vp, u64, u32, u32, u32, u32, u64, vp, vp]),
        "svd_workload_caves": (i32, [vp, u32, u64, u32, C.c_double, C.c_double, P(vp), P(u64)]),
        "svd_device_free": (i32, [vp, vp]),
        "svd_device_read": (i32, [vp, vp, vp, u64]),
        "svd_workload_balls": (i32, [vp, u32, vp, u32, P(vp), P(u64)]),
        "svd_workload_menger": (i32, [vp, u32, u32, P(vp), P(u64)]),
        "svd_workload_terrain": (i32, [vp, u32, u64, u32, C.c_double, C.c_double, C.c_double, P(vp), P(u64)]),
        "svd_batch_build": (i32, [vp, u32, vp, P(u64), P(u32), i32, P(vp)]),
        "svd_batch_free": (i32, [vp]),
        "svd_batch_scene_info": (i32, [vp, u32, P(u32), P(u64), P(u64), P(u64)]),
        "svd_batch_svdag": (i32, [vp, u32, P(vp), P(u64)]),
        "svd_batch_svdag_copy": (i32, [vp, u32, vp, u64]),
        "svd_batch_trace": (i32, [vp, u32, vp, P(u32), vp, vp]),
        "svd_device_copy": (i32, [vp, vp, vp, u64]),
        "svd_workload_materials": (i32, [vp, vp, u64, u32, u64, u32, C.c_double, vp]),
        "svd_codes_union": (i32, [vp, vp, u64, vp, u64, P(vp), P(u64)]),
        "svd_set_trace_tile": (i32, [vp, u32]),
        "svd_get_trace_tile": (i32, [vp, P(u32)]),
        "svd_autotune_trace": (i32, [vp, P(TraceParams), vp, vp, u32, P(u32), P(C.c_double)]),
        "svd_set_trace_mode": (i32, [vp, u32]),
        "svd_get_trace_mode": (i32, [vp, P(u32)]),
        "svd_autotune_trace_modes": (i32, [vp, P(TraceParams), vp, vp, u32, P(u32), P(u32), P(C.c_double)]),
        "svd_set_build_limits": (i32, [vp, P(u32)]),
        "svd_get_build_limits": (i32, [vp, P(u32)]),
        "svd_autotune_build": (i32, [vp, vp, u64, u32, u32, P(u32), P(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = ["svd_ctx_create", "svd_ctx_destroy", "svd_last_error", "svd_ctx_synchronize", "svd_ctx_stream",
            "svd_octree_from_morton", "svd_octree_from_svo", "svd_octree_from_svdag", "svd_octree_from_mesh", "svd_octree_from_mesh_subtrees",
            "svd_octree_from_points",
            "svd_octree_free", "svd_octree_from_levels", "svd_octree_copy", "svd_reduce_sym_pointers",
            "svd_reduce_dag",
            "svd_reduce_sym", "svd_cross_level_merge", "svd_octree_state", "svd_octree_levels",
            "svd_octree_counts", "svd_level_size", "svd_get_level", "svd_encode", "svd_encode_formats", "svd_encoded_copy",
            "svd_encoded_device", "svd_scene_create", "svd_scene_from_octree", "svd_scene_free", "svd_trace", "svd_trace_hits", "svd_trace_beam",
            "svd_render", "svd_trace_ao", "svd_render_noalloc",
            "svd_last_kernel_ms", "svd_refcount_order", "svd_refcount_order_ranges", "svd_attr_bittree_encode", "svd_lossy_merge",
            "svd_lossy_signatures", "svd_mlp_create", "svd_mlp_free", "svd_mlp_get_params", "svd_mlp_set_params",
            "svd_mlp_embed", "svd_mlp_loss_grad", "svd_mlp_train", "svd_knn", "svd_workload_caves",
            "svd_device_free", "svd_device_read", "svd_workload_balls", "svd_workload_menger",
            "svd_workload_terrain", "svd_batch_build", "svd_batch_free", "svd_batch_scene_info", "svd_batch_svdag",
            "svd_batch_svdag_copy", "svd_batch_trace", "svd_device_copy",
            "svd_set_trace_tile", "svd_get_trace_tile", "svd_autotune_trace", "svd_set_trace_mode",
            "svd_get_trace_mode", "svd_autotune_trace_modes", "svd_set_build_limits", "svd_get_build_limits",
            "svd_autotune_build", "svd_workload_materials",
            "svd_codes_union"]


class Context:
    """One device + one HIP stream (svd_ctx)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.svd_ctx_create(device, C.byref(h))
        if rc != OK:
            raise SvdError(rc, f"cannot create a context on device {device}")
        self.h = h

    def check(self, rc):
        if rc != OK:
            raise SvdError(rc, self.L.svd_last_error(self.h).decode())

    def sync(self):
        self.check(self.L.svd_ctx_synchronize(self.h))

    @property
    def stream(self) -> int:
        return self.L.svd_ctx_stream(self.h) or 0

    def kernel_ms(self, name: str) -> float:
        v = C.c_double()
        self.check(self.L.svd_last_kernel_ms(self.h, name.encode(), C.byref(v)))
        return v.value

    def refcount_order(self, idx, val):
        """Device emulation of std::sort by reference count (returns sorted copies)."""
        i = np.ascontiguousarray(idx, dtype=np.uint32).copy()
        v = np.ascontiguousarray(val, dtype=np.uint32).copy()
        self.check(self.L.svd_refcount_order(self.h, i.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                                             i.size))
        return i, v

    def refcount_order_ranges(self, idx, val, ranges):
        """refcount_order applied to each [first, last) range independently, in one device sort."""
        i = np.ascontiguousarray(idx, dtype=np.uint32).copy()
        v = np.ascontiguousarray(val, dtype=np.uint32).copy()
        f = np.ascontiguousarray([r[0] for r in ranges], dtype=np.uint64)
        l = np.ascontiguousarray([r[1] for r in ranges], dtype=np.uint64)
        self.check(self.L.svd_refcount_order_ranges(self.h, i.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                                                    i.size, f.ctypes.data_as(C.c_void_p),
                                                    l.ctypes.data_as(C.c_void_p), f.size))
        return i, v

    def attr_bittree(self, attrs, gray: bool) -> bytes:
        """Gray-code (or binary) attribute bit-tree of per-voxel u8 attributes in voxel Morton order."""
        if isinstance(attrs, tuple):                 # (device_ptr, n)
            ptr, n = attrs
            src, on_dev = C.c_void_p(ptr), 1
        else:
            a = np.ascontiguousarray(attrs, dtype=np.uint8)
            src, n, on_dev = a.ctypes.data_as(C.c_void_p), a.size, 0
        size = C.c_uint64()
        self.check(self.L.svd_attr_bittree_encode(self.h, src, n, int(gray), on_dev, None, 0, C.byref(size)))
        buf = np.zeros(size.value, np.uint8)
        self.check(self.L.svd_attr_bittree_encode(self.h, src, n, int(gray), on_dev, buf.ctypes.data_as(C.c_void_p),
                                                  size.value, C.byref(size)))
        return buf.tobytes()

    def knn(self, emb_ptr: int, n: int, k: int = 16, tables: int = 4, bits: int = 16, window: int = 32,
            seed: int = 7, out_ptr: int = 0, sig_ptr: int = 0):
        """k nearest neighbours of n 32-d fp32 points (device pointers); with
        sig_ptr the candidates are scored by signature Hamming distance."""
        self.check(self.L.svd_knn(self.h, C.c_void_p(emb_ptr), n, k, tables, bits, window, seed, C.c_void_p(out_ptr),
                                  C.c_void_p(sig_ptr) if sig_ptr else None))

    def caves(self, levels: int, seed: int = 3, octaves: int = 5, threshold: float = 0.1,
              base_cells: float = 4.0) -> "DeviceCodes":
        """Cave-volume surface voxels generated on the device (same definition as workloads.cave_codes)."""
        ptr, n = C.c_void_p(), C.c_uint64()
        self.check(self.L.svd_workload_caves(self.h, levels, seed, octaves, threshold, base_cells, C.byref(ptr),
                                             C.byref(n)))
        return DeviceCodes(self, ptr.value, n.value)

    def balls(self, levels: int, balls) -> "DeviceCodes":
        b = np.ascontiguousarray(balls, dtype=np.float64).reshape(-1, 4)
        ptr, n = C.c_void_p(), C.c_uint64()
        self.check(self.L.svd_workload_balls(self.h, levels, b.ctypes.data_as(C.c_void_p), b.shape[0], C.byref(ptr),
                                             C.byref(n)))
        return DeviceCodes(self, ptr.value, n.value)

    def menger(self, levels: int, order: int) -> "DeviceCodes":
        ptr, n = C.c_void_p(), C.c_uint64()
        self.check(self.L.svd_workload_menger(self.h, levels, order, C.byref(ptr), C.byref(n)))
        return DeviceCodes(self, ptr.value, n.value)

    def terrain(self, levels: int, seed: int = 1, octaves: int = 6, hmin: float = 0.2, hmax: float = 0.6,
                base_cells: float = 8.0) -> "DeviceCodes":
        ptr, n = C.c_void_p(), C.c_uint64()
        self.check(self.L.svd_workload_terrain(self.h, levels, seed, octaves, hmin, hmax, base_cells, C.byref(ptr),
                                               C.byref(n)))
        return DeviceCodes(self, ptr.value, n.value)

    def materials(self, codes: "DeviceCodes", levels: int, out_ptr: int, seed: int = 4, n_values: int = 16,
                  base_cells: float = 8.0):
        """Per-voxel u8 materials of device codes into a device buffer (workloads.material_attrs)."""
        self.check(self.L.svd_workload_materials(self.h, C.c_void_p(codes.ptr), codes.n, levels, seed, n_values,
                                                 base_cells, C.c_void_p(out_ptr)))

    def union(self, a: "DeviceCodes", b: "DeviceCodes") -> "DeviceCodes":
        ptr, n = C.c_void_p(), C.c_uint64()
        self.check(self.L.svd_codes_union(self.h, C.c_void_p(a.ptr), a.n, C.c_void_p(b.ptr), b.n, C.byref(ptr),
                                          C.byref(n)))
        return DeviceCodes(self, ptr.value, n.value)

    def spec_codes(self, spec) -> "DeviceCodes":
        """Device voxels of one batch scene (workloads.batch_specs)."""
        from .workloads import sphere_balls
        if spec["kind"] == "spheres":
            return self.balls(spec["levels"], sphere_balls(spec["n_spheres"], spec["seed"]))
        if spec["kind"] == "terrain":
            return self.terrain(spec["levels"], seed=spec["seed"])
        return self.menger(spec["levels"], spec["order"])

    @property
    def trace_tile(self) -> int:
        w = C.c_uint32()
        self.check(self.L.svd_get_trace_tile(self.h, C.byref(w)))
        return w.value

    @trace_tile.setter
    def trace_tile(self, w: int):
        self.check(self.L.svd_set_trace_tile(self.h, w))

    @property
    def trace_mode(self) -> str:
        """"tiled" (one workgroup per pixel tile) or "persistent" (ray queue)."""
        m = C.c_uint32()
        self.check(self.L.svd_get_trace_mode(self.h, C.byref(m)))
        return TRACE_MODES[m.value]

    @trace_mode.setter
    def trace_mode(self, mode: str):
        self.check(self.L.svd_set_trace_mode(self.h, TRACE_MODES.index(mode)))

    @property
    def build_limits(self) -> tuple:
        """One-workgroup level limits of the build: (SVO codes, toDAG nodes,
        toSDAG nodes, layout nodes)."""
        v = (C.c_uint32 * 4)()
        self.check(self.L.svd_get_build_limits(self.h, v))
        return tuple(int(x) for x in v)

    @build_limits.setter
    def build_limits(self, limits):
        v = (C.c_uint32 * 4)(*[int(x) for x in limits])
        self.check(self.L.svd_set_build_limits(self.h, v))

    def autotune_build(self, codes_ptr: int, n: int, levels: int, reps: int = 5) -> dict:
        """Time the lossless build of these device codes under each preset of
        the build limits, keep the fastest in the context; returns the chosen
        limits and every preset's median time (ms)."""
        best = (C.c_uint32 * 4)()
        ms = (C.c_double * 5)()
        self.check(self.L.svd_autotune_build(self.h, C.c_void_p(codes_ptr), n, levels, reps, best, ms))
        return {"limits": tuple(int(x) for x in best), "ms": [round(x, 4) for x in ms]}

    def autotune_trace(self, scene: "Scene", params: TraceParams, prim_ptr: int, shad_ptr: int | None,
                       reps: int = 3, cache: str | None = None, key: str | None = None) -> int:
        """Pick the fastest traversal launch (mode x tile shape) for this frame
        and keep it in the context; returns the tile width.  With a cache file
        (JSON) and a key, a stored choice is reused and a new one is stored."""
        import json
        if cache and key and os.path.exists(cache):
            try:
                stored = json.load(open(cache)).get(key)
            except (OSError, ValueError):
                stored = None
            if stored and stored.get("mode") in TRACE_MODES:
                self.trace_tile = int(stored["tile_w"])
                self.trace_mode = stored["mode"]
                self.last_autotune = stored
                return self.trace_tile
        best, mode = C.c_uint32(), C.c_uint32()
        ms = (C.c_double * 8)()
        self.check(self.L.svd_autotune_trace_modes(scene.h, C.byref(params), C.c_void_p(prim_ptr),
                                                   C.c_void_p(shad_ptr) if shad_ptr else None, reps, C.byref(best),
                                                   C.byref(mode), ms))
        self.last_autotune = {"tile_w": best.value, "mode": TRACE_MODES[mode.value],
                              "ms": {m: {w: round(ms[4 * i + j], 4) for j, w in enumerate((8, 16, 32, 64))}
                                     for i, m in enumerate(TRACE_MODES)}}
        if cache and key:
            data = {}
            if os.path.exists(cache):
                try:
                    data = json.load(open(cache))
                except (OSError, ValueError):
                    data = {}
            data[key] = self.last_autotune
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(data, f, indent=1)
        return best.value

    def close(self):
        if getattr(self, "h", None):
            self.L.svd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCodes:
    """Sorted Morton codes in device memory owned by the library."""

    def __init__(self, ctx: Context, ptr: int, n: int):
        self.ctx, self.ptr, self.n = ctx, ptr, n

    def numpy(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint64)
        if self.n:
            self.ctx.check(self.ctx.L.svd_device_read(self.ctx.h, out.ctypes.data_as(C.c_void_p),
                                                      C.c_void_p(self.ptr), self.n * 8))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.L.svd_device_free(self.ctx.h, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many scenes reduced to DAGs and encoded as .svdag images in shared launches (svd_batch)."""

    def __init__(self, ctx: Context, codes, offsets, levels, on_device: bool = False):
        """codes: concatenated per-scene sorted codes (numpy uint64, or a device
        pointer with on_device=True); offsets: n_scenes + 1 starts; levels: depths."""
        self.ctx = ctx
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        lev = np.ascontiguousarray(levels, dtype=np.uint32)
        self.S = lev.size
        h = C.c_void_p()
        if on_device:
            src = C.c_void_p(codes)
        else:
            self._codes = np.ascontiguousarray(codes, dtype=np.uint64)
            src = self._codes.ctypes.data_as(C.c_void_p)
        ctx.check(ctx.L.svd_batch_build(ctx.h, self.S, src, off.ctypes.data_as(P_U64), lev.ctypes.data_as(P_U32),
                                        int(on_device), C.byref(h)))
        self.h = h
        self._codes = None

    @classmethod
    def from_device_scenes(cls, ctx: Context, scenes, levels):
        """scenes: list of DeviceCodes; packs them into one device buffer first."""
        import torch
        sizes = [d.n for d in scenes]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        buf = torch.empty(int(off[-1]), dtype=torch.int64, device="cuda:0")
        for d, o in zip(scenes, off[:-1]):
            ctx.check(ctx.L.svd_device_copy(ctx.h, C.c_void_p(buf.data_ptr() + 8 * int(o)), C.c_void_p(d.ptr),
                                            8 * d.n))
        b = cls(ctx, buf.data_ptr(), off, levels, on_device=True)
        del buf
        return b

    def info(self, s: int):
        L, nv, nn, nb = C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.svd_batch_scene_info(self.h, s, C.byref(L), C.byref(nv), C.byref(nn), C.byref(nb)))
        return dict(levels=L.value, voxels=nv.value, nodes=nn.value, bytes=nb.value)

    def svdag_bytes(self, s: int) -> bytes:
        n = self.info(s)["bytes"]
        buf = np.zeros(n, np.uint8)
        self.ctx.check(self.ctx.L.svd_batch_svdag_copy(self.h, s, buf.ctypes.data_as(C.c_void_p), n))
        return buf.tobytes()

    def trace(self, cams, cam_scene, prim_ptr: int, shad_ptr: int | None):
        """cams: list of TraceParams (one resolution); results camera-major in device buffers."""
        arr = (TraceParams * len(cams))(*cams)
        cs = np.ascontiguousarray(cam_scene, dtype=np.uint32)
        self.ctx.check(self.ctx.L.svd_batch_trace(self.h, len(cams), C.cast(arr, C.c_void_p),
                                                  cs.ctypes.data_as(P_U32), C.c_void_p(prim_ptr),
                                                  C.c_void_p(shad_ptr) if shad_ptr else None))

    def free(self):
        if getattr(self, "h", None) and self.ctx.h:      # a closed context released everything
            self.ctx.L.svd_batch_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


P_U64 = C.POINTER(C.c_uint64)
P_U32 = C.POINTER(C.c_uint32)


class Octree:
    """GeomOctree on the device (svd_octree)."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.h = ctx, handle

    @classmethod
    def from_morton(cls, ctx: Context, codes, levels: int):
        """codes: sorted unique uint64 numpy array, or a device pointer (int) with count in codes_n."""
        h = C.c_void_p()
        if isinstance(codes, tuple):          # (device_ptr, n)
            ptr, n = codes
            ctx.check(ctx.L.svd_octree_from_morton(ctx.h, C.c_void_p(ptr), n, levels, 1, C.byref(h)))
        else:
            codes = np.ascontiguousarray(codes, dtype=np.uint64)
            ctx.check(ctx.L.svd_octree_from_morton(ctx.h, codes.ctypes.data_as(C.c_void_p), codes.size, levels, 0,
                                                   C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_mesh(cls, ctx: Context, tris, levels: int, bbox_min, bbox_max):
        """buildSVO on the device: tris float32 [T, 3, 3] (host array) or (device_ptr, T)."""
        h = C.c_void_p()
        bb = (C.c_double * 6)(*[float(x) for x in list(bbox_min) + list(bbox_max)])
        if isinstance(tris, tuple):
            ptr, n = tris
            ctx.check(ctx.L.svd_octree_from_mesh(ctx.h, C.c_void_p(ptr), n, levels, bb, 1, C.byref(h)))
        else:
            t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
            ctx.check(ctx.L.svd_octree_from_mesh(ctx.h, t.ctypes.data_as(C.c_void_p), t.shape[0], levels, bb, 0,
                                                 C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_points(cls, ctx: Context, points, levels: int, bbox_min, bbox_max):
        """buildSVOFromPoints on the device (geom_octree.cpp:62-169): points float64 [N, 3]
        (host array, the file's order) or (device_ptr, N)."""
        h = C.c_void_p()
        bb = (C.c_double * 6)(*[float(x) for x in list(bbox_min) + list(bbox_max)])
        if isinstance(points, tuple):
            ptr, n = points
            ctx.check(ctx.L.svd_octree_from_points(ctx.h, C.c_void_p(ptr), n, levels, bb, 1, C.byref(h)))
        else:
            pt = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
            ctx.check(ctx.L.svd_octree_from_points(ctx.h, pt.ctypes.data_as(C.c_void_p), pt.shape[0], levels, bb, 0,
                                                   C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_mesh_subtrees(cls, ctx: Context, tris, levels: int, step_level: int, bbox_min, bbox_max):
        """buildDAG's joined SVO (top + per-voxel subtrees) on the device; to_dag() completes buildDAG."""
        h = C.c_void_p()
        bb = (C.c_double * 6)(*[float(x) for x in list(bbox_min) + list(bbox_max)])
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        ctx.check(ctx.L.svd_octree_from_mesh_subtrees(ctx.h, t.ctypes.data_as(C.c_void_p), t.shape[0], levels,
                                                      step_level, bb, 0, C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_svdag(cls, ctx: Context, file):
        """Decode an .svdag file image (bytes, or (device_ptr, n_bytes)) into a DAG-state octree."""
        h = C.c_void_p()
        if isinstance(file, tuple):
            ptr, n = file
            ctx.check(ctx.L.svd_octree_from_svdag(ctx.h, C.c_void_p(ptr), n, 1, C.byref(h)))
        else:
            buf = np.frombuffer(bytes(file), dtype=np.uint8)
            ctx.check(ctx.L.svd_octree_from_svdag(ctx.h, buf.ctypes.data_as(C.c_void_p), buf.size, 0, C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_svo(cls, ctx: Context, level_words, bbox, root_side: float):
        """level_words: list of (n_l, 9) uint32 arrays (word0 = mask, words 1..8 = children)."""
        L = len(level_words)
        arrs = [np.ascontiguousarray(w, dtype=np.uint32).reshape(-1, 9) for w in level_words]
        sizes = (C.c_uint64 * L)(*[a.shape[0] for a in arrs])
        ptrs = (C.c_void_p * L)(*[a.ctypes.data_as(C.c_void_p).value for a in arrs])
        bb = (C.c_float * 6)(*[float(x) for x in bbox])
        h = C.c_void_p()
        ctx.check(ctx.L.svd_octree_from_svo(ctx.h, L, sizes, ptrs, bb, float(root_side), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_levels(cls, ctx: Context, level_words, bbox, root_side: float, n_nodes: int, n_voxels: int):
        """GeomOctree(NodeData, bbox, rootSide, levels, Stats): the given DAG
        levels as a DAG-state octree reporting the given counts."""
        L = len(level_words)
        arrs = [np.ascontiguousarray(w, dtype=np.uint32).reshape(-1, 9) for w in level_words]
        sizes = (C.c_uint64 * L)(*[a.shape[0] for a in arrs])
        ptrs = (C.c_void_p * L)(*[a.ctypes.data_as(C.c_void_p).value for a in arrs])
        bb = (C.c_float * 6)(*[float(x) for x in bbox])
        h = C.c_void_p()
        ctx.check(ctx.L.svd_octree_from_levels(ctx.h, L, sizes, ptrs, bb, float(root_side), int(n_nodes),
                                               int(n_voxels), C.byref(h)))
        return cls(ctx, h)

    def copy(self) -> "Octree":
        """A deep copy (the GeomOctree copy constructor)."""
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.svd_octree_copy(self.h, C.byref(h)))
        return Octree(self.ctx, h)

    def free(self):
        if self.h and self.ctx.h:
            self.ctx.L.svd_octree_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def to_dag(self):
        self.ctx.check(self.ctx.L.svd_reduce_dag(self.h))

    def to_sdag(self, skip_symmetry: bool = False):
        """toSDAG; skip_symmetry = the reference's pointer-only variant."""
        if skip_symmetry:
            self.ctx.check(self.ctx.L.svd_reduce_sym_pointers(self.h))
        else:
            self.ctx.check(self.ctx.L.svd_reduce_sym(self.h))

    def merge_cross_levels(self) -> int:
        n = C.c_uint64()
        self.ctx.check(self.ctx.L.svd_cross_level_merge(self.h, C.byref(n)))
        return n.value

    @property
    def state(self) -> int:
        s = C.c_int()
        self.ctx.check(self.ctx.L.svd_octree_state(self.h, C.byref(s)))
        return s.value

    @property
    def levels(self) -> int:
        v = C.c_uint32()
        self.ctx.check(self.ctx.L.svd_octree_levels(self.h, C.byref(v)))
        return v.value

    def counts(self):
        a, b = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.svd_octree_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def level_size(self, view: int, lev: int) -> int:
        n = C.c_uint64()
        self.ctx.check(self.ctx.L.svd_level_size(self.h, view, lev, C.byref(n)))
        return n.value

    def level(self, view: int, lev: int, with_inv=False, with_child_levels=False):
        n = self.level_size(view, lev)
        w = np.zeros((n, 9), np.uint32)
        inv = np.zeros(n, np.uint8) if with_inv else None
        cl = np.zeros((n, 8), np.uint32) if with_child_levels else None
        if n:
            self.ctx.check(self.ctx.L.svd_get_level(self.h, view, lev, w.ctypes.data_as(C.c_void_p),
                                                    inv.ctypes.data_as(C.c_void_p) if inv is not None else None,
                                                    cl.ctypes.data_as(C.c_void_p) if cl is not None else None))
        out = [w]
        if with_inv:
            out.append(inv)
        if with_child_levels:
            out.append(cl)
        return out[0] if len(out) == 1 else tuple(out)

    def encode(self, fmt: int) -> int:
        n = C.c_uint64()
        self.ctx.check(self.ctx.L.svd_encode(self.h, fmt, C.byref(n)))
        return n.value

    def encode_many(self, fmts) -> list:
        """Several formats in one call (passes shared); returns their sizes."""
        f = np.ascontiguousarray(list(fmts), dtype=np.int32)
        out = np.zeros(f.size, np.uint64)
        self.ctx.check(self.ctx.L.svd_encode_formats(self.h, f.ctypes.data_as(C.c_void_p), f.size,
                                                     out.ctypes.data_as(C.c_void_p)))
        return [int(x) for x in out]

    def encoded_bytes(self, fmt: int) -> bytes:
        ptr, n = C.c_void_p(), C.c_uint64()
        self.ctx.check(self.ctx.L.svd_encoded_device(self.h, fmt, C.byref(ptr), C.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        self.ctx.check(self.ctx.L.svd_encoded_copy(self.h, fmt, buf.ctypes.data_as(C.c_void_p), n.value))
        return buf.tobytes()

    def encoded_device(self, fmt: int):
        ptr, n = C.c_void_p(), C.c_uint64()
        self.ctx.check(self.ctx.L.svd_encoded_device(self.h, fmt, C.byref(ptr), C.byref(n)))
        return ptr.value, n.value

    def lossy_merge(self, lev_lo: int, lev_hi: int, max_diff: int = 0, diff_factor: float = 1.0, max_refs: int = 1,
                    list_cap: int = 32, model: "MLP | None" = None, k: int = 16, tables: int = 8, bits: int = 16,
                    window: int = 64, refine: int = 2, knn_seed: int = 7, record_knn: bool = False,
                    ranked: bool = False):
        """GeomOctree::toLossyDag on the device.  model=None: exhaustive
        proposer (the first list_cap earlier nodes within the bound, or the
        list_cap nearest of them when ranked); else the learned one.  Returns
        per-level stats (and the kNN lists used, concatenated level by level,
        when record_knn)."""
        prop = PROPOSER_LEARNED if model is not None else (PROPOSER_RANKED if ranked else PROPOSER_EXHAUSTIVE)
        p = LossyParams(lev_lo, lev_hi, max_diff, max_refs, list_cap, diff_factor,
                        prop, k, tables, bits, window, refine,
                        knn_seed, model.h if model is not None else None, None, 0)
        rec = None
        if record_knn and model is not None:
            cap = sum(self.level_size(VIEW_DAG, l) for l in range(lev_lo, lev_hi + 1)) * k
            rec = np.zeros(max(cap, 1), np.uint32)
            p.knn_record = rec.ctypes.data_as(C.c_void_p).value
            p.knn_record_cap = cap
        st = np.zeros(lev_hi - lev_lo + 1, LOSSY_STATS_DTYPE)
        self.ctx.check(self.ctx.L.svd_lossy_merge(self.h, C.byref(p), st.ctypes.data_as(C.c_void_p), st.size))
        return (st, rec) if record_knn else st

    def lossy_signatures(self, lev: int, sig_ptr: int):
        """512-bit occupancies (16 u32 per node) of DAG level lev into a device buffer."""
        self.ctx.check(self.ctx.L.svd_lossy_signatures(self.h, lev, C.c_void_p(sig_ptr)))


class MLP:
    """The learned candidate encoder (svd_mlp): 512 -> 256 -> 128 -> 32, bf16 MFMA."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(ctx.L.svd_mlp_create(ctx.h, seed, C.byref(h)))
        self.h = h

    def free(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.L.svd_mlp_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, np.float32)
        self.ctx.check(self.ctx.L.svd_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p), p.size))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.ctx.check(self.ctx.L.svd_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p), p.size))

    def embed(self, sig_ptr: int, n: int, emb_ptr: int):
        self.ctx.check(self.ctx.L.svd_mlp_embed(self.h, C.c_void_p(sig_ptr), n, C.c_void_p(emb_ptr)))

    def loss_grad(self, sig_ptr: int, labels_ptr: int, npairs: int):
        loss = C.c_float()
        g = np.zeros(MLP_PARAMS, np.float32)
        self.ctx.check(self.ctx.L.svd_mlp_loss_grad(self.h, C.c_void_p(sig_ptr), C.c_void_p(labels_ptr), npairs,
                                                    C.byref(loss), g.ctypes.data_as(C.c_void_p)))
        return loss.value, g

    def train(self, octree: Octree, lev_lo: int, lev_hi: int, pairs_per_step: int = 2048, steps: int = 1024,
              lr: float = 1e-3, seed: int = 11, near_pct: int = 50) -> TrainStats:
        p = TrainParams(lev_lo, lev_hi, pairs_per_step, steps, lr, near_pct, seed)
        st = TrainStats()
        self.ctx.check(self.ctx.L.svd_mlp_train(self.h, octree.h, C.byref(p), C.byref(st)))
        return st


class Scene:
    """An encoded file resident on the device, ready for traversal (svd_scene)."""

    def __init__(self, ctx: Context, handle, fmt: int, keep=None):
        self.ctx, self.h, self.fmt, self._keep = ctx, handle, fmt, keep

    @classmethod
    def from_bytes(cls, ctx: Context, fmt: int, data: bytes):
        buf = np.frombuffer(data, np.uint8)
        h = C.c_void_p()
        ctx.check(ctx.L.svd_scene_create(ctx.h, fmt, buf.ctypes.data_as(C.c_void_p), buf.size, 0, C.byref(h)))
        return cls(ctx, h, fmt)

    @classmethod
    def from_octree(cls, octree: Octree, fmt: int):
        h = C.c_void_p()
        octree.ctx.check(octree.ctx.L.svd_scene_from_octree(octree.h, fmt, C.byref(h)))
        return cls(octree.ctx, h, fmt, keep=octree)

    def free(self):
        if self.h and self.ctx.h:
            self.ctx.L.svd_scene_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def trace_beam_device(self, params: TraceParams, depth_projection_factor: float, prim_ptr: int,
                          shad_ptr: int | None, min_depth_ptr: int | None = None):
        """Beam-optimised frame (depth pre-pass at 1/8 resolution) into device buffers."""
        self.ctx.check(self.ctx.L.svd_trace_beam(self.h, C.byref(params), float(depth_projection_factor),
                                                 C.c_void_p(prim_ptr), C.c_void_p(shad_ptr) if shad_ptr else None,
                                                 C.c_void_p(min_depth_ptr) if min_depth_ptr else None))

    def render_device(self, params: TraceParams, view: ViewParams, rgb_ptr: int, beam_projection_factor: float = 0.0,
                      prim_ptr: int | None = None):
        """The viewer's colour per pixel (3 floats) into a device buffer; beam pre-pass when the factor is > 0."""
        self.ctx.check(self.ctx.L.svd_render(self.h, C.byref(params), C.byref(view), float(beam_projection_factor),
                                             C.c_void_p(rgb_ptr), C.c_void_p(prim_ptr) if prim_ptr else None))

    def ao_device(self, params: TraceParams, num_rays: int, vis_ptr: int, prim_ptr: int | None = None):
        """AO visibility per pixel (float, -1 without a primary hit) into a device buffer."""
        self.ctx.check(self.ctx.L.svd_trace_ao(self.h, C.byref(params), int(num_rays), C.c_void_p(vis_ptr),
                                               C.c_void_p(prim_ptr) if prim_ptr else None))

    def trace_hits_device(self, params: TraceParams, prim_ptr: int, shad_ptr: int | None):
        """Trace one frame without instrumentation: 4 floats (t, level, iters, node) per
        ray; shadow entries are written only where a shadow ray is cast."""
        self.ctx.check(self.ctx.L.svd_trace_hits(self.h, C.byref(params), C.c_void_p(prim_ptr),
                                                 C.c_void_p(shad_ptr) if shad_ptr else None))

    def trace_device(self, params: TraceParams, prim_ptr: int, shad_ptr: int | None):
        """Trace one frame into device buffers (e.g. torch tensors' data_ptr())."""
        self.ctx.check(self.ctx.L.svd_trace(self.h, C.byref(params), C.c_void_p(prim_ptr),
                                            C.c_void_p(shad_ptr) if shad_ptr else None))
