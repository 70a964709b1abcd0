"""ctypes wrapper of liboracle_svdag.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference path (svdag_oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
it, and only as the checker / CPU baseline, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_svdag.so")
NULL_NODE = 0xFFFFFFFE
_lib = None


def build() -> None:
    out = subprocess.run(["make", "-C", _HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


class TraceParams(C.Structure):
    _fields_ = [("view_inv", C.c_float * 16), ("proj_inv", C.c_float * 16),
                ("width", C.c_uint32), ("height", C.c_uint32), ("max_iters", C.c_uint32),
                ("draw_level", C.c_uint32), ("projection_factor", C.c_float),
                ("light_pos", C.c_float * 3), ("shadows", C.c_uint32), ("pad", C.c_uint32)]


class LossyParams(C.Structure):
    """Leading fields of svd_lossy_params (include/svdag_amd.h)."""
    _fields_ = [("lev_lo", C.c_uint32), ("lev_hi", C.c_uint32), ("max_diff", C.c_uint32),
                ("max_refs", C.c_uint32), ("list_cap", C.c_uint32), ("diff_factor", C.c_float),
                ("proposer", C.c_uint32)]


LOSSY_STATS_DTYPE = np.dtype([(f, "<u8") for f in ("level", "bound", "nodes_before", "nodes_after", "candidates",
                                                   "listed", "verified", "merged", "error_bound")])

class ViewParams(C.Structure):
    """Layout of svd_view_params (render mode, normal samples, selection, colour switches)."""
    _fields_ = [("render_mode", C.c_int32), ("norm_samples", C.c_int32), ("selected_voxel", C.c_uint32),
                ("random_colors", C.c_uint32), ("freq_colors", C.c_uint32), ("pad", C.c_uint32)]


RAY_DTYPE = np.dtype([("t", "<f4"), ("level", "<f4"), ("iters", "<f4"), ("node", "<f4"),
                      ("path", "<u4"), ("bytes", "<u4")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        sig = {
            "orc_new": (vp, []), "orc_free": (None, [vp]),
            "orc_build_morton": (i32, [vp, vp, u64, u32]),
            "orc_build_mesh": (i32, [vp, vp, u64, u32, vp]),
            "orc_build_points": (i32, [vp, vp, u64, u32, vp]),
            "orc_build_dag_mesh": (i32, [vp, vp, u64, u32, u32, vp]),
            "orc_to_dag": (i32, [vp]), "orc_to_sdag": (i32, [vp]),
            "orc_tri_box": (i32, [vp, C.c_double, vp]),
            "orc_build_dag_morton": (i32, [vp, vp, u64, u32, u32]),
            "orc_omp_threads": (i32, []),
            "orc_sdag_map": (C.c_int64, [vp, u32, vp, u64]),
            "orc_init_child_levels": (None, [vp]), "orc_merge_cross_levels": (i32, [vp]),
            "orc_levels": (u32, [vp]), "orc_state": (i32, [vp]), "orc_level_size": (u64, [vp, u32]),
            "orc_nnodes": (u64, [vp]), "orc_nvoxels": (u64, [vp]), "orc_nnodes_svo": (u64, [vp]),
            "orc_cross_merged": (u64, [vp]), "orc_root_side": (C.c_double, [vp]),
            "orc_get_level": (i32, [vp, u32, vp, vp, vp]),
            "orc_set_level": (i32, [vp, u32, vp, u64]),
            "orc_init_levels": (i32, [vp, u32, i32]),
            "orc_encode": (C.c_int64, [vp, i32, vp, u64]),
            "orc_std_sort_refcount": (None, [vp, vp, u64]),
            "orc_trace_frame": (i32, [vp, u64, i32, C.POINTER(TraceParams), vp, vp, u32, u32]),
            "orc_trace_frame_beam": (i32, [vp, u64, i32, C.POINTER(TraceParams), C.c_float, vp, vp, u32, u32, vp]),
            "orc_render_frame": (i32, [vp, u64, i32, C.POINTER(TraceParams), C.POINTER(ViewParams), C.c_float,
                                       vp, vp, u32, u32]),
            "orc_trace_ao": (i32, [vp, u64, i32, C.POINTER(TraceParams), u32, vp, vp, u32, u32]),
            "orc_hemisphere_samples": (None, [vp]),
            "orc_vm_math": (None, [vp, u64, vp, vp, vp]),
            "orc_attr_bittree_encode": (C.c_int64, [vp, u64, i32, vp, u64]),
            "orc_attr_bittree_decode": (i32, [vp, u64, vp, u64]),
            "orc_lossy_merge": (i32, [vp, C.POINTER(LossyParams), vp, u32, vp, u32]),
            "orc_lossy_signatures": (i32, [vp, u32, vp]),
            "orc_ranked_lists": (i32, [vp, u64, u32, u32, vp, vp]),
            "orc_effective_refcounts": (i32, [vp, u32, vp]),
            "orc_subtree_diff": (u64, [vp, u32, u32, u32, u64]),
            "orc_knn_lsh": (None, [vp, u64, u32, u32, u32, u32, u32, u64, vp, vp]),
            "orc_knn_refine": (None, [vp, u64, u32, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleOctree:
    """GeomOctree restated on the CPU (std::map dedup, libstdc++ std::sort)."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.orc_new()

    def __del__(self):
        try:
            if self.h:
                self.L.orc_free(self.h)
                self.h = None
        except Exception:
            pass

    @classmethod
    def from_morton(cls, codes: np.ndarray, levels: int):
        o = cls()
        codes = np.ascontiguousarray(codes, dtype=np.uint64)
        rc = o.L.orc_build_morton(o.h, _p(codes), codes.size, levels)
        assert rc == 0, rc
        return o

    @classmethod
    def from_mesh(cls, tris: np.ndarray, levels: int, bbox_min, bbox_max):
        """tris: float32 [T, 3, 3]; bbox in double (svbuilder/main.cpp:150-155)."""
        o = cls()
        t = np.ascontiguousarray(tris, dtype=np.float32)
        bb = np.array(list(bbox_min) + list(bbox_max), dtype=np.float64)
        rc = o.L.orc_build_mesh(o.h, _p(t), t.shape[0], levels, _p(bb))
        assert rc == 0, rc
        return o

    @classmethod
    def from_points(cls, xyz: np.ndarray, levels: int, bbox_min, bbox_max):
        """buildSVOFromPoints (geom_octree.cpp:62-169): points float64 [N, 3] in file order."""
        o = cls()
        pt = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        bb = np.array(list(bbox_min) + list(bbox_max), dtype=np.float64)
        rc = o.L.orc_build_points(o.h, _p(pt), pt.shape[0], levels, _p(bb))
        assert rc == 0, rc
        return o

    @classmethod
    def from_mesh_subtrees(cls, tris: np.ndarray, levels: int, step_level: int, bbox_min, bbox_max):
        """GeomOctree::buildDAG (geom_octree.cpp:289-435): subtree builder, ends in the DAG state."""
        o = cls()
        t = np.ascontiguousarray(tris, dtype=np.float32)
        bb = np.array(list(bbox_min) + list(bbox_max), dtype=np.float64)
        rc = o.L.orc_build_dag_mesh(o.h, _p(t), t.shape[0], levels, step_level, _p(bb))
        assert rc == 0, rc
        return o

    @classmethod
    def from_morton_subtrees(cls, codes: np.ndarray, levels: int, step_level: int):
        """GeomOctree::buildDAG's OpenMP subtree builder on a voxel list
        (geom_octree.cpp:289-435): ends in the DAG state."""
        o = cls()
        codes = np.ascontiguousarray(codes, dtype=np.uint64)
        rc = o.L.orc_build_dag_morton(o.h, _p(codes), codes.size, levels, step_level)
        assert rc == 0, rc
        return o

    @classmethod
    def from_levels(cls, level_words, state: int):
        o = cls()
        o.L.orc_init_levels(o.h, len(level_words), state)
        for l, w in enumerate(level_words):
            w = np.ascontiguousarray(w, dtype=np.uint32).reshape(-1, 9)
            o.L.orc_set_level(o.h, l, _p(w), w.shape[0])
        return o

    def to_dag(self):
        rc = self.L.orc_to_dag(self.h)
        assert rc == 0, rc

    def to_sdag(self):
        rc = self.L.orc_to_sdag(self.h)
        assert rc == 0, rc

    def init_child_levels(self):
        self.L.orc_init_child_levels(self.h)

    def merge_cross_levels(self) -> int:
        return self.L.orc_merge_cross_levels(self.h)

    @property
    def levels(self):
        return self.L.orc_levels(self.h)

    @property
    def n_nodes(self):
        return self.L.orc_nnodes(self.h)

    @property
    def n_voxels(self):
        return self.L.orc_nvoxels(self.h)

    @property
    def cross_merged(self):
        return self.L.orc_cross_merged(self.h)

    def level_size(self, lev):
        return self.L.orc_level_size(self.h, lev)

    def level(self, lev, with_inv=False, with_child_levels=False):
        n = self.level_size(lev)
        w = np.zeros((n, 9), np.uint32)
        inv = np.zeros(n, np.uint8)
        cl = np.zeros((n, 8), np.uint32)
        if n:
            self.L.orc_get_level(self.h, lev, _p(w), _p(inv), _p(cl))
        out = [w]
        if with_inv:
            out.append(inv)
        if with_child_levels:
            out.append(cl)
        return out[0] if len(out) == 1 else tuple(out)

    def lossy_merge(self, lev_lo, lev_hi, max_diff=0, diff_factor=1.0, max_refs=1, list_cap=32,
                    proposals=None, k=16, ranked=False):
        """Lossy DAG with the exhaustive proposer (ranked: the list_cap nearest
        earlier nodes within the bound), or with given per-level kNN lists
        (concatenated n_l x k arrays, levels hi..lo).  Returns per-level stats."""
        p = LossyParams(lev_lo, lev_hi, max_diff, max_refs, list_cap, diff_factor, 2 if ranked else 0)
        st = np.zeros(64, LOSSY_STATS_DTYPE)
        prop = None if proposals is None else np.ascontiguousarray(proposals, dtype=np.uint32)
        rc = self.L.orc_lossy_merge(self.h, C.byref(p), _p(prop) if prop is not None else None, k, _p(st), 64)
        assert rc == 0, rc
        return st[: lev_hi - lev_lo + 1]

    def sdag_map(self, lev):
        """toSDAG's correspondence of every DAG node of level lev: (sdag id, mirror xyz bits)."""
        n = self.L.orc_sdag_map(self.h, lev, None, 0)
        assert n >= 0, n
        m = np.zeros(n, np.uint32)
        if n:
            self.L.orc_sdag_map(self.h, lev, _p(m), n)
        return m & np.uint32(0x1FFFFFFF), (m >> np.uint32(29)).astype(np.uint8)

    def lossy_signatures(self, lev):
        sig = np.zeros((self.level_size(lev), 16), np.uint32)
        assert self.L.orc_lossy_signatures(self.h, lev, _p(sig)) == 0
        return sig

    def effective_refcounts(self, lev):
        r = np.zeros(self.level_size(lev), np.uint64)
        assert self.L.orc_effective_refcounts(self.h, lev, _p(r)) == 0
        return r

    def subtree_diff(self, lev, a, b, cap):
        return int(self.L.orc_subtree_diff(self.h, lev, a, b, cap))

    def levels_words(self):
        return [self.level(l) for l in range(self.levels)]

    def encode(self, fmt: int) -> bytes:
        n = self.L.orc_encode(self.h, fmt, None, 0)
        assert n >= 0, n
        buf = np.zeros(n, np.uint8)
        n2 = self.L.orc_encode(self.h, fmt, _p(buf), n)
        assert n2 == n
        return buf.tobytes()


def omp_threads() -> int:
    """Threads of the oracle's OpenMP loops (OMP_NUM_THREADS or all cores)."""
    return int(lib().orc_omp_threads())


def tri_box(center, half: float, tri) -> bool:
    """testTriBox (test_triangle_box.cpp:105-184): tri = 3 x 3 float32, cube centre (double) and half side."""
    c = np.ascontiguousarray(center, dtype=np.float64)
    t = np.ascontiguousarray(tri, dtype=np.float32).reshape(9)
    return bool(lib().orc_tri_box(_p(c), float(half), _p(t)))


def std_sort_refcount(idx: np.ndarray, val: np.ndarray):
    """libstdc++ std::sort of (idx, val) pairs by val descending (encoded_ssvdag.cpp:273-275)."""
    i = np.ascontiguousarray(idx, dtype=np.uint32).copy()
    v = np.ascontiguousarray(val, dtype=np.uint32).copy()
    lib().orc_std_sort_refcount(_p(i), _p(v), i.size)
    return i, v


def trace_frame(file_bytes: bytes, fmt: int, params, shadows: bool = True, row0: int = 0, nrows: int | None = None):
    """CPU port of the octree_dda shader over an encoded file (fmt 0/1/2)."""
    p = TraceParams()
    C.memmove(C.addressof(p), C.addressof(params), C.sizeof(TraceParams))
    n = p.width * p.height
    prim = np.full(n, 0, RAY_DTYPE)
    shad = np.full(n, 0, RAY_DTYPE) if shadows else None
    buf = np.frombuffer(file_bytes, np.uint8)
    f = 2 if fmt == 3 else fmt
    rc = lib().orc_trace_frame(_p(buf), buf.size, f, C.byref(p), _p(prim), _p(shad) if shad is not None else None,
                               row0, p.height if nrows is None else nrows)
    assert rc == 0, rc
    return prim, shad


def trace_frame_beam(file_bytes: bytes, fmt: int, params, depth_projection_factor: float, shadows: bool = True):
    """trace_frame with the viewer's minimum-depth pre-pass (octree_dda_renderer.cpp:389-408):
    returns (primary, shadow, min-depth texture [height/8, width/8])."""
    p = TraceParams()
    C.memmove(C.addressof(p), C.addressof(params), C.sizeof(TraceParams))
    n = p.width * p.height
    prim = np.full(n, 0, RAY_DTYPE)
    shad = np.full(n, 0, RAY_DTYPE) if shadows else None
    tex = np.zeros((p.height // 8, p.width // 8), np.float32)
    buf = np.frombuffer(file_bytes, np.uint8)
    f = 2 if fmt == 3 else fmt
    rc = lib().orc_trace_frame_beam(_p(buf), buf.size, f, C.byref(p), C.c_float(depth_projection_factor), _p(prim),
                                    _p(shad) if shad is not None else None, 0, p.height, _p(tex))
    assert rc == 0, rc
    return prim, shad, tex


def _copy_params(params) -> TraceParams:
    p = TraceParams()
    C.memmove(C.addressof(p), C.addressof(params), C.sizeof(TraceParams))
    return p


def render_frame(file_bytes: bytes, fmt: int, params, view, beam_projection_factor: float = 0.0,
                 row0: int = 0, nrows: int | None = None):
    """VIEWER_MODE colour of rows [row0, row0 + nrows) (octree_dda.frag.glsl:671-858):
    returns (rgb [rows, width, 3] float32, primary rays)."""
    p = _copy_params(params)
    v = ViewParams()
    C.memmove(C.addressof(v), C.addressof(view), C.sizeof(ViewParams))
    rows = p.height - row0 if nrows is None else min(nrows, p.height - row0)
    rgb = np.zeros((rows, p.width, 3), np.float32)
    prim = np.zeros(rows * p.width, RAY_DTYPE)
    buf = np.frombuffer(file_bytes, np.uint8)
    f = 2 if fmt == 3 else fmt
    rc = lib().orc_render_frame(_p(buf), buf.size, f, C.byref(p), C.byref(v), C.c_float(beam_projection_factor),
                                _p(rgb), _p(prim), row0, rows)
    assert rc == 0, rc
    return rgb, prim


def trace_ao(file_bytes: bytes, fmt: int, params, num_rays: int, row0: int = 0, nrows: int | None = None):
    """AO_MODE visibility of rows [row0, row0 + nrows) over the primary-ray G-buffer:
    returns (visibility [rows, width] float32, primary rays)."""
    p = _copy_params(params)
    rows = p.height - row0 if nrows is None else min(nrows, p.height - row0)
    vis = np.zeros((rows, p.width), np.float32)
    prim = np.zeros(rows * p.width, RAY_DTYPE)
    buf = np.frombuffer(file_bytes, np.uint8)
    f = 2 if fmt == 3 else fmt
    rc = lib().orc_trace_ao(_p(buf), buf.size, f, C.byref(p), int(num_rays), _p(vis), _p(prim), row0, rows)
    assert rc == 0, rc
    return vis, prim


def hemisphere_samples() -> np.ndarray:
    """The 256 cosine-weighted Halton directions of OctreeDDARenderer::generateHSSamples, [256, 3]."""
    out = np.zeros((256, 3), np.float32)
    lib().orc_hemisphere_samples(_p(out))
    return out


def vm_math(x: np.ndarray):
    """The oracle's sin, cos and log2 definitions over a float32 array."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    sn, cs, lg = (np.zeros_like(x) for _ in range(3))
    lib().orc_vm_math(_p(x), x.size, _p(sn), _p(cs), _p(lg))
    return sn, cs, lg


def attr_bittree_encode(attrs: np.ndarray, gray: bool) -> bytes:
    """Gray-code / binary attribute bit-tree of a per-voxel u8 sequence (voxel Morton order)."""
    a = np.ascontiguousarray(attrs, dtype=np.uint8)
    n = lib().orc_attr_bittree_encode(_p(a), a.size, int(gray), None, 0)
    buf = np.zeros(n, np.uint8)
    lib().orc_attr_bittree_encode(_p(a), a.size, int(gray), _p(buf), n)
    return buf.tobytes()


def attr_bittree_decode(blob: bytes, n: int) -> np.ndarray:
    b = np.frombuffer(blob, np.uint8)
    out = np.zeros(max(n, 1), np.uint8)
    rc = lib().orc_attr_bittree_decode(_p(b), b.size, _p(out), out.size)
    assert rc == 0, rc
    return out[:n]


def knn_lsh(emb: np.ndarray, k: int = 16, tables: int = 4, bits: int = 16, window: int = 32, seed: int = 7,
            sig: np.ndarray | None = None):
    """The learned proposer's LSH neighbour search, restated sequentially.
    With sig (n x 16 u32), candidates are scored by signature Hamming distance."""
    e = np.ascontiguousarray(emb, dtype=np.float32)
    n, dim = e.shape
    out = np.zeros((n, k), np.uint32)
    s = None if sig is None else np.ascontiguousarray(sig, dtype=np.uint32)
    lib().orc_knn_lsh(_p(e), n, dim, k, tables, bits, window, seed, _p(out), _p(s) if s is not None else None)
    return out


def knn_refine(sig: np.ndarray, lists: np.ndarray, rounds: int = 1) -> np.ndarray:
    """Neighbour-of-neighbour rounds on the signature Hamming metric."""
    s = np.ascontiguousarray(sig, dtype=np.uint32)
    cur = np.ascontiguousarray(lists, dtype=np.uint32)
    n, k = cur.shape
    for _ in range(rounds):
        out = np.zeros_like(cur)
        lib().orc_knn_refine(_p(s), n, k, _p(cur), _p(out))
        cur = out
    return cur


def ranked_lists(sig: np.ndarray, bound: int, cap: int):
    """The ranked proposer's lists of a signature array (n x 16 uint32): per
    node its cap nearest earlier nodes within the bound by (Hamming, index)."""
    sig = np.ascontiguousarray(sig, dtype=np.uint32)
    n = sig.shape[0]
    lists = np.full((max(n, 1), cap), 0xFFFFFFFF, np.uint32)
    counts = np.zeros(max(n, 1), np.uint32)
    assert lib().orc_ranked_lists(_p(sig), n, bound, cap, _p(lists), _p(counts)) == 0
    return lists[:n], counts[:n]
