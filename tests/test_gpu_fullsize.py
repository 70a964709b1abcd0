"""Parity at the workloads' full sizes, on the GPU, against the oracle run on
the host of the same machine.

  * 2048^3 fBm terrain (11 levels, 13.7M voxels): every DAG and SDAG level and
    all four files (.svdag, .esvdag, .ussvdag, .ssvdag) byte for byte
    (encoded_svdag.cpp:76-199, encoded_ussvdag.cpp:60-170,
    encoded_ssvdag.cpp:194-466), and the whole 1920x1080 frame with its shadow
    rays, plain and with the 1/8 depth pre-pass, ray for ray
    (octree_dda.frag.glsl:539-581, 803-820).
  * 4096^3 terrain union caves (12 levels, 148M voxels): every level and all
    four files byte for byte; the cross-level merge (levels, child levels and
    the -multi .svdag, geom_octree_extension.cpp:1192-1544); the 3840x2160
    frame with shadow rays, ray for ray.
  * 4096^3 caves (12 levels): the lossy merge of levels 6-9 at a 10-voxel
    bound changes the voxel set by no more than its accounted error
    (sum of voxel difference x effective references over the merges).

Each test builds its scene on the device and reads the voxel codes back for
the oracle.  The scenes are large, so every test announces its case first.
"""
import numpy as np
import pytest

import oracle as O
import svdag_amd as S
from decode import decode_words

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

FILES = ((S.FMT_SVDAG, 0), (S.FMT_ESVDAG, 2))
SYM_FILES = ((S.FMT_USSVDAG, 1), (S.FMT_SSVDAG, 2))


def say(msg):
    print(msg, flush=True)


def compare_levels(g, o, view, L, what, with_inv=False):
    for l in range(L):
        if with_inv:
            w_o, inv_o = o.level(l, with_inv=True)
            w_g, inv_g = g.level(view, l, with_inv=True)
            if l > 0:                                     # the root is never reduced
                np.testing.assert_array_equal(inv_g, inv_o, err_msg=f"{what} level {l} invariance bits")
        else:
            w_o, w_g = o.level(l), g.level(view, l)
        np.testing.assert_array_equal(w_g, w_o, err_msg=f"{what} level {l}")


def lossless_parity(ctx, dev, L, name):
    """GPU octree from device codes vs the oracle from the same codes read
    back: DAG levels + .svdag/.esvdag, then SDAG levels + .ussvdag/.ssvdag."""
    codes = dev.numpy()
    say(f"[case] {name}: {codes.size} voxels, {L} levels")
    o = O.OracleOctree.from_morton(codes, L)
    g = S.Octree.from_morton(ctx, (dev.ptr, dev.n), L)
    assert g.counts()[1] == o.n_voxels == codes.size
    o.to_dag()
    o.init_child_levels()
    g.to_dag()
    assert g.counts()[0] == o.n_nodes
    compare_levels(g, o, S.VIEW_DAG, L, "DAG")
    for fmt, ofmt in FILES:
        g.encode(fmt)
        assert g.encoded_bytes(fmt) == o.encode(ofmt), f"{name}: {S.FMT_EXT[fmt]} differs"
        say(f"  {S.FMT_EXT[fmt]} identical")
    o.to_sdag()
    g.to_sdag()
    assert g.counts()[0] == o.n_nodes
    compare_levels(g, o, S.VIEW_SDAG, L, "SDAG", with_inv=True)
    files = {}
    for fmt, ofmt in SYM_FILES:
        g.encode(fmt)
        files[fmt] = g.encoded_bytes(fmt)
        assert files[fmt] == o.encode(ofmt), f"{name}: {S.FMT_EXT[fmt]} differs"
        say(f"  {S.FMT_EXT[fmt]} identical")
    # the four images again in one call with shared passes (the bench's build step)
    want = {f: g.encoded_bytes(f) for f in (S.FMT_SVDAG, S.FMT_ESVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG)}
    g.encode_many(list(want))
    for f, data in want.items():
        assert g.encoded_bytes(f) == data, f"{name}: {S.FMT_EXT[f]} differs when encoded with the others"
    say("  shared-pass encode of the four images identical")
    return g, files


def frame_parity(ctx, g, fmt, file_bytes, p, name, beam_factor=0.0):
    """The GPU frame over the octree's encoded buffer vs the shader port over
    the same file, every primary and shadow ray."""
    import torch
    n = p.width * p.height
    sc = S.Scene.from_octree(g, fmt)
    prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    shad = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    if beam_factor > 0:
        sc.trace_beam_device(p, beam_factor, prim.data_ptr(), shad.data_ptr())
        want_p, want_s, _ = O.trace_frame_beam(file_bytes, fmt, p, beam_factor)
    else:
        sc.trace_device(p, prim.data_ptr(), shad.data_ptr())
        want_p, want_s = O.trace_frame(file_bytes, fmt, p, shadows=True)
    ctx.sync()
    got_p = prim.cpu().numpy().view(S.RAY_DTYPE)
    got_s = shad.cpu().numpy().view(S.RAY_DTYPE)
    if beam_factor <= 0:
        hits_parity(sc, p, want_p, want_s, name)
    sc.free()
    hits = int((want_p["t"] >= 0).sum())
    say(f"  {name}: {n} primary rays, {hits} hits / shadow rays")
    assert hits > 0.05 * n
    np.testing.assert_array_equal(got_p, want_p, err_msg=f"{name}: primary rays")
    np.testing.assert_array_equal(got_s, want_s, err_msg=f"{name}: shadow rays")


SENTINEL = 0x7FC0DEAD   # a NaN pattern no traversal result takes


def hits_parity(sc, p, want_p, want_s, name):
    """The uninstrumented frame (svd_trace_hits): the same (t, level, iters,
    node) per primary ray, the shadow vec4 exactly where a shadow ray was cast
    and the caller's contents elsewhere."""
    import torch
    n = p.width * p.height
    prim = torch.full((n * 4,), SENTINEL, dtype=torch.int32, device="cuda")
    shad = torch.full((n * 4,), SENTINEL, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sc.trace_hits_device(p, prim.data_ptr(), shad.data_ptr())
    sc.ctx.sync()
    gp = prim.cpu().numpy().reshape(n, 4)
    gs = shad.cpu().numpy().reshape(n, 4)
    f4 = lambda r: np.stack([r[k] for k in ("t", "level", "iters", "node")], axis=1).astype(np.float32).view(np.int32)
    np.testing.assert_array_equal(gp, f4(want_p), err_msg=f"{name}: uninstrumented primary rays")
    cast = want_p["t"] >= 0 if p.shadows else np.zeros(n, bool)
    np.testing.assert_array_equal(gs[cast], f4(want_s)[cast], err_msg=f"{name}: uninstrumented shadow rays")
    assert (gs[~cast] == SENTINEL).all(), f"{name}: shadow entries written without a shadow ray"


@pytest.fixture(scope="module")
def terrain2048(ctx):
    L = 11
    dev = ctx.terrain(L, seed=1)
    assert dev.n == 13703073                       # the headline workload's voxel count
    g, files = lossless_parity(ctx, dev, L, "terrain 2048^3")
    yield L, g, files
    g.free()
    dev.free()


def test_terrain_2048_levels_and_files_byte_exact(terrain2048):
    L, g, files = terrain2048
    assert g.state == S.S_SDAG and len(files[S.FMT_SSVDAG]) > 0


def test_terrain_2048_frame_1080p_every_ray(ctx, terrain2048):
    import svdag_amd.workloads as W
    L, g, files = terrain2048
    p = W.trace_params(1920, 1080, seed=1, levels=L, shadows=True, light=(1.0, 1.0, 1.0))
    frame_parity(ctx, g, S.FMT_SSVDAG, files[S.FMT_SSVDAG], p, "1920x1080 on the 2048^3 .ssvdag")
    frame_parity(ctx, g, S.FMT_SSVDAG, files[S.FMT_SSVDAG], p, "1920x1080 with the depth pre-pass",
                 beam_factor=W.projection_factor(1080, screen_divisor=8.0))


def _union4096(ctx):
    L = 12
    t, c = ctx.terrain(L, seed=1), ctx.caves(L, seed=3)
    u = ctx.union(t, c)
    t.free()
    c.free()
    return L, u


def test_union_4096_levels_and_files_byte_exact(ctx):
    import svdag_amd.workloads as W
    L, u = _union4096(ctx)
    g, files = lossless_parity(ctx, u, L, "terrain + caves 4096^3")
    # the 3840x2160 frame of the same scene, on the file just shown identical
    p = W.trace_params(3840, 2160, seed=5, levels=L, shadows=True)
    frame_parity(ctx, g, S.FMT_SSVDAG, files[S.FMT_SSVDAG], p, "3840x2160 on the 4096^3 .ssvdag")
    g.free()
    u.free()


def test_union_4096_cross_level_merge_byte_exact(ctx):
    L, u = _union4096(ctx)
    codes = u.numpy()
    say(f"[case] cross-level merge, terrain + caves 4096^3: {codes.size} voxels")
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    assert o.merge_cross_levels() == 0
    g = S.Octree.from_morton(ctx, (u.ptr, u.n), L)
    g.to_dag()
    assert g.merge_cross_levels() == o.cross_merged > 0
    for l in range(L):
        w_o, cl_o = o.level(l, with_child_levels=True)
        w_g, cl_g = g.level(S.VIEW_DAG, l, with_child_levels=True)
        np.testing.assert_array_equal(w_g, w_o, err_msg=f"level {l}")
        if l < L - 1:
            np.testing.assert_array_equal(cl_g, cl_o, err_msg=f"childLevels {l}")
    g.encode(S.FMT_SVDAG)
    assert g.encoded_bytes(S.FMT_SVDAG) == o.encode(0)
    say(f"  {o.cross_merged} nodes eliminated; -multi .svdag identical")
    g.free()
    u.free()


def test_caves_4096_lossy_error_within_accounted_bound(ctx):
    """The 4096^3 cave volume, levels 6-9 merged at a 10-voxel bound on nodes
    with at most one effective reference (exhaustive proposer): the lossy
    DAG's voxel set differs from the input by at most the sum over merges of
    (voxel difference x effective references), and the merge did something."""
    L = 12
    dev = ctx.caves(L, seed=3)
    codes = dev.numpy()
    say(f"[case] lossy caves 4096^3: {codes.size} voxels")
    g = S.Octree.from_morton(ctx, (dev.ptr, dev.n), L)
    g.to_dag()
    before = g.counts()[0]
    st = g.lossy_merge(L - 6, L - 3, max_diff=10)
    merged, bound = int(st["merged"].sum()), int(st["error_bound"].sum())
    g.encode(S.FMT_SVDAG)
    vox = decode_words(g.encoded_bytes(S.FMT_SVDAG), False)
    diff = int(np.setxor1d(vox, codes, assume_unique=True).size)
    say(f"  {before} -> {g.counts()[0]} nodes, {merged} merged, voxel difference {diff} <= {bound}")
    assert merged > 100000 and g.counts()[0] < before
    assert 0 < diff <= bound
    g.free()
    dev.free()
