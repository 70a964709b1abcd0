"""GPU parity: the HIP path (through the C ABI) against the oracle.

Integer / byte work must be bit-exact: SVO / DAG / SDAG level arrays,
invariant bits, refcount order, and all four encoded file images.  The
traversal must return, for every ray, the same (t, level, iterations, node,
descent-path hash, fetched bytes) as the CPU port of the shader; t is FP32
and is required to be bit-identical (tolerance 0, stricter than the 1e-5
relative bound), because both sides execute the same FP32 operations with no
contraction.
"""
import numpy as np
import pytest

import oracle as O
from decode import decode_words, decode_ssvdag

import svdag_amd as S
import svdag_amd.workloads as W
from test_oracle import clustered_codes, random_codes

pytestmark = pytest.mark.gpu

SCENES = {
    "terrain7": lambda: (W.terrain_codes(7, seed=1), 7),
    "random6": lambda: (random_codes(6, 20000, 3), 6),
    "clustered7": lambda: (clustered_codes(7, 2), 7),
    "caves6": lambda: (W.cave_codes(6, seed=3), 6),
    "single": lambda: (W.sorted_codes(np.array([5]), np.array([9]), np.array([2])), 5),
    "full8": lambda: (W.sorted_codes(*[a.ravel() for a in np.meshgrid(np.arange(8), np.arange(8), np.arange(8))]), 3),
}


def oracle_pipeline(codes, L):
    o = O.OracleOctree.from_morton(codes, L)
    svo = o.levels_words()
    o.to_dag()
    dag = [o.level(l) for l in range(L)]
    o.init_child_levels()
    files = {S.FMT_SVDAG: o.encode(0)}
    if L >= 3:
        files[S.FMT_ESVDAG] = o.encode(2)
    o.to_sdag()
    sdag = [o.level(l, with_inv=True) for l in range(L)]
    files[S.FMT_USSVDAG] = o.encode(1)
    if L >= 3:
        files[S.FMT_SSVDAG] = o.encode(2)
    return svo, dag, sdag, files


def gpu_pipeline(ctx, codes, L):
    g = S.Octree.from_morton(ctx, codes, L)
    svo = [g.level(S.VIEW_SVO, l) for l in range(L)]
    g.to_dag()
    dag = [g.level(S.VIEW_DAG, l) for l in range(L)]
    files = {}
    g.encode(S.FMT_SVDAG)
    files[S.FMT_SVDAG] = g.encoded_bytes(S.FMT_SVDAG)
    if L >= 3:
        g.encode(S.FMT_ESVDAG)
        files[S.FMT_ESVDAG] = g.encoded_bytes(S.FMT_ESVDAG)
    g.to_sdag()
    sdag = [g.level(S.VIEW_SDAG, l, with_inv=True) for l in range(L)]
    g.encode(S.FMT_USSVDAG)
    files[S.FMT_USSVDAG] = g.encoded_bytes(S.FMT_USSVDAG)
    if L >= 3:
        g.encode(S.FMT_SSVDAG)
        files[S.FMT_SSVDAG] = g.encoded_bytes(S.FMT_SSVDAG)
    return g, svo, dag, sdag, files


@pytest.mark.parametrize("name", list(SCENES))
def test_build_levels_and_files_bit_exact(ctx, name):
    codes, L = SCENES[name]()
    print(f"[case] {name} L={L} voxels={codes.size}", flush=True)
    svo_o, dag_o, sdag_o, files_o = oracle_pipeline(codes, L)
    g, svo_g, dag_g, sdag_g, files_g = gpu_pipeline(ctx, codes, L)
    for l in range(L):
        np.testing.assert_array_equal(svo_g[l], svo_o[l], err_msg=f"SVO level {l}")
        np.testing.assert_array_equal(dag_g[l], dag_o[l], err_msg=f"DAG level {l}")
        np.testing.assert_array_equal(sdag_g[l][0], sdag_o[l][0], err_msg=f"SDAG level {l}")
        if l > 0:
            np.testing.assert_array_equal(sdag_g[l][1], sdag_o[l][1], err_msg=f"SDAG inv level {l}")
    for fmt, want in files_o.items():
        got = files_g[fmt]
        assert len(got) == len(want), (fmt, len(got), len(want))
        assert got == want, f"format {fmt} differs at byte {next(i for i in range(len(got)) if got[i] != want[i])}"


def test_refcount_order_unfinished_batch_reruns_checked(ctx, monkeypatch):
    """The SSVDAG-layout encoder runs the refcount order's device rounds
    without reading the list size and finds out with its layout read whether
    they were enough; if not it encodes again with the checked loop.  With no
    rounds past the size bound (SVD_REFSORT_EXTRA_ROUNDS=0, read when a
    context is created) the first pass is unfinished on these scenes, and the
    files must still equal the oracle's; so must those of encode_many, which
    reads the word encoders' sizes with the same read."""
    codes, L = W.terrain_codes(9, seed=2), 9
    _, _, _, files_o = oracle_pipeline(codes, L)
    monkeypatch.setenv("SVD_REFSORT_EXTRA_ROUNDS", "0")
    short = S.Context(0)
    monkeypatch.delenv("SVD_REFSORT_EXTRA_ROUNDS")
    try:
        for c in (ctx, short):
            _, _, _, _, files_g = gpu_pipeline(c, codes, L)
            for f in (S.FMT_SVDAG, S.FMT_ESVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG):
                assert files_g[f] == files_o[f], S.FMT_EXT[f]
            g = S.Octree.from_morton(c, codes, L)
            g.to_dag()
            g.to_sdag()
            fm = [S.FMT_SVDAG, S.FMT_ESVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG]
            g.encode_many(fm)
            for f in fm:
                assert g.encoded_bytes(f) == files_o[f], S.FMT_EXT[f]
            g.free()
    finally:
        short.close()


@pytest.mark.parametrize("limits", ["8192,4096,1024,4096", "1,1,1,1"])
def test_one_workgroup_level_limits(ctx, monkeypatch, limits):
    """The top levels of each pass go to one-workgroup kernels up to a size
    limit per pass (SVD_SMALL_LEVELS, read when a context is created); at the
    compiled maxima and with every such kernel off (limit 1) the levels and
    files equal the oracle's."""
    codes, L = W.terrain_codes(9, seed=2), 9
    _, dag_o, sdag_o, files_o = oracle_pipeline(codes, L)
    monkeypatch.setenv("SVD_SMALL_LEVELS", limits)
    c = S.Context(0)
    monkeypatch.delenv("SVD_SMALL_LEVELS")
    try:
        g, _, dag_g, sdag_g, files_g = gpu_pipeline(c, codes, L)
        g.free()
        for l in range(L):
            np.testing.assert_array_equal(dag_g[l], dag_o[l], err_msg=f"DAG level {l}")
            np.testing.assert_array_equal(sdag_g[l][0], sdag_o[l][0], err_msg=f"SDAG level {l}")
        for f in (S.FMT_SVDAG, S.FMT_ESVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG):
            assert files_g[f] == files_o[f], S.FMT_EXT[f]
    finally:
        c.close()


def test_build_limits_and_autotune(ctx):
    """svd_set/get_build_limits round-trip; svd_autotune_build keeps one of its
    presets, and the build under the chosen limits gives the oracle's files."""
    import torch
    keep = ctx.build_limits
    try:
        ctx.build_limits = (1, 2, 3, 4)
        assert ctx.build_limits == (1, 2, 3, 4)
        codes, L = SCENES["terrain7"]()
        d = torch.from_numpy(codes.view(np.int64)).to("cuda")
        r = ctx.autotune_build(d.data_ptr(), codes.size, L, reps=2)
        presets = [(2048, 1024, 512, 2048), (8192, 4096, 1024, 4096), (4096, 2048, 1024, 4096),
                   (1024, 512, 256, 1024), (2048, 1024, 1024, 2048)]
        assert r["limits"] in presets and ctx.build_limits == r["limits"], r
        assert len(r["ms"]) == 5 and all(x > 0 for x in r["ms"]), r
        _, _, _, files_o = oracle_pipeline(codes, L)
        g, _, _, _, files_g = gpu_pipeline(ctx, codes, L)
        for f, want in files_o.items():
            assert files_g[f] == want, f
        g.free()
    finally:
        ctx.build_limits = keep


def test_tosdag_checked_rounds_equal_speculative(ctx, monkeypatch):
    """toSDAG runs two resolution rounds per level without host reads and
    checks convergence once at the end; SVD_TOSDAG_CHECKED=1 (read when a
    context is created) checks after every round instead.  Both give the
    oracle's SDAG and files."""
    codes, L = W.terrain_codes(9, seed=2), 9
    _, _, sdag_o, files_o = oracle_pipeline(codes, L)
    monkeypatch.setenv("SVD_TOSDAG_CHECKED", "1")
    checked = S.Context(0)
    monkeypatch.delenv("SVD_TOSDAG_CHECKED")
    try:
        for c in (ctx, checked):
            g, _, _, sdag_g, files_g = gpu_pipeline(c, codes, L)
            g.free()
            for l in range(L):
                np.testing.assert_array_equal(sdag_g[l][0], sdag_o[l][0], err_msg=f"SDAG level {l}")
                if l > 0:
                    np.testing.assert_array_equal(sdag_g[l][1], sdag_o[l][1], err_msg=f"SDAG inv level {l}")
            for f in (S.FMT_USSVDAG, S.FMT_SSVDAG):
                assert files_g[f] == files_o[f]
    finally:
        checked.close()


def test_mesh_svo_reduction_bit_exact(ctx):
    """Config-1 style input: icosphere mesh voxelized in the reference's
    triangle-major order (oracle buildSVO), reduced and encoded on the GPU."""
    V, F = W.spheres_mesh(n_spheres=16, subdiv=2, seed=0)
    bmin, bmax = W.mesh_bbox(V)
    L = 7
    o = O.OracleOctree.from_mesh(V[F], L, bmin.astype(np.float64), bmax.astype(np.float64))
    svo = o.levels_words()
    rs = O.lib().orc_root_side(o.h)
    g = S.Octree.from_svo(ctx, svo, list(bmin) + list(bmax), rs)
    o.to_dag(); o.init_child_levels()
    g.to_dag()
    for l in range(L):
        np.testing.assert_array_equal(g.level(S.VIEW_DAG, l), o.level(l), err_msg=f"DAG level {l}")
    for fmt_o, fmt_g in ((0, S.FMT_SVDAG), (2, S.FMT_ESVDAG)):
        g.encode(fmt_g)
        assert g.encoded_bytes(fmt_g) == o.encode(fmt_o)
    o.to_sdag(); g.to_sdag()
    for fmt_o, fmt_g in ((1, S.FMT_USSVDAG), (2, S.FMT_SSVDAG)):
        g.encode(fmt_g)
        assert g.encoded_bytes(fmt_g) == o.encode(fmt_o)


def _killer(n):
    val = [-1] * n
    state = {"solid": 0, "cand": 0}
    import functools

    def cmp(x, y):
        if val[x] < 0 and val[y] < 0:
            if x == state["cand"]:
                val[x] = state["solid"]
            else:
                val[y] = state["solid"]
            state["solid"] += 1
        if val[x] < 0:
            state["cand"] = x
        elif val[y] < 0:
            state["cand"] = y
        return -1 if (val[x] >= 0 and (val[y] < 0 or val[x] > val[y])) else 1
    sorted(range(n), key=functools.cmp_to_key(cmp))
    return np.array([v if v >= 0 else state["solid"] for v in val], np.uint32)


@pytest.mark.parametrize("case", ["ones", "small", "mixed", "sorted", "big", "huge", "tiny"])
def test_refcount_order_equals_std_sort(ctx, case):
    rng = np.random.default_rng(hash(case) % 1000)
    n = {"ones": 300000, "small": 5000, "mixed": 777777, "sorted": 200000, "big": 2_500_000, "huge": 3000,
         "tiny": 17}[case]
    if case == "ones":
        vals = np.ones(n, np.uint32)
    elif case == "small":
        vals = rng.integers(0, 4, n).astype(np.uint32)
    elif case == "sorted":
        vals = np.sort(rng.integers(0, 9, n)).astype(np.uint32)
    elif case == "huge":
        vals = rng.integers(0, 2 ** 31, n).astype(np.uint32)
    else:
        vals = np.minimum(rng.geometric(0.55, n), 60).astype(np.uint32)
    idx = np.arange(n, dtype=np.uint32)
    wi, wv = O.std_sort_refcount(idx, vals)
    gi, gv = ctx.refcount_order(idx, vals)
    np.testing.assert_array_equal(gv, wv)
    np.testing.assert_array_equal(gi, wi)


def test_refcount_order_heap_fallback(ctx):
    """McIlroy adversary input drives libstdc++ into its heapsort fallback."""
    for n in (200, 3000, 9000):
        vals = _killer(n)
        idx = np.arange(n, dtype=np.uint32)
        wi, _ = O.std_sort_refcount(idx, vals)
        gi, _ = ctx.refcount_order(idx, vals)
        np.testing.assert_array_equal(gi, wi)


def test_refcount_order_ranges_equal_separate_sorts(ctx):
    """Ranges of every class (leaf, workgroup, device-wide, heap fallback, empty,
    single) sorted together equal one std::sort per range."""
    rng = np.random.default_rng(12)
    parts = [np.ones(70000, np.uint32),
             np.minimum(rng.geometric(0.55, 300000), 60).astype(np.uint32),
             rng.integers(0, 4, 1500).astype(np.uint32),
             _killer(9000),
             np.zeros(0, np.uint32),
             np.array([5], np.uint32),
             rng.integers(0, 3, 16).astype(np.uint32),
             np.minimum(rng.geometric(0.3, 5000), 9).astype(np.uint32),
             np.full(2049, 2, np.uint32)]
    vals = np.concatenate(parts)
    idx = rng.permutation(vals.size).astype(np.uint32)
    bounds = np.cumsum([0] + [p.size for p in parts])
    ranges = list(zip(bounds[:-1], bounds[1:]))
    gi, gv = ctx.refcount_order_ranges(idx, vals, ranges)
    for a, b in ranges:
        wi, wv = O.std_sort_refcount(idx[a:b], vals[a:b])
        np.testing.assert_array_equal(gv[a:b], wv)
        np.testing.assert_array_equal(gi[a:b], wi)


def test_refcount_order_ranges_many(ctx):
    """More ranges than the initial lists can take from kernel arguments (48):
    the lists are then uploaded by copies.  Each range still equals its own
    std::sort."""
    rng = np.random.default_rng(13)
    sizes = list(rng.integers(0, 3000, 66)) + [20000, 1, 17, 2048]
    parts = [np.minimum(rng.geometric(0.4, int(s)), 30).astype(np.uint32) for s in sizes]
    vals = np.concatenate(parts)
    idx = rng.permutation(vals.size).astype(np.uint32)
    bounds = np.cumsum([0] + [p.size for p in parts])
    ranges = list(zip(bounds[:-1], bounds[1:]))
    assert len(ranges) > 48
    gi, gv = ctx.refcount_order_ranges(idx, vals, ranges)
    for a, b in ranges:
        wi, wv = O.std_sort_refcount(idx[a:b], vals[a:b])
        np.testing.assert_array_equal(gv[a:b], wv)
        np.testing.assert_array_equal(gi[a:b], wi)


def test_refcount_order_ranges_bad_range(ctx):
    idx = np.arange(10, dtype=np.uint32)
    with pytest.raises(S.SvdError):
        ctx.refcount_order_ranges(idx, idx, [(0, 11)])
    with pytest.raises(S.SvdError):
        ctx.refcount_order_ranges(idx, idx, [(6, 5)])


@pytest.mark.parametrize("fmt", [S.FMT_SVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG])
def test_trace_matches_shader_port(ctx, fmt):
    codes, L = W.terrain_codes(8, seed=1), 8
    _, _, _, files = oracle_pipeline(codes, L)
    buf = files[fmt]
    p = W.trace_params(256, 192, seed=4, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
    want_p, want_s = O.trace_frame(buf, fmt, p, shadows=True)
    sc = S.Scene.from_bytes(ctx, fmt, buf)
    import torch
    n = p.width * p.height
    prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    shad = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sc.trace_device(p, prim.data_ptr(), shad.data_ptr())
    ctx.sync()
    got_p = prim.cpu().numpy().view(S.RAY_DTYPE)
    got_s = shad.cpu().numpy().view(S.RAY_DTYPE)
    assert (want_p["t"] >= 0).sum() > n // 8, "camera sees too little of the scene"
    for k in ("t", "level", "iters", "node", "path", "bytes"):
        np.testing.assert_array_equal(got_p[k], want_p[k], err_msg=f"primary {k}")
        np.testing.assert_array_equal(got_s[k], want_s[k], err_msg=f"shadow {k}")
    from test_gpu_fullsize import hits_parity
    hits_parity(sc, p, want_p, want_s, S.FMT_EXT[fmt])
    p.shadows = 0
    want_p0, _ = O.trace_frame(buf, fmt, p, shadows=False)
    hits_parity(sc, p, want_p0, want_s, S.FMT_EXT[fmt] + " without shadows")


@pytest.mark.parametrize("fmt", [S.FMT_SVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG])
@pytest.mark.parametrize("size", [(256, 192), (203, 121)])
def test_trace_beam_matches_shader_port(ctx, fmt, size):
    """The viewer's minimum-depth optimisation: the 1/8-resolution depth
    pre-pass texture and every primary/shadow ray equal the CPU port's, also
    when the frame is not a multiple of 8 (edge texels read as 0)."""
    import torch
    codes, L = W.terrain_codes(8, seed=1), 8
    _, _, _, files = oracle_pipeline(codes, L)
    buf = files[fmt]
    w, h = size
    p = W.trace_params(w, h, seed=4, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
    dp = W.projection_factor(h, screen_divisor=8.0)
    want_p, want_s, want_tex = O.trace_frame_beam(buf, fmt, p, dp, shadows=True)
    sc = S.Scene.from_bytes(ctx, fmt, buf)
    n = w * h
    prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    shad = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    tex = torch.zeros((h // 8) * (w // 8), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    sc.trace_beam_device(p, dp, prim.data_ptr(), shad.data_ptr(), tex.data_ptr())
    ctx.sync()
    np.testing.assert_array_equal(tex.cpu().numpy().reshape(h // 8, w // 8), want_tex)
    assert (want_tex > 0).mean() > 0.1
    got_p = prim.cpu().numpy().view(S.RAY_DTYPE)
    got_s = shad.cpu().numpy().view(S.RAY_DTYPE)
    for k in ("t", "level", "iters", "node", "path", "bytes"):
        np.testing.assert_array_equal(got_p[k], want_p[k], err_msg=f"primary {k}")
        np.testing.assert_array_equal(got_s[k], want_s[k], err_msg=f"shadow {k}")
    # the pre-pass only moves the start of primary rays: fewer iterations
    plain_p, _ = O.trace_frame(buf, fmt, p, shadows=False)
    assert want_p["iters"].mean() < plain_p["iters"].mean()


def test_trace_no_lod_all_formats(ctx):
    codes, L = clustered_codes(6, 8), 6
    _, _, _, files = oracle_pipeline(codes, L)
    p = W.trace_params(128, 96, seed=2, levels=L, shadows=True)
    p.projection_factor = 1e30
    import torch
    n = p.width * p.height
    for fmt in (S.FMT_SVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG):
        want_p, want_s = O.trace_frame(files[fmt], fmt, p, shadows=True)
        sc = S.Scene.from_bytes(ctx, fmt, files[fmt])
        prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
        shad = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        sc.trace_device(p, prim.data_ptr(), shad.data_ptr())
        ctx.sync()
        np.testing.assert_array_equal(prim.cpu().numpy().view(S.RAY_DTYPE), want_p)
        np.testing.assert_array_equal(shad.cpu().numpy().view(S.RAY_DTYPE), want_s)


@pytest.mark.parametrize("fmt", [S.FMT_SVDAG, S.FMT_USSVDAG])
def test_trace_flags_pointers_outside_the_scene(ctx, fmt):
    """A malformed file whose root child pointers lie past the node array:
    every fetch is bounds-checked (the word guard of the traversal), so a ray
    that follows such a pointer ends with t = -6 and the others are unchanged."""
    import torch
    codes, L = W.terrain_codes(7, seed=1), 7
    _, _, _, files = oracle_pipeline(codes, L)
    p = W.trace_params(160, 120, seed=4, levels=L, shadows=True)
    want_p, _ = O.trace_frame(files[fmt], fmt, p, shadows=True)
    buf = bytearray(files[fmt])
    words = np.frombuffer(buf, dtype=np.uint32, offset=44).copy()
    k = bin(int(words[0]) & 0xFF).count("1")
    assert k > 0
    words[1:1 + k] = np.uint32(0xFFFFFFF0)
    bad = bytes(buf[:44]) + words.tobytes()
    sc = S.Scene.from_bytes(ctx, fmt, bad)
    n = p.width * p.height
    prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    shad = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sc.trace_device(p, prim.data_ptr(), shad.data_ptr())
    ctx.sync()
    got = prim.cpu().numpy().view(S.RAY_DTYPE)
    flagged = got["t"] == -6.0
    assert flagged.sum() >= (want_p["t"] >= 0).sum() > 0, "every ray that hit the scene must reach a bad pointer"
    np.testing.assert_array_equal(got["t"][~flagged], want_p["t"][~flagged])
    sc.free()


def test_full_size_terrain_properties(ctx):
    """2048^3 fBm terrain (11 levels): size-independent properties.
    Every file decodes back to the input voxel set, and the per-level DAG node
    counts equal the number of distinct (mask, children) rows found by an
    independent sort-based dedup (np.unique) of the GPU's own SVO."""
    L = 11
    codes = W.terrain_codes(L, seed=1)
    print(f"[case] terrain L={L} voxels={codes.size}", flush=True)
    g = S.Octree.from_morton(ctx, codes, L)
    remap = None
    counts = []
    for l in range(L - 1, 0, -1):
        w = g.level(S.VIEW_SVO, l)
        if remap is not None:
            ch = w[:, 1:].copy()
            m = ch != S.NULL_NODE
            ch[m] = remap[ch[m]]
            w = np.ascontiguousarray(np.concatenate([w[:, :1], ch], 1))
        rows = w.view(np.dtype((np.void, 36))).ravel()
        u, first, inv = np.unique(rows, return_index=True, return_inverse=True)
        order = np.argsort(first)
        rank = np.empty(len(u), np.uint32)
        rank[order] = np.arange(len(u), dtype=np.uint32)
        remap = rank[inv.ravel()]
        counts.append((l, len(u)))
        print(f"  level {l}: svo {len(w)} distinct {len(u)}", flush=True)
    g.to_dag()
    for l, c in counts:
        assert g.level_size(S.VIEW_DAG, l) == c, (l, g.level_size(S.VIEW_DAG, l), c)
    for fmt in (S.FMT_SVDAG, S.FMT_ESVDAG):
        g.encode(fmt)
    np.testing.assert_array_equal(decode_words(g.encoded_bytes(S.FMT_SVDAG), False), codes)
    np.testing.assert_array_equal(decode_ssvdag(g.encoded_bytes(S.FMT_ESVDAG)), codes)
    g.to_sdag()
    g.encode(S.FMT_SSVDAG)
    np.testing.assert_array_equal(decode_ssvdag(g.encoded_bytes(S.FMT_SSVDAG)), codes)


def test_errors_are_reported(ctx):
    with pytest.raises(S.SvdError):
        S.Octree.from_morton(ctx, np.array([5, 3], np.uint64), 4)     # not ascending
    with pytest.raises(S.SvdError):
        S.Octree.from_morton(ctx, np.array([1 << 20], np.uint64), 4)  # out of range
    g = S.Octree.from_morton(ctx, np.array([1, 2, 3], np.uint64), 4)
    with pytest.raises(S.SvdError):
        g.encode(S.FMT_SVDAG)                                          # not a DAG yet
    with pytest.raises(S.SvdError):
        g.to_sdag()


MERGE_SCENES = dict(SCENES, terrain9=lambda: (W.terrain_codes(9, seed=1), 9),
                    caves8=lambda: (W.cave_codes(8, seed=3), 8),
                    terrain10=lambda: (W.terrain_codes(10, seed=1), 10))


@pytest.mark.parametrize("name", ["terrain7", "clustered7", "caves6", "random6", "single", "full8", "terrain9",
                                  "caves8", "terrain10"])
def test_cross_level_merge_bit_exact(ctx, name):
    codes, L = MERGE_SCENES[name]()
    print(f"[case] csvdag {name} L={L}", flush=True)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    assert o.merge_cross_levels() == 0
    g = S.Octree.from_morton(ctx, codes, L)
    g.to_dag()
    assert g.merge_cross_levels() == o.cross_merged
    for l in range(L):
        w_o, cl_o = o.level(l, with_child_levels=True)
        w_g, cl_g = g.level(S.VIEW_DAG, l, with_child_levels=True)
        np.testing.assert_array_equal(w_g, w_o, err_msg=f"level {l}")
        if l < L - 1:
            np.testing.assert_array_equal(cl_g, cl_o, err_msg=f"childLevels {l}")
    assert g.counts()[0] == o.n_nodes
    g.encode(S.FMT_SVDAG)
    assert g.encoded_bytes(S.FMT_SVDAG) == o.encode(0)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 63, 64, 65, 1000, 2048, 2049, 65536, 300001, 3_000_017])
@pytest.mark.parametrize("gray", [False, True])
def test_attr_bittree_bit_exact(ctx, n, gray):
    rng = np.random.default_rng(n + gray)
    a = np.repeat(rng.integers(0, 16, max(n // 7, 1)), 7)[:n].astype(np.uint8)
    assert ctx.attr_bittree(a, gray) == O.attr_bittree_encode(a, gray)


@pytest.mark.parametrize("kind", ["random", "constant", "ramp", "sparse"])
def test_attr_bittree_patterns(ctx, kind):
    """Fully mixed planes (random bytes), a constant, a ramp and sparse flips."""
    rng = np.random.default_rng(7)
    n = 100_003
    if kind == "random":
        a = rng.integers(0, 256, n).astype(np.uint8)
    elif kind == "constant":
        a = np.full(n, 0xA5, np.uint8)
    elif kind == "ramp":
        a = (np.arange(n) // 391 % 256).astype(np.uint8)
    else:
        a = np.zeros(n, np.uint8)
        a[rng.integers(0, n, 40)] = rng.integers(1, 256, 40).astype(np.uint8)
    for gray in (False, True):
        assert ctx.attr_bittree(a, gray) == O.attr_bittree_encode(a, gray)


def test_attr_bittree_full_size(ctx):
    """2048^3 terrain with quantized value-noise materials: GPU blob decodes
    (with the oracle's decoder) back to the attributes, Gray and binary."""
    L = 11
    codes = W.terrain_codes(L, seed=1)
    a = W.material_attrs(codes, L, seed=4)
    for gray in (True, False):
        blob = ctx.attr_bittree(a, gray)
        np.testing.assert_array_equal(O.attr_bittree_decode(blob, a.size), a)


@pytest.mark.parametrize("name", list(SCENES))
def test_encode_formats_shared_passes_equal_oracle(ctx, name):
    """svd_encode_formats: the four images in one call (the .ssvdag / .esvdag
    pair in shared passes with one segmented refcount order, the .svdag /
    .ussvdag pair with one size read) equal the oracle's; subsets too."""
    codes, L = SCENES[name]()
    if L < 3:
        pytest.skip("SSVDAG layouts need 3 levels")
    print(f"[case] encode_formats {name} L={L}", flush=True)
    _, _, _, files_o = oracle_pipeline(codes, L)
    g = S.Octree.from_morton(ctx, codes, L)
    g.to_dag()
    sizes = g.encode_many([S.FMT_ESVDAG, S.FMT_SVDAG])            # DAG state: the DAG's two images
    assert sizes == [len(files_o[S.FMT_ESVDAG]), len(files_o[S.FMT_SVDAG])]
    assert g.encoded_bytes(S.FMT_ESVDAG) == files_o[S.FMT_ESVDAG]
    assert g.encoded_bytes(S.FMT_SVDAG) == files_o[S.FMT_SVDAG]
    with pytest.raises(S.SvdError):
        g.encode_many([S.FMT_SSVDAG])                               # needs the SDAG
    g.to_sdag()
    allf = [S.FMT_SSVDAG, S.FMT_USSVDAG, S.FMT_ESVDAG, S.FMT_SVDAG]
    sizes = g.encode_many(allf)
    for f, n in zip(allf, sizes):
        assert n == len(files_o[f]) and g.encoded_bytes(f) == files_o[f], f
    with pytest.raises(S.SvdError):
        g.encode_many([S.FMT_SVDAG, S.FMT_SVDAG])                   # a format twice
    g.free()
