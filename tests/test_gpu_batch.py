"""Batched multi-scene build + multi-camera traversal, and the device scene
generators, against the oracle / the NumPy definitions / single-scene runs."""
import numpy as np
import pytest
import torch

import oracle as O

import svdag_amd as S
import svdag_amd.workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.mark.parametrize("kind,levels,arg", [("balls", 6, 20), ("balls", 5, 8), ("menger", 5, 3), ("menger", 6, 4),
                                             ("terrain", 6, 1), ("terrain", 7, 5)])
def test_generators_match_numpy(ctx, kind, levels, arg):
    if kind == "balls":
        b = W.sphere_balls(arg, 3)
        d, want = ctx.balls(levels, b), W.balls_codes(levels, b)
    elif kind == "menger":
        d, want = ctx.menger(levels, arg), W.menger_surface_codes(levels, arg)
    else:
        d, want = ctx.terrain(levels, seed=arg), W.terrain_codes(levels, seed=arg)
    got = d.numpy()
    d.free()
    assert want.size > 0
    np.testing.assert_array_equal(got, want)


def _scenes(n, seed, depth):
    specs = W.batch_specs(n, seed, depth)
    return specs, [W.spec_codes(s) for s in specs]


def _pack(codes):
    off = np.concatenate([[0], np.cumsum([c.size for c in codes])]).astype(np.uint64)
    return np.concatenate(codes).astype(np.uint64), off


def test_batch_build_equals_oracle(ctx):
    specs, codes = _scenes(24, 1, (3, 6))
    # uneven extremes: a 2-level scene and a single-voxel scene
    codes += [W.sorted_codes(np.array([0, 3]), np.array([1, 2]), np.array([3, 0])),
              W.sorted_codes(np.array([5]), np.array([9]), np.array([2]))]
    levels = [s["levels"] for s in specs] + [2, 5]
    allc, off = _pack(codes)
    print(f"[case] batch of {len(codes)} scenes, {allc.size} voxels, depths {sorted(set(levels))}", flush=True)
    b = S.Batch(ctx, allc, off, levels)
    for s, (c, L) in enumerate(zip(codes, levels)):
        o = O.OracleOctree.from_morton(c, L)
        o.to_dag()
        o.init_child_levels()
        info = b.info(s)
        assert info["nodes"] == o.n_nodes and info["voxels"] == c.size and info["levels"] == L
        assert b.svdag_bytes(s) == o.encode(0), f"scene {s} ({L} levels)"


def test_batch_many_tiny_scenes_equal_oracle(ctx):
    """Hundreds of scenes of 1-40 voxels each: every 8192-code tile of the leaf
    pass spans many scenes, so most codes find their scene by bisecting the
    scene offsets (not from the tile's record of its first two scenes), and
    scene boundaries fall inside and across tiles."""
    rng = np.random.default_rng(11)
    codes, levels = [], []
    for s in range(1000):
        L = int(rng.integers(2, 6))
        k = int(rng.integers(1, 41))
        side = 1 << L
        xyz = rng.integers(0, side, size=(k, 3))
        c = W.sorted_codes(xyz[:, 0], xyz[:, 1], xyz[:, 2])
        codes.append(c)
        levels.append(L)
    allc, off = _pack(codes)
    print(f"[case] batch of {len(codes)} tiny scenes, {allc.size} voxels", flush=True)
    assert allc.size > 2 * 8192, allc.size
    b = S.Batch(ctx, allc, off, levels)
    for s, (c, L) in enumerate(zip(codes, levels)):
        o = O.OracleOctree.from_morton(c, L)
        o.to_dag()
        o.init_child_levels()
        assert b.info(s)["nodes"] == o.n_nodes, f"scene {s}"
        assert b.svdag_bytes(s) == o.encode(0), f"scene {s} ({L} levels, {c.size} voxels)"


def test_batch_device_scenes_equal_single_builds(ctx):
    specs = W.batch_specs(16, 2, (6, 8))
    dev = [ctx.spec_codes(sp) for sp in specs]
    levels = [sp["levels"] for sp in specs]
    b = S.Batch.from_device_scenes(ctx, dev, levels)
    for s, (d, L) in enumerate(zip(dev, levels)):
        g = S.Octree.from_morton(ctx, (d.ptr, d.n), L)
        g.to_dag()
        g.encode(S.FMT_SVDAG)
        assert b.svdag_bytes(s) == g.encoded_bytes(S.FMT_SVDAG), f"scene {s} {specs[s]}"
        g.free()
    for d in dev:
        d.free()


def test_batch_trace_equals_single_traces(ctx):
    specs, codes = _scenes(6, 3, (5, 7))
    levels = [s["levels"] for s in specs]
    allc, off = _pack(codes)
    b = S.Batch(ctx, allc, off, levels)
    W_, H_ = 96, 64
    cams, cam_scene = [], []
    for k in range(10):
        s = k % len(codes)
        cams.append(W.trace_params(W_, H_, seed=40 + k, levels=levels[s], shadows=True))
        cam_scene.append(s)
    n = W_ * H_
    prim = torch.zeros(len(cams) * n * 6, dtype=torch.int32, device=DEV)
    shad = torch.zeros(len(cams) * n * 6, dtype=torch.int32, device=DEV)
    b.trace(cams, cam_scene, prim.data_ptr(), shad.data_ptr())
    pr = prim.cpu().numpy().view(S.RAY_DTYPE).reshape(len(cams), n)
    sh = shad.cpu().numpy().view(S.RAY_DTYPE).reshape(len(cams), n)
    one_p = torch.zeros(n * 6, dtype=torch.int32, device=DEV)
    one_s = torch.zeros(n * 6, dtype=torch.int32, device=DEV)
    for k, s in enumerate(cam_scene):
        sc = S.Scene.from_bytes(ctx, S.FMT_SVDAG, b.svdag_bytes(s))
        sc.trace_device(cams[k], one_p.data_ptr(), one_s.data_ptr())
        ctx.sync()
        np.testing.assert_array_equal(pr[k], one_p.cpu().numpy().view(S.RAY_DTYPE), err_msg=f"camera {k}")
        np.testing.assert_array_equal(sh[k], one_s.cpu().numpy().view(S.RAY_DTYPE), err_msg=f"camera {k}")
        sc.free()
    assert (pr["t"] >= 0).mean() > 0.05


def test_batch_rejects_bad_input(ctx):
    c = W.terrain_codes(5, seed=1)
    with pytest.raises(S.SvdError):                     # an empty scene
        S.Batch(ctx, c, np.array([0, c.size, c.size], np.uint64), [5, 5])
    with pytest.raises(S.SvdError):                     # codes beyond the scene's depth
        S.Batch(ctx, c, np.array([0, c.size], np.uint64), [4])
    with pytest.raises(S.SvdError):                     # not ascending within a scene
        S.Batch(ctx, c[::-1].copy(), np.array([0, c.size], np.uint64), [5])


@pytest.mark.parametrize("fmt", [S.FMT_SVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG])
def test_trace_launch_shapes_give_identical_rays(ctx, tmp_path, fmt):
    """The autotuner's choices never change results: both launch modes (tiled,
    persistent ray queue) and every tile shape trace the same rays bit for bit,
    on a frame whose size is not a multiple of any tile."""
    codes, L = W.terrain_codes(7, seed=1), 7
    g = S.Octree.from_morton(ctx, codes, L)
    g.to_dag()
    if fmt != S.FMT_SVDAG:
        g.to_sdag()
    g.encode(fmt)
    sc = S.Scene.from_octree(g, fmt)
    p = W.trace_params(200, 121, seed=3, levels=L, shadows=True)
    n = 200 * 121
    prim = torch.zeros(n * 6, dtype=torch.int32, device=DEV)
    shad = torch.zeros(n * 6, dtype=torch.int32, device=DEV)
    ref = None
    for mode in S.TRACE_MODES:
        ctx.trace_mode = mode
        for w in (8, 16, 32, 64):
            ctx.trace_tile = w
            prim.fill_(-7)
            shad.fill_(-7)
            sc.trace_device(p, prim.data_ptr(), shad.data_ptr())
            ctx.sync()
            got = (prim.cpu().numpy().copy(), shad.cpu().numpy().copy())
            if ref is None:
                ref = got
            else:
                np.testing.assert_array_equal(got[0], ref[0], err_msg=f"{mode} tile {w}")
                np.testing.assert_array_equal(got[1], ref[1], err_msg=f"{mode} tile {w}")
    for r in ref:   # every pixel written
        assert not (r.reshape(n, 6) == -7).all(axis=1).any()
    if fmt != S.FMT_SSVDAG:
        ctx.trace_mode = "tiled"
        return
    cache = str(tmp_path / "autotune.json")
    best = ctx.autotune_trace(sc, p, prim.data_ptr(), shad.data_ptr(), reps=2, cache=cache, key="terrain7-200x121")
    assert best in (8, 16, 32, 64) and ctx.trace_tile == best
    assert set(ctx.last_autotune["ms"]) == set(S.TRACE_MODES)
    assert all(set(v) == {8, 16, 32, 64} for v in ctx.last_autotune["ms"].values())
    chosen = ctx.trace_mode
    ctx.trace_tile = 16
    ctx.trace_mode = "tiled" if chosen == "persistent" else "persistent"
    assert ctx.autotune_trace(sc, p, prim.data_ptr(), shad.data_ptr(), cache=cache, key="terrain7-200x121") == best
    assert ctx.trace_mode == chosen
    with pytest.raises(S.SvdError):
        ctx.trace_tile = 12
    with pytest.raises(ValueError):
        ctx.trace_mode = "other"
    ctx.trace_tile = 16
    ctx.trace_mode = "tiled"
