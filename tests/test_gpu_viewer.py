"""Viewer shading (svd_render, VIEWER_MODE main of octree_dda.frag.glsl:671-858)
and the AO pass (svd_trace_ao, AO_MODE main :884-919) on the GPU, bit for bit
against the oracle's restatement on the same scenes and cameras."""
import numpy as np
import pytest

import oracle as O
import svdag_amd as S
import svdag_amd.workloads as W

pytestmark = pytest.mark.gpu

FORMATS = [S.FMT_SVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG, S.FMT_ESVDAG]


@pytest.fixture(scope="module")
def scene_files():
    codes, L = W.terrain_codes(8, seed=1), 8
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    files = {S.FMT_SVDAG: o.encode(0), S.FMT_ESVDAG: o.encode(2)}
    o.to_sdag()
    files[S.FMT_USSVDAG] = o.encode(1)
    files[S.FMT_SSVDAG] = o.encode(2)
    return L, files


def views(sel):
    V = S.ViewParams.defaults
    return {
        "iterations": V(),
        "iterations-selected": V(selected_voxel=sel),
        "depth-plain": V(render_mode=1, random_colors=0, freq_colors=0),
        "levels": V(render_mode=2, freq_colors=0),
        "pretty": V(render_mode=3),
        "pretty-samples": V(render_mode=3, norm_samples=6, random_colors=0),
        "unknown-mode": V(render_mode=7),
    }


def gpu_render(ctx, fmt, buf, p, v, beam):
    import torch
    sc = S.Scene.from_bytes(ctx, fmt, buf)
    n = p.width * p.height
    rgb = torch.full((n * 3,), float("nan"), dtype=torch.float32, device="cuda")
    prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sc.render_device(p, v, rgb.data_ptr(), beam, prim.data_ptr())
    ctx.sync()
    return rgb.cpu().numpy().reshape(p.height, p.width, 3), prim.cpu().numpy().view(S.RAY_DTYPE)


@pytest.mark.parametrize("fmt", FORMATS)
def test_render_matches_oracle(ctx, scene_files, fmt):
    L, files = scene_files
    buf = files[fmt]
    p = W.trace_params(160, 112, seed=4, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
    want_prim, _ = O.trace_frame(buf, fmt, p, shadows=False)
    hits = want_prim["t"] >= 0
    assert hits.mean() > 0.2
    nodes, counts = np.unique(want_prim["node"][hits], return_counts=True)
    sel = int(nodes[np.argmax(counts)])
    for name, v in views(sel).items():
        want, wprim = O.render_frame(buf, fmt, p, v)
        got, gprim = gpu_render(ctx, fmt, buf, p, v, 0.0)
        print(f"[render] fmt {fmt} view {name}", flush=True)
        np.testing.assert_array_equal(gprim, wprim, err_msg=f"{name}: primary rays")
        np.testing.assert_array_equal(got, want, err_msg=f"{name}: colour")
    # the frequency colours (SSVDAG / ESVDAG) change the colour of some hit
    if fmt in (S.FMT_SSVDAG, S.FMT_ESVDAG):
        a, _ = O.render_frame(buf, fmt, p, S.ViewParams.defaults())
        b, _ = O.render_frame(buf, fmt, p, S.ViewParams.defaults(freq_colors=0, random_colors=0))
        assert (a != b).any()


@pytest.mark.parametrize("fmt", [S.FMT_SVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG])
def test_render_iteration_limit_and_draw_level(ctx, scene_files, fmt):
    """maxIters small enough that rays run out (-3: the red fragments of
    octree_dda.frag.glsl:837-840) and a drawLevel above the leaves (hits at
    coarser levels): colours still equal the oracle's."""
    L, files = scene_files
    buf = files[fmt]
    p = W.trace_params(144, 96, seed=9, levels=L, shadows=True, light=(1.0, 0.6, 1.0), max_iters=24)
    want, wprim = O.render_frame(buf, fmt, p, S.ViewParams.defaults(render_mode=3, norm_samples=2))
    assert (wprim["t"] == -3).sum() > 50
    got, gprim = gpu_render(ctx, fmt, buf, p, S.ViewParams.defaults(render_mode=3, norm_samples=2), 0.0)
    np.testing.assert_array_equal(gprim, wprim)
    np.testing.assert_array_equal(got, want)
    red = (wprim["t"] == -3).reshape(p.height, p.width)
    np.testing.assert_array_equal(got[red], np.tile(np.float32([1, 0, 0]), (red.sum(), 1)))
    p = W.trace_params(144, 96, seed=9, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
    p.draw_level = L - 3
    for v in (S.ViewParams.defaults(render_mode=2), S.ViewParams.defaults(render_mode=0, random_colors=0)):
        want, wprim = O.render_frame(buf, fmt, p, v)
        hit = wprim["t"] >= 0
        assert hit.any() and wprim["level"][hit].max() <= L - 4
        got, gprim = gpu_render(ctx, fmt, buf, p, v, 0.0)
        np.testing.assert_array_equal(gprim, wprim)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("fmt", [S.FMT_SVDAG, S.FMT_SSVDAG])
@pytest.mark.parametrize("size", [(160, 112), (203, 121)])
def test_render_beam_matches_oracle(ctx, scene_files, fmt, size):
    L, files = scene_files
    buf = files[fmt]
    w, h = size
    p = W.trace_params(w, h, seed=6, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
    dp = W.projection_factor(h, screen_divisor=8.0)
    v = S.ViewParams.defaults(render_mode=3, norm_samples=3)
    want, wprim = O.render_frame(buf, fmt, p, v, dp)
    got, gprim = gpu_render(ctx, fmt, buf, p, v, dp)
    np.testing.assert_array_equal(gprim, wprim)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("nrays", [16, 5, 0])
def test_ao_matches_oracle(ctx, scene_files, fmt, nrays):
    import torch
    L, files = scene_files
    buf = files[fmt]
    p = W.trace_params(160, 112, seed=8, levels=L, shadows=False)
    want, wprim = O.trace_ao(buf, fmt, p, nrays)
    sc = S.Scene.from_bytes(ctx, fmt, buf)
    n = p.width * p.height
    vis = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    prim = torch.zeros(n * 6, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sc.ao_device(p, nrays, vis.data_ptr(), prim.data_ptr())
    ctx.sync()
    print(f"[ao] fmt {fmt} rays {nrays}", flush=True)
    np.testing.assert_array_equal(prim.cpu().numpy().view(S.RAY_DTYPE), wprim)
    got = vis.cpu().numpy().reshape(p.height, p.width)
    np.testing.assert_array_equal(got, want)
    if nrays:
        hit = wprim["t"].reshape(p.height, p.width) >= 0
        assert (got[hit] < 1).any(), "no pixel is occluded"


def test_render_and_ao_reject_bad_arguments(ctx, scene_files):
    import torch
    L, files = scene_files
    sc = S.Scene.from_bytes(ctx, S.FMT_SVDAG, files[S.FMT_SVDAG])
    p = W.trace_params(16, 16, seed=1, levels=L)
    buf = torch.zeros(16 * 16 * 3, dtype=torch.float32, device="cuda")
    with pytest.raises(S.SvdError):
        sc.render_device(p, S.ViewParams.defaults(norm_samples=65), buf.data_ptr())
    with pytest.raises(S.SvdError):
        sc.render_device(p, S.ViewParams.defaults(norm_samples=-1), buf.data_ptr())
    with pytest.raises(S.SvdError):
        sc.ao_device(p, 257, buf.data_ptr())
    p.draw_level = 0
    with pytest.raises(S.SvdError):
        sc.ao_device(p, 4, buf.data_ptr())
