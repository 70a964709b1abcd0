"""svd_render_noalloc (the allocation-free variant of svd_render) against
svd_render itself, through the C ABI: two frames in a row, the second
smaller, with and without the minimum-depth pre-pass whose texture lives in
the library's static scratch array; and the size limit of that array."""
import ctypes as C

import numpy as np
import pytest

import svdag_amd as S
import svdag_amd.workloads as W

pytestmark = pytest.mark.gpu

SCRATCH_TEXELS = 131072   # kRenderScratchTexels


@pytest.fixture(scope="module")
def scenes(ctx):
    L = 9
    codes = W.terrain_codes(L, seed=1)
    g = S.Octree.from_morton(ctx, codes, L)
    g.to_dag()
    g.encode(S.FMT_SVDAG)
    g.to_sdag()
    g.encode(S.FMT_SSVDAG)
    out = {fmt: S.Scene.from_octree(g, fmt) for fmt in (S.FMT_SVDAG, S.FMT_SSVDAG)}
    yield L, out
    for sc in out.values():
        sc.free()
    g.free()


def frame(ctx, fn, sc, p, v, beam):
    """One call of `fn` (svd_render or svd_render_noalloc) into fresh buffers."""
    import torch
    n = p.width * p.height
    rgb = torch.full((n * 3,), float("nan"), dtype=torch.float32, device="cuda")
    prim = torch.full((n * 6,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(fn(sc.h, C.byref(p), C.byref(v), C.c_float(beam), C.c_void_p(rgb.data_ptr()),
                 C.c_void_p(prim.data_ptr())))
    ctx.check(ctx.L.svd_ctx_synchronize(ctx.h))
    return rgb.cpu().numpy(), prim.cpu().numpy().view(S.RAY_DTYPE)


@pytest.mark.parametrize("fmt", [S.FMT_SSVDAG, S.FMT_SVDAG])
@pytest.mark.parametrize("beam", [True, False])
def test_noalloc_render_matches_render(ctx, scenes, fmt, beam):
    L, sc = scenes
    v = S.ViewParams.defaults(render_mode=3, norm_samples=2)
    # the second frame is smaller: the scratch keeps the first frame's texels
    # beyond the second texture, which the second frame must never read
    for w, h, seed in ((640, 360, 3), (200, 120, 8)):
        p = W.trace_params(w, h, seed=seed, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
        bpf = W.projection_factor(h, screen_divisor=8.0) if beam else 0.0
        print(f"[noalloc] fmt {fmt} beam {beam} {w}x{h}", flush=True)
        want_rgb, want_prim = frame(ctx, ctx.L.svd_render, sc[fmt], p, v, bpf)
        got_rgb, got_prim = frame(ctx, ctx.L.svd_render_noalloc, sc[fmt], p, v, bpf)
        assert (want_prim["t"] >= 0).mean() > 0.05
        np.testing.assert_array_equal(got_prim, want_prim, err_msg=f"{w}x{h}: primary rays")
        np.testing.assert_array_equal(got_rgb, want_rgb, err_msg=f"{w}x{h}: colour")


def test_noalloc_render_rejects_oversized_prepass(ctx, scenes):
    """(width/8) * (height/8) above the scratch size fails on the host before
    any launch; the largest texture that fits gives svd_render's frame."""
    L, sc = scenes
    v = S.ViewParams.defaults()
    import torch
    p = W.trace_params(4104, 2048, seed=3, levels=L, shadows=False)
    assert (p.width // 8) * (p.height // 8) > SCRATCH_TEXELS
    rgb = torch.zeros(p.width * p.height * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = ctx.L.svd_render_noalloc(sc[S.FMT_SSVDAG].h, C.byref(p), C.byref(v), C.c_float(1.0),
                                  C.c_void_p(rgb.data_ptr()), None)
    assert rc == -1                                    # SVD_ERR_ARG
    assert b"scratch" in ctx.L.svd_last_error(ctx.h)
    # the largest frame that fits still runs
    p = W.trace_params(4096, 2048, seed=3, levels=L, shadows=False)
    assert (p.width // 8) * (p.height // 8) == SCRATCH_TEXELS
    bpf = W.projection_factor(p.height, screen_divisor=8.0)
    want_rgb, want_prim = frame(ctx, ctx.L.svd_render, sc[S.FMT_SSVDAG], p, v, bpf)
    got_rgb, got_prim = frame(ctx, ctx.L.svd_render_noalloc, sc[S.FMT_SSVDAG], p, v, bpf)
    np.testing.assert_array_equal(got_prim, want_prim)
    np.testing.assert_array_equal(got_rgb, want_rgb)
