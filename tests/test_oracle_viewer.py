"""CPU checks of the oracle's viewer shading (VIEWER_MODE main, octree_dda.frag.glsl:671-858)
and AO pass (AO_MODE main, :884-919).

What pins them:
  - sin / cos / log2: against float64 numpy (bounded ulp error, exact on the
    special values);
  - hemisphere samples: against the radical inverse in exact rationals
    (halton.cpp:24-40) and the cosine mapping (octree_dda_renderer.cpp:510-531);
  - the shading: against the oracle's traversal port (render modes 0 and 1 are
    closed forms of the primary ray; the shadow factor is 0.8 exactly where the
    port's own shadow ray hits; the beam primary equals the beam port's), and
    against geometry (on a flat floor the smoothed normal of the pretty mode
    is the floor normal whatever the sample count; AO on a wall is 1 above
    the reach of its 0.5-long rays and < 1 next to the floor).
"""
from fractions import Fraction
import math

import numpy as np
import pytest

import oracle as O
import svdag_amd.workloads as W

f32 = np.float32


def view(**kw):
    v = O.ViewParams(0, 0, 0xFFFFFFFF, 1, 1, 0)
    for k, x in kw.items():
        setattr(v, k, x)
    return v


def ulps(got, ref):
    ref32 = np.abs(ref).astype(np.float32)
    return np.abs(got.astype(np.float64) - ref) / np.spacing(np.maximum(ref32, np.float32(1e-30))).astype(np.float64)


# --------------------------------------------------------------------------- math
def test_sin_cos_against_float64():
    x = np.concatenate([np.linspace(-10, 10, 400001, dtype=np.float32),
                        np.float32([0.0, -0.0, 1e-20, math.pi / 4, math.pi / 2, math.pi, 3 * math.pi / 2])])
    sn, cs, _ = O.vm_math(x)
    xs = x.astype(np.float64)
    for got, ref in ((sn, np.sin(xs)), (cs, np.cos(xs))):
        assert np.abs(got - ref).max() < 1.2e-7
        big = np.abs(ref) > 1e-3
        assert ulps(got[big], ref[big]).max() <= 2.0
    # odd / even symmetry (to 1 ulp: floor(y + 0.5) breaks quadrant ties
    # upwards for x and -x alike) and sin^2 + cos^2 = 1
    sn2, cs2, _ = O.vm_math(-x)
    assert np.abs(sn2 + sn).max() <= 6e-8
    assert np.abs(cs2 - cs).max() <= 6e-8
    assert np.abs(sn.astype(np.float64) ** 2 + cs.astype(np.float64) ** 2 - 1).max() < 3e-7


def test_log2_against_float64():
    rng = np.random.default_rng(0)
    x = np.exp(rng.uniform(-80, 80, 300000)).astype(np.float32)
    x = np.concatenate([x, np.float32([0.5, 0.7071067, 0.7071068, 1.0000001, 0.9999999, 3.0, 1e-40])])
    _, _, lg = O.vm_math(x)
    ref = np.log2(x.astype(np.float64))
    assert ulps(lg, ref).max() <= 3.0
    p = np.float32(2.0) ** np.arange(-126, 128, dtype=np.float32)
    _, _, lp = O.vm_math(p)
    np.testing.assert_array_equal(lp, np.arange(-126, 128, dtype=np.float32))
    _, _, sp = O.vm_math(np.float32([0.0, -1.0, np.inf, np.nan]))
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf and np.isnan(sp[3])


def test_hemisphere_samples_are_cosine_weighted_halton():
    hs = O.hemisphere_samples()

    def radical_inverse(i, b):
        r, f = Fraction(0), Fraction(1, b)
        while i:
            r += f * (i % b)
            i //= b
            f /= b
        return r

    for i in range(1, 257):
        h2, h3 = float(radical_inverse(i, 2)), float(radical_inverse(i, 3))
        y = math.sqrt(h2)
        l = math.sqrt(1 - y * y)
        want = (l * math.cos(2 * math.pi * h3), y, l * math.sin(2 * math.pi * h3))
        np.testing.assert_allclose(hs[i - 1], want, atol=6e-7, err_msg=f"sample {i}")
    assert (hs[:, 1] > 0).all()
    np.testing.assert_allclose(np.linalg.norm(hs.astype(np.float64), axis=1), 1.0, atol=4e-7)
    # cosine weighting: E[y] = 2/3 over the hemisphere
    assert abs(hs[:, 1].astype(np.float64).mean() - 2 / 3) < 0.01


# --------------------------------------------------------------------------- shading
@pytest.fixture(scope="module")
def terrain():
    codes, L = W.terrain_codes(6, seed=1), 6
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    svdag = o.encode(0)
    o.to_sdag()
    return L, {0: svdag, 2: o.encode(2)}


def test_modes_0_and_1_are_closed_forms_of_the_primary_ray(terrain):
    L, files = terrain
    p = W.trace_params(72, 48, seed=3, levels=L, shadows=False)
    prim, _ = O.trace_frame(files[0], 0, p, shadows=False)
    hit = prim["t"] >= 0
    assert hit.mean() > 0.2
    rgb0, prim0 = O.render_frame(files[0], 0, p, view(render_mode=0, random_colors=0, freq_colors=0))
    np.testing.assert_array_equal(prim0, prim)
    rgb0 = rgb0.reshape(-1, 3)
    np.testing.assert_array_equal(rgb0[hit, 0], f32(1) - prim["iters"][hit] / f32(p.max_iters))
    assert (rgb0[:, 0] == rgb0[:, 1]).all() and (rgb0[:, 1] == rgb0[:, 2]).all()
    miss = np.isin(prim["t"], [-1.0, -2.0, -4.0])
    assert miss.any() and (rgb0[miss] == 1).all()
    rgb1, _ = O.render_frame(files[0], 0, p, view(render_mode=1, random_colors=0, freq_colors=0))
    ext = f32(1) - f32(0)
    diag = np.sqrt(f32(ext * ext) + f32(ext * ext) + f32(ext * ext), dtype=np.float32)
    np.testing.assert_array_equal(rgb1.reshape(-1, 3)[hit, 0], prim["t"][hit] / diag)


def test_selected_voxel_and_random_colours(terrain):
    L, files = terrain
    p = W.trace_params(72, 48, seed=3, levels=L, shadows=False)
    prim, _ = O.trace_frame(files[0], 0, p, shadows=False)
    nodes, counts = np.unique(prim["node"][prim["t"] >= 0], return_counts=True)
    sel = int(nodes[np.argmax(counts)])
    rgb, _ = O.render_frame(files[0], 0, p, view(render_mode=0, selected_voxel=sel))
    rgb = rgb.reshape(-1, 3)
    is_sel = (prim["t"] >= 0) & (prim["node"] == sel)
    t = f32(1) - prim["iters"] / f32(p.max_iters)
    np.testing.assert_array_equal(rgb[is_sel, 2], 1.0)
    np.testing.assert_array_equal(rgb[is_sel, 0], t[is_sel] * f32(0.5))
    other = (prim["t"] >= 0) & ~is_sel
    n = prim["node"][other].astype(np.uint32)
    rc = np.stack([(n % 100).astype(np.float32) / f32(100), ((3 * n) % 200).astype(np.float32) / f32(200),
                   ((2 * n) % 300).astype(np.float32) / f32(300)], 1)
    rc = rc / np.linalg.norm(rc.astype(np.float64), axis=1, keepdims=True)
    np.testing.assert_allclose(rgb[other], t[other, None] * rc, rtol=2e-6, atol=1e-7)


def test_shadow_factor_matches_the_ports_shadow_ray(terrain):
    """Pretty mode with no normal samples casts the same shadow ray as the
    traversal port's shadow pass: the colour is scaled by exactly 0.8 where
    that ray hits."""
    L, files = terrain
    p = W.trace_params(72, 48, seed=5, levels=L, shadows=True, light=(1.0, 0.6, 1.0))
    prim, shad = O.trace_frame(files[2], 2, p, shadows=True)
    lit, _ = O.render_frame(files[2], 2, p, view(render_mode=3), 0.0)
    p.shadows = 0
    dark, _ = O.render_frame(files[2], 2, p, view(render_mode=3), 0.0)
    lit, dark = lit.reshape(-1, 3), dark.reshape(-1, 3)
    occl = shad["t"] > 0
    assert occl.sum() > 10 and ((prim["t"] >= 0) & ~occl).sum() > 10
    np.testing.assert_array_equal(lit[occl], dark[occl] * f32(0.8))
    np.testing.assert_array_equal(lit[~occl], dark[~occl])


def test_beam_primary_equals_the_beam_port(terrain):
    L, files = terrain
    p = W.trace_params(96, 64, seed=2, levels=L, shadows=False)
    dp = W.projection_factor(p.height, screen_divisor=8.0)
    want, _, _ = O.trace_frame_beam(files[2], 2, p, dp, shadows=False)
    _, got = O.render_frame(files[2], 2, p, view(), dp)
    np.testing.assert_array_equal(got, want)


# --------------------------------------------------------------------------- geometry invariants
def slab_files(L, axis, thickness):
    n = 1 << L
    g = np.arange(n)
    a, b = np.meshgrid(g, g, indexing="ij")
    parts = []
    for k in range(thickness):
        c = np.full(a.size, k)
        xyz = [a.ravel(), b.ravel()]
        xyz.insert(axis, c)
        parts.append(W.morton_encode(*[np.asarray(v, dtype=np.uint64) for v in xyz]))
    return np.unique(np.concatenate(parts))


def encode_svdag(codes, L):
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    return o.encode(0)


def camera(eye, center, w, h, L):
    view_m = W._look_at(np.array(eye, float), np.array(center, float), np.array([0.0, 1.0, 0.0]))
    proj = W._perspective(math.radians(60.0), w / h, 0.01, 100.0)
    p = W.trace_params(w, h, seed=0, levels=L, shadows=False)
    p.view_inv[:] = [float(x) for x in np.linalg.inv(view_m).astype(np.float32).T.ravel()]
    p.proj_inv[:] = [float(x) for x in np.linalg.inv(proj).astype(np.float32).T.ravel()]
    p.projection_factor = 1e30          # no level of detail: every hit is a finest-level voxel
    return p


def hit_points(p, t):
    """float64 hit positions o + t d of the frame's primary rays (computeCameraRay)."""
    vi = np.array(p.view_inv, np.float64).reshape(4, 4).T
    pi = np.array(p.proj_inv, np.float64).reshape(4, 4).T
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    sx = (xs.ravel() + 0.5) / p.width * 2 - 1
    sy = (ys.ravel() + 0.5) / p.height * 2 - 1
    one = np.ones_like(sx)
    w0 = pi @ np.stack([sx, sy, 0 * one, one])
    w1 = pi @ np.stack([sx, sy, one, one])
    rd = w1[:3] / w1[3] - w0[:3] / w0[3]
    rd /= np.linalg.norm(rd, axis=0)
    o = vi @ np.array([0, 0, 0, 1.0])
    e = vi @ np.vstack([rd, one])
    d = e[:3] - o[:3, None]
    d /= np.linalg.norm(d, axis=0)
    return (o[:3, None] + t[None, :] * d).T


def test_pretty_mode_normal_on_a_floor_is_the_floor_normal():
    """Every sample ray that lands on the floor adds (0, 1, 0), so the smoothed
    normal of an interior floor pixel normalises back to (0, 1, 0): the colour
    does not depend on the number of normal samples."""
    L = 6
    buf = encode_svdag(slab_files(L, 1, 4), L)
    p = camera((1.3, 0.9, 1.1), (0.5, 0.0, 0.5), 80, 60, L)
    base, prim = O.render_frame(buf, 0, p, view(render_mode=3, random_colors=0))
    hp = hit_points(p, prim["t"].astype(np.float64))
    top = (prim["t"] >= 0) & (np.abs(hp[:, 1] - 4 / 64) < 1e-3) & \
          (hp[:, 0] > 0.15) & (hp[:, 0] < 0.85) & (hp[:, 2] > 0.15) & (hp[:, 2] < 0.85)
    assert top.sum() > 200
    for ns in (1, 5, 16):
        got, _ = O.render_frame(buf, 0, p, view(render_mode=3, random_colors=0, norm_samples=ns))
        np.testing.assert_array_equal(got.reshape(-1, 3)[top], base.reshape(-1, 3)[top], err_msg=f"{ns} samples")


def test_ao_on_a_wall_over_a_floor():
    """Wall x in [0, 4/64) on floor y in [0, 4/64): a wall point higher than
    the 0.5 reach of the AO rays above the floor sees nothing (visibility 1);
    one close to the floor loses the downward half of its hemisphere.  Floor
    pixels (normal +y) take the degenerate tangent of the reference's fixed
    rvec = (0, 1, 0): every direction is NaN, no ray counts, visibility 1."""
    L = 6
    codes = np.unique(np.concatenate([slab_files(L, 0, 4), slab_files(L, 1, 4)]))
    buf = encode_svdag(codes, L)
    p = camera((1.6, 0.55, 0.5), (0.0, 0.45, 0.5), 80, 60, L)
    vis, prim = O.trace_ao(buf, 0, p, 16)
    vis = vis.ravel()
    hit = prim["t"] >= 0
    np.testing.assert_array_equal(vis[~hit], -1.0)
    hp = hit_points(p, prim["t"].astype(np.float64))
    interior_z = (hp[:, 2] > 0.2) & (hp[:, 2] < 0.8)
    wall = hit & (np.abs(hp[:, 0] - 4 / 64) < 1e-3) & interior_z
    floor = hit & (np.abs(hp[:, 1] - 4 / 64) < 1e-3) & (hp[:, 0] > 0.1) & interior_z
    high = wall & (hp[:, 1] > 4 / 64 + 0.55)
    low = wall & (hp[:, 1] < 4 / 64 + 0.08)
    assert high.sum() > 20 and low.sum() > 20 and floor.sum() > 20
    np.testing.assert_array_equal(vis[high], 1.0)
    assert (vis[low] < 1.0).all()
    np.testing.assert_array_equal(vis[floor], 1.0)
    assert ((vis[hit] >= 0) & (vis[hit] <= 1)).all()
