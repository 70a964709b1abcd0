"""Device point voxelization (buildSVOFromPoints, geom_octree.cpp:62-169, with
the points in memory instead of a LAS file) against the oracle's literal
restatement: the same SVO node for node in the reference's creation order,
and the same files after toDAG/toSDAG."""
import numpy as np
import pytest

import oracle as O
import svdag_amd as S

pytestmark = pytest.mark.gpu


def _compare(ctx, pts, L, bmin, bmax, device=False):
    o = O.OracleOctree.from_points(pts, L, bmin, bmax)
    if device:
        import torch
        d = torch.from_numpy(np.ascontiguousarray(pts, np.float64).reshape(-1)).to("cuda")
        g = S.Octree.from_points(ctx, (d.data_ptr(), pts.shape[0]), L, bmin, bmax)
    else:
        g = S.Octree.from_points(ctx, pts, L, bmin, bmax)
    want = o.levels_words()
    for lev in range(L):
        np.testing.assert_array_equal(g.level(S.VIEW_SVO, lev), want[lev], err_msg=f"level {lev}")
    assert g.counts() == (o.n_nodes, o.n_voxels)
    return o, g


def _cloud(rng, n, bmin, bmax, L):
    """Random points in and around the box, some on the double grid of voxel
    faces (where the half-open tests decide) and some on the max faces."""
    bmin, bmax = np.asarray(bmin, np.float64), np.asarray(bmax, np.float64)
    ext = bmax - bmin
    pts = rng.uniform(bmin - 0.05 * ext, bmax + 0.05 * ext, size=(n, 3))
    grid = bmin + ext * rng.integers(0, 2 ** L + 1, size=(n // 8, 3)) / 2 ** L
    face = rng.uniform(bmin, bmax, size=(n // 32, 3))
    ax = rng.integers(0, 3, face.shape[0])
    face[np.arange(face.shape[0]), ax] = bmax[ax]
    pts = np.concatenate([pts, grid, face])
    return np.ascontiguousarray(pts[rng.permutation(len(pts))])


@pytest.mark.parametrize("L,n,seed", [(5, 2000, 0), (8, 60000, 1), (10, 300000, 2)])
def test_points_svo_equals_oracle(ctx, L, n, seed):
    rng = np.random.default_rng(seed)
    bmin, bmax = (0.1, -3.7, 2.2), (5.3, 1.1, 9.9)
    pts = _cloud(rng, n, bmin, bmax, L)
    print(f"[case] points L={L} n={len(pts)}", flush=True)
    o, g = _compare(ctx, pts, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    g.to_dag()
    g.encode(S.FMT_SVDAG)
    assert g.encoded_bytes(S.FMT_SVDAG) == o.encode(0)
    o.to_sdag()
    g.to_sdag()
    g.encode(S.FMT_SSVDAG)
    assert g.encoded_bytes(S.FMT_SSVDAG) == o.encode(2)


def test_points_dyadic_box_device_input(ctx):
    """A dyadic box (exact arithmetic) with points on every kind of face, the
    points given on the device."""
    rng = np.random.default_rng(7)
    L = 7
    pts = _cloud(rng, 40000, (-8.0, -8.0, -8.0), (8.0, 8.0, 8.0), L)
    _compare(ctx, pts, L, (-8.0, -8.0, -8.0), (8.0, 8.0, 8.0), device=True)


def test_points_empty_and_outside(ctx):
    L = 4
    for pts in (np.zeros((0, 3)), np.array([[2.0, 0.5, 0.5], [1.0, 1.0, 1.0], [-1e-9, 0.2, 0.2]])):
        o, g = _compare(ctx, pts, L, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
        assert g.counts()[1] == 0
    with pytest.raises(S.SvdError):
        S.Octree.from_points(ctx, np.zeros((1, 3)), 1, (0, 0, 0), (1, 1, 1))


def test_points_large_cloud_files_equal_oracle(ctx):
    """2M points of a non-cubic box at 11 levels, the point order shuffled:
    the four files equal the oracle's."""
    rng = np.random.default_rng(11)
    L = 11
    bmin, bmax = (-17.25, 3.5, 0.0), (40.0, 22.125, 9.75)
    t = rng.uniform(0, 2 * np.pi, 2_000_000)
    r = rng.uniform(0.2, 1.0, t.size)
    pts = np.stack([11.0 + 25.0 * r * np.cos(t), 12.8 + 9.0 * r * np.sin(t),
                    4.8 + 4.5 * np.sin(3 * t) * r], axis=1)
    print(f"[case] points large L={L} n={len(pts)}", flush=True)
    o, g = _compare(ctx, pts, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    g.to_dag()
    files = {S.FMT_SVDAG: 0, S.FMT_ESVDAG: 2}
    for f, of in files.items():
        g.encode(f)
        assert g.encoded_bytes(f) == o.encode(of), S.FMT_EXT[f]
    o.to_sdag()
    g.to_sdag()
    for f, of in {S.FMT_USSVDAG: 1, S.FMT_SSVDAG: 2}.items():
        g.encode(f)
        assert g.encoded_bytes(f) == o.encode(of), S.FMT_EXT[f]
