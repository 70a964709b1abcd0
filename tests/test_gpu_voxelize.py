"""Device mesh voxelization (buildSVO + cleanEmptyNodes, geom_octree.cpp:171-280,
437-456) against the oracle's literal restatement: same SVO node for node, in
the reference's creation order, and the same encoded files after toDAG/toSDAG."""
import numpy as np
import pytest

import oracle as O
import svdag_amd as S
import svdag_amd.workloads as W

pytestmark = pytest.mark.gpu


def _compare_svo(ctx, tris, L, bmin, bmax):
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    g = S.Octree.from_mesh(ctx, tris, L, bmin, bmax)
    want = o.levels_words()
    for l in range(L):
        got = g.level(S.VIEW_SVO, l)
        np.testing.assert_array_equal(got, want[l], err_msg=f"level {l}")
    assert g.counts() == (o.n_nodes, o.n_voxels)
    return o, g


@pytest.mark.parametrize("L,n,subdiv", [(5, 4, 1), (7, 16, 2), (8, 24, 2)])
def test_mesh_svo_equals_oracle(ctx, L, n, subdiv):
    V, F = W.spheres_mesh(n_spheres=n, subdiv=subdiv, seed=0)
    bmin, bmax = W.mesh_bbox(V)
    print(f"[case] voxelize spheres n={n} L={L}", flush=True)
    o, g = _compare_svo(ctx, V[F], L, bmin.astype(np.float64), bmax.astype(np.float64))
    o.to_dag()
    o.init_child_levels()
    g.to_dag()
    g.encode(S.FMT_SVDAG)
    assert g.encoded_bytes(S.FMT_SVDAG) == o.encode(0)
    o.to_sdag()
    g.to_sdag()
    g.encode(S.FMT_SSVDAG)
    assert g.encoded_bytes(S.FMT_SSVDAG) == o.encode(2)


def test_mesh_svo_boundary_triangles(ctx):
    """Triangles lying on cell faces and edges and a degenerate one, where the
    FP64 box test decides ties: the orders and masks must still match."""
    L = 6
    tris = np.array([
        [[0.0, 0.0, 0.5], [1.0, 0.0, 0.5], [0.0, 1.0, 0.5]],        # on a mid plane
        [[0.25, 0.25, 0.25], [0.75, 0.25, 0.25], [0.25, 0.75, 0.25]],
        [[0.5, 0.5, 0.0], [0.5, 0.5, 1.0], [0.5, 0.50001, 0.5]],     # along a cell edge
        [[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [0.9, 0.8, 0.7]],        # degenerate (a segment)
        [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, 0.0, 1.0]],
        [[0.3, 0.3, 0.3], [0.3001, 0.3, 0.3], [0.3, 0.3001, 0.3]],  # tiny
    ], np.float32)
    _compare_svo(ctx, tris, L, np.zeros(3), np.ones(3))


def test_mesh_svo_non_cubic_box_and_device_input(ctx):
    import torch
    V, F = W.spheres_mesh(n_spheres=6, subdiv=2, seed=3)
    V = V * np.array([1.0, 0.4, 0.7], np.float32)
    bmin, bmax = W.mesh_bbox(V)
    tris = np.ascontiguousarray(V[F], np.float32)
    o, g = _compare_svo(ctx, tris, 7, bmin.astype(np.float64), bmax.astype(np.float64))
    d = torch.from_numpy(tris.reshape(-1)).to("cuda")
    g2 = S.Octree.from_mesh(ctx, (d.data_ptr(), tris.shape[0]), 7, bmin.astype(np.float64), bmax.astype(np.float64))
    for l in range(7):
        np.testing.assert_array_equal(g2.level(S.VIEW_SVO, l), g.level(S.VIEW_SVO, l))


def test_mesh_svo_empty_and_bad_args(ctx):
    g = S.Octree.from_mesh(ctx, np.zeros((0, 3, 3), np.float32), 4, np.zeros(3), np.ones(3))
    assert g.counts()[1] == 0
    with pytest.raises(S.SvdError):
        S.Octree.from_mesh(ctx, np.zeros((1, 3, 3), np.float32), 1, np.zeros(3), np.ones(3))


@pytest.mark.parametrize("L,step", [(7, 1), (7, 2), (8, 3), (6, 4)])
def test_mesh_subtrees_equal_oracle_builddag(ctx, L, step):
    """buildDAG (geom_octree.cpp:289-435): top SVO + per-voxel subtrees, each
    voxelized in its own box, joined and reduced: the same DAG and counts as
    the oracle's literal restatement (step L - 2: subtrees of one level)."""
    V, F = W.spheres_mesh(n_spheres=16, subdiv=2, seed=0)
    bmin, bmax = W.mesh_bbox(V)
    bmin, bmax = bmin.astype(np.float64), bmax.astype(np.float64)
    tris = np.ascontiguousarray(V[F], np.float32)
    print(f"[case] builddag L={L} step={step}", flush=True)
    o = O.OracleOctree.from_mesh_subtrees(tris, L, step, bmin, bmax)
    g = S.Octree.from_mesh_subtrees(ctx, tris, L, step, bmin, bmax)
    assert g.counts()[1] == o.n_voxels
    g.to_dag()
    for l in range(L):
        np.testing.assert_array_equal(g.level(S.VIEW_DAG, l), o.level(l), err_msg=f"level {l}")
    assert g.counts()[0] == o.n_nodes
    o.init_child_levels()
    g.encode(S.FMT_SVDAG)
    assert g.encoded_bytes(S.FMT_SVDAG) == o.encode(0)


def test_mesh_subtrees_reject_step_without_subtrees(ctx):
    tris = np.array([[[0.1, 0.2, 0.3], [0.8, 0.2, 0.3], [0.1, 0.9, 0.6]]], np.float32)
    for step in (0, 5, 6):
        with pytest.raises(S.SvdError):
            S.Octree.from_mesh_subtrees(ctx, tris, 6, step, np.zeros(3), np.ones(3))
