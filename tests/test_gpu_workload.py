"""The device scene generator reproduces the NumPy definition bit for bit."""
import numpy as np
import pytest

import svdag_amd.workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("levels,seed,threshold", [(5, 3, 0.1), (6, 3, 0.1), (7, 3, 0.1), (6, 9, 0.0)])
def test_caves_match_numpy(ctx, levels, seed, threshold):
    d = ctx.caves(levels, seed=seed, threshold=threshold)
    got = d.numpy()
    d.free()
    want = W.cave_codes(levels, seed=seed, threshold=threshold)
    assert got.size > 0
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [5, 7])
def test_materials_match_numpy(ctx, levels):
    import torch
    d = ctx.caves(levels, seed=3)
    out = torch.zeros(d.n, dtype=torch.uint8, device="cuda:0")
    ctx.materials(d, levels, out.data_ptr(), seed=4, n_values=16)
    codes = d.numpy()
    d.free()
    np.testing.assert_array_equal(out.cpu().numpy(), W.material_attrs(codes, levels, seed=4, n_values=16))


def test_codes_union_matches_numpy(ctx):
    a, b = ctx.caves(6, seed=3), ctx.terrain(6, seed=1)
    u = ctx.union(a, b)
    want = np.union1d(a.numpy(), b.numpy())
    np.testing.assert_array_equal(u.numpy(), want)
    for x in (a, b, u):
        x.free()
