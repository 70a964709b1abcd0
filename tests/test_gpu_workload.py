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
