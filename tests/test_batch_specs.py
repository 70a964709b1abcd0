"""CPU checks of the multi-scene batch definition and its NumPy scene generators."""
import collections

import numpy as np

import svdag_amd.workloads as W


def test_batch_specs_distribution():
    sp = W.batch_specs(256, 0)
    kinds = collections.Counter(s["kind"] for s in sp)
    depths = collections.Counter(s["levels"] for s in sp)
    assert set(kinds) == {"spheres", "terrain", "menger"} and min(kinds.values()) > 50
    assert set(depths) == {6, 7, 8, 9, 10}
    for s in sp:
        if s["kind"] == "spheres":
            assert 8 <= s["n_spheres"] <= 128
        if s["kind"] == "menger":
            assert 3 <= s["order"] <= 5
    assert sp == W.batch_specs(256, 0)                 # deterministic


def test_menger_surface_closed_form():
    """Order 1 on a 3^k-aligned grid is the 20-cube frame; at 2^L the cut
    positions follow floor(3x/n).  Every surface voxel is solid, and a voxel
    whose 6 neighbours are solid is never listed."""
    L = 4
    n = 1 << L
    c = W.menger_surface_codes(L, 1)
    x, y, z = W.morton_decode(c)
    third = lambda v: (v.astype(np.int64) * 3) // n == 1
    assert not np.any((third(x) & third(y)) | (third(y) & third(z)) | (third(x) & third(z)))
    full = W.menger_codes(L, 1)
    assert np.isin(c, full).all() and c.size < full.size


def test_balls_surface_is_a_shell():
    b = np.array([[0.5, 0.5, 0.5, 0.3]])
    L = 5
    c = W.balls_codes(L, b)
    x, y, z = (v.astype(np.float64) for v in W.morton_decode(c))
    n = 1 << L
    r = np.sqrt(((x + 0.5) / n - 0.5) ** 2 + ((y + 0.5) / n - 0.5) ** 2 + ((z + 0.5) / n - 0.5) ** 2)
    assert (r <= 0.3).all() and (r > 0.3 - 2.0 / n).all()
