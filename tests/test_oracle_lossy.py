"""CPU pins of the lossy-DAG oracle and of the learned encoder's oracle.

Each check compares the oracle with something computed another way:
dense voxel expansion of subtrees (exact XOR counts, occupancy prefixes),
SVO node instances (effective reference counts), brute-force kNN, central
finite differences (backward formulas) and closed forms.
"""
import functools

import numpy as np
import pytest

import mlp_oracle as M
import oracle as O

import svdag_amd.workloads as W
from test_oracle import clustered_codes


def dag_of(codes, L):
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    return o


def expander(levels, L):
    """Voxels of a subtree as sorted local Morton codes (3*(L-lev) bits)."""
    @functools.lru_cache(maxsize=None)
    def sub(lev, n):
        w = levels[lev][n]
        m = int(w[0]) & 0xFF
        if lev == L - 1:
            return tuple(c for c in range(8) if (m >> c) & 1)
        sh = 3 * (L - lev - 1)
        out = []
        for c in range(8):
            if (m >> c) & 1:
                out.extend((c << sh) | s for s in sub(lev + 1, int(w[1 + c])))
        return tuple(out)
    return sub


SCENES = {
    "caves6": lambda: (W.cave_codes(6, seed=3), 6),
    "terrain6": lambda: (W.terrain_codes(6, seed=1), 6),
    "clustered6": lambda: (clustered_codes(6, 2), 6),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_subtree_diff_is_exact_voxel_xor(name):
    codes, L = SCENES[name]()
    o = dag_of(codes, L)
    levels = o.levels_words()
    sub = expander(levels, L)
    rng = np.random.default_rng(1)
    for lev in range(1, L):
        n = o.level_size(lev)
        for _ in range(30):
            a, b = (int(x) for x in rng.integers(0, n, 2))
            d = len(set(sub(lev, a)) ^ set(sub(lev, b)))
            for cap in (0, 3, 10, 10 ** 9):
                assert o.subtree_diff(lev, a, b, cap) == min(d, cap + 1), (lev, a, b, cap, d)


@pytest.mark.parametrize("name", list(SCENES))
def test_signature_is_occupancy_three_levels_down(name):
    codes, L = SCENES[name]()
    o = dag_of(codes, L)
    levels = o.levels_words()
    sub = expander(levels, L)
    rng = np.random.default_rng(2)
    for lev in range(1, L - 1):
        sig = o.lossy_signatures(lev)
        depth = min(3, L - lev)
        for n in rng.integers(0, sig.shape[0], 40):
            n = int(n)
            pref = {c >> (3 * (L - lev - depth)) for c in sub(lev, n)}
            bits = {b for b in range(512) if (sig[n, b >> 5] >> (b & 31)) & 1}
            assert bits == pref
        # Hamming of the signatures is a lower bound of the voxel difference
        for _ in range(30):
            a, b = (int(x) for x in rng.integers(0, sig.shape[0], 2))
            h = int(M.hamming(sig[a:a + 1], sig[b:b + 1])[0])
            assert h <= o.subtree_diff(lev, a, b, 10 ** 9)


@pytest.mark.parametrize("name", list(SCENES))
def test_effective_refcounts_count_svo_instances(name):
    """sortByEffectiveRefCount (geom_octree_extension.cpp:521-560): the number
    of SVO nodes a DAG node stands for, found here by matching subtrees."""
    codes, L = SCENES[name]()
    s = O.OracleOctree.from_morton(codes, L)
    svo = s.levels_words()
    o = dag_of(codes, L)
    dag = o.levels_words()
    sub_s, sub_d = expander(svo, L), expander(dag, L)
    for lev in range(1, L):
        counts = {}
        for n in range(len(svo[lev])):
            k = sub_s(lev, n)
            counts[k] = counts.get(k, 0) + 1
        refs = o.effective_refcounts(lev)
        for n in range(len(dag[lev])):
            assert refs[n] == counts[sub_d(lev, n)]


def test_lossy_zero_bound_is_identity():
    codes, L = W.cave_codes(6, seed=3), 6
    o = dag_of(codes, L)
    before = o.levels_words()
    st = o.lossy_merge(1, L - 2, max_diff=0, diff_factor=0.0)
    assert (st["bound"] == 0).all() and st["merged"].sum() == 0
    for a, b in zip(before, o.levels_words()):
        np.testing.assert_array_equal(a, b)


def test_lossy_bound_formula():
    """maxAllowedDiff = round((L - lev)^2 * factor) (geom_octree_extension.cpp:901), with the
    factor a float (main.cpp:111): 25 * 0.7f = 17.4999997 rounds to 17."""
    codes, L = W.terrain_codes(7, seed=1), 7
    o = dag_of(codes, L)
    st = o.lossy_merge(1, L - 2, max_diff=0, diff_factor=0.7)
    assert st["level"].tolist() == list(range(L - 2, 0, -1))
    f = float(np.float32(0.7))
    assert st["bound"].tolist() == [int(np.floor((L - l) ** 2 * f + 0.5)) for l in range(L - 2, 0, -1)]
    assert st["bound"].tolist()[3] == 17


@pytest.mark.parametrize("name", list(SCENES))
def test_lossy_error_within_accounted_bound(name):
    codes, L = SCENES[name]()
    o = dag_of(codes, L)
    st = o.lossy_merge(1, L - 2, max_diff=10)
    levels = o.levels_words()
    vox = np.array(expander(levels, L)(0, 0), np.uint64)
    diff = np.setxor1d(vox, codes).size
    assert st["merged"].sum() > 0
    assert 0 < diff <= int(st["error_bound"].sum())
    # the result is a reduced DAG: another toDAG changes nothing
    o.to_dag()
    for a, b in zip(levels, o.levels_words()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", list(SCENES))
def test_all_pairs_proposals_equal_exhaustive(name):
    """The learned path's rule on the complete proposal set is the exhaustive rule."""
    codes, L = SCENES[name]()
    ref = dag_of(codes, L)
    st_ref = ref.lossy_merge(1, L - 2, max_diff=10, list_cap=8)
    o = dag_of(codes, L)
    sizes = [o.level_size(l) for l in range(L - 2, 0, -1)]
    K = max(sizes) - 1
    props = []
    for n in sizes:
        p = np.full((n, K), 0xFFFFFFFF, np.uint32)
        for x in range(n):
            others = [y for y in range(n) if y != x]
            p[x, :len(others)] = others
        props.append(p.ravel())
    st = o.lossy_merge(1, L - 2, max_diff=10, list_cap=8, proposals=np.concatenate(props), k=K)
    for f in ("nodes_after", "candidates", "listed", "verified", "merged", "error_bound"):
        np.testing.assert_array_equal(st[f], st_ref[f], err_msg=f)
    for a, b in zip(ref.levels_words(), o.levels_words()):
        np.testing.assert_array_equal(a, b)


def test_lossy_rejects_bad_arguments():
    codes, L = W.cave_codes(6, seed=3), 6
    o = dag_of(codes, L)
    p = O.LossyParams(0, 3, 4, 1, 8, 1.0)
    st = np.zeros(8, O.LOSSY_STATS_DTYPE)
    assert O.lib().orc_lossy_merge(o.h, p, None, 16, st.ctypes.data_as(O.C.c_void_p), 8) == -1
    s = O.OracleOctree.from_morton(codes, L)
    assert O.lib().orc_lossy_merge(s.h, O.LossyParams(1, 3, 4, 1, 8, 1.0), None, 16,
                                   st.ctypes.data_as(O.C.c_void_p), 8) == -3


# ---------------------------------------------------------------- kNN restatement
def _f32_sqdist(e):
    n = e.shape[0]
    d = np.zeros((n, n), np.float32)
    for k in range(e.shape[1]):
        t = e[:, None, k] - e[None, :, k]
        d = d + t * t
    np.fill_diagonal(d, np.inf)
    return d


def test_knn_lsh_full_window_is_brute_force():
    rng = np.random.default_rng(3)
    e = rng.standard_normal((60, 32)).astype(np.float32)
    got = O.knn_lsh(e, k=8, tables=2, bits=6, window=96, seed=5)
    np.testing.assert_array_equal(got, np.argsort(_f32_sqdist(e), axis=1, kind="stable")[:, :8])


def test_knn_lsh_hamming_full_window_is_brute_force():
    """Signature re-ranking: with the window covering every point the result
    is the exact Hamming top-k (ties by index)."""
    rng = np.random.default_rng(8)
    e = rng.standard_normal((50, 32)).astype(np.float32)
    sig = (rng.integers(0, 2 ** 32, (50, 16), dtype=np.uint64)
           & rng.integers(0, 2 ** 32, (50, 16), dtype=np.uint64)).astype(np.uint32)
    got = O.knn_lsh(e, k=8, tables=2, bits=4, window=96, seed=3, sig=sig)
    H = np.array([[M.hamming(sig[i:i + 1], sig[j:j + 1])[0] for j in range(50)] for i in range(50)])
    np.fill_diagonal(H, 10 ** 6)
    np.testing.assert_array_equal(got, np.argsort(H, axis=1, kind="stable")[:, :8])


def test_knn_refine_improves_monotonically_and_is_exact_on_complete_lists():
    rng = np.random.default_rng(9)
    n, k = 60, 8
    sig = (rng.integers(0, 2 ** 32, (n, 16), dtype=np.uint64)
           & rng.integers(0, 2 ** 32, (n, 16), dtype=np.uint64)).astype(np.uint32)
    H = np.array([[M.hamming(sig[i:i + 1], sig[j:j + 1])[0] for j in range(n)] for i in range(n)])
    np.fill_diagonal(H, 10 ** 6)
    lists = np.array([rng.choice([j for j in range(n) if j != i], k, replace=False) for i in range(n)], np.uint32)
    before = np.sort(np.take_along_axis(H, lists.astype(np.int64), 1), axis=1)
    after = O.knn_refine(sig, lists, 2)
    d_after = np.take_along_axis(H, after.astype(np.int64), 1)
    assert (np.diff(d_after, axis=1) >= 0).all()                 # ascending
    assert (d_after <= before).all()                              # never worse at any rank
    full = np.array([[j for j in range(n) if j != i] for i in range(n)], np.uint32)
    # complete lists (k = n - 1): one round returns the exact order
    got = O.knn_refine(sig, full, 1)
    np.testing.assert_array_equal(got, np.argsort(H, axis=1, kind="stable")[:, :n - 1])


def test_knn_lsh_recall_on_clustered_points():
    rng = np.random.default_rng(4)
    centers = rng.standard_normal((40, 32)) * 4
    e = (centers[rng.integers(0, 40, 2000)] + rng.standard_normal((2000, 32)) * 0.3).astype(np.float32)
    got = O.knn_lsh(e, k=16, tables=4, bits=16, window=32, seed=7)
    want = np.argsort(_f32_sqdist(e), axis=1, kind="stable")[:, :16]
    recall = np.mean([len(set(g) & set(w)) / 16 for g, w in zip(got, want)])
    assert recall > 0.8, recall
    assert (got != np.arange(2000)[:, None]).all()


# ---------------------------------------------------------------- MLP restatement
def test_bf16_rounding_closed_forms():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 2 ** -8 + 2 ** -20, 1 + 3 * 2 ** -8, -(1 + 3 * 2 ** -8), 0.0],
                 np.float32)
    np.testing.assert_array_equal(M.bf16(x), np.array([1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -(1 + 2 ** -6), 0.0],
                                                      np.float32))


def test_unpack_bits_layout():
    sig = np.zeros((1, 16), np.uint32)
    sig[0, 3] = 1 << 5          # bit 3*32 + 5
    assert M.unpack_bits(sig)[0].nonzero()[0].tolist() == [101]


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(5).astype(np.float64)
    sig = rng.integers(0, 2 ** 32, (16, 16), dtype=np.uint64).astype(np.uint32)
    sig &= rng.integers(0, 2 ** 32, (16, 16), dtype=np.uint64).astype(np.uint32)
    X = M.unpack_bits(sig)
    H = M.hamming(sig[0::2], sig[1::2])
    _, g = M.loss_and_grads(p, X, H, f64=True)
    eps = 1e-5
    idx = np.concatenate([rng.choice(M.N_PARAMS, 40, replace=False),
                          np.arange(M.N_PARAMS - 32, M.N_PARAMS)])       # b3 too
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_and_grads(pp, X, H, f64=True)[0] - M.loss_and_grads(pm, X, H, f64=True)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_init_is_glorot_uniform():
    p = M.init_params(5)
    off = 0
    for i in range(3):
        fi, fo = M.DIMS[i], M.DIMS[i + 1]
        w = p[off:off + fi * fo]
        a = np.sqrt(6.0 / (fi + fo))
        assert np.abs(w).max() <= a and np.abs(w).max() > 0.99 * a
        assert abs(w.mean()) < 0.02 * a
        assert (p[off + fi * fo:off + fi * fo + fo] == 0).all()
        off += fi * fo + fo


def test_mlp_adam_step_closed_form():
    """First Adam step: m^ = g, v^ = g^2, so the move is -lr * g / (|g| + eps)."""
    p = np.zeros(8, np.float32)
    g = np.array([1, -2, 3e-3, -4e-5, 0, 5, -6, 7], np.float32)
    m, v = np.zeros(8, np.float32), np.zeros(8, np.float32)
    M.adam(p, g, m, v, 1, 0.01)
    want = -0.01 * g.astype(np.float64) / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p, want, rtol=1e-4, atol=1e-7)


# ---------------------------------------------------------------- ranked proposer
def _hamming_all(sig):
    """Pairwise Hamming distances of 512-bit signatures (n x 16 uint32)."""
    bits = np.unpackbits(np.ascontiguousarray(sig).view(np.uint8), axis=1).astype(np.int32)
    return (bits[:, None, :] != bits[None, :, :]).sum(axis=2)


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("bound,cap", [(0, 4), (6, 3), (20, 8), (40, 16)])
def test_ranked_lists_brute_force(name, bound, cap):
    """The ranked proposer's list of node j is its cap nearest earlier nodes
    within the bound by (Hamming distance, index): a brute force over all
    pairs from numpy's bit unpacking, on real signatures (ties included)."""
    codes, L = SCENES[name]()
    o = dag_of(codes, L)
    lev = L - 3
    sig = o.lossy_signatures(lev)
    D = _hamming_all(sig)
    lists, counts = O.ranked_lists(sig, bound, cap)
    for j in range(sig.shape[0]):
        cand = sorted((int(D[i, j]), i) for i in range(j) if D[i, j] <= bound)[:cap]
        want = [i for _, i in cand]
        assert counts[j] == len(want), j
        assert list(lists[j, :counts[j]]) == want, j


@pytest.mark.parametrize("name", list(SCENES))
def test_ranked_zero_bound_is_exhaustive(name):
    """Within a zero bound all listed nodes are at distance 0, so the nearest
    first order is the index order: the same merge as the exhaustive one."""
    codes, L = SCENES[name]()
    a, b = dag_of(codes, L), dag_of(codes, L)
    sa = a.lossy_merge(1, L - 2, max_diff=0, diff_factor=0.0, list_cap=8)
    sb = b.lossy_merge(1, L - 2, max_diff=0, diff_factor=0.0, list_cap=8, ranked=True)
    for f in ("nodes_after", "listed", "verified", "merged"):
        np.testing.assert_array_equal(sa[f], sb[f], err_msg=f)
    for x, y in zip(a.levels_words(), b.levels_words()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("name", list(SCENES))
def test_ranked_merge_error_within_bound(name):
    codes, L = SCENES[name]()
    o = dag_of(codes, L)
    st = o.lossy_merge(1, L - 2, max_diff=10, list_cap=8, ranked=True)
    vox = np.array(expander(o.levels_words(), L)(0, 0), np.uint64)
    diff = np.setxor1d(vox, codes).size
    assert st["merged"].sum() > 0 and 0 < diff <= int(st["error_bound"].sum())
