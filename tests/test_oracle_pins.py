"""CPU pins of the oracle's functions that the round-trip checks leave loose.

Each test checks the oracle (oracle/svdag_oracle.cpp) against something
computed without it, chosen so that a plausible slip in the restatement fails:

  * testTriBox (test_triangle_box.cpp:105-184)  <- the closed separating-axis
    test in exact integer arithmetic, on dyadic triangles (where double
    arithmetic is exact) including face- and edge-aligned ones, and on
    general float triangles away from exact ties;
  * buildSVO's node order (geom_octree.cpp:171-280) <- the closed form of its
    creation order: (first triangle reaching the node, Morton order of the
    parent, 7 - slot), over exact cell/triangle overlaps;
  * buildDAG (geom_octree.cpp:289-435) <- equal per-level DAG counts and the
    same voxel set as buildSVO + toDAG when the voxel sets agree; step 0 is
    rejected (the reference indexes an empty leavesCenters there);
  * mergeAcrossAllLevels (geom_octree_extension.cpp:1192-1544) <- brute force
    over every (1 <= levB < levA, j) pair with subtree equality taken from the
    voxel sets (Morton codes), the greedy top-down frontier, and the final
    compacted levels with their cross-level pointers;
  * toSDAG's representative and transform (geom_octree.cpp:551-697) <- for
    every node whose children have no mirror symmetry (where the reference's
    keys and voxel sets correspond one to one): the representative is the
    first node of its reflection orbit and the transform is the first of
    [I, X, Y, Z, XY, XZ, YZ, XYZ] that maps it there;
  * trace_ray's loop (octree_dda.frag.glsl:539-581) <- the iteration counter's
    closed form (a ray that needs k iterations gives the same record for any
    maxIters > k and (-3, maxIters) otherwise), and the LOD stop: a ray with
    LOD on takes the same steps as the full-resolution ray until it stops,
    and it stops only where cell_size * projection_factor < t.
"""
from functools import lru_cache

import numpy as np
import pytest

import oracle as O
import svdag_amd.workloads as W
from decode import decode_ssvdag, decode_words
from test_oracle import clustered_codes, random_codes, subtree_keys


# ============================================================== triangle / box
AXES = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def sat_overlap(c, h, tri):
    """Closed separating-axis test of a triangle and a cube in exact
    arithmetic (ints or Fractions): the 3 box normals, the triangle normal and
    the 9 edge x box-axis products.  Touching counts as overlap."""
    v = [tuple(p[i] - c[i] for i in range(3)) for p in tri]
    e = [tuple(v[(k + 1) % 3][i] - v[k][i] for i in range(3)) for k in range(3)]
    cand = list(AXES) + [_cross(e[0], e[1])] + [_cross(ek, ax) for ek in e for ax in AXES]
    for a in cand:
        p = [a[0] * q[0] + a[1] * q[1] + a[2] * q[2] for q in v]
        r = h * (abs(a[0]) + abs(a[1]) + abs(a[2]))
        if min(p) > r or max(p) < -r:
            return False
    return True


def sat_slack(c, h, tri):
    """Smallest |projection - radius| over the axes, relative to the scale of
    the numbers involved: how close an exact decision is to a tie."""
    v = [tuple(p[i] - c[i] for i in range(3)) for p in tri]
    e = [tuple(v[(k + 1) % 3][i] - v[k][i] for i in range(3)) for k in range(3)]
    cand = list(AXES) + [_cross(e[0], e[1])] + [_cross(ek, ax) for ek in e for ax in AXES]
    best = None
    for a in cand:
        s = abs(a[0]) + abs(a[1]) + abs(a[2])
        if s == 0:
            continue
        p = [a[0] * q[0] + a[1] * q[1] + a[2] * q[2] for q in v]
        r = h * s
        scale = s * (h + max(abs(x) for q in v for x in q))
        for d in (min(p) - r, max(p) + r):
            rel = abs(d) / scale
            best = rel if best is None else min(best, rel)
    return best


def test_tri_box_dyadic_exact():
    """Dyadic coordinates (multiples of 2^-7 in [0,1], cubes of half side
    2^-k): every double operation of the test is exact, so its answer must be
    the exact closed separating-axis answer, ties included."""
    rng = np.random.default_rng(5)
    S = 128
    checked = touching = 0
    for trial in range(3000):
        k = int(rng.integers(2, 6))                     # half side 2^-k
        cell = 1 << (7 - k)                             # cube edge in grid units of 2^-7 ... /2 below
        ci = rng.integers(0, S // cell, 3) * cell + cell // 2
        kind = trial % 3
        if kind == 0:                                   # anywhere
            tri = rng.integers(0, S + 1, (3, 3))
        elif kind == 1:                                 # on grid lines of the cube size (face / edge aligned)
            tri = rng.integers(0, S // cell + 1, (3, 3)) * cell
        else:                                           # small, near the cube
            tri = ci + rng.integers(-2 * cell, 2 * cell + 1, (3, 3))
        want = sat_overlap([int(x) * 2 for x in ci], cell, [[int(x) * 2 for x in p] for p in tri])
        got = O.tri_box(ci / S, (cell / 2) / S, (tri / S).astype(np.float32))
        assert got == want, (trial, ci.tolist(), cell, tri.tolist())
        checked += 1
        if want:
            # a touching contact: the cube shrunk by a hair no longer meets the triangle
            touching += not sat_overlap([int(x) * 4 for x in ci], 2 * cell - 1, [[int(x) * 4 for x in p] for p in tri])
    assert checked == 3000 and touching > 50


def test_tri_box_general_floats_away_from_ties():
    from fractions import Fraction as F
    rng = np.random.default_rng(9)
    agree = skipped = 0
    for trial in range(1500):
        tri = rng.random((3, 3)).astype(np.float32)
        if trial % 2:
            tri = (tri * 0.2 + rng.random(3) * 0.8).astype(np.float32)
        c = rng.random(3)
        h = float(2.0 ** -int(rng.integers(1, 6)))
        ex_tri = [[F(float(x)) for x in p] for p in tri]
        ex_c = [F(float(x)) for x in c]
        want = sat_overlap(ex_c, F(h), ex_tri)
        got = O.tri_box(c, h, tri)
        if got != want:
            assert sat_slack(ex_c, F(h), ex_tri) < 1e-12, (trial, tri.tolist(), c.tolist(), h)
            skipped += 1
        else:
            agree += 1
    assert agree >= 1490


# ============================================================== buildSVO order
def _dyadic_mesh(rng, n_tris, grid):
    """Triangles with vertices on a grid of 1/grid in [0,1]^3 (float32-exact)."""
    out = []
    for t in range(n_tris):
        if t % 4 == 3:       # axis-aligned: all vertices on one grid plane
            a = int(rng.integers(0, 3))
            tri = rng.integers(0, grid + 1, (3, 3))
            tri[:, a] = rng.integers(0, grid + 1)
        else:
            base = rng.integers(0, grid + 1, 3)
            tri = np.clip(base + rng.integers(-grid // 4, grid // 4 + 1, (3, 3)), 0, grid)
        out.append(tri)
    return np.array(out, np.int64)


def _overlap_cells(tri_int, grid, L):
    """For every level l (1..L), the cells of the 2^l grid that the triangle
    overlaps, by exact descent (a closed triangle meeting a cell meets its
    parent), as Morton codes of the cell at that level."""
    # integer frame: coordinates * 2 * 2^L * grid... scale so cell centres and half sides are integers
    S = 2 * (1 << L) * grid
    tri = [[int(x) * (S // grid) for x in p] for p in tri_int]
    out = [set() for _ in range(L + 1)]
    stack = [(0, 0, 0, 0)]          # level, ix, iy, iz of the cell
    while stack:
        l, ix, iy, iz = stack.pop()
        if l == L:
            continue
        for c in range(8):
            jx, jy, jz = 2 * ix + (c >> 2 & 1), 2 * iy + (c >> 1 & 1), 2 * iz + (c & 1)
            edge = S >> (l + 1)
            centre = (jx * edge + edge // 2, jy * edge + edge // 2, jz * edge + edge // 2)
            if sat_overlap(centre, edge // 2, tri):
                code = int(W.morton_encode(np.array([jx]), np.array([jy]), np.array([jz]))[0])
                if code not in out[l + 1]:
                    out[l + 1].add(code)
                    stack.append((l + 1, jx, jy, jz))
    return out


def _oracle_cells(o, L):
    """Cell (Morton code at its level) of every node of every SVO level, by
    walking the oracle's child pointers from the root."""
    cells = [np.zeros(1, np.int64)]
    masks = []
    for l in range(L):
        w = o.level(l)
        masks.append(w[:, 0] & 0xFF)
        if l == L - 1:
            break
        nxt = np.full(o.level_size(l + 1), -1, np.int64)
        for i in range(w.shape[0]):
            for c in range(8):
                if w[i, 0] >> c & 1:
                    nxt[w[i, 1 + c]] = cells[l][i] * 8 + c
        assert (nxt >= 0).all()
        cells.append(nxt)
    return cells, masks


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_buildsvo_creation_order_closed_form(seed):
    rng = np.random.default_rng(seed)
    L, grid = 5, 64
    tris = _dyadic_mesh(rng, 24, grid)
    per_tri = [_overlap_cells(t, grid, L) for t in tris]
    o = O.OracleOctree.from_mesh((tris / grid).astype(np.float32).reshape(-1, 3, 3), L, (0.0,) * 3, (1.0,) * 3)
    cells, masks = _oracle_cells(o, L)
    for l in range(1, L):
        first = {}
        for t, cs in enumerate(per_tri):
            for code in cs[l]:
                first.setdefault(code, t)
        want = sorted(first, key=lambda x: (first[x], x >> 3, 7 - (x & 7)))
        assert cells[l].tolist() == want, f"level {l}"
    # masks: the overlapped children of each node (leaves: the voxels)
    for l in range(L):
        occ = set().union(*[cs[l + 1] for cs in per_tri])
        want = [sum(1 << c for c in range(8) if (int(p) * 8 + c) in occ) for p in cells[l]]
        assert masks[l].tolist() == want, f"level {l} masks"


# ============================================================== buildDAG
@pytest.mark.parametrize("step", [1, 2, 3])
def test_builddag_counts_equal_buildsvo(step):
    """Appendix A.8 of the reference's behaviour: buildDAG joins subtrees in
    another order, but on the same voxel set its DAG has the same per-level
    node counts as buildSVO + toDAG.  Dyadic triangles make the subtree boxes'
    float path exact, so the voxel sets agree."""
    rng = np.random.default_rng(11)
    L, grid = 5, 64
    tris = (_dyadic_mesh(rng, 40, grid) / grid).astype(np.float32).reshape(-1, 3, 3)
    a = O.OracleOctree.from_mesh(tris, L, (0.0,) * 3, (1.0,) * 3)
    a.to_dag()
    b = O.OracleOctree.from_mesh_subtrees(tris, L, step, (0.0,) * 3, (1.0,) * 3)
    assert a.n_voxels == b.n_voxels > 0
    a.init_child_levels()
    b.init_child_levels()
    np.testing.assert_array_equal(decode_words(a.encode(0), False), decode_words(b.encode(0), False))
    assert [a.level_size(l) for l in range(L)] == [b.level_size(l) for l in range(L)]
    assert a.n_nodes == b.n_nodes


def test_builddag_rejects_step_zero_and_last():
    tris = np.array([[[0.1, 0.1, 0.1], [0.9, 0.2, 0.3], [0.4, 0.8, 0.7]]] * 6, np.float32)
    bb = np.array([0, 0, 0, 1, 1, 1], np.float64)
    lib = O.lib()
    for step in (0, 4):
        h = lib.orc_new()
        try:
            rc = lib.orc_build_dag_mesh(h, O._p(tris), tris.shape[0], 5, step, O._p(bb))
        finally:
            lib.orc_free(h)
        assert rc == -2, step


# ============================================================== CSVDAG brute force
class VoxelDAG:
    """The DAG of a Morton voxel set, from the codes alone: node i of level
    lev is the i-th distinct subtree in first-occurrence order (the toDAG
    order), each with its voxel offsets and the Morton prefix of its first
    instance."""

    def __init__(self, codes, L):
        self.L = L
        self.sets, self.first, self.node_of = [], [], []
        for lev in range(L):
            shift = np.uint64(3 * (L - lev))
            pre = (codes >> shift).astype(np.int64)
            rel = codes & ((np.uint64(1) << shift) - np.uint64(1))
            starts = np.flatnonzero(np.r_[True, pre[1:] != pre[:-1]])
            ends = np.r_[starts[1:], len(codes)]
            idx, sets, first, node_of = {}, [], [], {}
            for s, e in zip(starts, ends):
                k = frozenset(rel[s:e].tolist())
                if k not in idx:
                    idx[k] = len(sets)
                    sets.append(k)
                    first.append(int(pre[s]))
                node_of[int(pre[s])] = idx[k]
            self.sets.append(sets)
            self.first.append(first)
            self.node_of.append(node_of)

    def size(self, lev):
        return len(self.sets[lev])

    def child(self, lev, i, c):
        return self.node_of[lev + 1].get(self.first[lev][i] * 8 + c)

    @lru_cache(maxsize=None)
    def cells(self, lev, i, depth):
        """Occupied cells `depth` levels below node (lev, i)."""
        sh = 3 * (self.L - lev - depth)
        return frozenset(v >> sh for v in self.sets[lev][i])


def csvdag_bruteforce(dag):
    """mergeAcrossAllLevels restated on voxel sets: A at levA equals B at
    levB < levA (compareSubtrees) iff B's occupied cells L - levA levels down
    are A's voxels; a frontier node takes the lexicographically first
    (levB, j) with levB >= 1 (the candidate maps never hold the root:
    buildMultiMap(depth, ...) hashes levels 1..L-1, geom_octree.hpp:180-181);
    the matched subtree's nodes map to B's descendants on the same
    slot paths (last visit wins inside one comparison, the first recorded
    match wins across comparisons, a frontier match overwrites)."""
    L = dag.L
    corr = [dict() for _ in range(L)]
    frontier = set(range(dag.size(1)))
    for levA in range(1, L):
        nxt = set()
        for a in sorted(frontier):
            want = dag.sets[levA][a]
            m = None
            for levB in range(1, levA):
                for j in range(dag.size(levB)):
                    if dag.cells(levB, j, L - levA) == want:
                        m = (levB, j)
                        break
                if m:
                    break
            if m is None:
                if levA < L - 1:
                    nxt |= {dag.child(levA, a, c) for c in range(8) if dag.child(levA, a, c) is not None}
                continue
            corr[levA][a] = m
            sub = [dict() for _ in range(L)]

            def walk(la, x, lb, y):
                if la == L - 1:
                    return
                for c in range(8):
                    cx = dag.child(la, x, c)
                    if cx is None:
                        continue
                    cy = dag.child(lb, y, c)
                    sub[la + 1][cx] = (lb + 1, cy)
                    walk(la + 1, cx, lb + 1, cy)

            walk(levA, a, m[0], m[1])
            for lev in range(L):
                for k, v in sub[lev].items():
                    corr[lev].setdefault(k, v)
        frontier = nxt
    return corr


def csvdag_expected_levels(dag, corr):
    """The compacted levels: kept nodes in order; each child pointer as
    (level, index) after compaction."""
    L = dag.L
    newidx = []
    for lev in range(L):
        m, k = {}, 0
        for i in range(dag.size(lev)):
            if i not in corr[lev]:
                m[i] = k
                k += 1
        newidx.append(m)
    out = []
    for lev in range(L - 1):
        rows = []
        for i in range(dag.size(lev)):
            if i in corr[lev]:
                continue
            ptrs = {}
            for c in range(8):
                ch = dag.child(lev, i, c)
                if ch is None:
                    continue
                if ch in corr[lev + 1]:
                    lb, j = corr[lev + 1][ch]
                    assert j in newidx[lb], "a match target was removed"
                    ptrs[c] = (lb, newidx[lb][j])
                else:
                    ptrs[c] = (lev + 1, newidx[lev + 1][ch])
            rows.append(ptrs)
        out.append(rows)
    return out, [len(m) for m in newidx]


CSVDAG_SCENES = {
    "menger5-2": lambda: (W.menger_codes(5, 2), 5),
    "menger-surface5-3": lambda: (W.menger_surface_codes(5, 3), 5),
    "menger-surface4-1": lambda: (W.menger_surface_codes(4, 1), 4),
    "terrain5": lambda: (W.terrain_codes(5, seed=1), 5),
    "clustered5": lambda: (clustered_codes(5, 3), 5),
    "random4": lambda: (random_codes(4, 500, 2), 4),
}


@pytest.mark.parametrize("name", list(CSVDAG_SCENES))
def test_cross_level_merge_bruteforce(name):
    codes, L = CSVDAG_SCENES[name]()
    dag = VoxelDAG(codes, L)
    corr = csvdag_bruteforce(dag)
    want_levels, want_sizes = csvdag_expected_levels(dag, corr)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    assert [o.level_size(l) for l in range(L)] == [dag.size(l) for l in range(L)]
    o.init_child_levels()
    assert o.merge_cross_levels() == 0
    eliminated = sum(len(c) for c in corr)
    assert o.cross_merged == eliminated
    if name.startswith("menger"):
        assert eliminated > 0
    assert [o.level_size(l) for l in range(L)] == want_sizes
    for lev in range(L - 1):
        w, cl = o.level(lev, with_child_levels=True)
        assert w.shape[0] == len(want_levels[lev])
        for i, ptrs in enumerate(want_levels[lev]):
            got = {c: (int(cl[i, c]), int(w[i, 1 + c])) for c in range(8) if w[i, 0] >> c & 1}
            assert got == ptrs, (lev, i)


# ============================================================== toSDAG representative / transform
ORDER = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1))


def _reflect_set(s, t, side):
    a = np.fromiter(s, np.uint64, len(s))
    x, y, z = (v.astype(np.int64) for v in W.morton_decode(a))
    if t[0]:
        x = side - 1 - x
    if t[1]:
        y = side - 1 - y
    if t[2]:
        z = side - 1 - z
    return frozenset(W.morton_encode(x, y, z).tolist())


def clean_nodes(dag):
    """Per level, the nodes below which no node has a child with a mirror
    symmetry.  For them the reference's keys (mask, child ids, child mirror
    bits) and their voxel sets correspond one to one: a child's mirror bits
    are fixed by its voxel set, and so, level by level from the leaves, is its
    representative.  (A symmetric child admits two keys for one voxel set,
    and invertInvs only normalises single-axis symmetries.)"""
    L = dag.L
    clean = [None] * L
    clean[L - 1] = [True] * dag.size(L - 1)            # leaf keys are their masks
    for lev in range(L - 2, 0, -1):
        cs = 1 << (L - lev - 1)
        ok = [clean[lev + 1][j] and all(_reflect_set(s, t, cs) != s for t in ORDER[1:])
              for j, s in enumerate(dag.sets[lev + 1])]
        clean[lev] = [all(ok[ch] for ch in (dag.child(lev, i, c) for c in range(8)) if ch is not None)
                      for i in range(dag.size(lev))]
    return clean


def sdag_expectations(dag, lev, clean, order=ORDER):
    """For the clean nodes of DAG level lev: node -> (representative DAG node,
    transform bits x | y << 1 | z << 2)."""
    side = 1 << (dag.L - lev)
    refl = [[_reflect_set(s, t, side) for t in ORDER] for s in dag.sets[lev]]
    first_in_orbit, out = {}, {}
    pos = {t: k for k, t in enumerate(ORDER)}
    for i in range(dag.size(lev)):
        key = min(tuple(sorted(r)) for r in refl[i])
        rep = first_in_orbit.setdefault(key, i)
        if not clean[lev][i]:
            continue
        t = next(t for t in order if refl[i][pos[t]] == dag.sets[lev][rep])
        out[i] = (rep, t[0] | t[1] << 1 | t[2] << 2)
    return out


SDAG_SCENES = {
    "clustered6": lambda: (clustered_codes(6, 3), 6),
    "terrain6": lambda: (W.terrain_codes(6, seed=1), 6),
    "random5": lambda: (random_codes(5, 3000, 7), 5),
    "menger-surface5-2": lambda: (W.menger_surface_codes(5, 2), 5),
}


@pytest.mark.parametrize("name", list(SDAG_SCENES))
def test_tosdag_representative_and_transform(name):
    codes, L = SDAG_SCENES[name]()
    dag = VoxelDAG(codes, L)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.to_sdag()
    clean = clean_nodes(dag)
    checked = symmetric_choice = inner = 0
    for lev in range(1, L):
        ids, bits = o.sdag_map(lev)
        assert ids.size == dag.size(lev)
        exp = sdag_expectations(dag, lev, clean)
        inner += len(exp) if lev < L - 1 else 0
        for i, (rep, t) in exp.items():
            if rep == i:
                assert bits[i] == 0 and ids[i] not in set(ids[:i].tolist()), (lev, i)
            else:
                assert ids[i] == ids[rep] and bits[i] == t, (lev, i, rep, t, int(bits[i]))
            checked += 1
        # nodes where more than one transform reaches the representative: the order decides
        for alt in (ORDER[:1] + (ORDER[2], ORDER[1]) + ORDER[3:],             # Y before X
                    ORDER[:2] + (ORDER[3], ORDER[2]) + ORDER[4:],             # Z before Y
                    ORDER[:1] + (ORDER[4],) + ORDER[1:4] + ORDER[5:]):        # XY before X
            other = sdag_expectations(dag, lev, clean, alt)
            symmetric_choice += sum(other[i] != exp[i] for i in exp)
    print(f"[tosdag] {name}: {checked} nodes checked, {inner} inner, {symmetric_choice} order-sensitive")
    assert checked > 100
    if name in ("clustered6", "terrain6", "random5"):
        assert inner > 20                 # (the sponge's inner nodes all have symmetric children)
        assert symmetric_choice > 0       # the scene tells the reference order from a permuted one


# ============================================================== traversal loop
def _frame(L=6, seed=13):
    codes = clustered_codes(L, seed)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.to_sdag()
    return codes, o.encode(2)


def test_trace_iteration_counter_closed_form():
    """trace_ray's do-while increments once per pass and stops at maxIters: a
    ray whose result took k passes gives the same record for any maxIters > k,
    and (-3, iters = maxIters) for maxIters <= k."""
    _, buf = _frame()
    p = W.trace_params(40, 30, seed=3, levels=6, shadows=False)
    base, _ = O.trace_frame(buf, 2, p, shadows=False)
    assert (base["t"] != -3).all()
    for M in (1, 2, 3, 5, 8, 13, 21, 40):
        p.max_iters = M
        got, _ = O.trace_frame(buf, 2, p, shadows=False)
        same = base["iters"] < M
        assert same.any() and (~same).any() or M in (1, 40)
        np.testing.assert_array_equal(got[same], base[same])
        assert (got["t"][~same] == -3).all() and (got["iters"][~same] == M).all()


def _world_rays(p):
    """Camera rays of every pixel in float64 (origin, unit direction)."""
    vi, pi = np.array(p.view_inv).reshape(4, 4).T, np.array(p.proj_inv).reshape(4, 4).T
    orig = (vi @ np.array([0, 0, 0, 1.0]))[:3]
    dirs = np.zeros((p.height * p.width, 3))
    for py in range(p.height):
        for px in range(p.width):
            sx = (px + 0.5) / p.width * 2 - 1
            sy = (py + 0.5) / p.height * 2 - 1
            a = pi @ np.array([sx, sy, 0, 1.0])
            b = pi @ np.array([sx, sy, 1, 1.0])
            dc = b[:3] / b[3] - a[:3] / a[3]
            dw = (vi @ np.append(dc / np.linalg.norm(dc), 1.0))[:3] - orig
            dirs[py * p.width + px] = dw / np.linalg.norm(dw)
    return orig, dirs


def test_trace_lod_is_a_prefix_of_the_full_traversal():
    """With LOD on, a ray takes the full-resolution ray's steps until it stops
    on a full cell whose size times the projection factor is below t: every
    record that is not such a stop equals the full-resolution one, and every
    stop satisfies the criterion at its level, comes no later than the full
    ray, and lies in an occupied cell of its level."""
    L = 6
    codes, buf = _frame(L)
    p = W.trace_params(64, 48, seed=7, levels=L, shadows=False)
    p.projection_factor = W.projection_factor(48, pixel_tolerance=6.0)
    pf = np.float32(p.projection_factor)
    lod, _ = O.trace_frame(buf, 2, p, shadows=False)
    p.projection_factor = 1e30
    full, _ = O.trace_frame(buf, 2, p, shadows=False)
    max_level = L - 2                             # SSVDAG: INNER_LEVELS = L - 2 < drawLevel - 1
    stop = (lod["t"] >= 0) & (lod["level"] < max_level)
    assert stop.sum() > 20 and (~stop).sum() > 20
    for f in ("t", "level", "iters", "node"):
        np.testing.assert_array_equal(lod[f][~stop], full[f][~stop], err_msg=f)
    lev = lod["level"][stop].astype(np.int64)
    cell = (np.float32(0.5) / (np.float32(2.0) ** lev.astype(np.float32))).astype(np.float32)
    t = lod["t"][stop]                            # scale 2 * rootHalfSide = 1 for the unit box
    assert (cell * pf < t).all()
    assert (full["iters"][stop] >= lod["iters"][stop]).all()
    fh = full["t"][stop]
    assert ((fh < 0) | (fh >= t)).all()
    orig, dirs = _world_rays(p)
    occupied = ok = 0
    for k, r in zip(np.flatnonzero(stop), range(int(stop.sum()))):
        g = 1 << (int(lev[r]) + 1)
        q = orig + (float(t[r]) + 1e-3 / g) * dirs[k]
        c = np.floor(q * g).astype(np.int64)
        if ((c < 0) | (c >= g)).any():
            continue
        ok += 1
        code = int(W.morton_encode(c[:1], c[1:2], c[2:3])[0])
        occupied += code in set((codes >> np.uint64(3 * (L - int(lev[r]) - 1))).tolist())
    assert occupied >= 0.99 * ok and ok > 0.9 * stop.sum()


# ============================================================== OpenMP subtree builder (the CPU baseline)
@pytest.mark.parametrize("name,L,step", [("terrain", 7, 1), ("terrain", 7, 3), ("clustered", 6, 2),
                                         ("random", 6, 4)])
def test_builddag_morton_subtrees_equal_plain_build(name, L, step):
    """buildDAG's OpenMP subtree builder on a voxel list joins subtrees in
    (i ascending, j = 7..0) order, so its node order differs, but it reduces
    the same voxel set: per-level DAG counts and the decoded voxels equal the
    single-pass build's.  toSDAG then depends on that order where invertInvs
    lets two representatives share a reflection orbit, so its counts may
    differ by a few nodes; its files still decode to the same voxels."""
    codes = {"terrain": lambda: W.terrain_codes(L, seed=1), "clustered": lambda: clustered_codes(L, 3),
             "random": lambda: random_codes(L, 5000, 4)}[name]()
    a = O.OracleOctree.from_morton(codes, L)
    a.to_dag()
    b = O.OracleOctree.from_morton_subtrees(codes, L, step)
    assert b.n_voxels == a.n_voxels == codes.size
    assert [b.level_size(l) for l in range(L)] == [a.level_size(l) for l in range(L)]
    assert b.n_nodes == a.n_nodes
    b.init_child_levels()
    np.testing.assert_array_equal(decode_words(b.encode(0), False), codes)
    a.to_sdag()
    b.to_sdag()
    sa, sb = [a.level_size(l) for l in range(L)], [b.level_size(l) for l in range(L)]
    assert all(abs(x - y) <= max(2, x // 100) for x, y in zip(sa, sb)), (sa, sb)
    np.testing.assert_array_equal(decode_ssvdag(b.encode(2)), codes)
    np.testing.assert_array_equal(decode_words(b.encode(1), True), codes)


# ---------------------------------------------------------------- buildSVOFromPoints
def _morton_of(ix, iy, iz, L):
    """Morton codes of integer voxel coordinates (x most significant within
    each level's triple, so code & 7 = child slot 4x + 2y + z)."""
    code = np.zeros(ix.shape, np.uint64)
    for b in range(L - 1, -1, -1):
        code = (code << np.uint64(3)) | (((ix >> b) & 1) << 2).astype(np.uint64) \
            | (((iy >> b) & 1) << 1).astype(np.uint64) | ((iz >> b) & 1).astype(np.uint64)
    return code


def _levels_equal(a, b, L, what):
    for lev in range(L):
        np.testing.assert_array_equal(a.level(lev), b.level(lev), err_msg=f"{what}: level {lev}")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_buildsvo_points_dyadic_box_is_quantization_in_point_order(seed):
    """buildSVOFromPoints (geom_octree.cpp:62-169) in a dyadic box, where
    every double operation of its DFS is exact: a point inside [min, max)
    lands in the voxel floor((p - min) / voxel), points outside or on the max
    faces add nothing, and every level's node order is the order of first
    arrival in the point sequence, i.e. the Morton builder (pinned by brute
    force) fed the points' codes in point order."""
    rng = np.random.default_rng(seed)
    L, lo, hi = 6, -4.0, 4.0
    vox = (hi - lo) / 2 ** L
    pts = rng.uniform(lo - 0.5, hi + 0.5, size=(3000, 3))
    on_grid = lo + vox * rng.integers(0, 2 ** L + 1, size=(400, 3))     # exact voxel faces, max face included
    pts = np.concatenate([pts, on_grid])
    pts = pts[rng.permutation(len(pts))]
    inside = np.all((pts >= lo) & (pts < hi), axis=1)
    assert inside.sum() > 2000 and (~inside).sum() > 100
    q = np.floor((pts[inside] - lo) / vox).astype(np.int64)
    codes = _morton_of(q[:, 0], q[:, 1], q[:, 2], L)
    o = O.OracleOctree.from_points(pts, L, (lo, lo, lo), (hi, hi, hi))
    ref = O.OracleOctree.from_morton(codes, L)
    _levels_equal(o, ref, L, "points vs Morton in point order")
    assert o.n_voxels == np.unique(codes).size


@pytest.mark.parametrize("seed", [3, 4])
def test_buildsvo_points_general_box_away_from_faces(seed):
    """A box whose bounds are not dyadic (float32 root side, double centre,
    geom_octree.cpp:66-73): every point farther than 1e-7 voxel from every
    voxel face in exact rational arithmetic lands in the voxel the exact
    quantization gives, in point order as above."""
    from fractions import Fraction as F
    rng = np.random.default_rng(seed)
    L = 5
    bmin = np.array([0.1, -3.7, 2.2])
    bmax = np.array([5.3, 1.1, 9.9])
    side = max(float(np.float32((np.float32(b) - np.float32(a)) * np.float32(0.5)) * np.float32(2.0))
               for a, b in zip(bmin, bmax))
    centre = (bmin + bmax) * 0.5
    vox = F(side) / 2 ** L
    root_lo = [F(float(c)) - F(side) / 2 for c in centre]
    pts = rng.uniform(bmin - 0.2, bmax + 0.2, size=(1500, 3))
    keep, cells = [], []
    for p in pts:
        t = [(F(float(p[a])) - root_lo[a]) / vox for a in range(3)]
        fl = [x.numerator // x.denominator for x in t]
        frac = [x - f for x, f in zip(t, fl)]
        if all(0 <= f < 2 ** L for f in fl) and all(F(1, 10 ** 7) < r < 1 - F(1, 10 ** 7) for r in frac):
            keep.append(p)
            cells.append(fl)
    keep, cells = np.array(keep), np.array(cells, np.int64)
    assert len(keep) > 500
    codes = _morton_of(cells[:, 0], cells[:, 1], cells[:, 2], L)
    o = O.OracleOctree.from_points(keep, L, bmin, bmax)
    ref = O.OracleOctree.from_morton(codes, L)
    _levels_equal(o, ref, L, "points vs exact quantization")
