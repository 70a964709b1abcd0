"""CPU tests: pin the oracle (the CPU restatement in oracle/) against brute
force, closed forms and invariants, and check the C ABI surface.

What each check pins:
  * toDAG level counts and node order  <- brute-force distinct-subtree
    enumeration in first-occurrence (Morton) order;
  * encoders + mirror semantics        <- expanding every file format gives
    back the input voxel set (decoders in tests/decode.py follow the shader's
    read side, not the oracle's write side);
  * SVDAG sizes                        <- closed form of the word count;
  * toSDAG                             <- S_l <= D_l and S_l >= number of
    reflection orbits of distinct subtrees (brute force);
  * mesh voxelizer                     <- every sampled surface point lies in a
    set voxel;
  * traversal port                     <- the hit voxel with LOD disabled is the
    first set voxel along the ray in a float64 dense-grid DDA;
  * refcount order                     <- the segment-parallel introsort
    restatement (the algorithm the GPU runs) equals libstdc++ std::sort.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from decode import decode_words, decode_ssvdag, parse_words

import svdag_amd
import svdag_amd.workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def random_codes(levels, n, seed):
    rng = np.random.default_rng(seed)
    side = 1 << levels
    pts = rng.integers(0, side, size=(n, 3))
    return W.sorted_codes(pts[:, 0], pts[:, 1], pts[:, 2])


def clustered_codes(levels, seed):
    """A voxel set with many repeated and mirrored subtrees."""
    rng = np.random.default_rng(seed)
    side = 1 << levels
    blk = rng.random((4, 4, 4)) < 0.35
    pts = []
    for bx in range(0, side, 4):
        for by in range(0, side, 4):
            for bz in range(0, side, 4):
                r = rng.integers(0, 5)
                if r == 0:
                    continue
                b = blk
                if r == 2:
                    b = b[::-1]
                if r == 3:
                    b = b[:, ::-1, ::-1]
                if r == 4:
                    b = rng.random((4, 4, 4)) < 0.2
                x, y, z = np.nonzero(b)
                pts.append(np.stack([x + bx, y + by, z + bz], 1))
    p = np.concatenate(pts)
    return W.sorted_codes(p[:, 0], p[:, 1], p[:, 2])


SCENES = {
    "terrain6": lambda: (W.terrain_codes(6, seed=1), 6),
    "random5": lambda: (random_codes(5, 3000, 7), 5),
    "clustered6": lambda: (clustered_codes(6, 3), 6),
    "caves5": lambda: (W.cave_codes(5, seed=3), 5),
}


# ---------------------------------------------------------------- toDAG
def subtree_keys(codes, levels, lev):
    """For every level-`lev` node (Morton prefix, in ascending order), the
    frozenset of voxel offsets below it."""
    shift = np.uint64(3 * (levels - lev))
    pre = codes >> shift
    rel = codes & ((np.uint64(1) << shift) - np.uint64(1))
    keys, order = {}, []
    starts = np.flatnonzero(np.r_[True, pre[1:] != pre[:-1]])
    ends = np.r_[starts[1:], len(codes)]
    for s, e in zip(starts, ends):
        order.append(frozenset(rel[s:e].tolist()))
    return order


@pytest.mark.parametrize("name", ["random5", "clustered6", "terrain6"])
def test_todag_counts_and_order_bruteforce(name):
    codes, L = SCENES[name]()
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    for lev in range(1, L):
        subs = subtree_keys(codes, L, lev)
        first = []
        seen = set()
        for k in subs:
            if k not in seen:
                seen.add(k)
                first.append(k)
        assert o.level_size(lev) == len(first), (lev, o.level_size(lev), len(first))
        # node i of the DAG level expands to the i-th distinct subtree in first-occurrence order
        words = o.level(lev)
        if lev == L - 1:
            for i, k in enumerate(first):
                assert words[i, 0] & 0xFF == sum(1 << int(v) for v in k)
        elif lev == L - 2:
            leaf = o.level(L - 1)[:, 0] & 0xFF
            for i, k in enumerate(first):
                got = set()
                for c in range(8):
                    if words[i, 0] >> c & 1:
                        m = leaf[words[i, 1 + c]]
                        got |= {c * 8 + b for b in range(8) if m >> b & 1}
                assert got == set(k)


# ---------------------------------------------------------------- encoders round trip
@pytest.mark.parametrize("name", list(SCENES))
def test_all_formats_roundtrip(name):
    codes, L = SCENES[name]()
    o = O.OracleOctree.from_morton(codes, L)
    assert o.n_voxels == codes.size
    o.to_dag()
    o.init_child_levels()
    svdag = o.encode(0)
    esvdag = o.encode(2)
    o.to_sdag()
    ussvdag = o.encode(1)
    ssvdag = o.encode(2)
    np.testing.assert_array_equal(decode_words(svdag, False), codes)
    np.testing.assert_array_equal(decode_words(ussvdag, True), codes)
    np.testing.assert_array_equal(decode_ssvdag(esvdag), codes)
    np.testing.assert_array_equal(decode_ssvdag(ssvdag), codes)


def test_svdag_size_closed_form():
    codes, L = SCENES["terrain6"]()
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    f = parse_words(o.encode(0))
    inner = sum(int(np.sum(1 + np.array([bin(m).count("1") for m in (o.level(l)[:, 0] & 0xFF)])))
                for l in range(L - 1))
    D = [o.level_size(l) for l in range(L)]
    # _firstLeafPtr is assigned after the loop over all levels (encoded_svdag.cpp:129-137)
    assert len(f["words"]) == inner + D[L - 1]
    assert f["first_leaf"] == len(f["words"])
    assert f["n_nodes"] == 1 + sum(D[1:])
    assert f["levels"] == L


# ---------------------------------------------------------------- toSDAG bounds
def _reflect(k, flip, side):
    out = set()
    for v in k:
        x, y, z = W.morton_decode(np.array([v], np.uint64))
        x, y, z = int(x[0]), int(y[0]), int(z[0])
        if flip & 4:
            x = side - 1 - x
        if flip & 2:
            y = side - 1 - y
        if flip & 1:
            z = side - 1 - z
        out.add(int(W.morton_encode(np.array([x]), np.array([y]), np.array([z]))[0]))
    return frozenset(out)


@pytest.mark.parametrize("name", ["clustered6", "random5"])
def test_tosdag_bounds_bruteforce(name):
    codes, L = SCENES[name]()
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    D = [o.level_size(l) for l in range(L)]
    o.to_sdag()
    S = [o.level_size(l) for l in range(L)]
    for lev in range(max(1, L - 3), L):
        distinct = set(subtree_keys(codes, L, lev))
        side = 1 << (L - lev)
        orbits = {min(tuple(sorted(_reflect(k, f, side))) for f in range(8)) for k in distinct}
        assert len(orbits) <= S[lev] <= D[lev], (lev, len(orbits), S[lev], D[lev])
    assert o.n_nodes == sum(S[1:])


# ---------------------------------------------------------------- mesh voxelizer
def test_mesh_voxelizer_covers_surface():
    V, F = W.spheres_mesh(n_spheres=6, subdiv=2, seed=0)
    tris = V[F]
    bmin, bmax = W.mesh_bbox(V)
    L = 6
    o = O.OracleOctree.from_mesh(tris, L, bmin.astype(np.float64), bmax.astype(np.float64))
    o.to_dag()
    o.init_child_levels()
    codes = decode_words(o.encode(0), False)
    rs = float(max(bmax - bmin))
    center = (bmin.astype(np.float64) + bmax.astype(np.float64)) / 2
    lo = center - rs / 2
    rng = np.random.default_rng(0)
    u = rng.random((len(tris), 2))
    m = u.sum(1) > 1
    u[m] = 1 - u[m]
    pts = tris[:, 0] + u[:, :1] * (tris[:, 1] - tris[:, 0]) + u[:, 1:] * (tris[:, 2] - tris[:, 0])
    g = (pts.astype(np.float64) - lo) / rs * (1 << L)
    frac = g - np.floor(g)
    ok = np.all((frac > 1e-3) & (frac < 1 - 1e-3), axis=1)      # skip points on voxel faces
    vi = np.floor(g[ok]).astype(np.int64)
    want = W.morton_encode(vi[:, 0], vi[:, 1], vi[:, 2])
    assert np.isin(want, codes).all()


# ---------------------------------------------------------------- traversal port
def dense_first_hit(occ, o, d, L):
    """float64 Amanatides-Woo DDA over the dense 2^L grid in unit-cube units."""
    n = 1 << L
    t0, t1 = 0.0, np.inf
    for a in range(3):
        if abs(d[a]) < 1e-12:
            if o[a] < 0 or o[a] > 1:
                return None
            continue
        ta, tb = (0 - o[a]) / d[a], (1 - o[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    if t0 > t1:
        return None
    p = o + (t0 + 1e-9) * d
    cell = np.clip(np.floor(p * n).astype(np.int64), 0, n - 1)
    step = np.sign(d).astype(np.int64)
    tnext = np.array([((cell[a] + (step[a] > 0)) / n - o[a]) / d[a] if step[a] else np.inf for a in range(3)])
    tdelta = np.array([abs(1.0 / (n * d[a])) if step[a] else np.inf for a in range(3)])
    while np.all((cell >= 0) & (cell < n)):
        if occ[cell[0], cell[1], cell[2]]:
            return tuple(cell)
        a = int(np.argmin(tnext))
        cell[a] += step[a]
        tnext[a] += tdelta[a]
    return None


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_trace_port_hits_first_voxel(fmt):
    L = 5
    codes = clustered_codes(L, 11)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    if fmt == 0:
        buf = o.encode(0)
    else:
        o.to_sdag()
        buf = o.encode(fmt)
    p = W.trace_params(48, 40, seed=5, levels=L, shadows=False)
    p.projection_factor = 1e30            # LOD off: descend to voxels
    prim, _ = O.trace_frame(buf, fmt, p, shadows=False)
    n = 1 << L
    x, y, z = W.morton_decode(codes)
    occ = np.zeros((n, n, n), bool)
    occ[x.astype(int), y.astype(int), z.astype(int)] = True
    vi, pi = np.array(p.view_inv).reshape(4, 4).T, np.array(p.proj_inv).reshape(4, 4).T
    agree = total = 0
    for py in range(p.height):
        for px in range(p.width):
            sx = (px + 0.5) / p.width * 2 - 1
            sy = (py + 0.5) / p.height * 2 - 1
            a = pi @ np.array([sx, sy, 0, 1.0]); b = pi @ np.array([sx, sy, 1, 1.0])
            dcam = b[:3] / b[3] - a[:3] / a[3]
            dcam /= np.linalg.norm(dcam)
            orig = (vi @ np.array([0, 0, 0, 1.0]))[:3]
            dw = (vi @ np.append(dcam, 1.0))[:3] - orig
            dw /= np.linalg.norm(dw)
            want = dense_first_hit(occ, orig.astype(np.float64), dw, L)
            r = prim[py * p.width + px]
            total += 1
            if r["t"] >= 0:
                assert int(r["level"]) == L - 1
                hit = orig + (r["t"] + 1e-6) * dw
                got = tuple(np.floor(hit * n).astype(int))
                agree += want is not None and got == want
            else:
                agree += want is None
    assert agree >= 0.99 * total, (agree, total)


def test_trace_formats_agree_without_lod():
    L = 5
    codes = clustered_codes(L, 5)
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    f0 = o.encode(0)
    o.to_sdag()
    f1, f2 = o.encode(1), o.encode(2)
    p = W.trace_params(40, 32, seed=9, levels=L, shadows=True)
    p.projection_factor = 1e30
    r0, s0 = O.trace_frame(f0, 0, p)
    r1, s1 = O.trace_frame(f1, 1, p)
    r2, s2 = O.trace_frame(f2, 2, p)
    np.testing.assert_array_equal(r0["t"], r1["t"])     # same tree, mirrors only change storage
    hit = r0["t"] >= 0
    assert np.array_equal(hit, r2["t"] >= 0)


# ---------------------------------------------------------------- refcount order
def eq_depth_ok(s, d):
    """refsort.hip: the introsort of s equal values at depth budget d reaches
    no heap sort (the larger half of a segment has ceil(s/2) elements)."""
    while s > 16:
        if d <= 0:
            return False
        s, d = (s + 1) // 2, d - 1
    return True


def eq_dest(x, s):
    """refsort.hip: final place of the element at place x among s equal
    values: the median of three is the middle, every element stops both
    scanners (pairs f+1+k <-> l-1-k swap for k < (s-1)//2), the cut is
    f+1+(s-1)//2."""
    f, l = 0, s
    while l - f > 16:
        sz = l - f
        mid, K = f + sz // 2, (sz - 1) // 2
        if x == f:
            x = mid
        elif x == mid:
            x = f
        if x > f:
            if x - (f + 1) < K:
                x = l - 1 - (x - (f + 1))
            elif l - 1 - x < K:
                x = f + 1 + (l - 1 - x)
        cut = f + 1 + K
        if x < cut:
            l = cut
        else:
            f = cut
    return x


def introsort_model(vals, eq_shortcut=False):
    """The segment-parallel restatement run by refsort.hip, in plain Python
    (eq_shortcut: segments of equal values placed in closed form, as its
    workgroup phase does)."""
    a = [(i, v) for i, v in enumerate(vals)]
    n = len(a)
    if n == 0:
        return []
    comp = lambda x, y: x[1] > y[1]

    def adjust(f, hole, ln, val):
        top, sc = hole, hole
        while sc < (ln - 1) // 2:
            sc = 2 * (sc + 1)
            if comp(a[f + sc], a[f + sc - 1]):
                sc -= 1
            a[f + hole] = a[f + sc]; hole = sc
        if ln % 2 == 0 and sc == (ln - 2) // 2:
            sc = 2 * (sc + 1); a[f + hole] = a[f + sc - 1]; hole = sc - 1
        parent = (hole - 1) // 2
        while hole > top and comp(a[f + parent], val):
            a[f + hole] = a[f + parent]; hole = parent; parent = (hole - 1) // 2
        a[f + hole] = val

    def heapsort(f, ln):
        if ln >= 2:
            parent = (ln - 2) // 2
            while True:
                adjust(f, parent, ln, a[f + parent])
                if parent == 0:
                    break
                parent -= 1
        last = ln
        while last > 1:
            last -= 1
            v = a[f + last]; a[f + last] = a[f]; adjust(f, 0, last, v)

    active, leaves = [(0, n, 2 * (n.bit_length() - 1))], []
    while active:
        nxt = []
        for first, last, depth in active:
            if last - first <= 16:
                leaves.append((first, last)); continue
            if eq_shortcut and len({x[1] for x in a[first:last]}) == 1 and eq_depth_ok(last - first, depth):
                seg = a[first:last]                  # the workgroup phase's closed form
                for x in range(last - first):
                    a[first + eq_dest(x, last - first)] = seg[x]
                continue
            if depth == 0:
                heapsort(first, last - first); continue
            mid = first + (last - first) // 2
            A, B, Cc = first + 1, mid, last - 1
            if comp(a[A], a[B]):
                m = B if comp(a[B], a[Cc]) else (Cc if comp(a[A], a[Cc]) else A)
            else:
                m = A if comp(a[A], a[Cc]) else (Cc if comp(a[B], a[Cc]) else B)
            a[first], a[m] = a[m], a[first]
            p = a[first][1]
            Lp = [i for i in range(first + 1, last) if a[i][1] <= p]
            Rp = [i for i in range(last - 1, first - 1, -1) if a[i][1] >= p]
            K = 0
            while K < len(Lp) and K < len(Rp) and Lp[K] < Rp[K]:
                K += 1
            for k in range(K):
                a[Lp[k]], a[Rp[k]] = a[Rp[k]], a[Lp[k]]
            rprev = Rp[K - 1] if K else last
            cut = Lp[K] if (K < len(Lp) and Lp[K] < rprev) else rprev
            nxt += [(cut, last, depth - 1), (first, cut, depth - 1)]
        active = nxt
    for first, last in leaves:
        seg = a[first:last]
        seg.sort(key=lambda x: -x[1])     # Python's sort is stable = insertion sort's result
        a[first:last] = seg
    return [x[0] for x in a]


@pytest.mark.parametrize("seed", range(12))
def test_introsort_model_equals_std_sort(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 4000)) if seed else 17
    hi = [2, 3, 5, 50, 10 ** 6][seed % 5]
    vals = rng.integers(0, hi, n).astype(np.uint32)
    if seed % 4 == 3:
        vals = np.sort(vals)
    idx, _ = O.std_sort_refcount(np.arange(n, dtype=np.uint32), vals)
    assert introsort_model(vals.tolist()) == idx.tolist()
    assert introsort_model(vals.tolist(), eq_shortcut=True) == idx.tolist()


def test_equal_values_closed_form_equals_std_sort():
    """The closed-form place of every element among s equal values is where
    libstdc++ std::sort puts it, for every s up to 3000; and the shortcut
    never applies where the recursion would reach a heap sort."""
    for s in range(1, 3001):
        idx, _ = O.std_sort_refcount(np.arange(s, dtype=np.uint32), np.full(s, 7, np.uint32))
        want = np.arange(s) if s <= 16 else np.array([eq_dest(int(i), s) for i in idx])
        np.testing.assert_array_equal(want, np.arange(s), err_msg=f"s={s}")
        assert eq_depth_ok(s, 2 * (s.bit_length() - 1))
    assert not eq_depth_ok(17, 0) and eq_depth_ok(16, 0) and not eq_depth_ok(33, 1)


# ---------------------------------------------------------------- C ABI surface
def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "svdag_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void\*|const char\*|void)\s+\**(svd_\w+)\s*\(", hdr, re.M))
    assert declared and "svd_trace" in declared
    so = svdag_amd.LIB_PATH
    if not os.path.exists(so):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(svd_\w+)\b", out))
    assert declared <= exported, declared - exported
    assert set(svdag_amd.EXPORTED) <= exported


def test_library_loads_without_gpu():
    if not os.path.exists(svdag_amd.LIB_PATH):
        pytest.skip("library not built")
    L = svdag_amd.lib()
    for name in svdag_amd.EXPORTED:
        assert hasattr(L, name)


# ---------------------------------------------------------------- CSVDAG (oracle)
@pytest.mark.parametrize("name", ["terrain6", "clustered6", "caves5", "random5"])
def test_cross_level_merge_is_lossless(name):
    """Cross-level pointers must not change the voxel set read by the shader
    (a leaf mask equals the mask of the deeper-level node it points to)."""
    codes, L = SCENES[name]()
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    before = [o.level_size(l) for l in range(L)]
    assert o.merge_cross_levels() == 0
    after = [o.level_size(l) for l in range(L)]
    assert o.cross_merged == sum(before) - sum(after)
    assert o.n_nodes == 1 + sum(after[1:])
    np.testing.assert_array_equal(decode_words(o.encode(0), False), codes)


# ---------------------------------------------------------------- Gray-code attribute bit-tree (oracle)
def _plane_bits(blob):
    import struct
    return struct.unpack_from("<8I", blob, 16)


@pytest.mark.parametrize("n", [0, 1, 2, 5, 64, 1000, 4097])
@pytest.mark.parametrize("gray", [False, True])
def test_bittree_roundtrip(n, gray):
    a = np.random.default_rng(n).integers(0, 256, n).astype(np.uint8)
    assert np.array_equal(O.attr_bittree_decode(O.attr_bittree_encode(a, gray), n), a)


def test_bittree_closed_forms():
    # constant sequence: every plane is one uniform root -> 2 bits; a single voxel -> 1 bit
    assert _plane_bits(O.attr_bittree_encode(np.full(100, 9, np.uint8), False)) == (2,) * 8
    assert _plane_bits(O.attr_bittree_encode(np.array([3], np.uint8), True)) == (1,) * 8
    # alternating plane 0 over P = 128: every internal node mixed -> 2P - 1 bits
    alt = (np.arange(128) % 2).astype(np.uint8)
    assert _plane_bits(O.attr_bittree_encode(alt, False))[0] == 255
    # a step of +-1 in a Gray-coded ramp changes exactly one plane: with one
    # element per value, planes k = 0..3 flip 8, 4, 2, 1 times (15 in total);
    # a uniform plane costs 2 bits, so planes 4..7 cost 8 bits together
    ramp = np.arange(16, dtype=np.uint8)
    g = _plane_bits(O.attr_bittree_encode(ramp, True))
    assert g[4:] == (2, 2, 2, 2)
    gcode = ramp ^ (ramp >> 1)
    flips = [int(np.sum(((gcode[1:] >> k) & 1) != ((gcode[:-1] >> k) & 1))) for k in range(4)]
    assert flips == [8, 4, 2, 1]


def test_bittree_gray_beats_binary_on_coherent_materials():
    codes = W.terrain_codes(6, seed=1)
    a = W.material_attrs(codes, 6, seed=4)
    g = sum(_plane_bits(O.attr_bittree_encode(a, True)))
    b = sum(_plane_bits(O.attr_bittree_encode(a, False)))
    assert g < b, (g, b)
