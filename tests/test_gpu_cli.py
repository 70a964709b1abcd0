"""The drop-in command lines (svbuilder, svviewer) against the oracle.

svbuilder is run exactly like the reference's (`svbuilder model.obj L step [-c]`);
its output files must be byte-identical to the oracle's restatement of the
reference svbuilder path (mesh voxelized in triangle-major order, toDAG,
encoders, toSDAG, encoders; buildDAG for step > 0; mergeAcrossAllLevels for -c).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import svdag_amd as S
import svdag_amd.workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "svdag-compression_amd", "bin")


def run(args, cwd):
    print("[case]", " ".join(args), flush=True)
    out = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    return out.stdout


@pytest.fixture(scope="module")
def mesh(tmp_path_factory):
    d = tmp_path_factory.mktemp("mesh")
    V, F = W.spheres_mesh(n_spheres=16, subdiv=2, seed=0)
    W.write_obj(str(d / "spheres.obj"), V, F)
    W.write_ply(str(d / "spheresply.ply"), V, F)
    bmin, bmax = W.mesh_bbox(V)
    return d, V[F], bmin.astype(np.float64), bmax.astype(np.float64)


def oracle_files(o):
    o.init_child_levels()
    f = {".svdag": o.encode(0), ".esvdag": o.encode(2)}
    o.to_sdag()
    f[".ussvdag"] = o.encode(1)
    f[".ssvdag"] = o.encode(2)
    return f


@pytest.mark.parametrize("step", [0, 2])
def test_svbuilder_files_equal_oracle(mesh, step):
    d, tris, bmin, bmax = mesh
    L = 7
    out = run([os.path.join(BIN, "svbuilder"), "spheres.obj", str(L), str(step)], d)
    assert "RESULTS" in out
    if step == 0:
        o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
        o.to_dag()
    else:
        o = O.OracleOctree.from_mesh_subtrees(tris, L, step, bmin, bmax)
    want = oracle_files(o)
    for ext, data in want.items():
        got = open(d / f"spheres_{L}{ext}", "rb").read()
        assert got == data, ext
    assert os.path.exists(d / "stats.txt")


@pytest.mark.parametrize("largs,infix,factor,refs", [([], "-l_2.0_1.0_1", 1.0, 1),
                                                     (["1.5", "0.5", "2"], "-l_1.5_0.5_2", 0.5, 2)])
def test_svbuilder_lossy_equals_oracle(mesh, largs, infix, factor, refs):
    """svbuilder model.obj L 0 -l [infl factor refs] (main.cpp:116-137, 225-264)."""
    d, tris, bmin, bmax = mesh
    L = 7
    out = run([os.path.join(BIN, "svbuilder"), "spheres.obj", str(L), "0", "-l"] + largs, d)
    assert "Lossy compressed from" in out
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    assert open(d / f"spheres_{L}.esvdag", "rb").read() == o.encode(2)      # saved before the lossy pass
    st = o.lossy_merge(1, L - 2, max_diff=0, diff_factor=factor, max_refs=refs, list_cap=32)
    assert st["merged"].sum() > 0
    want = oracle_files(o)
    for ext, data in want.items():
        assert open(d / f"spheres_{L}{infix}{ext}", "rb").read() == data, ext


def test_svbuilder_ply_input_equals_obj(mesh):
    d, tris, bmin, bmax = mesh
    run([os.path.join(BIN, "svbuilder"), "spheresply.ply", "6", "0"], d)
    run([os.path.join(BIN, "svbuilder"), "spheres.obj", "6", "0"], d)
    for ext in (".svdag", ".esvdag", ".ussvdag", ".ssvdag"):
        assert open(d / f"spheresply_6{ext}", "rb").read() == open(d / f"spheres_6{ext}", "rb").read()


def test_svbuilder_cross_level(mesh):
    d, tris, bmin, bmax = mesh
    L = 7
    run([os.path.join(BIN, "svbuilder"), "spheres.obj", str(L), "0", "-c"], d)
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    plain = o.encode(0)
    assert o.merge_cross_levels() == 0
    assert open(d / f"spheres_{L}.svdag", "rb").read() == plain
    assert open(d / f"spheres_{L}-multi.svdag", "rb").read() == o.encode(0)


def test_svviewer_headless(mesh):
    d, *_ = mesh
    run([os.path.join(BIN, "svbuilder"), "spheres.obj", "6", "0"], d)
    for ext in (".svdag", ".ussvdag", ".ssvdag", ".esvdag"):
        out = run([os.path.join(BIN, "svviewer"), f"spheres_6{ext}", "--width", "320", "--height", "240",
                   "--frames", "2", "--image", f"img{ext}.ppm"], d)
        assert "Rays/s" in out
        assert os.path.getsize(d / f"img{ext}.ppm") > 320 * 240 * 3


def read_ppm(path):
    data = open(path, "rb").read()
    head, w, h, mx = data.split(maxsplit=3)[:4]
    assert head == b"P6" and mx.startswith(b"255")
    w, h = int(w), int(h)
    return np.frombuffer(data[-w * h * 3:], np.uint8).reshape(h, w, 3)


def viewer_params(buf, w, h, seed):
    """svviewer's frame: orbit camera around the file's box, renderer defaults, light at the box max."""
    bb = np.frombuffer(buf[:24], np.float32).astype(np.float64)
    levels = int(np.frombuffer(buf[28:32], np.uint32)[0])
    center = tuple((bb[:3] + bb[3:]) / 2)
    radius = 1.5 * float(np.sqrt(((bb[3:] - bb[:3]) ** 2).sum())) / 2.0
    p = W.trace_params(w, h, seed=seed, levels=levels, shadows=True, light=tuple(float(x) for x in bb[3:].astype(np.float32)),
                       center=center, radius=radius)
    return p


def test_svviewer_image_is_the_viewer_frame(mesh):
    """svviewer --image: the viewer's colour (pretty mode, normal samples,
    shadows, beam pre-pass on as in the reference) equals the oracle's frame
    after the framebuffer's clamp to [0, 1] and 8-bit rounding; with --ao the
    image is the AO pass's visibility."""
    d, *_ = mesh
    run([os.path.join(BIN, "svbuilder"), "spheres.obj", "6", "0"], d)
    w, h = 160, 120
    for ext, fmt in ((".svdag", S.FMT_SVDAG), (".ssvdag", S.FMT_SSVDAG)):
        buf = open(d / f"spheres_6{ext}", "rb").read()
        p = viewer_params(buf, w, h, seed=1)
        run([os.path.join(BIN, "svviewer"), f"spheres_6{ext}", "--width", str(w), "--height", str(h), "--frames", "1",
             "--mode", "3", "--norm-samples", "2", "--image", f"view{ext}.ppm"], d)
        rgb, prim = O.render_frame(buf, fmt, p, S.ViewParams.defaults(render_mode=3, norm_samples=2),
                                   W.projection_factor(h, screen_divisor=8.0))
        assert (prim["t"] >= 0).mean() > 0.1
        q = np.rint(np.clip(np.nan_to_num(rgb, nan=0.0), 0, 1) * 255).astype(np.uint8)[::-1]
        np.testing.assert_array_equal(read_ppm(d / f"view{ext}.ppm"), q, err_msg=ext)
        run([os.path.join(BIN, "svviewer"), f"spheres_6{ext}", "--width", str(w), "--height", str(h), "--frames", "1",
             "--ao", "8", "--image", f"ao{ext}.ppm"], d)
        vis, _ = O.trace_ao(buf, fmt, p, 8)
        q = np.rint(np.clip(vis, 0, 1) * 255).astype(np.uint8)[::-1]
        np.testing.assert_array_equal(read_ppm(d / f"ao{ext}.ppm"), np.repeat(q[..., None], 3, axis=2), err_msg=ext)


def test_svbuilder_svdag_input(mesh, tmp_path):
    """svbuilder model.svdag L 0 (main.cpp:97-101, 147): the decoded DAG skips
    the build and goes through the same passes as a mesh input; with -c the
    cross-level merge runs on the decoded DAG."""
    d, tris, bmin, bmax = mesh
    L = 7
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    (tmp_path / "in.svdag").write_bytes(o.encode(0))
    out = run([os.path.join(BIN, "svbuilder"), "in.svdag", str(L), "0"], tmp_path)
    assert "Decoding nodes" in out and "RESULTS" in out
    want = oracle_files(o)
    for ext, data in want.items():
        assert open(tmp_path / f"in_{L}{ext}", "rb").read() == data, ext
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    assert o.merge_cross_levels() == 0
    run([os.path.join(BIN, "svbuilder"), "in.svdag", str(L), "0", "-c"], tmp_path)
    assert open(tmp_path / f"in_{L}-multi.svdag", "rb").read() == o.encode(0)


def parse_obj_like_reference(text):
    """Triangles of an OBJ text by the reference's rules (scene.cpp:97-216):
    a face line is split on spaces (runs of spaces give no empty tokens), its
    last token loses its final character, each of the first three tokens is
    read as "%d//%d" (a vertex index, negative = from the end), and a second
    triangle (tokens 2, 3, 0) exists only when more than 4 tokens remain."""
    import re
    verts, tris = [], []
    num = re.compile(r"\s*([+-]?\d+)")
    for line in text.splitlines(keepends=True):
        if line.startswith("v") and line[1:2] not in ("n", "t"):
            xyz = np.array([float(x) for x in line[1:].split()[:3]], np.float32)
            verts.append(xyz)
        elif line.startswith("f"):
            toks = [t for t in line[1:].split(" ") if t]
            toks[-1] = toks[-1][:-1]
            cur = [None, None, None]

            def idx(t):
                m = num.match(t)
                v = int(m.group(1))
                return len(verts) + v + 1 if v < 0 else v
            for i in range(3):
                cur[i] = idx(toks[i]) - 1
            tris.append(list(cur))
            if len(toks) > 4:
                for i in range(3):
                    cur[i] = idx(toks[(i + 2) % 4]) - 1
                tris.append(list(cur))
    V = np.array(verts, np.float32)
    good = [t for t in tris if all(0 <= k < len(V) for k in t)]
    return V, V[np.array(good)]


def test_svbuilder_obj_quirks_and_bincache(tmp_path):
    """OBJ faces written as "a b c", "a//n b//n c//n", "a/t/n ...", negative
    indices, quads with and without a trailing token, "vn"/"vt" lines and
    comments: svbuilder's files equal the oracle's on the triangles the
    reference's parser produces.  The first run writes <obj>.bincache in the
    reference's layout; the second run loads it and writes the same files."""
    V, F = W.icosphere(2)
    V = (V * 0.4 + 0.5).astype(np.float32)
    lines = ["# quirks\n"]
    lines += [f"v {x:.6f} {y:.6f} {z:.6f}\n" for x, y, z in V]
    lines += ["vn 0 0 1\n", "vt 0.5 0.5\n"]
    n = len(V)
    for k, (a, b, c) in enumerate(F):
        a, b, c = a + 1, b + 1, c + 1
        form = k % 5
        if form == 0:
            lines.append(f"f {a} {b} {c}\n")
        elif form == 1:
            lines.append(f"f {a}//1 {b}//1 {c}//1\n")
        elif form == 2:
            lines.append(f"f {a}/1/1 {b}/1/1 {c}/1/1\n")
        elif form == 3:
            lines.append(f"f {a - n - 1} {b - n - 1} {c - n - 1}\n")         # negative: from the end
        else:
            lines.append(f"f {a} {b} {c}  \n")                                # runs of spaces
    # a quad with a trailing space (two triangles) and one without (only the first)
    base = len(V)
    quad = np.array([[0.1, 0.1, 0.9], [0.9, 0.1, 0.9], [0.9, 0.9, 0.9], [0.1, 0.9, 0.9]], np.float32)
    lines += [f"v {x:.6f} {y:.6f} {z:.6f}\n" for x, y, z in quad]
    lines.append(f"f {base + 1} {base + 2} {base + 3} {base + 4} \n")
    quad2 = quad.copy()
    quad2[:, 2] = 0.1
    lines += [f"v {x:.6f} {y:.6f} {z:.6f}\n" for x, y, z in quad2]
    lines.append(f"f {base + 5} {base + 6} {base + 7} {base + 8}\n")
    text = "".join(lines)
    (tmp_path / "quirks.obj").write_text(text)
    verts, tris = parse_obj_like_reference(text)
    assert len(tris) == len(F) + 3
    bmin, bmax = verts.min(0).astype(np.float64), verts.max(0).astype(np.float64)
    L = 6
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    o.to_dag()
    want = oracle_files(o)
    out1 = run([os.path.join(BIN, "svbuilder"), "quirks.obj", str(L), "0"], tmp_path)
    assert "Saving" in out1 and "bincache" in out1
    cache = (tmp_path / "quirks.obj.bincache").read_bytes()
    nv, nn, nm, nt = np.frombuffer(cache[:32], np.uint64)
    assert (nv, nn, nm, nt) == (len(verts), 0, 0, len(F) + 3)
    assert len(cache) == 56 + 12 * nv + 56 * nt
    np.testing.assert_array_equal(np.frombuffer(cache[56:56 + 12 * int(nv)], np.float32).reshape(-1, 3), verts)
    for ext, data in want.items():
        assert open(tmp_path / f"quirks_{L}{ext}", "rb").read() == data, ext
        os.remove(tmp_path / f"quirks_{L}{ext}")
    out2 = run([os.path.join(BIN, "svbuilder"), "quirks.obj", str(L), "0"], tmp_path)
    assert "Binary cache found" in out2
    for ext, data in want.items():
        assert open(tmp_path / f"quirks_{L}{ext}", "rb").read() == data, ext


def read_stats_blocks(path):
    return [b.strip().splitlines() for b in open(path).read().split("\n\n\n") if b.strip()]


def test_svbuilder_lossy_stats_block(mesh, tmp_path):
    """stats.txt after a lossy run carries the reference's lossy block
    (main.cpp:337-344): the construction times, the parameters, and
    TotalVoxDifference = the accounted error, #NodesIn - #ClustersOut = the
    merged nodes, #edges, and the .svdag sizes before and after."""
    d, tris, bmin, bmax = mesh
    L = 7
    (tmp_path / "spheres.obj").write_bytes(open(d / "spheres.obj", "rb").read())
    out = run([os.path.join(BIN, "svbuilder"), "spheres.obj", str(L), "0", "-l", "1.5", "0.5", "2"], tmp_path)
    blk = read_stats_blocks(tmp_path / "stats.txt")[-1]
    assert blk[0] == f"spheres, {L}, lossy: 1, cross-level: 0"
    assert blk[5] == "Construction times:"
    assert blk[6] == ",SVDAG, LSVDAG, Total, Hashing, SimNodes, Clustering"
    assert blk[7].startswith("time (ms), ") and len(blk[7].split(",")) == 7
    assert blk[8] == "Lossy stats, 1.50, 0.50, 2"
    assert blk[9] == "TotalVoxDifference, #NodesIn, #ClustersOut, #edges, svdag mem, lsvdag mem"
    diff, nodes_in, clusters, edges, mem0, mem1 = (int(x) for x in blk[10].split(","))
    o = O.OracleOctree.from_mesh(tris, L, bmin, bmax)
    o.to_dag()
    o.init_child_levels()
    lossless = o.encode(0)
    st = o.lossy_merge(1, L - 2, max_diff=0, diff_factor=0.5, max_refs=2, list_cap=32)
    assert diff == int(st["error_bound"].sum()) > 0
    assert nodes_in - clusters == int(st["merged"].sum())
    assert edges == int(st["verified"].sum())
    assert mem0 == len(lossless) - 44 and mem1 == len(o.encode(0)) - 44
    # the same block is printed on the console
    assert "Lossy stats, 1.50, 0.50, 2" in out


def test_cpp_api_travnode_copy_nodedata_pointer_sdag(tmp_path):
    """The C++ drop-in API through a program written against it
    (tests/cpp/symvox_api_test.cpp): the TravNode API enumerates exactly the
    input voxels in all four formats; the copy constructor, the NodeData
    constructor and decode() by value give the octree's own .svdag; the
    pointer-only SDAG (toSDAG skipSymmetry) is the DAG with no mirror bits, so
    its files are the DAG's own except the node count; the SSVDAG data
    pointers are the file's arrays."""
    from decode import decode_ssvdag, decode_words, parse_words
    L = 7
    codes = W.terrain_codes(L, seed=4)
    (tmp_path / "codes.bin").write_bytes(codes.astype("<u8").tobytes())
    out = run([os.path.join(BIN, "symvox_api_test"), "codes.bin", str(L), str(tmp_path)], tmp_path)
    kv = dict(line.split(" ", 1) for line in out.splitlines() if " " in line and not line.startswith("*"))
    for fmt in ("svdag", "esvdag", "ussvdag", "ssvdag"):
        got = np.fromfile(tmp_path / f"{fmt}.trav", np.uint64)
        np.testing.assert_array_equal(got, codes, err_msg=fmt)
        assert int(kv[f"{fmt}_voxels"]) == codes.size
    o = O.OracleOctree.from_morton(codes, L)
    o.to_dag()
    o.init_child_levels()
    dag_file, esvdag = o.encode(0), o.encode(2)
    D = [o.level_size(l) for l in range(L)]
    for name in ("dag", "copy", "nodedata", "decoded"):
        assert (tmp_path / f"{name}.svdag").read_bytes() == dag_file, name
    assert int(kv["nodedata_state"]) == 2 and int(kv["decoded_state"]) == 2
    assert int(kv["nodedata_nodes"]) == o.n_nodes == int(kv["assigned_nodes"])
    assert int(kv["svdag_nvoxels"]) == codes.size and int(kv["loaded_nvoxels"]) == 0
    assert kv["svdag_center"] == "0.500 0.500 0.500"
    # pointer-only SDAG: every DAG node its own representative
    assert int(kv["pointer_sdag_nodes"]) == sum(D[1:])
    ptr_u, ptr_s = (tmp_path / "ptr.ussvdag").read_bytes(), (tmp_path / "ptr.ssvdag").read_bytes()
    np.testing.assert_array_equal(decode_words(ptr_u, True), codes)
    np.testing.assert_array_equal(decode_ssvdag(ptr_s), codes)
    assert parse_words(ptr_u)["n_nodes"] == sum(D[1:])
    assert ptr_u[:32] == dag_file[:32] and ptr_u[36:] == dag_file[36:]
    assert ptr_s[:32] == esvdag[:32] and ptr_s[36:] == esvdag[36:]
    o.to_sdag()
    assert (tmp_path / "sym.ssvdag").read_bytes() == o.encode(2)
    assert kv["inner_ptr_ok"] == "1" and kv["leaf_ptr_ok"] == "1" and kv["leaf_child_default"] == "1"
    assert "WARNING! Asking for a node child" in out


def write_las(path, raw, scale, offset):
    """A LAS 1.2 file, point data format 0 (20-byte records), raw int32 X Y Z."""
    import struct
    n = raw.shape[0]
    xyz = raw * np.asarray(scale) + np.asarray(offset)
    hdr = bytearray(227)
    hdr[0:4] = b"LASF"
    hdr[24], hdr[25] = 1, 2
    struct.pack_into("<HIIBHI", hdr, 94, 227, 227, 0, 0, 20, n)
    struct.pack_into("<3d3d", hdr, 131, *scale, *offset)
    struct.pack_into("<6d", hdr, 179, xyz[:, 0].max(), xyz[:, 0].min(), xyz[:, 1].max(), xyz[:, 1].min(),
                     xyz[:, 2].max(), xyz[:, 2].min())
    rec = np.zeros(n, dtype=[("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("rest", "V8")])
    rec["x"], rec["y"], rec["z"] = raw[:, 0], raw[:, 1], raw[:, 2]
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(rec.tobytes())
    return xyz


def test_cpp_api_build_svo_from_las_points(tmp_path):
    """GeomOctree::buildSVOFromPoints(fileName, ...) of the drop-in
    (geom_octree.cpp:62-169) on a LAS file: its points are libLAS's raw *
    scale + offset, and the DAG's .svdag equals the oracle's from the same
    points in file order."""
    rng = np.random.default_rng(21)
    L = 8
    raw = np.concatenate([rng.integers(-40000, 40000, size=(50000, 3)),
                          rng.integers(-5, 5, size=(5000, 3)) * 8000]).astype(np.int64)
    raw = raw[rng.permutation(len(raw))]
    scale, offset = (0.001, 0.0025, 0.0005), (100.5, -20.25, 3.0)
    xyz = write_las(tmp_path / "cloud.las", raw, scale, offset)
    bmin = xyz.min(axis=0) - 0.01
    bmax = xyz.max(axis=0) - 1.0          # some points fall outside the box
    args = [os.path.join(BIN, "symvox_api_test"), "--las", "cloud.las", str(L), str(tmp_path)]
    args += [repr(float(v)) for v in list(bmin) + list(bmax)]
    out = run(args, tmp_path)
    o = O.OracleOctree.from_points(xyz, L, bmin, bmax)
    assert f"points_voxels {o.n_voxels}" in out
    o.to_dag()
    o.init_child_levels()
    assert (tmp_path / "points.svdag").read_bytes() == o.encode(0)
