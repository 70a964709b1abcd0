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
