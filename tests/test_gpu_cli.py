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
