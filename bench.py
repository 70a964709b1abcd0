"""Benchmark of the sparse-voxel-DAG hot path on one MI355X.

Workload (the GPU configuration of the headline metric): 2048^3 fBm terrain
(11 levels, ~1.4e7 surface voxels, seed 1), voxel list resident in HBM as
sorted Morton codes.  One step =
    build: voxel list -> SVO -> DAG -> SDAG -> .svdag + .esvdag + .ussvdag + .ssvdag
           images in HBM (the svbuilder lossless path),
    trace: 1920x1080 primary rays + 1 shadow ray per hit over the .ssvdag
           (seeded orbit camera, FOV 60, maxIters 300, pixel tolerance 1).
value = build Gvox/s (voxels / build time); traversal Mrays/s is reported
beside it.  The roofline object is for the traversal kernel (the dominant
kernel): achieved = algorithmic bytes fetched by the shader's fetch functions
(counted per ray by the kernel itself) / its event-timed duration.
build_roofline: the build's algorithmic bytes (voxel list, level records and
remaps of the DAG and symmetry passes, leaf masks, the four files) / build time.
cpu_baseline: the reference's OpenMP CPU builder (buildDAG over subtrees in
parallel, then the serial toDAG / toSDAG / encoders) restated in the oracle,
on this host's cores, over the same workload and region.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "svdag-compression_amd")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    spec = importlib.util.spec_from_file_location("svdag_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["svdag_amd"] = m
    spec.loader.exec_module(m)
    return m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_step(S, ctx, dcodes_ptr, n, L):
    """The lossless build: voxel list -> SVO -> toDAG -> toSDAG -> the four
    images (.svdag, .esvdag of the DAG; .ussvdag, .ssvdag of the SDAG), the
    images in one call whose passes the formats share."""
    g = S.Octree.from_morton(ctx, (dcodes_ptr, n), L)
    g.to_dag()
    g.to_sdag()
    g.encode_many([S.FMT_SVDAG, S.FMT_ESVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG])
    return g


CPU_STEP_LEVEL = 3


def host_cores():
    """Logical CPUs of this host and physical cores (lscpu), for the baseline's record."""
    import subprocess
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        phys = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count(), phys


def cpu_build(codes, L, step_level=CPU_STEP_LEVEL):
    """The reference's OpenMP CPU builder on the oracle (svbuilder's levelStep
    path): buildDAG over subtrees in parallel, the final toDAG, then the serial
    .svdag / .esvdag encoders, toSDAG and the .ussvdag / .ssvdag encoders.
    Same region as the GPU step: voxel list -> four encoded images.  Returns
    the wall time in seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.perf_counter()
    o = O.OracleOctree.from_morton_subtrees(codes, L, step_level)
    o.init_child_levels()
    o.encode(0)
    o.encode(2)
    o.to_sdag()
    o.encode(1)
    o.encode(2)
    t = time.perf_counter() - t0
    del o
    return t


def cpu_baseline(codes, L):
    """The OpenMP CPU builder timed on this host's cores over the full
    headline workload (the same region and voxel list as the GPU step)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = cpu_build(codes, L)
    nproc, phys = host_cores()
    return dict(value=codes.size / t / 1e9, unit="Gvox/s", cores=O.omp_threads(), kind="port",
                sample=f"the whole workload: terrain {1 << L}^3 ({L} levels, {codes.size} voxels, seed 1), "
                       f"voxel list -> buildDAG (levelStep {CPU_STEP_LEVEL}: subtrees SVO + toDAG in an OpenMP "
                       f"parallel for, schedule(dynamic,2), {O.omp_threads()} threads) -> toDAG -> .svdag + "
                       f".esvdag -> toSDAG -> .ussvdag + .ssvdag (std::map / std::sort, the reference's "
                       f"algorithms), {t:.2f} s; host: {nproc} logical CPUs, {phys} physical cores")


def build_bytes(S, g, L, V):
    """Algorithmic HBM bytes of one lossless build: the 8-byte voxel list; per
    inner level a 36-byte SVO record read, a 4-byte remap and a 36-byte DAG
    record written, then a 36-byte DAG record read, a 4-byte remap and a
    36-byte SDAG record written; 4 + 1 bytes per SVO leaf mask and 4 + 1 per
    DAG leaf mask; plus the four encoded files written."""
    N = [g.level_size(S.VIEW_SVO, l) for l in range(L)]
    D = [g.level_size(S.VIEW_DAG, l) for l in range(L)]
    Q = [g.level_size(S.VIEW_SDAG, l) for l in range(L)]
    F = sum(int(g.encoded_device(f)[1]) for f in (S.FMT_SVDAG, S.FMT_ESVDAG, S.FMT_USSVDAG, S.FMT_SSVDAG))
    B = 8 * V
    B += sum(40 * N[l] + 36 * D[l] for l in range(L - 1)) + 5 * N[L - 1] + D[L - 1]
    B += sum(40 * D[l] + 36 * Q[l] for l in range(L - 1)) + 5 * D[L - 1] + Q[L - 1]
    return int(B + F), int(F)


LOSSY_BYTES_PER_NODE = 1284   # 64 B occupancy + 64 B embedding write + 64 B read + 16 x (64 + 4) B verify + 4 B remap
MFMA_BF16_PEAK_TFLOPS = 2500.0
MLP_FWD_FLOP_PER_ROW = 2 * (512 * 256 + 256 * 128 + 128 * 32)
# forward + activation gradients of layers 3 and 2 + the three weight gradients
MLP_TRAIN_FLOP_PER_ROW = 2 * MLP_FWD_FLOP_PER_ROW + 2 * (128 * 32 + 256 * 128)


def mfma_rate(flop, ms):
    t = flop / (ms * 1e-3) / 1e12
    return {"tflops": round(t, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "frac": round(t / MFMA_BF16_PEAK_TFLOPS, 5),
            "flop": int(flop)}


def lossy_bench(S, ctx, levels, log):
    """Lossy DAG on the cave volume: learned proposer (MLP embeddings + kNN,
    every candidate verified exactly) against the exhaustive proposers (first
    fit, and ranked: the nearest earlier nodes by Hamming distance), same
    voxel bound, levels L-6..L-3, nodes with at most one effective reference;
    and the learned kNN's recall@16 on the first merged level."""
    lo, hi, bound = levels - 6, levels - 3, 10
    t0 = time.perf_counter()
    codes = ctx.caves(levels, seed=3)
    log(f"[bench] caves {1 << levels}^3: {codes.n} voxels ({time.perf_counter() - t0:.1f} s)")

    def fresh():
        g = S.Octree.from_morton(ctx, (codes.ptr, codes.n), levels)
        g.to_dag()
        return g

    g = fresh()
    nodes_dag = g.counts()[0]
    m = S.MLP(ctx, seed=1)
    pairs_per_step, steps = 4096, 1024
    tr = m.train(g, lo, hi, pairs_per_step=pairs_per_step, steps=steps, lr=2e-3, seed=5, near_pct=80)
    log(f"[bench] trained on {tr.pairs} pairs: loss {tr.first_loss:.2f} -> {tr.last_loss:.2f} ({tr.ms_train:.0f} ms)")
    g.free()
    out = {"workload": f"caves {1 << levels}^3 ({levels} levels, seed 3, {codes.n} voxels): lossy DAG merge of "
                       f"levels {lo}-{hi}, bound {bound} voxels, nodes with <= 1 effective reference; "
                       f"learned = MLP 512-256-128-32 embeddings -> LSH orderings (8 tables x 16 bits, +-64 window) "
                       f"+ 2 neighbour-of-neighbour rounds, candidates ranked and verified exactly (XOR-popcount, "
                       f"voxel difference)",
           "nodes_dag": int(nodes_dag),
           "training": {"pairs": int(tr.pairs), "ms": round(tr.ms_train, 1), "first_loss": round(tr.first_loss, 3),
                        "last_loss": round(tr.last_loss, 3),
                        # whole training steps (sampling, loss, Adam included) over the GEMM work
                        "mfma": mfma_rate(2 * tr.pairs * MLP_TRAIN_FLOP_PER_ROW, tr.ms_train)}}
    g = fresh()
    sig_hi = level_signatures(S, g, hi)
    g.free()
    # exhaustive = the first list_cap earlier nodes within the bound;
    # ranked_exhaustive = the list_cap nearest of all of them (nearest first)
    for tag, model, ranked in (("learned", m, False), ("exhaustive", None, False), ("ranked_exhaustive", None, True)):
        g = fresh()
        r = lossy_run(g, lo, hi, bound, model, ranked=ranked, record=model is not None)
        if model is not None:
            out[tag], knn = r
            out["embedding_mfma"] = mfma_rate(out[tag]["nodes_embedded"] * MLP_FWD_FLOP_PER_ROW,
                                              out[tag]["ms_phases"]["ms_embed"])
            out["recall_at_16"] = knn_recall(sig_hi, knn[: sig_hi.shape[0] * 16].reshape(-1, 16), 16, 4096, seed=9)
            log(f"[bench] learned recall@16 on level {hi}: {out['recall_at_16']}")
        else:
            out[tag] = r
        log(f"[bench] lossy {tag}: {out[tag]}")
        g.free()
    codes.free()
    le, ex, rk = out["learned"], out["exhaustive"], out["ranked_exhaustive"]
    out["compression_gap"] = round((le["nodes_after"] - ex["nodes_after"]) / ex["nodes_after"], 5)
    out["speedup_vs_exhaustive"] = round(le["nodes_s"] / ex["nodes_s"], 2)
    out["compression_gap_vs_ranked"] = round((le["nodes_after"] - rk["nodes_after"]) / rk["nodes_after"], 5)
    out["speedup_vs_ranked"] = round(le["nodes_s"] / rk["nodes_s"], 2)
    achieved = le["candidates"] * LOSSY_BYTES_PER_NODE / (le["ms"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 5), "bytes_per_node": LOSSY_BYTES_PER_NODE}
    return out


def lossy_run(g, lo, hi, bound, model, ranked, record):
    """One lossy merge; its counts, device time, throughput and phases."""
    r = g.lossy_merge(lo, hi, max_diff=bound, model=model, ranked=ranked, record_knn=record)
    st, knn = r if record else (r, None)
    cand = int(st["candidates"].sum())
    ms = float(st["ms_total"].sum())
    res = {"nodes_after": int(g.counts()[0]), "merged": int(st["merged"].sum()), "candidates": cand,
           "ms": round(ms, 3), "nodes_s": round(cand / (ms * 1e-3), 1),
           "ms_phases": {k: round(float(st[k].sum()), 3) for k in
                         ("ms_embed", "ms_knn", "ms_lists", "ms_verify", "ms_merge")}}
    if model is not None:
        res["nodes_embedded"] = int(st["nodes_before"].sum())
    return (res, knn) if record else res


def level_signatures(S, g, lev):
    """The 512-bit occupancies of a DAG level, as a device tensor (n x 8 int64)."""
    import torch
    n = g.level_size(S.VIEW_DAG, lev)
    sig = torch.zeros(max(n, 1) * 16, dtype=torch.int32, device="cuda:0")
    g.lossy_signatures(lev, sig.data_ptr())
    torch.cuda.synchronize()
    return sig.view(torch.int64).view(-1, 8)[:n]


def _popcount64(x):
    # SWAR popcount on int64 (arithmetic shifts are masked off)
    x = x - ((x >> 1) & 0x5555555555555555)
    x = (x & 0x3333333333333333) + ((x >> 2) & 0x3333333333333333)
    x = (x + (x >> 4)) & 0x0F0F0F0F0F0F0F0F
    return (x * 0x0101010101010101) >> 56


def knn_recall(sig, knn, k, sample, seed):
    """Recall@k of proposed neighbour lists against the true k nearest nodes by
    Hamming distance of the signatures (exact, on the device), over a sample
    of nodes: the share of a node's k true-nearest distance classes that its
    proposals reach (a proposal counts when its distance is at most the k-th
    smallest true distance, so ties are not penalised)."""
    import torch
    n = sig.shape[0]
    rng = np.random.default_rng(seed)
    q = np.sort(rng.choice(n, size=min(sample, n), replace=False))
    prop = torch.from_numpy(knn[q].astype(np.int64)).to(sig.device)
    hits, tot = 0, 0
    for a in range(0, q.size, 16):
        qa = torch.from_numpy(q[a:a + 16]).to(sig.device)
        d = _popcount64(sig[qa][:, None, :] ^ sig[None, :, :]).sum(dim=2)      # [c, n]
        d[torch.arange(qa.numel(), device=sig.device), qa] = 1 << 20            # not itself
        kth = torch.kthvalue(d, k, dim=1).values                                # k-th smallest true distance
        p = prop[a:a + 16]
        valid = (p >= 0) & (p < n) & (p != qa[:, None])
        pd = torch.where(valid, d.gather(1, p.clamp(0, n - 1)), torch.full_like(p, 1 << 20))
        # distinct proposals only
        srt, _ = torch.sort(torch.where(valid, p, torch.full_like(p, -1)), dim=1)
        dup = torch.zeros_like(valid)
        dup[:, 1:] = (srt[:, 1:] == srt[:, :-1]) & (srt[:, 1:] >= 0)
        good = (pd <= kth[:, None]).sum(dim=1) - dup.sum(dim=1)
        hits += int(torch.clamp(good, max=k).sum())
        tot += k * qa.numel()
    return round(hits / tot, 4)


def batch_bench(S, W, ctx, n_scenes, log):
    """Config-3 batch: n_scenes scenes of uneven depth (6..10), kinds drawn per
    scene (icosphere balls, fBm terrain, Menger sponge surfaces), generated on
    the device; one batched build (SVO -> DAG -> .svdag images for all scenes)
    against the same scenes built one by one, then 64 cameras of 640x360
    spread over the scenes traced in one launch."""
    import torch
    specs = W.batch_specs(n_scenes, 0)
    t0 = time.perf_counter()
    dev = [ctx.spec_codes(sp) for sp in specs]
    levels = [sp["levels"] for sp in specs]
    off = np.concatenate([[0], np.cumsum([d.n for d in dev])]).astype(np.uint64)
    nvox = int(off[-1])
    log(f"[bench] batch: {n_scenes} scenes, {nvox} voxels generated ({time.perf_counter() - t0:.1f} s)")
    buf = torch.empty(nvox, dtype=torch.int64, device="cuda:0")
    for d, o in zip(dev, off[:-1]):
        ctx.check(ctx.L.svd_device_copy(ctx.h, S.C.c_void_p(buf.data_ptr() + 8 * int(o)), S.C.c_void_p(d.ptr),
                                        8 * d.n))
    tb = []
    b = None
    for r in range(3):
        if b is not None:
            b.free()
        torch.cuda.synchronize()
        t = time.perf_counter()
        b = S.Batch(ctx, buf.data_ptr(), off, levels, on_device=True)
        ctx.sync()
        tb.append(time.perf_counter() - t)
    t = time.perf_counter()
    for d, L in zip(dev, levels):
        g = S.Octree.from_morton(ctx, (d.ptr, d.n), L)
        g.to_dag()
        g.encode(S.FMT_SVDAG)
        g.free()
    ctx.sync()
    t_seq = time.perf_counter() - t
    for d in dev:
        d.free()
    n_cam, cw, ch = 64, 640, 360
    cam_scene = [k * n_scenes // n_cam for k in range(n_cam)]
    cams = [W.trace_params(cw, ch, seed=100 + k, levels=levels[s], shadows=True) for k, s in enumerate(cam_scene)]
    nr = n_cam * cw * ch
    prim = torch.zeros(nr * 6, dtype=torch.int32, device="cuda:0")
    shad = torch.zeros(nr * 6, dtype=torch.int32, device="cuda:0")
    b.trace(cams, cam_scene, prim.data_ptr(), shad.data_ptr())          # warm-up
    tt = []
    for _ in range(3):
        b.trace(cams, cam_scene, prim.data_ptr(), shad.data_ptr())
        tt.append(ctx.kernel_ms("trace"))
    pr = prim.cpu().numpy().view(S.RAY_DTYPE)
    rays = nr + int((pr["t"] >= 0).sum())
    t_b = float(np.median(tb))
    out = {"workload": f"{n_scenes} scenes, depth 6-10, {nvox} voxels (spheres / fBm terrain / Menger sponge "
                       f"surfaces); {n_cam} cameras x {cw}x{ch} primary + shadow rays",
           "build_ms": round(t_b * 1e3, 2), "build_gvox_s": round(nvox / t_b / 1e9, 4),
           "sequential_build_ms": round(t_seq * 1e3, 2), "speedup_vs_sequential": round(t_seq / t_b, 2),
           "trace_ms": round(float(np.median(tt)), 3), "rays": rays,
           "trace_mrays_s": round(rays / (float(np.median(tt)) * 1e-3) / 1e6, 1)}
    b.free()
    log(f"[bench] batch: {out}")
    return out


def mesh_bench(S, W, ctx, levels, log, cpu=True):
    """Mesh input (the svbuilder case): 64 icospheres (subdivision 3, 81,920
    triangles, seed 0) voxelized at 2^levels by the device buildSVO, reduced and
    encoded to the four images.  The oracle runs the same region (buildSVO in
    triangle-major DFS order + toDAG + encoders + toSDAG + encoders) on one
    core when cpu is set (this section's CPU baseline leg)."""
    V, F = W.spheres_mesh(n_spheres=64, subdiv=3, seed=0)
    tris = np.ascontiguousarray(V[F], np.float32)
    bmin, bmax = W.mesh_bbox(V)
    bmin, bmax = bmin.astype(np.float64), bmax.astype(np.float64)

    def gpu_once():
        t0 = time.perf_counter()
        g = S.Octree.from_mesh(ctx, tris, levels, bmin, bmax)
        t1 = time.perf_counter()
        g.to_dag()
        for f in (S.FMT_SVDAG, S.FMT_ESVDAG):
            g.encode(f)
        g.to_sdag()
        for f in (S.FMT_USSVDAG, S.FMT_SSVDAG):
            g.encode(f)
        ctx.sync()
        t2 = time.perf_counter()
        return g, t1 - t0, t2 - t0

    gpu_once()
    runs = [gpu_once() for _ in range(3)]
    g = runs[-1][0]
    nvox = g.counts()[1]
    t_vox = float(np.median([r[1] for r in runs]))
    t_all = float(np.median([r[2] for r in runs]))
    out = {"workload": f"64 icospheres (subdivision 3, {tris.shape[0]} triangles, seed 0) at {1 << levels}^3 "
                       f"({levels} levels, {nvox} voxels): buildSVO + toDAG + .svdag/.esvdag, toSDAG + .ussvdag/.ssvdag",
           "voxels": nvox, "svo_ms": round(t_vox * 1e3, 3), "total_ms": round(t_all * 1e3, 3)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        o = O.OracleOctree.from_mesh(tris, levels, bmin, bmax)
        t1 = time.perf_counter()
        o.to_dag()
        o.init_child_levels()
        o.encode(0)
        o.encode(2)
        o.to_sdag()
        o.encode(1)
        o.encode(2)
        t2 = time.perf_counter()
        out["cpu_oracle_svo_ms"] = round((t1 - t0) * 1e3, 1)
        out["cpu_oracle_total_ms"] = round((t2 - t0) * 1e3, 1)
        out["speedup_total"] = round((t2 - t0) / t_all, 1)
    log(f"[bench] mesh: {out}")
    return out


def attr_bench(S, W, ctx, levels, log, cpu=True):
    """Config-5 scene: fBm terrain (seed 1) union caves (seed 3) at 2^levels,
    per-voxel u8 materials (value noise seed 4, 16 levels).  Cross-level merge
    of the DAG, Gray vs binary attribute bit-trees, and a 3840x2160 frame
    (primary + shadow rays) on the scene's lossless SSVDAG."""
    import torch
    t0 = time.perf_counter()
    terr, cav = ctx.terrain(levels, seed=1), ctx.caves(levels, seed=3)
    u = ctx.union(terr, cav)
    terr.free()
    cav.free()
    log(f"[bench] attr scene {1 << levels}^3: {u.n} voxels ({time.perf_counter() - t0:.1f} s)")
    out = {"workload": f"fBm terrain (seed 1) + caves (seed 3) {1 << levels}^3 ({levels} levels, {u.n} voxels), "
                       f"u8 materials (value noise seed 4, 16 levels)", "voxels": int(u.n)}
    # the full lossless build (the headline step's build region) at this size
    tb = []
    for _ in range(4):
        torch.cuda.synchronize()
        t = time.perf_counter()
        g = build_step(S, ctx, u.ptr, u.n, levels)
        ctx.sync()
        tb.append(time.perf_counter() - t)
        g.free()
    bt = float(np.median(tb[1:]))
    out["lossless_build_ms"] = round(bt * 1e3, 2)
    out["lossless_build_gvox_s"] = round(u.n / bt / 1e9, 3)
    if cpu:
        # the OpenMP CPU builder on this scene, same region (voxel list -> four images)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        host = u.numpy()
        t_cpu = cpu_build(host, levels)
        del host
        out["cpu_build_s"] = round(t_cpu, 2)
        out["cpu_build_gvox_s"] = round(u.n / t_cpu / 1e9, 5)
        out["cpu_build_threads"] = O.omp_threads()
        out["lossless_build_vs_cpu_same_scene"] = round(t_cpu / bt, 1)
        log(f"[bench] attr: OpenMP CPU builder {t_cpu:.2f} s on {O.omp_threads()} threads")
    g = S.Octree.from_morton(ctx, (u.ptr, u.n), levels)
    g.to_dag()
    out["nodes_dag"] = int(g.counts()[0])
    g.encode(S.FMT_SVDAG)
    out["svdag_bytes"] = int(g.encoded_device(S.FMT_SVDAG)[1])
    torch.cuda.synchronize()
    t = time.perf_counter()
    elim = g.merge_cross_levels()
    ctx.sync()
    out["cross_level_ms"] = round((time.perf_counter() - t) * 1e3, 2)
    out["cross_level_eliminated"] = int(elim)
    g.encode(S.FMT_SVDAG)
    out["csvdag_bytes"] = int(g.encoded_device(S.FMT_SVDAG)[1])
    g.free()
    mat = torch.zeros(max(u.n, 1), dtype=torch.uint8, device="cuda:0")
    ctx.materials(u, levels, mat.data_ptr(), seed=4, n_values=16)
    for tag, gray in (("gray", True), ("binary", False)):
        blob = ctx.attr_bittree((mat.data_ptr(), u.n), gray)        # warm-up
        ms = []
        for _ in range(3):
            blob = ctx.attr_bittree((mat.data_ptr(), u.n), gray)
            ms.append(ctx.kernel_ms("attr"))
        out[f"bittree_{tag}_bytes"] = len(blob)
        out[f"bittree_{tag}_bits_per_voxel"] = round(8 * len(blob) / u.n, 4)
        out[f"bittree_{tag}_ms"] = round(float(np.median(ms)), 3)
    # traversal of the lossless SSVDAG of the same scene
    g = S.Octree.from_morton(ctx, (u.ptr, u.n), levels)
    g.to_dag()
    g.to_sdag()
    g.encode(S.FMT_SSVDAG)
    sc = S.Scene.from_octree(g, S.FMT_SSVDAG)
    w, h = 3840, 2160
    p = W.trace_params(w, h, seed=5, levels=levels, shadows=True)
    prim = torch.zeros(w * h * 6, dtype=torch.int32, device="cuda:0")
    shad = torch.zeros(w * h * 6, dtype=torch.int32, device="cuda:0")
    tile = ctx.autotune_trace(sc, p, prim.data_ptr(), shad.data_ptr(), reps=2)
    tk = []
    for _ in range(3):
        sc.trace_device(p, prim.data_ptr(), shad.data_ptr())
        ctx.sync()
        tk.append(ctx.kernel_ms("trace"))
    pr = prim.cpu().numpy().view(S.RAY_DTYPE)
    rays = w * h + int((pr["t"] >= 0).sum())
    out["trace_4k_ms"] = round(float(np.median(tk)), 3)
    out["trace_4k_mrays_s"] = round(rays / (float(np.median(tk)) * 1e-3) / 1e6, 1)
    out["trace_tile"] = f"{tile}x{256 // tile} (autotuned: {ctx.last_autotune['ms']})"
    sc.free()
    g.free()
    u.free()
    log(f"[bench] attr: {out}")
    return out


def measured_traffic(L, args):
    """HBM bytes per trace launch from the committed counter summary (two
    rocprofv3 --pmc runs of this workload, FETCH_SIZE and WRITE_SIZE, each
    scaled by the same counter over a plain copy of known size), or None when
    no summary matches this workload."""
    path = os.path.join(ROOT, "profiles", "k_trace_ssvdag_terrain2048_pmc.md")
    if L != 11 or (args.width, args.height) != (1920, 1080) or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--levels", type=int, default=11)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mesh-levels", type=int, default=8,
                    help="mesh (svbuilder input) section: 64 icospheres at 2^L; 0 = skip")
    ap.add_argument("--attr-levels", type=int, default=12,
                    help="grid of the cross-level + attribute bit-tree + 4K trace measurement; 0 skips it")
    ap.add_argument("--batch-scenes", type=int, default=256,
                    help="scenes of the multi-scene batch measurement; 0 skips it")
    ap.add_argument("--lossy-levels", type=int, default=12,
                    help="grid of the lossy-DAG measurement (caves 2^L); 0 skips it")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(each GPU box has one MI355X); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        nb = args.copy_probe << 20
        a = torch.ones(nb // 4, dtype=torch.int32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb, "launches": 3}), flush=True)
        return
    S = load_pkg()
    import svdag_amd.workloads as W

    L = args.levels
    log(f"[bench] generating terrain {1 << L}^3 ...")
    codes = W.terrain_codes(L, seed=1)
    n = int(codes.size)
    log(f"[bench] {n} voxels")
    dcodes = torch.from_numpy(codes.view(np.int64)).to("cuda:0")
    ctx = S.Context(0)
    params = W.trace_params(args.width, args.height, seed=1, levels=L, shadows=True, light=(1.0, 1.0, 1.0))
    nray = args.width * args.height
    prim = torch.zeros(nray * 6, dtype=torch.int32, device="cuda:0")
    shad = torch.zeros(nray * 6, dtype=torch.int32, device="cuda:0")
    hprim = torch.zeros(nray * 4, dtype=torch.float32, device="cuda:0")
    hshad = torch.zeros(nray * 4, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()

    def step():
        t0 = time.perf_counter()
        g = build_step(S, ctx, dcodes.data_ptr(), n, L)
        ctx.sync()
        t1 = time.perf_counter()
        sc = S.Scene.from_octree(g, S.FMT_SSVDAG)
        ctx.sync()
        t2 = time.perf_counter()
        sc.trace_hits_device(params, hprim.data_ptr(), hshad.data_ptr())
        ctx.sync()
        t3 = time.perf_counter()
        trace_ms = ctx.kernel_ms("trace")
        sc.free()
        g.free()
        return t1 - t0, t3 - t2, trace_ms

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    tb, tt, tk = [], [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        b, t, k = step()
        tb.append(b)
        tt.append(t)
        tk.append(k)
        log(f"[bench] step {i + 1}/{args.steps}: build {b * 1e3:.2f} ms, trace {t * 1e3:.2f} ms (kernel {k:.3f} ms)")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start

    # per-ray stats of the same frame from one untimed run of the instrumented
    # kernel (descent-path hash and fetched bytes per ray); its records equal
    # the timed kernel's
    g = build_step(S, ctx, dcodes.data_ptr(), n, L)
    b_build, f_bytes = build_bytes(S, g, L, n)
    sc = S.Scene.from_octree(g, S.FMT_SSVDAG)
    sc.trace_device(params, prim.data_ptr(), shad.data_ptr())
    ctx.sync()
    instr_ms = ctx.kernel_ms("trace")
    pr = prim.cpu().numpy().view(S.RAY_DTYPE)
    sh = shad.cpu().numpy().view(S.RAY_DTYPE)
    hp = hprim.cpu().numpy().reshape(-1, 4)
    same = all(np.array_equal(hp[:, i], pr[k]) for i, k in enumerate(("t", "level", "iters", "node")))
    if not same:
        raise SystemExit("bench: the uninstrumented frame differs from the instrumented one")
    n_shadow = int((pr["t"] >= 0).sum())
    rays = nray + n_shadow
    bytes_fetched = int(pr["bytes"].astype(np.int64).sum() + sh["bytes"].astype(np.int64).sum())
    build_s = float(np.median(tb))
    trace_kernel_ms = float(np.median(tk))
    gvox = n / build_s / 1e9
    mrays = rays / (trace_kernel_ms * 1e-3) / 1e6
    achieved = bytes_fetched / (trace_kernel_ms * 1e-3) / 1e9
    build_achieved = b_build / build_s / 1e9

    # the viewer's default frame: minimum-depth pre-pass at 1/8 resolution, then the same rays
    dpf = W.projection_factor(args.height, screen_divisor=8.0)
    tbm = []
    for _ in range(1 + max(args.steps, 3)):
        sc.trace_beam_device(params, dpf, prim.data_ptr(), shad.data_ptr())
        ctx.sync()
        tbm.append(ctx.kernel_ms("trace"))
    pb = prim.cpu().numpy().view(S.RAY_DTYPE)
    rays_b = nray + int((pb["t"] >= 0).sum())
    beam_ms = float(np.median(tbm[1:]))
    beam = {"ms": round(beam_ms, 4), "mrays_s": round(rays_b / (beam_ms * 1e-3) / 1e6, 2),
            "prepass": f"{args.width // 8}x{args.height // 8}",
            "mean_primary_iters": round(float(pb["iters"].mean()), 3),
            "mean_primary_iters_no_prepass": round(float(pr["iters"].mean()), 3),
            "same_hit_t_frac": round(float((pb["t"] == pr["t"]).mean()), 6)}
    log(f"[bench] beam: {beam}")

    # the viewer's shaded frame (pretty mode, 4 normal samples, shadows, beam pre-pass) and its AO pass
    rgb = torch.empty(nray * 3, dtype=torch.float32, device="cuda")
    vis = torch.empty(nray, dtype=torch.float32, device="cuda")
    vp = S.ViewParams.defaults(render_mode=3, norm_samples=4)
    trm, tao = [], []
    for _ in range(1 + max(args.steps, 3)):
        sc.render_device(params, vp, rgb.data_ptr(), dpf)
        ctx.sync()
        trm.append(ctx.kernel_ms("trace"))
        sc.ao_device(params, 16, vis.data_ptr())
        ctx.sync()
        tao.append(ctx.kernel_ms("trace"))
    hits_b = int((pb["t"] >= 0).sum())
    render_ms, ao_ms = float(np.median(trm[1:])), float(np.median(tao[1:]))
    viewer = {"frame": f"{args.width}x{args.height} pretty mode, 4 normal samples, shadows, 1/8 depth pre-pass",
              "ms": round(render_ms, 4), "fps": round(1e3 / render_ms, 1),
              "rays_upper_bound": nray + hits_b * 5, "ao_rays_per_pixel": 16, "ao_ms": round(ao_ms, 4)}
    log(f"[bench] viewer: {viewer}")
    sc.free()
    g.free()

    out = {
        "metric": "SSVDAG build Gvox/s, traversal Mrays/s; learned vs exhaustive lossy @ same err",
        "value": round(gvox, 4),
        "unit": "Gvox/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"fBm terrain {1 << L}^3 ({L} levels, seed 1, {n} voxels): lossless "
                               f"SVDAG+ESVDAG+USSVDAG+SSVDAG build, then {args.width}x{args.height} primary "
                               f"+ shadow rays on the SSVDAG",
                   "levels": L, "voxels": n, "rays_primary": nray, "rays_shadow": n_shadow},
        "build_ms": round(build_s * 1e3, 3),
        "traversal_mrays_s": round(mrays, 2),
        "trace_kernel_ms": round(trace_kernel_ms, 4),
        "trace_kernel_ms_instrumented": round(instr_ms, 4),
        "traversal_beam": beam,
        "viewer": viewer,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(L, args),
                     "kernel": "k_trace_hits<SSVDAG>", "algorithmic_bytes": bytes_fetched},
        "build_roofline": {"bound": "hbm", "achieved": round(build_achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(build_achieved / HBM_PEAK_GBS, 5), "algorithmic_bytes": b_build,
                           "file_bytes": f_bytes, "ms": round(build_s * 1e3, 3)},
    }
    if args.lossy_levels:
        out["lossy"] = lossy_bench(S, ctx, args.lossy_levels, log)
    if args.batch_scenes:
        out["batch"] = batch_bench(S, W, ctx, args.batch_scenes, log)
    if args.attr_levels:
        out["cross_level_attr"] = attr_bench(S, W, ctx, args.attr_levels, log, cpu=not args.no_cpu_baseline)
    if args.mesh_levels:
        out["mesh"] = mesh_bench(S, W, ctx, args.mesh_levels, log, cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (OpenMP builder on the oracle) ...")
        out["cpu_baseline"] = cpu_baseline(codes, L)
        log(f"[bench] cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
