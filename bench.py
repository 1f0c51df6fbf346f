"""Traversal benchmark of the Aperture MI355X core (one JSON line on stdout).

Workload (the traversal microbenchmark): a displaced subdivided tetrahedron of
1,048,576 FP32 triangles; 4096 x 4096 ray slots, half pinhole primary rays
and half cosine-weighted incoherent rays from random surface points; one
step = closest hit over all 16,777,216 rays, shadow-segment generation from
the hits toward (2, 2, 2), and any hit over those segments.  Inputs are
resident in HBM before the timed region.

value = (closest queries + any-hit queries) / second, in Grays/s (1e9 ray
queries per second).

roofline: the dominant kernel is the closest-hit traversal.  Its work model
is 16 ops per box test + 40 ops per triangle test (counted per ray by the
kernel's counting variant on the same rays), priced against the non-packed
VALU lane-op peak of MI355X (256 CUs x 64 lanes x 2.4 GHz = 39.3 Tops/s);
its compulsory HBM traffic is 32 B per ray read + 20 B per hit written.
The kernel's average duration comes from HIP events recorded on the stream
the kernel runs on, over the timed region.

cpu_baseline: the oracle's BVH2 traversal (the reference builder's binned
SAH with 12 bins and leaves of 8, the reference's sorted 128-entry stack, and
a slab test that also requires entry <= exit; the reference's own slab test
omits that check and visits ~1e5 boxes per ray on this mesh) on a bounded
sample of the same rays (every 2nd slot: 8,388,608 closest hits + their shadow
any hits), with OpenMP over the host cores available to this process.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
METRIC = "Grays/s traversal; paths/s at equal FP64-oracle relMSE, plain vs batched+MLP"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12      # 39.3 Tops/s
HBM_PEAK_GBPS = 8000.0
MFMA_BF16_PEAK = 2500.0                          # dense BF16 MFMA, TFLOP/s
LIGHT = (2.0, 2.0, 2.0)
TRAFFIC_FILE = ROOT / "profiles" / "k_trace_closest_bvh4_traversal_1M_tris_16M_rays.md"
_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--levels", type=int, default=9, help="subdivision levels (9 = 1,048,576 triangles)")
    ap.add_argument("--res", type=int, default=4096, help="ray grid side (4096 = 16,777,216 slots)")
    ap.add_argument("--leaf-size", type=int, default=2)
    ap.add_argument("--bins", type=int, default=16, help="SAH bins")
    ap.add_argument("--width", type=int, default=4, help="BVH branching factor (2, 4 or 8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--spectral-only", action="store_true", help="run only the spectral 1080p render bench")
    ap.add_argument("--spectral-spp", type=int, default=64)
    ap.add_argument("--batch-only", action="store_true", help="run only the batched multi-scene bench")
    ap.add_argument("--batch-scenes", type=int, default=48)
    ap.add_argument("--mlp-only", action="store_true", help="run only the guiding-network kernel bench")
    ap.add_argument("--output-only", action="store_true", help="run only the output-stage bench")
    ap.add_argument("--sdf-only", action="store_true", help="run only the SDF-object render bench")
    ap.add_argument("--smlp-only", action="store_true", help="run only the spectral basis network bench")
    ap.add_argument("--guiding-only", action="store_true", help="run only the equal-time guiding comparison")
    ap.add_argument("--guided-render", choices=["room_plain", "room_learned", "batch_learned"],
                    help="one render of the guiding workload (for kernel traces): the room at 1024 spp plain or "
                         "learned-guided, or the room and the spectral scene as one learned-guided batch")
    ap.add_argument("--cpu-sample-stride", type=int, default=2)
    ap.add_argument("--profile", action="store_true",
                    help="timed steps only, plus one 256 MiB device copy used to scale HBM counters")
    return ap.parse_args()


def read_traffic():
    """HBM bytes per closest-hit launch, from the committed rocprofv3 --pmc summary."""
    if not TRAFFIC_FILE.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_FILE.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(V, F, rays_np, stride, shadow_np):
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc
    cores = len(os.sched_getaffinity(0))
    bvh = orc.OracleBVH(V, F, leaf_size=8, bins=12)
    sub = rays_np[:, ::stride].copy()
    sh = shadow_np[:, ::stride].copy()
    n_any = int((sh[7] > 0).sum())
    t0 = time.perf_counter()
    bvh.closest(sub, mode=0, threads=cores)
    bvh.any(sh, mode=0, threads=cores)
    dt = time.perf_counter() - t0
    log(f"traversal CPU baseline: {cores} cores {dt:.1f} s; one core")
    # one core, and the reference's own rule (FastAABBHit as written, first
    # hit in traversal order: ~1e5 box tests per ray here, so a small sample)
    small = rays_np[:, ::stride * 64].copy()
    t0 = time.perf_counter()
    bvh.closest(small, mode=0, threads=1)
    dt1 = time.perf_counter() - t0
    log(f"traversal CPU baseline: one core {dt1:.1f} s; reference rule")
    tiny = rays_np[:, ::stride * 32768].copy()
    t0 = time.perf_counter()
    bvh.closest(tiny, mode=1, threads=1)
    dt1r = time.perf_counter() - t0
    return {"value": (sub.shape[1] + n_any) / dt / 1e9, "unit": "Grays/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th ray slot of the same workload: {sub.shape[1]} closest-hit rays + "
                      f"{n_any} shadow any-hit segments, oracle BVH2 traversal (reference builder: binned SAH, "
                      f"12 bins, leaf 8; sorted 128-entry stack; strict slab test), OpenMP, {dt:.2f} s",
            "one_core": {"value": small.shape[1] / dt1 / 1e9, "cores": 1,
                         "sample": f"every {stride * 64}th ray slot, {small.shape[1]} closest-hit rays, {dt1:.2f} s"},
            "one_core_reference_rule": {"value": tiny.shape[1] / dt1r / 1e9, "cores": 1,
                                        "sample": f"every {stride * 32768}th ray slot, {tiny.shape[1]} closest-hit "
                                                  "rays, with the reference's FastAABBHit as written and its "
                                                  f"first-hit-in-order rule, {dt1r:.2f} s"}}


SHADE_OPS_PER_VERTEX = 200


def integrator_roofline(counts, ms):
    """Ops model of an integrator run (16 per box test, 40 per triangle test
    of both traversals, 200 per shaded vertex) over its wall time, against the
    non-packed VALU lane-op peak."""
    ops = 16 * (counts["closest_box_tests"] + counts["shadow_box_tests"]) + \
        40 * (counts["closest_tri_tests"] + counts["shadow_tri_tests"]) + SHADE_OPS_PER_VERTEX * counts["vertices"]
    achieved = ops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 4), "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tops/s",
            "frac": round(achieved / VALU_PEAK_TOPS, 5), "traffic": None, "ops": ops, "counts": counts,
            "model": "16 per box test + 40 per triangle test (closest and shadow traversals) + 200 per shaded "
                     "vertex, counted by the counting render (ar_render_counters) of the same frame"}


def counted(ctx, fn):
    """Runs fn() once with the context's test counters on; returns the totals."""
    ctx.render_counters(True)
    fn()
    counts = ctx.render_counters_read()
    ctx.render_counters(False)
    return counts


def integrator_bench(ctx_cls, reps=5, cpu=True):
    """Plain integrator on the Cornell-box workload (32 triangles, 256x256,
    16 spp, max depth 5, NEE on, seed 1): GPU paths/s through ar_render, the
    oracle's FP32 integrator on the host cores on the same frame, and the
    relMSE between the two images (same seed: same paths)."""
    import torch
    from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box
    V, F, M, mats = cornell_box()
    cam = CORNELL_CAMERA
    kw = dict(cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"], seed=1, max_depth=5)
    w = h = 256
    spp = 16
    with ctx_cls(0) as ctx:
        ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        ctx.build_bvh()
        out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda:0")
        for _ in range(2):
            ctx.render(w, h, spp, out=out, **kw)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ctx.render(w, h, spp, out=out, **kw)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        img = out.cpu().numpy()
        counts = counted(ctx, lambda: ctx.render(w, h, spp, out=torch.empty_like(out), **kw))
        # accuracy against the FP64 oracle at 65,536 spp (its own noise subtracted)
        fix = ROOT / "tests" / "golden" / "cornell_fp64_256x256.npz"
        acc = None
        if fix.exists():
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle_lib as orc
            d = np.load(fix)
            ref64 = d["mean"].astype(np.float64)
            floor = float(np.mean(d["var_mean"].astype(np.float64) / (ref64 ** 2 + 0.01)))
            acc = {"reference": "FP64 oracle, 65536 spp (tests/golden/cornell_fp64_256x256.npz)",
                   "reference_noise": floor, f"relmse_{spp}spp": orc.rel_mse(img, ref64) - floor}
            for hs in (1024, 8192):
                hi = ctx.render(w, h, hs, **dict(kw, seed=2)).cpu().numpy()
                acc[f"relmse_{hs}spp"] = orc.rel_mse(hi, ref64) - floor
            acc["mean_rel_diff_8192spp"] = float(abs(hi.mean() - ref64.mean()) / ref64.mean())
    paths = w * h * spp
    res = {"workload": "Cornell box, 32 triangles, 256x256, 16 spp, max depth 5, NEE on, seed 1",
           "gpu_ms_per_frame": round(ms, 4), "gpu_paths_per_s": round(paths / (ms * 1e-3), 1),
           "roofline": integrator_roofline(counts, ms)}
    if acc is not None:
        res["accuracy_vs_fp64"] = acc
    if cpu:
        log("integrator_bench: CPU baseline")
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc
        cores = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        ref = orc.render(V, F, M, mats, w, h, spp, precision=32, threads=cores, **kw)
        dt = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.render(V, F, M, mats, w, h, 2, precision=32, threads=1, **kw)
        dt1 = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(paths / dt, 1), "unit": "paths/s", "cores": cores, "kind": "port",
                               "sample": f"the full frame ({paths} paths) with the oracle FP32 integrator, {dt:.2f} s",
                               "one_core": {"value": round(w * h * 2 / dt1, 1), "cores": 1,
                                            "sample": f"the frame at 2 spp ({w * h * 2} paths), {dt1:.2f} s"}}
        res["relmse_vs_oracle_same_seed"] = orc.rel_mse(img, ref)
    return res


def spectral_bench(ctx_cls, frames=1, spp=64, cpu=True):
    """Spectral render of the instanced scene: ~500k triangles (displaced
    icospheres and boxes above a floor), 64 triangle lights, Lambert/GGX,
    1920x1080, `spp` samples, max depth 8, four hero wavelengths, NEE on, sky
    off, CIE XYZ output.  GPU paths/s through ar_render (one frame per step);
    the oracle's FP32 integrator on the host cores at 8 spp on a 960x540
    image of the same view as the CPU sample."""
    import torch
    from aperturerendererv3_amd.workloads import SPECTRAL_CAMERA, spectral_scene
    V, F, M, mats = spectral_scene()
    cam = SPECTRAL_CAMERA
    kw = dict(cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"], seed=1, max_depth=8, spectral=4, output=0)
    w, h = 1920, 1080
    with ctx_cls(0) as ctx:
        ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        t0 = time.perf_counter()
        info = ctx.build_bvh()
        build_s = time.perf_counter() - t0
        out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda:0")
        ctx.render(w, h, 1, out=out, **kw)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(frames):
            ctx.render(w, h, spp, out=out, **kw)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / frames
        mean = float(out.mean().item())
        counts = counted(ctx, lambda: ctx.render(w, h, spp, out=torch.empty_like(out), **kw))
        # accuracy: the same frame in linear sRGB against the FP64 oracle fixture of its 320x180 centre window
        fix = ROOT / "tests" / "golden" / "guiding_spectral_crop_fp64_320x180.npz"
        acc = None
        if fix.exists():
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle_lib as orc
            d = np.load(fix)
            refw = d["mean"].astype(np.float64)
            floor = float(np.mean(d["var_mean"].astype(np.float64) / (refw ** 2 + 0.01)))
            srgb = ctx.render(w, h, spp, **dict(kw, output=1)).cpu().numpy()[450:630, 800:1120]
            acc = {"relmse_vs_fp64_window": orc.rel_mse(srgb, refw) - floor, "reference_noise": floor,
                   "window": "320x180 centre window (pixels 800..1119 x 450..629), linear sRGB, FP64 oracle at "
                             "16384 spp (tests/golden/guiding_spectral_crop_fp64_320x180.npz), its own noise "
                             "subtracted"}
    paths = w * h * spp
    res = {"workload": f"instanced spectral scene, {F.shape[0]} triangles, 64 triangle lights, Lambert/GGX, "
                       f"{w}x{h}, {spp} spp, max depth 8, 4 hero wavelengths, NEE on, sky off, seed 1",
           "bvh_build_s": round(build_s, 3), "bvh": {"nodes": info["node_count"], "depth": info["depth"]},
           "gpu_ms_per_frame": round(ms, 2), "gpu_paths_per_s": round(paths / (ms * 1e-3), 1),
           "image_mean_Y": mean, "roofline": integrator_roofline(counts, ms)}
    if acc is not None:
        res["accuracy"] = acc
    if cpu:
        log("spectral_bench: CPU baseline")
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc
        cores = len(os.sched_getaffinity(0))
        cw, ch, cspp = 960, 540, 8
        t0 = time.perf_counter()
        orc.render(V, F, M, mats, cw, ch, cspp, precision=32, threads=cores, **kw)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(cw * ch * cspp / dt, 1), "unit": "paths/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{cw}x{ch} at {cspp} spp ({cw * ch * cspp} paths) of the same view with the "
                                         f"oracle FP32 integrator, including its BVH build, {dt:.2f} s"}
    return res


def batch_bench(ctx_cls, count=48, cpu=True, reps=2):
    """Batched workload: `count` instanced spectral scenes (scene k seeded
    1 + k), triangles log-uniform in [1e3, 2e6], resolution uniform in
    {256^2, 512^2, 1024^2, 1920x1080}, spp uniform in {4, 16, 64}, depth 6,
    RGB, NEE on, sky off.  Autotunes (BVH width, tile) at 1 spp, then times
    the whole batch as one HIP graph against the same scenes rendered one
    after another with ar_render (same options, same images)."""
    import torch
    from aperturerendererv3_amd import autotune, make_batch
    from aperturerendererv3_amd.workloads import SPECTRAL_CAMERA, batch_scene, batch_specs
    specs = batch_specs(count)
    cam = dict(cam_pos=SPECTRAL_CAMERA["pos"], cam_dir=SPECTRAL_CAMERA["dir"], vfov=SPECTRAL_CAMERA["vfov"])
    t0 = time.perf_counter()
    ctxs, views, tris = [], [], 0
    for sp in specs:
        V, F, M, mats = batch_scene(sp)
        c = ctx_cls(0)
        c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        c.build_bvh()
        ctxs.append(c)
        views.append(dict(width=sp["width"], height=sp["height"], spp=sp["spp"], **cam))
        tris += F.shape[0]
    setup_s = time.perf_counter() - t0
    shared = dict(max_depth=6, seed=1)
    t0 = time.perf_counter()
    tune = autotune(ctxs, views, probe_spp=1, **shared)
    tune_s = time.perf_counter() - t0
    paths = sum(v["width"] * v["height"] * v["spp"] for v in views)
    # sequential: each scene on its own, one ar_render after another
    outs = [torch.empty((v["height"], v["width"], 3), dtype=torch.float32, device="cuda:0") for v in views]
    for c in ctxs:
        c.set_tile(tune["tile"])
    for c, v, o in zip(ctxs, views, outs):
        c.render_views([v], outs=[o], **shared)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        for c, v, o in zip(ctxs, views, outs):
            c.render_views([v], outs=[o], **shared)
    e1.record()
    torch.cuda.synchronize()
    seq_ms = e0.elapsed_time(e1) / reps
    seq_img = [o.cpu().numpy() for o in outs]
    b = make_batch(ctxs)
    b.set_tile(tune["tile"])
    g = b.capture(views, **shared)
    g.launch()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        g.launch()
    e1.record()
    torch.cuda.synchronize()
    batch_ms = e0.elapsed_time(e1) / reps
    same = all(np.array_equal(a.view(np.uint32), o.cpu().numpy().view(np.uint32)) for a, o in zip(seq_img, g.outputs))
    counts = counted(b, lambda: b.render_views(views, **shared))
    res = {"workload": f"{count} instanced spectral scenes, {tris} triangles in total (log-uniform 1e3..2e6 per scene), "
                       "resolutions in {256^2, 512^2, 1024^2, 1920x1080}, spp in {4, 16, 64}, max depth 6, NEE on, "
                       "sky off, RGB",
           "paths": paths, "autotune": tune, "setup_s": round(setup_s, 2), "autotune_s": round(tune_s, 2),
           "batched_graph_ms": round(batch_ms, 2), "batched_paths_per_s": round(paths / (batch_ms * 1e-3), 1),
           "sequential_ms": round(seq_ms, 2), "sequential_paths_per_s": round(paths / (seq_ms * 1e-3), 1),
           "speedup_vs_sequential": round(seq_ms / batch_ms, 3), "images_identical_to_sequential": bool(same),
           "roofline": integrator_roofline(counts, batch_ms)}
    if cpu:
        log("batch_bench: CPU baseline")
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc
        cores = len(os.sched_getaffinity(0))
        order = sorted(range(count), key=lambda k: specs[k]["tris"])[::max(1, count // 8)][:8]
        cp, ct = 0, 0.0
        for k in order:
            V, F, M, mats = batch_scene(specs[k])
            w, h = specs[k]["width"], specs[k]["height"]
            t0 = time.perf_counter()
            orc.render(V, F, M, mats, w, h, 6, precision=32, threads=cores, max_depth=6, seed=1, **cam)
            ct += time.perf_counter() - t0
            cp += 6 * w * h
        res["cpu_baseline"] = {"value": round(cp / ct, 1), "unit": "paths/s", "cores": cores, "kind": "port",
                               "sample": f"8 of the scenes spanning the triangle-count range, full resolution, 6 spp "
                                         f"({cp} paths), with the oracle FP32 integrator including its BVH builds, "
                                         f"{ct:.2f} s"}
    g.close()
    b.close()
    for c in ctxs:
        c.close()
    res["small_scenes"] = small_batch_bench(ctx_cls)
    return res


def small_batch_bench(ctx_cls, count=256, reps=3):
    """Where batching pays: `count` small instanced spectral scenes
    (triangles log-uniform in [1e3, 2e4], scene k seeded 1001 + k) at 128 x 128
    and 4 spp, depth 6, RGB, NEE on, sky off: 65,536 paths per scene, too few
    to fill the GPU alone.  The whole batch as one HIP graph against the
    scenes one after another, as ar_render calls and as one captured graph
    per scene (no host work between launches); the same images."""
    import math
    import torch
    from aperturerendererv3_amd import make_batch
    from aperturerendererv3_amd.workloads import SPECTRAL_CAMERA, spectral_scene
    cam = dict(cam_pos=SPECTRAL_CAMERA["pos"], cam_dir=SPECTRAL_CAMERA["dir"], vfov=SPECTRAL_CAMERA["vfov"])
    ctxs, views, tris = [], [], 0
    lo, hi = math.log(1e3), math.log(2e4)
    for k in range(count):
        rng = np.random.Generator(np.random.PCG64(1001 + k))
        V, F, M, mats = spectral_scene(target_tris=int(round(math.exp(lo + (hi - lo) * rng.random()))), seed=1001 + k)
        c = ctx_cls(0)
        c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        c.build_bvh()
        ctxs.append(c)
        views.append(dict(width=128, height=128, spp=4, **cam))
        tris += F.shape[0]
    shared = dict(max_depth=6, seed=1)
    paths = sum(v["width"] * v["height"] * v["spp"] for v in views)
    outs = [torch.empty((v["height"], v["width"], 3), dtype=torch.float32, device="cuda:0") for v in views]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    seq_ms = timed(lambda: [c.render_views([v], outs=[o], **shared) for c, v, o in zip(ctxs, views, outs)])
    seq_img = [o.cpu().numpy() for o in outs]
    solo_graphs = [c.capture([v], **shared) for c, v in zip(ctxs, views)]
    seq_graph_ms = timed(lambda: [gr.launch() for gr in solo_graphs])
    same_solo = all(np.array_equal(a.view(np.uint32), gr.outputs[0].cpu().numpy().view(np.uint32))
                    for a, gr in zip(seq_img, solo_graphs))
    for gr in solo_graphs:
        gr.close()
    b = make_batch(ctxs)
    g = b.capture(views, **shared)
    batch_ms = timed(g.launch)
    same = all(np.array_equal(a.view(np.uint32), o.cpu().numpy().view(np.uint32)) for a, o in zip(seq_img, g.outputs))
    g.close()
    b.close()
    for c in ctxs:
        c.close()
    return {"workload": f"{count} instanced spectral scenes, {tris} triangles in total (log-uniform 1e3..2e4 per "
                        "scene), 128x128, 4 spp, max depth 6, NEE on, sky off, RGB",
            "paths": paths, "batched_graph_ms": round(batch_ms, 3),
            "batched_paths_per_s": round(paths / (batch_ms * 1e-3), 1),
            "sequential_ms": round(seq_ms, 3), "sequential_graphs_ms": round(seq_graph_ms, 3),
            "speedup_vs_sequential": round(seq_ms / batch_ms, 2),
            "speedup_vs_sequential_graphs": round(seq_graph_ms / batch_ms, 2),
            "images_identical_to_sequential": bool(same and same_solo)}


GUIDE_FIXTURES = {
    "room": ("guiding_room_fp64_320x180.npz", (320, 180), None),
    "spectral": ("guiding_spectral_crop_fp64_320x180.npz", (1920, 1080), (800, 450, 320, 180)),
}


def guiding_bench(ctx_cls, room_spp=1024, spectral_spp=128):
    """Learned guiding against voxel guiding and plain rendering at equal
    time, scored by relMSE against the FP64 oracle fixtures at 16,384 spp
    (tests/golden/make_guiding_references.py): the occluded-light room
    (196,770 triangles, light only through a 0.5 m door gap) at 320x180, and
    the 500k-triangle spectral scene rendered at 1920x1080 and scored on its
    320x180 centre window.  Both: 4 hero wavelengths, linear sRGB, depth 8,
    NEE on.  Each relMSE has the fixture's own noise (its per-pixel variance
    estimate) subtracted.  Wall time per render with the GPU synchronised.

    Equal-time ratio of a method: t_m e_m / (t_plain e_plain) at equal spp
    (error ~ 1 / time for an unbiased estimator): the time the method needs
    for the plain render's error, as a fraction of plain's.  The combined
    form renders both scenes as one batch with one guiding network (scene
    index input) and is compared with the plain renders of the two scenes:
    t_batch / sum_k(t_plain,k e_plain,k / e_batch,k)."""
    import torch
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc
    from aperturerendererv3_amd import make_batch
    from aperturerendererv3_amd.workloads import OCCLUDED_CAMERA, SPECTRAL_CAMERA, occluded_room, spectral_scene
    scenes, refs = {}, {}
    for name, (fname, size, crop) in GUIDE_FIXTURES.items():
        path = ROOT / "tests" / "golden" / fname
        if not path.exists():
            continue
        d = np.load(path)
        ref = d["mean"].astype(np.float64)
        floor = float(np.mean(d["var_mean"].astype(np.float64) / (ref ** 2 + 0.01)))
        refs[name] = (ref, floor)
        V, F, M, mats = occluded_room() if name == "room" else spectral_scene()
        cam = OCCLUDED_CAMERA if name == "room" else SPECTRAL_CAMERA
        scenes[name] = (V, F, M, mats, cam, size, crop)
    out = {"workload": "occluded-light room 320x180 and spectral scene 1920x1080 (scored on its 320x180 centre "
                       "window), 4 hero wavelengths, sRGB, depth 8, NEE on; relMSE against the FP64 oracle at "
                       "16384 spp minus the fixture's own noise",
           "reference_noise": {k: round(v[1], 6) for k, v in refs.items()}}
    ctxs = {}
    for name, (V, F, M, mats, cam, size, crop) in scenes.items():
        c = ctx_cls(0)
        c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        c.build_bvh()
        ctxs[name] = c
    kws = {name: dict(cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"], max_depth=8, spectral=4, output=1)
           for name, (V, F, M, mats, cam, size, crop) in scenes.items()}

    def score(name, img):
        ref, floor = refs[name]
        crop = scenes[name][6]
        if crop is not None:
            x0, y0, w, h = crop
            img = img[y0:y0 + h, x0:x0 + w]
        return orc.rel_mse(img, ref) - floor

    def run(name, spp, seed, prep=None, **extra):
        c = ctxs[name]
        w, h = scenes[name][5]
        if prep:
            prep(c, seed)
        torch.cuda.synchronize()
        t = time.perf_counter()
        img = c.render(w, h, spp, seed=seed, **kws[name], **extra)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        return ms, score(name, img.cpu().numpy())

    def summary(runs, spp):
        ms = [r[0] for r in runs]
        e = [r[1] for r in runs]
        return {"spp": spp, "ms": round(float(np.mean(ms)), 2), "relmse": float(np.mean(e)),
                "relmse_sem": float(np.std(e, ddof=1) / np.sqrt(len(e))), "relmse_per_seed": [float(x) for x in e]}

    # learned: frames of 16, 32, 64, ... spp (growth 2), 32 Adam steps of
    # 2^17 records after each; voxel: the reference's probes in frames of
    # 64 (room) / 16 (spectral) spp.  Each method: several seeds (the network
    # seeded alike), means of time and relMSE: 8 for the room, whose error
    # varies by about 10 % from seed to seed, and 32 for the spectral window
    # and the batch, whose spectral error is dominated by a few caustic
    # fireflies and varies tenfold from seed to seed (8 seeds moved its mean
    # by up to 2x).
    mlp_cfg = dict(learning_rate=1e-2, iters_per_frame=32, batch=1 << 17)
    methods = {
        "plain": (None, {}),
        "voxel": (lambda c, s: c.guiding_reset(), dict(guiding=1)),
        "learned": (lambda c, s: c.mlp_configure(seed=s, **mlp_cfg), dict(guiding=2, frame_spp=16, frame_growth=2)),
    }
    frame = {"room": 64, "spectral": 16}
    spp_of = {"room": room_spp, "spectral": spectral_spp}
    nseeds = {"room": 8, "spectral": 32}
    for name in scenes:
        spp = spp_of[name]
        n = nseeds[name]
        seeds = {"plain": tuple(range(1, 1 + n)), "voxel": tuple(range(11, 11 + n)), "learned": tuple(range(11, 11 + n))}
        res = {}
        for mname, (prep, extra) in methods.items():
            log(f"guiding: {name} {mname} at {spp} spp, seeds {seeds[mname][0]}..{seeds[mname][-1]}")
            ex = dict(dict(frame_spp=frame[name]), **extra) if extra else {}
            run(name, spp, 99, prep=prep, **ex)   # untimed: allocates the method's buffers
            res[mname] = summary([run(name, spp, sd, prep=prep, **ex) for sd in seeds[mname]], spp)
        ratios = {}
        for mname in ("voxel", "learned"):
            p, m = res["plain"], res[mname]
            ratios[mname] = round(m["ms"] * m["relmse"] / (p["ms"] * p["relmse"]), 3)
        res["time_ratio_at_equal_relmse"] = ratios
        out[name] = res
    # the combined form: both scenes in one batch, one network
    if len(scenes) == 2:
        names = list(scenes)
        b = make_batch([ctxs[n] for n in names])
        views = []
        for n in names:
            w, h = scenes[n][5]
            kw = kws[n]
            views.append(dict(width=w, height=h, spp=spp_of[n], cam_pos=kw["cam_pos"], cam_dir=kw["cam_dir"],
                              vfov=kw["vfov"]))
        shared = dict(max_depth=8, spectral=4, output=1)
        log("guiding: both scenes batched, one network")
        b.mlp_configure(seed=99, **mlp_cfg)
        b.render_views(views, guiding=2, frame_spp=16, frame_growth=2, seed=99, **shared)   # untimed
        runs = []
        bseeds = tuple(range(11, 11 + max(nseeds.values())))
        for sd in bseeds:
            b.mlp_configure(seed=sd, **mlp_cfg)
            torch.cuda.synchronize()
            t = time.perf_counter()
            imgs = b.render_views(views, guiding=2, frame_spp=16, frame_growth=2, seed=sd, **shared)
            torch.cuda.synchronize()
            runs.append(((time.perf_counter() - t) * 1e3, [score(n, im.cpu().numpy()) for n, im in zip(names, imgs)]))
        bms = float(np.mean([r[0] for r in runs]))
        errs = {n: float(np.mean([r[1][k] for r in runs])) for k, n in enumerate(names)}
        plain_equiv = sum(out[n]["plain"]["ms"] * out[n]["plain"]["relmse"] / errs[n] for n in names)
        out["batched_learned"] = {"scenes": names, "spp": [v["spp"] for v in views], "frame_spp": 16, "frame_growth": 2,
                                  "seeds": [bseeds[0], bseeds[-1]], "ms": round(bms, 2), "relmse": errs,
                                  "relmse_per_seed": {n: [float(r[1][k]) for r in runs] for k, n in enumerate(names)},
                                  "plain_ms": round(sum(out[n]["plain"]["ms"] for n in names), 2),
                                  "time_ratio_at_equal_relmse": round(bms / plain_equiv, 3)}
        out["learned_vs_plain_time_at_equal_relmse"] = out["batched_learned"]["time_ratio_at_equal_relmse"]
        b.close()
    for c in ctxs.values():
        c.close()
    return out


def guided_render(ctx_cls, which):
    """One render of the guiding workload with the settings of guiding_bench
    (seed 7), after the same render untimed (so that no buffer is allocated
    in the timed one; the network is re-initialised in between): prints its
    wall time."""
    import torch
    from aperturerendererv3_amd import make_batch
    from aperturerendererv3_amd.workloads import OCCLUDED_CAMERA, SPECTRAL_CAMERA, occluded_room, spectral_scene
    scenes = [(occluded_room(), OCCLUDED_CAMERA, 320, 180, 1024)]
    if which == "batch_learned":
        scenes.append((spectral_scene(), SPECTRAL_CAMERA, 1920, 1080, 128))
    ctxs = []
    for (V, F, M, mats), _, _, _, _ in scenes:
        c = ctx_cls(0)
        c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        c.build_bvh()
        ctxs.append(c)
    views = [dict(width=w, height=h, spp=spp, cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"])
             for _, cam, w, h, spp in scenes]
    shared = dict(max_depth=8, spectral=4, output=1)
    r = make_batch(ctxs) if which == "batch_learned" else ctxs[0]
    extra = {}
    if which != "room_plain":
        extra = dict(guiding=2, frame_spp=16, frame_growth=2)
        r.mlp_configure(seed=7, learning_rate=1e-2, iters_per_frame=32, batch=1 << 17)
    # the same render untimed first: every buffer is allocated before the timed one
    r.render_views(views, seed=99, **shared, **extra)
    if extra:
        r.mlp_configure(seed=7, learning_rate=1e-2, iters_per_frame=32, batch=1 << 17)
    torch.cuda.synchronize()
    t = time.perf_counter()
    r.render_views(views, seed=7, **shared, **extra)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3
    for c in ([r] if r is not ctxs[0] else []) + ctxs:
        c.close()
    return {"guided_render": which, "ms": round(ms, 2),
            "paths": int(sum(v["width"] * v["height"] * v["spp"] for v in views))}


# algorithmic FLOPs per vertex of the guiding network 5 -> 64 -> 64 -> 64 -> 40
MLP_FWD_FLOP = 2 * (5 * 64 + 64 * 64 + 64 * 64 + 64 * 40)
# training: forward, backward through layers 3..1 (no input gradient), weight gradients
MLP_TRAIN_FLOP = 2 * MLP_FWD_FLOP + 2 * (64 * 64 + 64 * 64 + 64 * 40)


def mlp_bench(ctx_cls, reps=10):
    """Guiding-network kernels on synthetic records: one Adam step over 2^18
    training records (k_mlp_train + fixed-order reductions + Adam) and inference of
    2^21 vertices (k_mlp_forward), each timed with HIP events on the current
    stream; bound = BF16 dense MFMA."""
    import torch
    rng = np.random.default_rng(1)
    dev = torch.device("cuda:0")
    n, ni = 1 << 18, 1 << 21
    x = torch.from_numpy(rng.random((n, 5), dtype=np.float32)).to(dev)
    d = rng.normal(size=(n, 3))
    d = torch.from_numpy((d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)).to(dev)
    w = torch.from_numpy(rng.random(n, dtype=np.float32)).to(dev)
    xi = torch.from_numpy(rng.random((ni, 5), dtype=np.float32)).to(dev)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with ctx_cls(0) as ctx:
        ctx.mlp_configure(learning_rate=1e-3, seed=1)
        ctx.mlp_train(x, d, w, 2)
        torch.cuda.synchronize()
        e0.record(stream)
        ctx.mlp_train(x, d, w, reps)
        e1.record(stream)
        torch.cuda.synchronize()
        train_ms = e0.elapsed_time(e1) / reps
        ctx.mlp_forward(xi)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            ctx.mlp_forward(xi)
        e1.record(stream)
        torch.cuda.synchronize()
        fwd_ms = e0.elapsed_time(e1) / reps
    train_tf = n * MLP_TRAIN_FLOP / (train_ms * 1e-3) / 1e12
    fwd_tf = ni * MLP_FWD_FLOP / (fwd_ms * 1e-3) / 1e12
    return {"workload": "guiding network 5-64-64-64-40 (8 vMF lobes), BF16 MFMA with FP32 accumulation",
            "train": {"records": n, "ms_per_adam_step": round(train_ms, 4), "flop_per_record": MLP_TRAIN_FLOP,
                      "achieved_TFLOPs": round(train_tf, 2), "frac_bf16_peak": round(train_tf / MFMA_BF16_PEAK, 5)},
            "forward": {"vertices": ni, "ms": round(fwd_ms, 4), "flop_per_vertex": MLP_FWD_FLOP,
                        "achieved_TFLOPs": round(fwd_tf, 2), "frac_bf16_peak": round(fwd_tf / MFMA_BF16_PEAK, 5)}}


def output_bench(ctx_cls, reps=20, cpu=True):
    """Output stage at 3840x2160: the ACES tonemap kernel (12 B read + 4 B
    written per pixel), the luminance MSE of two HDR images (24 B read per
    pixel), progressive 1-spp frames with odd / even halves on the Cornell
    box at 1920x1080 against plain 1-spp renders, the same frames keeping
    the denoiser's records, and the bilateral denoiser on the 1080p frame.
    CPU baselines: the oracle tonemap on the same pixels (one core) and the
    oracle filter on a 480x270 crop (all cores)."""
    import torch
    from aperturerendererv3_amd import AR_TONE_ACES
    from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box
    w, h = 3840, 2160
    npix = w * h
    rng = np.random.default_rng(1)
    hdr_np = (rng.random((h, w, 3), dtype=np.float32) * 2.0) ** 2
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {"workload": f"{w}x{h} HDR frame (tonemap, luminance MSE); Cornell box 1920x1080 progressive 1-spp frames"}
    with ctx_cls(0) as ctx:
        hdr = torch.from_numpy(hdr_np).to(dev)
        hdr2 = hdr.flip(0).contiguous()
        ldr = torch.empty((h, w), dtype=torch.int32, device=dev)
        for _ in range(3):
            ctx.tonemap(hdr, exposure=1.0, tone=AR_TONE_ACES, out=ldr)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            ctx.tonemap(hdr, exposure=1.0, tone=AR_TONE_ACES, out=ldr)
        e1.record(stream)
        torch.cuda.synchronize()
        tm_ms = e0.elapsed_time(e1) / reps
        ctx.luminance_mse(hdr, hdr2)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.luminance_mse(hdr, hdr2)
        mse_ms = (time.perf_counter() - t0) * 1e3 / reps     # synchronous entry point: wall time per call
        tm_gbps = npix * 16 / (tm_ms * 1e-3) / 1e9
        out["tonemap_aces"] = {"ms": round(tm_ms, 4), "bytes_per_pixel": 16, "achieved_GBps": round(tm_gbps, 1),
                               "hbm_frac": round(tm_gbps / HBM_PEAK_GBPS, 4)}
        out["luminance_mse"] = {"ms_wall": round(mse_ms, 4), "bytes_per_pixel": 24,
                                "achieved_GBps": round(npix * 24 / (mse_ms * 1e-3) / 1e9, 1)}
        V, F, M, mats = cornell_box()
        cam = CORNELL_CAMERA
        kw = dict(cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"], seed=1, max_depth=5)
        ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        ctx.build_bvh()
        fw, fh, frames = 1920, 1080, 16
        acc, odd, even = (torch.empty((fh, fw, 3), dtype=torch.float32, device=dev) for _ in range(3))
        img = torch.empty((fh, fw, 3), dtype=torch.float32, device=dev)
        ctx.render_progressive(0, acc, odd, even, spp=1, **kw)
        ctx.render(fw, fh, 1, out=img, **kw)
        torch.cuda.synchronize()
        e0.record(stream)
        for f in range(frames):
            ctx.render_progressive(f, acc, odd, even, spp=1, **kw)
        e1.record(stream)
        torch.cuda.synchronize()
        prog_ms = e0.elapsed_time(e1) / frames
        e0.record(stream)
        for f in range(frames):
            ctx.render(fw, fh, 1, sample_offset=f, out=img, **kw)
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1) / frames
        out["progressive_1080p"] = {"frames": frames, "ms_per_frame": round(prog_ms, 4),
                                    "plain_1spp_ms": round(plain_ms, 4),
                                    "paths_per_s": round(fw * fh / (prog_ms * 1e-3), 1),
                                    "variance_1e6_mse": round(ctx.luminance_mse(odd, even) * 1e6, 3)}
        # denoiser: progressive frames that also keep the G-buffer and sample
        # statistics, then the bilateral filter of the 1080p frame
        gb = torch.empty((fh, fw, 16), dtype=torch.float32, device=dev)
        st = torch.empty((fh, fw, 12), dtype=torch.float32, device=dev)
        ctx.render_progressive(0, acc, odd, even, spp=1, gbuffer=gb, stats=st, **kw)
        torch.cuda.synchronize()
        e0.record(stream)
        for f in range(frames):
            ctx.render_progressive(f, acc, odd, even, spp=1, gbuffer=gb, stats=st, **kw)
        e1.record(stream)
        torch.cuda.synchronize()
        rec_ms = e0.elapsed_time(e1) / frames
        den = torch.empty_like(acc)
        ctx.denoise(acc, gb, st, out=den)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            ctx.denoise(acc, gb, st, out=den)
        e1.record(stream)
        torch.cuda.synchronize()
        den_ms = e0.elapsed_time(e1) / reps
        out["denoise_1080p"] = {"frames_with_records_ms": round(rec_ms, 4), "denoise_ms": round(den_ms, 4),
                                "pixels_per_s": round(fw * fh / (den_ms * 1e-3), 1), "taps_per_pixel": 225}
        crop = (acc[:270, :480].contiguous().cpu().numpy(), gb[:270, :480].contiguous().cpu().numpy(),
                st[:270, :480].contiguous().cpu().numpy())
    if cpu:
        log("output_bench: CPU baseline")
        sys.path.insert(0, str(ROOT / "tests"))
        import ctypes as C
        import oracle_lib as orc
        f = orc.fn("orc_tonemap", None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_int, C.c_int])
        o = np.zeros(npix, np.uint32)
        t0 = time.perf_counter()
        f(hdr_np.ctypes.data, o.ctypes.data, npix, 1.0, 2, 0)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(npix / dt, 1), "unit": "pixels/s (ACES tonemap)", "cores": 1,
                               "kind": "port", "sample": f"the full {w}x{h} frame with the oracle tonemap, {dt:.2f} s"}
        out["tonemap_gpu_pixels_per_s"] = round(npix / (tm_ms * 1e-3), 1)
        cores = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        orc.bilateral(*crop)
        dt = time.perf_counter() - t0
        out["denoise_1080p"]["cpu_baseline"] = {
            "value": round(crop[0].shape[0] * crop[0].shape[1] / dt, 1), "unit": "pixels/s (bilateral)",
            "cores": cores, "kind": "port",
            "sample": f"a 480x270 crop of the same frame and records with the oracle filter, {dt:.2f} s"}
    return out


def sdf_bench(ctx_cls, reps=3, cpu=True):
    """Signed-distance objects: the reference's glass stress grids
    (CreateGlassTorus, CreateGlassSDFVolume with a baked 64^3 diamond; 125
    glass objects each) inside the object Cornell box, 512x512, 16 spp,
    max depth 12, RGB, NEE on.  GPU paths/s; CPU baseline: the oracle FP32
    integrator on a 256x256 frame of the same view at 8 spp."""
    import torch
    from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_objects, glass_sdf_grid
    cam = CORNELL_CAMERA
    kw = dict(cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"], seed=1, max_depth=12)
    w = h = 512
    spp = 16
    out = {"workload": "object Cornell box + 125 glass SDF objects (CreateGlassTorus / CreateGlassSDFVolume grids, "
                       "half extent 0.05), 512x512, 16 spp, max depth 12, NEE on, seed 1"}
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with ctx_cls(0) as ctx:
        vidx = ctx.sdf_volume_bake("diamond", (64, 64, 64), (0.0, 0.0, 0.0, 0.5))
        vol = ctx.sdf_volume_read(vidx)
        for kind in ("torus", "volume"):
            objs = cornell_objects()[:6] + glass_sdf_grid(kind, half=2, scale=1.2, position=(0.0, -0.1, 0.0),
                                                          volume=vidx)
            ctx.set_meshes([])
            ctx.set_objects(objs)
            ctx.build_bvh()
            img = torch.empty((h, w, 3), dtype=torch.float32, device=dev)
            ctx.render(w, h, spp, out=img, **kw)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(reps):
                ctx.render(w, h, spp, out=img, **kw)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res = {"gpu_ms_per_frame": round(ms, 3), "gpu_paths_per_s": round(w * h * spp / (ms * 1e-3), 1)}
            if cpu:
                log("sdf_bench: CPU baseline")
                sys.path.insert(0, str(ROOT / "tests"))
                import oracle_lib as orc
                cores = len(os.sched_getaffinity(0))
                oobjs = [dict(o, volume_data=vol) if o.get("type") == "volume" else o for o in objs]
                cw = ch = 256
                cs = 8
                t0 = time.perf_counter()
                orc.render(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), None, [{}], cw, ch, cs,
                           precision=32, threads=cores, objects=oobjs, **kw)
                dt = time.perf_counter() - t0
                res["cpu_baseline"] = {"value": round(cw * ch * cs / dt, 1), "unit": "paths/s", "cores": cores,
                                       "kind": "port", "sample": f"{cw}x{ch} at {cs} spp of the same view with the "
                                                                 f"oracle FP32 integrator, {dt:.2f} s"}
            out["glass_" + ("tori" if kind == "torus" else "diamond_volumes")] = res
    return out


SMLP_TRAIN_FLOP = 89600   # per sample: forward, backward-data and weight gradients of both encoders
SMLP_EVAL_FLOP = 31232    # per row: RGBEncode + LambdaEncode
SMLP_RGB_FLOP = 16640     # per RGB-table entry: RGBEncode
FP32_MFMA_PEAK_TFLOPS = 157.3


def smlp_bench(ctx_cls, cpu=True):
    """Spectral basis network (the reference's RGB / wavelength encoders):
    retraining on the GPU at the reference's batch of 4096 and at 65536,
    evaluation of 4M rows, the reference-size uplift tables (512^3 RGB,
    840 wavelengths), table queries, a fit of the library's sigmoid uplift
    (10000 steps) and the 1080p spectral frame rendered through the tables.
    FLOP counts are the algorithmic ones (the training kernel's second
    wavelength forward is not counted).  CPU baseline: the oracle's FP32
    evaluation (all host cores) and FP64 training gradient (one core)."""
    import torch
    from aperturerendererv3_amd.workloads import SPECTRAL_CAMERA, spectral_scene
    out = {"workload": "spectral basis network 2 x (7 residual 32-wide blocks) + RGB input layer, 16096 parameters, "
                       "FP32 MFMA; dataset 2^20 rows of the sigmoid uplift"}
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    with ctx_cls(0) as ctx:
        x, y = ctx.spectral_uplift_dataset(1 << 20, seed=1)
        xv, yv = ctx.spectral_uplift_dataset(1 << 16, seed=2)
        train = {}
        for batch, steps in ((4096, 200), (65536, 40)):
            ctx.spectral_mlp_init(1)
            ms = timed(lambda: ctx.spectral_mlp_train(x, y, steps=steps, batch=batch, seed=1), 2) / steps
            train[f"batch_{batch}"] = {"ms_per_step": round(ms, 4),
                                       "samples_per_s": round(batch / (ms * 1e-3), 1),
                                       "achieved_TFLOPs": round(batch * SMLP_TRAIN_FLOP / (ms * 1e-3) / 1e12, 2)}
        out["train"] = train
        # fit of the sigmoid uplift
        ctx.spectral_mlp_init(7)
        t0 = time.perf_counter()
        hist = ctx.spectral_mlp_train(x, y, steps=10000, batch=4096, seed=3, lr=1e-3, warmup_ratio=0.05,
                                      min_lr_ratio=0.01)
        torch.cuda.synchronize()
        fit_s = time.perf_counter() - t0
        pred = ctx.spectral_mlp_eval(xv)
        err = (pred - yv).abs()
        h = hist.cpu().numpy()
        out["fit"] = {"steps": 10000, "batch": 4096, "seconds": round(fit_s, 3),
                      "loss_first_100": float(h[:100].mean()), "loss_last_100": float(h[-100:].mean()),
                      "heldout_mae": round(float(err.mean()), 5), "heldout_p99_abs_err": round(
                          float(torch.quantile(err[:1 << 16], 0.99)), 5)}
        # evaluation, tables, queries
        n = 1 << 22
        xe = x[:n] if x.shape[0] >= n else x.repeat(n // x.shape[0], 1)
        oute = torch.empty(n, dtype=torch.float32, device=dev)
        ms = timed(lambda: ctx.spectral_mlp_eval(xe, out=oute), 5)
        out["eval"] = {"rows": n, "ms": round(ms, 4), "rows_per_s": round(n / (ms * 1e-3), 1),
                       "achieved_TFLOPs": round(n * SMLP_EVAL_FLOP / (ms * 1e-3) / 1e12, 2),
                       "frac_fp32_mfma_peak": round(n * SMLP_EVAL_FLOP / (ms * 1e-3) / 1e12 / FP32_MFMA_PEAK_TFLOPS, 4)}
        ms = timed(lambda: ctx.spectral_mlp_build_tables(512, 840), 2)
        entries = 512 ** 3
        out["tables_512"] = {"ms": round(ms, 3), "entries": entries,
                             "achieved_TFLOPs": round(entries * SMLP_RGB_FLOP / (ms * 1e-3) / 1e12, 2),
                             "device_bytes": entries * 16 + 840 * 16}
        ms = timed(lambda: ctx.spectral_mlp_query(xe, out=oute), 5)
        out["query"] = {"rows": n, "ms": round(ms, 4), "rows_per_s": round(n / (ms * 1e-3), 1)}
        # the 1080p spectral frame through the tables vs the sigmoid uplift
        V, F, M, mats = spectral_scene()
        cam = SPECTRAL_CAMERA
        kw = dict(cam_pos=cam["pos"], cam_dir=cam["dir"], vfov=cam["vfov"], seed=1, max_depth=8, spectral=4,
                  output=0)
        ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        ctx.build_bvh()
        img = torch.empty((1080, 1920, 3), dtype=torch.float32, device=dev)
        frame = {}
        for net in (False, True):
            ctx.spectral_set_basis(net)
            ms = timed(lambda: ctx.render(1920, 1080, 16, out=img, **kw), 2)
            frame["network" if net else "sigmoid"] = {"ms": round(ms, 3),
                                                      "paths_per_s": round(1920 * 1080 * 16 / (ms * 1e-3), 1)}
        ctx.spectral_set_basis(False)
        out["spectral_1080p_16spp"] = frame
    if cpu:
        log("smlp_bench: CPU baseline")
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as orc
        cores = len(os.sched_getaffinity(0))
        w0 = np.load(ROOT / "tests" / "golden" / "spectral_mlp_reference.npz")["weights"]
        rng = np.random.default_rng(0)
        xr = np.empty((200000, 4), np.float32)
        xr[:, :3] = rng.random((200000, 3))
        xr[:, 3] = rng.uniform(360, 780, 200000)
        t0 = time.perf_counter()
        orc.smlp_eval(w0, xr)
        dt = time.perf_counter() - t0
        yr = rng.random(512).astype(np.float32)
        t1 = time.perf_counter()
        orc.smlp_loss_grad(w0, xr[:512], yr, noise=True, seed=1, step=0)
        dt1 = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(200000 / dt, 1), "unit": "rows/s (evaluation)", "cores": cores,
                               "kind": "port",
                               "sample": f"200000 rows through the oracle's FP32 forward, {dt:.2f} s",
                               "train_samples_per_s_1_core": round(512 / dt1, 1)}
    return out


def main():
    args = parse_args()
    if args.mlp_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(mlp_bench(Context)), flush=True)
        return
    if args.guiding_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(guiding_bench(Context)), flush=True)
        return
    if args.guided_render:
        from aperturerendererv3_amd import Context
        print(json.dumps(guided_render(Context, args.guided_render)), flush=True)
        return
    if args.output_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(output_bench(Context, cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.smlp_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(smlp_bench(Context, cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.sdf_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(sdf_bench(Context, cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.batch_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(batch_bench(Context, count=args.batch_scenes, cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.spectral_only:
        from aperturerendererv3_amd import Context
        print(json.dumps(spectral_bench(Context, frames=max(1, args.steps), spp=args.spectral_spp,
                                        cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the traversal shards by ray slots with the BVH replicated, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    from aperturerendererv3_amd import AR_MISS, Context
    from aperturerendererv3_amd.workloads import traversal_mesh, traversal_rays

    V, F = traversal_mesh(levels=args.levels)
    rays_np = traversal_rays(V, F, res=args.res)
    n = rays_np.shape[1]
    ctx = Context(0)
    ctx.set_meshes([{"vertices": V, "indices": F}])
    info = ctx.build_bvh(width=args.width, leaf_size=args.leaf_size, bin_count=args.bins)
    dev = torch.device("cuda:0")
    rays = torch.from_numpy(rays_np).to(dev)
    hits = ctx.alloc_hits(n, dev)
    shadow = torch.empty_like(rays)
    occ = torch.empty((n + 31) // 32, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        ctx.trace_closest(rays, hits)
        if ev is not None:
            ev[1].record(stream)
        ctx.make_shadow_rays(rays, hits, LIGHT, out=shadow)
        if ev is not None:
            ev[2].record(stream)
        ctx.trace_any(shadow, occ)
        if ev is not None:
            ev[3].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    n_hit = int((hits["prim"] != -1).sum().item())
    n_any = int((shadow[7] > 0).sum().item())
    queries = n + n_any

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    closest_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    shadow_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    any_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in evs]))
    ms_per_step = wall * 1e3 / args.steps
    value = queries / (wall / args.steps) / 1e9

    if args.profile:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile_run": True, "closest_ms": closest_ms, "any_ms": any_ms,
                          "copy_bytes_each_way": src.numel() * 4}), flush=True)
        ctx.close()
        return

    # ops model of the closest-hit kernel on these rays (counting variant, untimed)
    cb = torch.empty(n, dtype=torch.int32, device=dev)
    ct = torch.empty(n, dtype=torch.int32, device=dev)
    h2 = ctx.alloc_hits(n, dev)
    ctx.trace_closest(rays, h2, counters=(cb, ct))
    torch.cuda.synchronize()
    boxes = int(cb.to(torch.int64).sum().item())
    tris = int(ct.to(torch.int64).sum().item())
    assert torch.equal(h2["prim"], hits["prim"]), "counting variant disagrees with the timed kernel"
    ops = 16 * boxes + 40 * tris
    achieved_tops = ops / (closest_ms * 1e-3) / 1e12
    compulsory = n * 32 + n * 20
    hbm_gbps = compulsory / (closest_ms * 1e-3) / 1e9
    traffic = read_traffic()

    # per-class breakdown (untimed for value): primary vs incoherent closest hit
    yy, xx = np.divmod(np.arange(n), args.res)
    prim_idx = torch.from_numpy(np.nonzero(((xx + yy) & 1) == 0)[0]).to(dev)
    inc_idx = torch.from_numpy(np.nonzero(((xx + yy) & 1) == 1)[0]).to(dev)
    breakdown = {}
    for name, idx in (("primary", prim_idx), ("incoherent", inc_idx)):
        sub = rays[:, idx].contiguous()
        hs = ctx.alloc_hits(sub.shape[1], dev)
        for _ in range(2):
            ctx.trace_closest(sub, hs)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record(stream)
        for _ in range(reps):
            ctx.trace_closest(sub, hs)
        e1.record(stream)
        torch.cuda.synchronize()
        breakdown[f"{name}_closest_grays"] = sub.shape[1] / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    breakdown["closest_grays"] = n / (closest_ms * 1e-3) / 1e9
    breakdown["any_grays"] = n_any / (any_ms * 1e-3) / 1e9
    breakdown["closest_ms"] = closest_ms
    breakdown["shadow_gen_ms"] = shadow_ms
    breakdown["any_ms"] = any_ms
    breakdown["hit_fraction"] = n_hit / n
    breakdown["box_tests_per_ray"] = boxes / n
    breakdown["tri_tests_per_ray"] = tris / n

    log(f"traversal: {value:.3f} Grays/s, closest {closest_ms:.3f} ms, any {any_ms:.3f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        log("traversal CPU baseline (oracle)")
        cpu = cpu_baseline(V, F, rays_np, args.cpu_sample_stride, shadow.cpu().numpy())
    blocks = {}
    for key, fn in (("integrator", lambda: integrator_bench(Context, cpu=not args.no_cpu_baseline)),
                    ("spectral_1080p", lambda: spectral_bench(Context, frames=2, spp=args.spectral_spp,
                                                              cpu=not args.no_cpu_baseline)),
                    ("batched", lambda: batch_bench(Context, count=args.batch_scenes, cpu=not args.no_cpu_baseline)),
                    ("guiding", lambda: guiding_bench(Context)),
                    ("guiding_network", lambda: mlp_bench(Context)),
                    ("output", lambda: output_bench(Context, cpu=not args.no_cpu_baseline)),
                    ("sdf_objects", lambda: sdf_bench(Context, cpu=not args.no_cpu_baseline)),
                    ("spectral_network", lambda: smlp_bench(Context, cpu=not args.no_cpu_baseline))):
        log(f"{key}: start")
        blocks[key] = fn()
        log(f"{key}: done")

    valu_frac = achieved_tops / VALU_PEAK_TOPS
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "traversal microbenchmark: displaced subdivided tetrahedron, "
                               f"{F.shape[0]} triangles, {args.res}x{args.res} ray slots (50% pinhole primary, "
                               "50% cosine-incoherent), closest hit + any-hit shadow segments",
                   "triangles": int(F.shape[0]), "rays": n, "any_hit_queries": n_any,
                   "bvh": {"width": args.width, "leaf_size": args.leaf_size, "nodes": info["node_count"],
                           "depth": info["depth"], "build_s": round(info["build_seconds"], 3)},
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": "k_trace_closest", "achieved": round(achieved_tops, 4),
                     "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tops/s", "frac": round(valu_frac, 5),
                     "traffic": traffic, "compulsory_bytes": compulsory,
                     "hbm_achieved_GBps": round(hbm_gbps, 2), "hbm_frac": round(hbm_gbps / HBM_PEAK_GBPS, 5),
                     "ops_per_launch": ops, "avg_launch_ms": round(closest_ms, 4)},
        "cpu_baseline": cpu,
        "breakdown": {k: round(v, 4) for k, v in breakdown.items()},
        **blocks,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
