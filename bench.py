#!/usr/bin/env python3
"""Closest-hit throughput on one MI355X.

Workload: the random-triangle microbenchmark -- 1,000,000 triangles (centre
~ U[0,1)^3, vertices centre + 0.01 * U[-1,1)^3, seed 2) in a SAH 4-wide QBVH,
4,194,304 rays (origin ~ U[0,1)^3, direction uniform on S^2, tnear 0,
tfar +inf, seed 3).  One step = one closest-hit launch over all rays, inputs
resident in HBM.  Also reports any-hit (tfar ~ U[0.01, 2], seed 4).

Prints one JSON line.  The roofline uses the canonical algorithmic bytes
per ray of the FP32 4-wide QBVH
  B_ray = 32 (ray) + 16 (hit) + 128 * nodes/ray + 48 * prims/ray
with nodes/ray and prims/ray counted by the CPU oracle's traversal of the
same QBVH with FP32 boxes on a sample of the same rays.  The GPU fetches
64 B quantised nodes (lr_node_q8); the bytes of that traversal (the
oracle's count over the decoded quantised boxes, 64 B per node) are
reported beside it as device_bytes_per_ray_*.  The CPU baseline is that oracle
traversal (a multi-threaded C restatement of LiAR's QBVH query) timed on the
host cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s closest-hit at 1 GPU; time-to-equal-MSE guided vs plain GPU tracer"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense
# guide training, per row: forward, input gradients of layers 1..5, weight gradients
TRAIN_FLOPS_PER_ROW = 2 * (5 * 64 * 64 + 48 * 64) + 2 * (4 * 64 * 64 + 48 * 64) + 2 * (5 * 64 * 64 + 48 * 64)
NODE_BYTES = 128  # canonical: one 4-wide FP32 SoA node
DEVICE_NODE_BYTES = 64  # lr_node_q8, what the GPU fetches
N_TRIS = 1_000_000
N_RAYS = 1 << 22


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _traffic(kind: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy)."""
    f = ROOT / "profiles" / "traffic_random1M.json"
    try:
        return json.loads(f.read_text())[kind]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def _valu_issue():
    """VALU issue figures of the closest-hit launch from the committed SQ
    counter summary (busy share of the issue slots, active lanes)."""
    f = ROOT / "profiles" / "traffic_random1M.json"
    try:
        v = json.loads(f.read_text())["valu_issue"]["closest"]
        return {"valu_busy": v["valu_busy"], "lanes_active": v["lanes_active"]}
    except (OSError, KeyError, ValueError):
        return None


def pin_config(args, ctx) -> bool:
    """--trace-config: install the given traversal configuration in ctx and
    skip the autotuner (profiling runs time exactly one configuration)."""
    if not args.trace_config:
        return False
    ctx.set_trace_config(**args.trace_config)
    return True


def host_info() -> dict:
    """CPU of the host that runs the CPU baseline."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def _film_error(fx, fw, ox, ow, np) -> dict:
    """FP32 GPU film against the FP64 oracle's of the same paths: relative
    error of the image total (Y) and the largest per-pixel difference
    relative to the brightest pixel."""
    g = fx.cpu().numpy().astype(np.float64) / fw.cpu().numpy().astype(np.float64)
    o = ox / ow
    return {"pixels": int(o.shape[1] * o.shape[2]),
            "rel_total_y": float(f"{abs(g[1].sum() - o[1].sum()) / o[1].sum():.3e}"),
            "max_pixel_diff_rel_max": float(f"{np.abs(g - o).max() / max(o[1].max(), 1e-30):.3e}")}


def bench_render(args, la, oracle, torch, np):
    """Path-tracer leg: the Cornell-box workload (256x256, 16 spp, 4 stratified
    wavelengths, depth <= 8, Russian roulette from generation 3) rendered
    whole per step; CPU baseline = the oracle renderer on a bounded crop."""
    from liar_amd import engine as E
    from liar_amd import workloads as W

    W_, H_, SPP, NL = 256, 256, 16, 4
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    ctx = la.Context(0)
    gs = E.GpuScene(ctx, cs)
    camera = E.camera_struct(cam)
    film = torch.zeros((3, H_, W_), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((H_, W_), dtype=torch.float32, device="cuda")
    tuned = None
    if not pin_config(args, ctx) and not args.no_tune:
        from liar_amd.autotune import tune_render

        tuned = tune_render(gs, camera, E.render_params(W_, H_, SPP, NL, 8, 3, 0.5, (380.0, 780.0), 99))
    stats = []
    for k in range(args.warmup + args.render_steps):
        film.zero_()
        wgt.zero_()
        torch.cuda.synchronize()
        st = gs.render(camera, E.render_params(W_, H_, SPP, NL, 8, 3, 0.5, (380.0, 780.0), 1 + k), film, wgt)
        if k >= args.warmup:
            stats.append(st)
    secs = float(np.mean([s["seconds"] for s in stats]))
    paths = stats[0]["paths"]
    rays = float(np.mean([s["rays"] for s in stats]))
    out = {
        "workload": "Cornell box 36 tris, 256x256, 16 spp x 4 wavelengths, depth <= 8, RR from 3",
        "ms_per_frame": round(secs * 1e3, 3),
        "paths_per_s": round(paths / secs, 1),
        "camera_samples_per_s": round(paths / NL / secs, 1),
        "mrays_per_s": round(rays / secs / 1e6, 2),
        "rays_per_path": round(rays / paths, 3),
        "trace_config": ctx.trace_config() if tuned is None else tuned["best"],
    }
    if not args.no_cpu:
        cw, ch = 64, 64
        threads = min(oracle.default_threads(), 16)
        t0 = time.perf_counter()
        ox, ow = oracle.render(cs, cam, cw, ch, SPP, wavelengths=NL, max_depth=8, rr_depth=3, seed=1,
                               threads=threads)
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cw * ch * SPP * NL / t_cpu, 1), "unit": "paths/s",
                               "cores": threads, "kind": "port", "host": host_info(),
                               "sample": f"{cw}x{ch} pixels of the same frame geometry (FP64 oracle), "
                                         f"{t_cpu:.2f} s"}
        # the FP32 GPU render of the same frame and seed against the FP64 oracle
        fx = torch.zeros((3, ch, cw), dtype=torch.float32, device="cuda")
        fw = torch.zeros((ch, cw), dtype=torch.float32, device="cuda")
        gs.render(camera, E.render_params(cw, ch, SPP, NL, 8, 3, 0.5, (380.0, 780.0), 1), fx, fw)
        out["fp32_vs_fp64"] = _film_error(fx, fw, ox, ow, np)
    gs.close()
    ctx.close()
    return out


def bench_direct(args, la, oracle, torch, np):
    """DirectLighting leg: the Cornell-box workload (256x256, 16 spp, 4
    stratified wavelengths) with LiAR's DirectLighting defaults (generations
    <= 8, 9 emission samples on the area light, one secondary sample, glossy
    continuations); CPU baseline = the oracle's DirectLighting on a crop."""
    from liar_amd import engine as E
    from liar_amd import workloads as W

    W_, H_, SPP, NL = 256, 256, 16, 4
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    ctx = la.Context(0)
    gs = E.GpuScene(ctx, cs)
    camera = E.camera_struct(cam)
    film = torch.zeros((3, H_, W_), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((H_, W_), dtype=torch.float32, device="cuda")
    tuned = None
    if not pin_config(args, ctx) and not args.no_tune:
        # traversal configuration for this integrand, timed on a 4-spp frame
        from liar_amd.autotune import tune_render

        tuned = tune_render(gs, camera, E.render_params(W_, H_, 4, NL, 0, 8, 0.5, (380.0, 780.0), 99), direct=True)
    stats = []
    for k in range(args.warmup + args.render_steps):
        film.zero_()
        wgt.zero_()
        torch.cuda.synchronize()
        st = gs.render_direct(camera, E.render_params(W_, H_, SPP, NL, 0, 8, 0.5, (380.0, 780.0), 1 + k), film, wgt)
        if k >= args.warmup:
            stats.append(st)
    secs = float(np.mean([s["seconds"] for s in stats]))
    paths = stats[0]["paths"]
    rays = float(np.mean([s["rays"] for s in stats]))
    out = {
        "workload": "Cornell box 36 tris, 256x256, 16 spp x 4 wavelengths, DirectLighting (generations <= 8, "
                    "9 light samples + 9 BSDF samples per camera hit, shadow rays)",
        "ms_per_frame": round(secs * 1e3, 3),
        "camera_samples_per_s": round(paths / NL / secs, 1),
        "paths_per_s": round(paths / secs, 1),
        "mrays_per_s": round(rays / secs / 1e6, 2),
        "rays_per_path": round(rays / paths, 3),
        "trace_config": ctx.trace_config() if tuned is None else tuned["best"],
    }
    if not args.no_cpu:
        cw, ch = 128, 128
        threads = min(oracle.default_threads(), 16)
        t0 = time.perf_counter()
        ox, ow = oracle.render_direct(cs, cam, cw, ch, SPP, wavelengths=NL, seed=1, threads=threads)
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cw * ch * SPP * NL / t_cpu, 1), "unit": "paths/s",
                               "cores": threads, "kind": "port", "host": host_info(),
                               "sample": f"{cw}x{ch} pixels of the same frame geometry (FP64 oracle, brute-force "
                                         f"shadow rays), {t_cpu:.2f} s"}
        fx = torch.zeros((3, ch, cw), dtype=torch.float32, device="cuda")
        fw = torch.zeros((ch, cw), dtype=torch.float32, device="cuda")
        gs.render_direct(camera, E.render_params(cw, ch, SPP, NL, 0, 8, 0.5, (380.0, 780.0), 1), fx, fw)
        out["fp32_vs_fp64"] = _film_error(fx, fw, ox, ow, np)
    gs.close()
    ctx.close()
    return out


def bench_film(args, la, torch, np):
    """Film output leg: lr_tonemap of a 1920x1080 film (automatic exposure,
    Reinhard 2002 RGB, sRGB) to float and 8-bit RGBA, timed with HIP events
    on the launch stream.  Algorithmic bytes per pixel: 16 read (X, Y, Z,
    weight) + 16 + 4 written; the exposure pass re-reads Y and weight (8)."""
    import ctypes as C

    from liar_amd import _native as N

    w, h = 1920, 1080
    n = w * h
    g = torch.Generator(device="cuda").manual_seed(5)
    wgt = torch.randint(1, 64, (h, w), device="cuda", generator=g).float()
    xyz = torch.rand((3, h, w), device="cuda", generator=g) * wgt
    rgba = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    rgba8 = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    ctx = la.Context(0)
    prm = N.TonemapParams(la.sRGB.struct(), 4, 1, 0.0, 0.0, 0.184)
    s = torch.cuda.current_stream()

    def run():
        N.check(N.lib().lr_tonemap(ctx.handle, C.c_void_p(xyz.data_ptr()), C.c_void_p(wgt.data_ptr()), None, w, h,
                                   C.byref(prm), C.c_void_p(rgba.data_ptr()), C.c_void_p(rgba8.data_ptr()), None,
                                   C.c_void_p(s.cuda_stream)), "lr_tonemap")

    for _ in range(3):
        run()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    torch.cuda.synchronize()
    a.record(s)
    for _ in range(reps):
        run()
    b.record(s)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    algo = n * (16 + 16 + 4 + 8)
    ctx.close()
    return {"workload": "lr_tonemap 1920x1080 film, Reinhard2002RGB, automatic exposure, float + 8-bit RGBA",
            "ms_per_film": round(ms, 4), "gpixels_per_s": round(n / ms / 1e6, 3),
            "algorithmic_bytes": algo, "achieved_gbs": round(algo / (ms * 1e-3) / 1e9, 1),
            "frac_hbm": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "launches": "k_film_stats, k_exposure, k_tonemap"}


def bench_instanced(args, la, torch, np):
    """Instancing leg: 2,000 placements (random rotation, scale 0.5-2 per axis,
    translation in a 60-unit cube) of one 20,000-triangle object -- 40M
    instanced triangles in a 2-level tree -- hit by 4,194,304 rays from a
    surrounding shell towards random interior points; closest and any hit
    (lr_trace_*_instanced), HIP events on the launch stream."""
    from liar_amd.instancing import Tlas, random_rotation

    rng = np.random.default_rng(21)
    nt = 20000
    c = rng.uniform(-0.5, 0.5, (nt, 3))
    verts = (np.repeat(c, 3, axis=0) + rng.uniform(-0.01, 0.01, (3 * nt, 3))).astype(np.float32)
    tris = np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3)
    irng = np.random.default_rng(22)
    inst = []
    for _ in range(2000):
        R = random_rotation(irng)
        S = np.diag(irng.uniform(0.5, 2.0, 3))
        inst.append((0, np.concatenate([R @ S, irng.uniform(-30, 30, (3, 1))], 1)))
    ctx = la.Context(0)
    q = la.Qbvh(ctx, verts, tris)
    t0 = time.perf_counter()
    tl = Tlas(ctx, [q], inst)
    build_s = time.perf_counter() - t0
    n = N_RAYS
    o = rng.normal(size=(n, 3))
    o = 90.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = rng.uniform(-30, 30, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((8, n), np.float32)
    r[0:3], r[3:6], r[7] = o.T, d.T, np.inf
    rays = torch.from_numpy(r).cuda()
    prim = torch.empty(n, dtype=torch.int32, device="cuda")
    tuv = torch.empty((3, n), dtype=torch.float32, device="cuda")
    ii = torch.empty(n, dtype=torch.int32, device="cuda")
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream()

    def timed(fn, reps=5):
        fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(s)
        for _ in range(reps):
            fn()
        b.record(s)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    ms_c = timed(lambda: tl.trace_closest(rays, prim, tuv, ii))
    ms_a = timed(lambda: tl.trace_any(rays, occ))
    hit = float((prim != -1).float().mean())
    agree = bool(((occ != 0) == (prim != -1)).all())
    info = tl.info()
    tl.close()
    q.close()
    ctx.close()
    return {"workload": "2000 instances of a 20k-triangle object (40M instanced triangles), 4,194,304 rays",
            "closest_mrays_per_s": round(n / ms_c / 1e3, 1), "closest_ms": round(ms_c, 4),
            "any_mrays_per_s": round(n / ms_a / 1e3, 1), "any_ms": round(ms_a, 4), "hit_rate": round(hit, 4),
            "any_agrees_with_closest": agree, "tlas_nodes": info["n_nodes"], "tlas_depth": info["max_depth"],
            "build_s": round(build_s, 3)}


GUIDED_CFG = dict(selection=0.5, light_share=0.5, learned_selection=True, balance_heuristic=True, train_steps=512,
                  train_batch=8192, weight_clamp=20.0, record_depth=0)
# radiance-driven Russian roulette on the network's reflected-radiance
# estimate (trained with the lobes, weight 1): from generation 3 on, a path's
# roulette weight at most 5
GUIDED_RR = dict(learned_rr=True, rr_scale=1.0, rr_min_survival=0.2, rr_start=3)
GUIDED_RADIANCE_WEIGHT = 1.0
GUIDED_TRAIN_PASSES = 1


def bench_guided(args, la, torch, np):
    """Learned path guiding on the slit-room workload: the terrain workload
    (4,992,800 triangles + 10k spheres, Lambert / GGX gold / dielectric) in a
    closed box room lit by the sun and sky only through a 2 % ceiling slit;
    camera on the indirectly lit half; 1920x1080, depth 16.  Plain and guided
    renders accumulate 4-spp passes against a plain 4096-spp reference.  The
    guided clock includes its online training (the first pass).  Each
    method runs 5 times (render and guide seeds) and its MSE curve is the
    per-pass mean.  Reports the wall time at which the guided render first
    reaches the plain tracer's 64-spp MSE, by two measures: the MSE over all pixels, and over
    all but the 0.1 % of pixels with the largest error (the sun caustics of
    the glass and gold spheres make single pixels -- of the renders and of
    the reference -- carry nearly all of the plain MSE).  The same figures
    with every sample's luminance clamped at 0.5 (Raster's
    maxSampleLuminance, raster.cpp:567-570) come beside them: clamping moves
    each estimator's limit differently, so those are for comparison only."""
    from liar_amd import engine as E
    from liar_amd import workloads as W
    from liar_amd.guide import Guide

    Wd, Hd = args.guided_width, args.guided_width * 9 // 16
    log(f"[bench] guided: building the slit room ({Wd}x{Hd})")
    cs, cam = W.slit_room()
    ctx = la.Context(0)
    gs = E.GpuScene(ctx, cs)
    camera = E.camera_struct(cam)

    def film():
        return (torch.zeros((3, Hd, Wd), dtype=torch.float32, device="cuda"),
                torch.zeros((Hd, Wd), dtype=torch.float32, device="cuda"))

    def params(spp, seed, ml):
        return E.render_params(Wd, Hd, spp, 1, 16, 16, 0.5, (380.0, 780.0), seed, 0, ml)

    tuned = None
    if not pin_config(args, ctx) and not args.no_tune:
        from liar_amd.autotune import tune_render

        log("[bench] guided: autotuning traversal on the plain render")
        tuned = tune_render(gs, camera, params(1, 77, 0.0), repeats=1)
    keep = int(0.999 * Wd * Hd)

    def errors(img, ref):
        e = ((img - ref) ** 2).flatten()
        return float(e.mean()), float(torch.sort(e)[0][:keep].mean())

    def run(guide, ml, ref, passes=32, seed0=100, rr=True):
        cfg = dict(GUIDED_CFG, **GUIDED_RR) if rr else GUIDED_CFG
        fx, fw = film()
        # warm-up pass (allocation, code load) outside the clock
        if guide is None:
            gs.render(camera, params(1, 1, ml), *film())
        else:
            E.guided_render(gs, guide, camera, params(1, 1, ml), *film(), **cfg)
        torch.cuda.synchronize()
        elapsed, curve, rays = 0.0, [], 0
        for p in range(passes):
            t0 = time.perf_counter()
            if guide is not None:
                st = E.guided_render(gs, guide, camera, params(4, seed0 + p, ml), fx, fw,
                                     train=p < GUIDED_TRAIN_PASSES, **cfg)
            else:
                st = gs.render(camera, params(4, seed0 + p, ml), fx, fw)
            torch.cuda.synchronize()
            elapsed += time.perf_counter() - t0
            rays += st["rays"]
            curve.append((elapsed, 4 * (p + 1)) + errors(fx[1] / fw, ref))
        return curve, rays / (passes * 4 * Wd * Hd)

    def leg(ml, no_rr=False):
        fx, fw = film()
        ref_spp = args.guided_ref_spp
        log(f"[bench] guided: {ref_spp}-spp plain reference (clamp {ml or 'none'})")
        t0 = time.perf_counter()
        for k in range(ref_spp // 512):
            gs.render(camera, params(512, 10000 + k, ml), fx, fw)
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t0
        ref = (fx / fw)[1].clone()
        del fx, fw
        # each method run `reps` times (render seeds, guide seeds): one guide's
        # training is one draw of a noisy process, and a single caustic
        # outlier moves one run's MSE; the curves are averaged per pass
        reps = max(1, args.guided_repeats)
        plains, guideds, norrs = [], [], []
        rpp = {"plain": [], "guided": [], "guided_no_roulette": []}
        for r in range(reps):
            log(f"[bench] guided: plain curve {r + 1}/{reps}")
            c, k = run(None, ml, ref, seed0=100 + 1000 * r)
            plains.append(c)
            rpp["plain"].append(k)
            log(f"[bench] guided: guided curve {r + 1}/{reps}")
            g = Guide(ctx, seed=1 + r, learning_rate=1e-3, radiance_weight=GUIDED_RADIANCE_WEIGHT)
            c, k = run(g, ml, ref, seed0=100 + 1000 * r)
            guideds.append(c)
            rpp["guided"].append(k)
            g.close()
            if no_rr:
                log(f"[bench] guided: guided curve without roulette {r + 1}/{reps}")
                g = Guide(ctx, seed=1 + r, learning_rate=1e-3)
                c, k = run(g, ml, ref, seed0=100 + 1000 * r, rr=False)
                norrs.append(c)
                rpp["guided_no_roulette"].append(k)
                g.close()

        def mean_curve(cs):
            return [tuple(float(np.mean([c[p][k] for c in cs])) for k in range(4)) for p in range(len(cs[0]))]

        def compare(plain, guided, out, suffix=""):
            for k, name in ((2, "mse"), (3, "mse_trimmed")):
                target = plain[15][k]
                hit = next((c[0] for c in guided if c[k] <= target), None)
                if not suffix:
                    out[f"plain_64spp_{name}"] = target
                out[f"guided{suffix}_64spp_{name}"] = guided[15][k]
                out[f"equal_spp_{name}_ratio_plain_over_guided{suffix}"] = round(target / guided[15][k], 3)
                out[f"guided{suffix}_seconds_to_plain_64spp_{name}"] = None if hit is None else round(hit, 4)
                out[f"time_ratio_guided{suffix}_over_plain_{name}"] = None if hit is None else round(hit / plain[15][0], 3)

        plain, guided = mean_curve(plains), mean_curve(guideds)
        out = {"reference": f"plain {ref_spp} spp ({t_ref:.1f} s)", "runs_per_method": reps,
               "plain_64spp_seconds": round(plain[15][0], 4), "guided_64spp_seconds": round(guided[15][0], 4),
               "rays_per_path": {k: round(float(np.mean(v)), 3) for k, v in rpp.items() if v},
               "equal_spp_mse_trimmed_ratio_per_run": [round(a[15][3] / b[15][3], 3)
                                                       for a, b in zip(plains, guideds)]}
        compare(plain, guided, out)
        curves = {"plain": plain, "guided": guided}
        if norrs:
            norr = mean_curve(norrs)
            out["guided_no_roulette_64spp_seconds"] = round(norr[15][0], 4)
            compare(plain, norr, out, "_no_roulette")
            curves["guided_no_roulette"] = norr
        out["curves"] = {k: [(round(c[0], 4), c[1], c[2], c[3]) for c in v[::4]] for k, v in curves.items()}
        return out

    unclamped = leg(0.0, no_rr=True)
    clamped = leg(0.5)
    out = {
        "workload": f"slit room (terrain 4,992,800 tris + 10k spheres in a closed box, 2% ceiling slit, sun + sky), "
                    f"{Wd}x{Hd}, 1 wavelength per sample, depth 16, no sample clamp",
        "mse_trimmed": "MSE over all pixels but the 0.1% with the largest squared error",
        **{k: v for k, v in unclamped.items()},
        "clamped_0.5": clamped,
        "guided_config": "one-sample MIS of BSDF, light and guide sampling (balance heuristic) with the network's "
                         "per-vertex strategy probabilities, at every bounce; online training during the first "
                         f"{GUIDED_TRAIN_PASSES} pass(es): 512 Adam steps (lr 1e-3, batch 8192) per 4M-path chunk on "
                         "the records of every bounce, weights clamped at 20x their mean; radiance-driven Russian "
                         "roulette on the network's reflected-radiance estimate (radiance loss weight 1) from "
                         "generation 3 on, a path's roulette weight at most 5; 'no_roulette' figures: the same "
                         "without the roulette or the radiance loss",
        "trace_config": ctx.trace_config() if tuned is None else tuned["best"],
    }
    gs.close()
    ctx.close()
    return out


def bench_train(args, la, torch, np):
    """Guide training steps (k_train + k_adam) at the guided benchmark's batch
    of 4096 rows and at larger batches; MFMA rate by algorithmic flops."""
    from liar_amd.guide import Guide

    ctx = la.Context(0)
    g = Guide(ctx, seed=1, learning_rate=1e-3)
    out = {"workload": "lr_guide_train_step on random records (8 inputs with light / BSDF pdfs, unit "
                       "directions, weights U[0.2, 1.2))",
           "flops_per_row": TRAIN_FLOPS_PER_ROW,
           "flops_basis": "forward 2(5*64*64 + 48*64) + input gradients 2(4*64*64 + 48*64) + weight gradients "
                          "2(5*64*64 + 48*64)",
           "launches_per_step": 2, "steps": []}
    for n in (4096, 8192, 32768, 262144):
        rng = np.random.default_rng(n)
        x = np.zeros((n, 8), np.float32)
        x[:, :8] = rng.random((n, 8))
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        X = torch.from_numpy(x).cuda()
        D = torch.from_numpy(d.astype(np.float32)).cuda()
        W = torch.from_numpy((rng.random(n) + 0.2).astype(np.float32)).cuda()
        for _ in range(5):
            g.train_step(X, D, W)
        steps = 200 if n <= 32768 else 50
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            g.train_step(X, D, W)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / steps
        tf = n * TRAIN_FLOPS_PER_ROW / (us * 1e-6) / 1e12
        out["steps"].append({"rows": n, "us_per_step": round(us, 2), "tflops": round(tf, 1),
                             "frac_bf16_peak": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)})
        log(f"[train] {n} rows: {us:.1f} us/step, {tf:.1f} TFLOP/s")
    g.close()
    ctx.close()
    return out


def bench_terrain(args, la, torch, np):
    """The plain path tracer at scale: the terrain workload (1581x1581 fBm
    heightfield, 4,992,800 triangles + 10k spheres; Lambert / GGX gold /
    dielectric; 0.5-degree sun + sky) at 1920x1080, 64 spp, depth 16, one
    wavelength per sample; traversal autotuned on a 480x270 4-spp render,
    the queue tile (paths per chunk) on a 16-spp frame."""
    from liar_amd import engine as E
    from liar_amd import workloads as W

    log("[bench] terrain: building")
    cs, cam = W.terrain()
    ctx = la.Context(0)
    gs = E.GpuScene(ctx, cs)
    camera = E.camera_struct(cam)
    tuned = None
    if not pin_config(args, ctx) and not args.no_tune:
        from liar_amd.autotune import tune_render

        log("[bench] terrain: autotuning traversal")
        tuned = tune_render(gs, camera, E.render_params(480, 270, 4, 1, 16, 16, 0.5, (380.0, 780.0), 5), repeats=2)
    Wd, Hd, SPP = 1920, 1080, 64
    chunk, chunk_table = 0, None
    if not args.no_tune:
        from liar_amd.autotune import tune_chunk

        # timed on the benchmark's own frame: at fewer spp the larger tiles
        # hold the whole frame and tie, while at 64 spp they differ
        log("[bench] terrain: autotuning the queue tile (paths per chunk)")
        tc = tune_chunk(gs, camera, E.render_params(Wd, Hd, SPP, 1, 16, 16, 0.5, (380.0, 780.0), 6))
        chunk, chunk_table = tc["best"], [(c, round(ms, 3)) for c, ms in tc["table"]]
    fx = torch.zeros((3, Hd, Wd), dtype=torch.float32, device="cuda")
    fw = torch.zeros((Hd, Wd), dtype=torch.float32, device="cuda")
    log(f"[bench] terrain: {Wd}x{Hd} x {SPP} spp frames")
    stats = []
    for k in range(1 + args.terrain_frames):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = gs.render(camera, E.render_params(Wd, Hd, SPP, 1, 16, 16, 0.5, (380.0, 780.0), 1 + k, chunk), fx, fw)
        torch.cuda.synchronize()
        if k >= 1:
            stats.append((time.perf_counter() - t0, st))
    secs = float(np.mean([t for t, _ in stats]))
    paths = stats[0][1]["paths"]
    rays = float(np.mean([st["rays"] for _, st in stats]))
    out = {
        "workload": "terrain 4,992,800 tris + 10k spheres, 1920x1080, 64 spp, 1 wavelength per sample, depth 16",
        "ms_per_frame": round(secs * 1e3, 2),
        "samples_per_s": round(paths / secs, 1),
        "mrays_per_s": round(rays / secs / 1e6, 1),
        "rays_per_path": round(rays / paths, 3),
        "trace_config": ctx.trace_config() if tuned is None else tuned["best"],
        "autotune_default_ms": None if tuned is None or tuned["default_ms"] is None else round(tuned["default_ms"], 3),
        "autotune_best_ms": None if tuned is None else round(tuned["best_ms"], 3),
        "chunk_paths": chunk or (1 << 25),
        "chunk_autotune_ms": chunk_table,
    }
    gs.close()
    ctx.close()
    return out


def bench_batch(args, la, torch, np):
    """The batched uneven workload: 64 independent scenes (triangle counts
    log-uniform in 10k..2M, random-cube and terrain generators alternately;
    square images of 128..1024; 4/16/64 spp; max depth 4..16) rendered by one
    lr_render_batch call, its launches captured in a HIP graph and the
    traversal configuration autotuned on the batch's camera rays.  Reports
    rays/s and the per-scene completion latency p50 / p99 (device clock, from
    the call's first launch), next to the same scenes rendered one by one."""
    from liar_amd import engine as E
    from liar_amd import workloads as W

    log("[bench] batch: generating and building 64 scenes")
    specs = W.batched_scenes(64)
    ctx = la.Context(0)
    t0 = time.perf_counter()
    gss = [E.GpuScene(ctx, s["scene"]) for s in specs]
    t_build = time.perf_counter() - t0
    batch = E.Batch(ctx, gss)

    def items(chunk=0):
        out = []
        for s in specs:
            prm = E.render_params(s["width"], s["height"], s["spp"], 1, s["max_depth"], 3, 0.5, (380.0, 780.0),
                                  s["seed"], chunk)
            out.append((E.camera_struct(s["camera"]), prm,
                        torch.zeros((3, s["height"], s["width"]), device="cuda"),
                        torch.zeros((s["height"], s["width"]), device="cuda")))
        return out

    # autotune the traversal on the largest scene's camera rays (one bounce)
    big = max(range(len(specs)), key=lambda k: specs[k]["n_tris"])
    cam = specs[big]["camera"]
    n_tune = 1 << 20
    tune_rays = _camera_rays(np, torch, cam, n_tune)
    if pin_config(args, ctx):
        res = {"best": ctx.trace_config()}
    else:
        res = gss[big].qbvh.autotune(tune_rays, repeats=2)
        ctx.set_trace_config(**res["best"])
    # the queue tile (paths per chunk; chunks span scenes): the fastest batch
    # call of 4M, 16M, 32M and 64M paths, each after a warm-up call (code
    # load, workspace, graph instantiation)
    chunk_table = []
    for chunk in (1 << 22, 1 << 24, 1 << 25, 1 << 26):
        batch.render(items(chunk), graph=True)
        it = items(chunk)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = batch.render(it, graph=True)
        torch.cuda.synchronize()
        chunk_table.append((chunk, time.perf_counter() - t0, r))
    chunk, wall, r = min(chunk_table, key=lambda c: c[1])
    ms = np.array(r["scene_ms"])
    it = items()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (c, prm, fx, fw), gs in zip(it, gss):
        gs.render(c, prm, fx, fw)
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t0
    out = {
        "workload": "64 scenes, triangles log-uniform 10k..2M (random cube / terrain), images 128^2..1024^2, "
                    "spp 4/16/64, depth 4..16, one wavelength per sample",
        "triangles_total": int(sum(s["n_tris"] for s in specs)),
        "paths": int(r["paths"]),
        "rays": int(r["rays"]),
        "chunks": int(r["chunks"]),
        "seconds": round(wall, 4),
        "mrays_per_s": round(r["rays"] / wall / 1e6, 1),
        "scene_latency_ms_p50": round(float(np.percentile(ms, 50)), 2),
        "scene_latency_ms_p99": round(float(np.percentile(ms, 99)), 2),
        "scene_latency_ms_max": round(float(ms.max()), 2),
        "sequential_seconds": round(wall_seq, 4),
        "sequential_mrays_per_s": round(r["rays"] / wall_seq / 1e6, 1),
        "trace_config": res["best"],
        "chunk_paths": chunk,
        "chunk_seconds": [(c, round(w, 4)) for c, w, _ in chunk_table],
        "build_seconds": round(t_build, 2),
    }
    batch.close()
    for gs in gss:
        gs.close()
    ctx.close()
    return out


def _camera_rays(np, torch, cam, n):
    """(8, n) primary rays of a PerspectiveCamera over a square grid."""
    side = int(np.sqrt(n))
    d = np.asarray(cam.direction, np.float64)
    d /= np.linalg.norm(d)
    sky = np.asarray(cam.sky, np.float64)
    r = np.cross(d, sky)
    r /= np.linalg.norm(r)
    u = np.cross(r, d)
    t = np.tan(cam.fovAngle / 2)
    ys, xs = np.meshgrid((np.arange(side) + 0.5) / side * 2 - 1, (np.arange(side) + 0.5) / side * 2 - 1,
                         indexing="ij")
    dirs = d[None, :] + t * (xs.ravel()[:, None] * r[None, :] - ys.ravel()[:, None] * u[None, :])
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    m = side * side
    rays = np.zeros((8, m), np.float32)
    rays[0:3] = np.asarray(cam.position, np.float32)[:, None]
    rays[3:6] = dirs.T
    rays[6] = 1e-6
    rays[7] = np.inf
    return torch.from_numpy(rays).cuda()


def profile_workload(args, la, torch):
    """A short fixed run for profilers (rocprofv3 --kernel-trace / --pmc):
    the guided benchmark's slit room at its 1920x1080, 4-spp passes: 2 plain,
    then 2 guided with online training and 2 without.  Prints one line per
    pass, no JSON."""
    from liar_amd import engine as E
    from liar_amd import workloads as W
    from liar_amd.guide import Guide

    Wd, Hd = args.guided_width, args.guided_width * 9 // 16
    cs, cam = W.slit_room()
    ctx = la.Context(0)
    pin_config(args, ctx)
    gs = E.GpuScene(ctx, cs)
    camera = E.camera_struct(cam)
    fx = torch.zeros((3, Hd, Wd), device="cuda")
    fw = torch.zeros((Hd, Wd), device="cuda")
    g = Guide(ctx, seed=1, learning_rate=1e-3, radiance_weight=GUIDED_RADIANCE_WEIGHT)
    for p in range(6):
        prm = E.render_params(Wd, Hd, 4, 1, 16, 16, 0.5, (380.0, 780.0), 100 + p)
        t0 = time.perf_counter()
        if p < 2:
            st = gs.render(camera, prm, fx, fw)
        else:
            st = E.guided_render(gs, g, camera, prm, fx, fw, train=p < 4, **GUIDED_CFG, **GUIDED_RR)
        torch.cuda.synchronize()
        kind = "plain" if p < 2 else ("guided+train" if p < 4 else "guided")
        log(f"[profile] pass {p} {kind}: {1e3 * (time.perf_counter() - t0):.2f} ms, {st['rays']} rays")
    g.close()
    gs.close()
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--fp64-sample", type=int, default=2048,
                    help="rays whose FP32 hits are compared with an FP64 brute force over the scene")
    ap.add_argument("--count-sample", type=int, default=1 << 18,
                    help="rays used for the oracle's nodes/prims-per-ray counts and parity")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracer leg")
    ap.add_argument("--no-guided", action="store_true", help="skip the learned-guiding leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default traversal configuration")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene leg")
    ap.add_argument("--no-terrain", action="store_true", help="skip the terrain path-tracer leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the DirectLighting leg")
    ap.add_argument("--no-film", action="store_true", help="skip the film output (tone mapping) leg")
    ap.add_argument("--no-instanced", action="store_true", help="skip the instancing leg")
    ap.add_argument("--no-train", action="store_true", help="skip the guide-training leg")
    ap.add_argument("--terrain-frames", type=int, default=2)
    ap.add_argument("--profile-workload", action="store_true",
                    help="run only a short slit-room plain + guided sequence for profilers")
    ap.add_argument("--guided-width", type=int, default=1920)
    ap.add_argument("--guided-ref-spp", type=int, default=4096)
    ap.add_argument("--guided-repeats", type=int, default=5,
                    help="plain and guided curves per method (averaged)")
    ap.add_argument("--render-steps", type=int, default=5)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy (PMC byte-counter calibration)")
    ap.add_argument("--trace-config", default=None,
                    help="block,lds_stack,blocks_per_cu,refill_min,schedule: use this traversal "
                         "configuration everywhere instead of autotuning (for profiler runs)")
    ap.add_argument("--only", default=None,
                    choices=["render", "direct", "film", "instanced", "terrain", "guided", "batch", "train"],
                    help="run only this leg (no headline) and print its JSON object")
    args = ap.parse_args()
    if args.trace_config:
        b, l, bpc, r, sch = (int(x) for x in args.trace_config.split(","))
        args.trace_config = {"block": b, "lds_stack": l, "blocks_per_cu": bpc, "refill_min": r, "schedule": sch}

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        log("bench.py measures one GPU only: there is no multi-GPU path "
            "(the ray query would shard by rays over a replicated QBVH).")
        return 2

    import numpy as np
    import torch

    import liar_amd as la
    import oracle

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.profile_workload:
        return profile_workload(args, la, torch)
    if args.only:
        legs = {"render": lambda: bench_render(args, la, oracle, torch, np),
                "direct": lambda: bench_direct(args, la, oracle, torch, np),
                "film": lambda: bench_film(args, la, torch, np),
                "instanced": lambda: bench_instanced(args, la, torch, np),
                "terrain": lambda: bench_terrain(args, la, torch, np),
                "guided": lambda: bench_guided(args, la, torch, np),
                "batch": lambda: bench_batch(args, la, torch, np),
                "train": lambda: bench_train(args, la, torch, np)}
        print(json.dumps({args.only: legs[args.only]()}), flush=True)
        return 0
    dev = torch.device("cuda:0")
    ctx = la.Context(0)

    log("[bench] generating scene and rays")
    verts, tris = la.random_triangles(N_TRIS, seed=2)
    bvh = la.Qbvh(ctx, verts, tris)
    info = bvh.info()
    rays_h = la.random_rays(N_RAYS, seed=3)
    rays_any_h = la.random_rays(N_RAYS, seed=4, tfar_lo=0.01, tfar_hi=2.0)
    rays = torch.from_numpy(rays_h).to(dev)
    rays_any = torch.from_numpy(rays_any_h).to(dev)
    prim = torch.empty(N_RAYS, dtype=torch.int32, device=dev)
    tuv = torch.empty((3, N_RAYS), dtype=torch.float32, device=dev)
    occ = torch.empty(N_RAYS, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in evs:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern_ms = [a.elapsed_time(b) for a, b in evs]
        return wall, float(np.mean(kern_ms))

    if args.calibrate_copy:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    default_cfg = ctx.trace_config()

    def tune(r, any_hit):
        # traversal block size, LDS stack depth, residency and refill threshold
        # from the autotuner, on the rays about to be timed (outside the timed region)
        if pin_config(args, ctx) or args.no_tune:
            return None
        res = bvh.autotune(r, any_hit=any_hit, repeats=2)
        d = next((ms for c, ms in res["table"] if c == default_cfg), None)
        return {"config": res["best"], "ms": round(res["best_ms"], 4),
                "default_config_ms": None if d is None else round(d, 4), "configs_timed": len(res["table"])}

    log("[bench] autotuning closest-hit traversal")
    tuned_c = tune(rays, False)
    log(f"[bench] closest-hit: {args.warmup} warmup + {args.steps} timed launches of {N_RAYS} rays")
    wall_c, kern_c = timed(lambda: bvh.trace_closest(rays, prim, tuv), args.steps, args.warmup)
    log("[bench] autotuning any-hit traversal")
    tuned_a = tune(rays_any, True)
    log("[bench] any-hit")
    wall_a, kern_a = timed(lambda: bvh.trace_any(rays_any, occ), args.steps, args.warmup)
    ctx.set_trace_config(**(tuned_c["config"] if tuned_c else default_cfg))

    mrays = N_RAYS * args.steps / wall_c / 1e6
    mrays_any = N_RAYS * args.steps / wall_a / 1e6

    # this box's measured HBM bandwidth: a 1 GiB device-to-device copy (read +
    # write), beside the 8 TB/s spec the roofline uses
    cp_a = torch.zeros(1 << 28, dtype=torch.float32, device=dev)
    cp_b = torch.empty_like(cp_a)
    cp_b.copy_(cp_a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        cp_b.copy_(cp_a)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * cp_a.numel() * 4 * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del cp_a, cp_b

    # oracle counts + parity on a sample of the same rays
    nodes, prims = bvh.export()
    S = min(args.count_sample, N_RAYS)
    log(f"[bench] oracle counts / parity on {S} rays")
    t0 = time.perf_counter()
    ref_p, ref_t, ref_u, ref_v, tied, ref_nodes, ref_prims = oracle.qbvh_closest_ties(nodes, prims, rays_h[:, :S])
    ref = (ref_p, ref_t, ref_u, ref_v, ref_nodes, ref_prims)
    t_count = time.perf_counter() - t0
    ref_any = oracle.qbvh_any(nodes, prims, rays_any_h[:, :S])
    # node / primitive counts of the traversal the GPU runs: the quantised boxes
    dec = oracle.decode_q8(bvh.export_q8())
    cnt = oracle.qbvh_closest(dec, prims, rays_h[:, :S])
    cnt_any = oracle.qbvh_any(dec, prims, rays_any_h[:, :S])
    g_prim = prim[:S].cpu().numpy().view(np.uint32)
    g_tuv = tuv[:, :S].cpu().numpy()
    mism = int((g_prim != ref[0]).sum()) + sum(
        int((g_tuv[k].view(np.uint32) != ref[1 + k].view(np.uint32)).sum()) for k in range(3))
    mism_any = int((occ[:S].cpu().numpy() != ref_any[0]).sum())
    # numerically tied hits (two or more primitives at the hit distance): the
    # GPU must break them as the oracle does, by the smaller primitive id
    is_tie = (tied & 1) != 0
    n_tied = int(is_tie.sum())
    mism_tied = int((g_prim[is_tie] != ref[0][is_tie]).sum())
    # rays whose answer LiAR's interval quirk decided (the mesh's nearest
    # triangle within near (1 + 1e-6): reported as no hit)
    n_quirk = int(((tied & 2) != 0).sum())
    nodes_per_ray = ref[4] / S
    prims_per_ray = ref[5] / S
    b_ray = 32 + 16 + NODE_BYTES * nodes_per_ray + 48 * prims_per_ray
    any_nodes = ref_any[1] / S
    any_prims = ref_any[2] / S
    b_ray_any = 32 + 1 + NODE_BYTES * any_nodes + 48 * any_prims
    dev_b_ray = 32 + 16 + DEVICE_NODE_BYTES * cnt[4] / S + 48 * cnt[5] / S
    dev_b_ray_any = 32 + 1 + DEVICE_NODE_BYTES * cnt_any[1] / S + 48 * cnt_any[2] / S
    achieved = N_RAYS * b_ray / (kern_c * 1e-3) / 1e9
    achieved_any = N_RAYS * b_ray_any / (kern_a * 1e-3) / 1e9
    # the FP32 results' error against an FP64 restatement of the same
    # candidates (brute force over every primitive, all arithmetic in FP64)
    F = min(S, args.fp64_sample)
    log(f"[bench] FP32 vs FP64: brute force over {N_TRIS} triangles for {F} rays")
    fp, ft, fu, fv = oracle.brute_closest_f64(verts, tris, np.zeros((0, 4), np.float32), rays_h[:, :F])
    both = (g_prim[:F] == fp) & (fp != oracle.MISS)
    rel_t = np.abs(g_tuv[0][:F][both].astype(np.float64) - ft[both]) / np.abs(ft[both])
    err_uv = np.maximum(np.abs(g_tuv[1][:F][both] - fu[both]), np.abs(g_tuv[2][:F][both] - fv[both]))
    fp64 = {"sample": F, "id_mismatches": int((g_prim[:F] != fp).sum()), "hits": int(both.sum()),
            "max_rel_t": float(f"{rel_t.max():.3e}"), "median_rel_t": float(f"{np.median(rel_t):.3e}"),
            "max_abs_uv": float(f"{err_uv.max():.3e}")}
    # compulsory traffic: every ray and hit once plus the whole scene once
    # (the device's 64 B nodes and 48 B primitive records)
    compulsory = N_RAYS * (32 + 16) + info["n_nodes"] * DEVICE_NODE_BYTES + len(prims) * 48

    cpu = None
    if not args.no_cpu:
        # every core of this GPU's CPU share (16 per GPU on the boxes this
        # runs on; fewer when the process is pinned to fewer)
        threads = min(oracle.default_threads(), 16)
        rate0 = S / max(t_count, 1e-6)  # rays/s with default threads
        n_cpu = int(min(N_RAYS, max(1 << 16, rate0 * args.cpu_seconds * threads / oracle.default_threads())))
        log(f"[bench] CPU baseline: oracle QBVH closest-hit on {n_cpu} rays, {threads} threads")
        t0 = time.perf_counter()
        oracle.qbvh_closest(nodes, prims, rays_h[:, :n_cpu], threads=threads)
        t_cpu = time.perf_counter() - t0
        cpu = {
            "value": round(n_cpu / t_cpu / 1e6, 4),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {n_cpu} of the {N_RAYS} closest-hit rays, same QBVH, "
                      f"{t_cpu:.1f} s on {threads} host threads",
            "host": host_info(),
        }

    render = None if args.no_render else bench_render(args, la, oracle, torch, np)
    direct = None if args.no_direct else bench_direct(args, la, oracle, torch, np)
    film = None if args.no_film else bench_film(args, la, torch, np)
    instanced = None if args.no_instanced else bench_instanced(args, la, torch, np)
    terrain = None if args.no_terrain else bench_terrain(args, la, torch, np)
    guided = None if args.no_guided else bench_guided(args, la, torch, np)
    batch = None if args.no_batch else bench_batch(args, la, torch, np)
    train = None if args.no_train else bench_train(args, la, torch, np)

    out = {
        "metric": METRIC,
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_c / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "random triangle soup, closest-hit",
            "triangles": N_TRIS,
            "rays": N_RAYS,
            "scene": "centre U[0,1)^3, vertices centre + 0.01*U[-1,1)^3, seed 2",
            "ray_gen": "origin U[0,1)^3, direction uniform on S^2, tnear 0, tfar inf, seed 3",
            "qbvh": f"SAH 4-wide, leaf<=4, {info['n_nodes']} nodes, depth {info['max_depth']}",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic("closest"),
            # achieved counts the canonical bytes per ray (128 B FP32 nodes,
            # oracle visit counts): an effective rate, part of it served by
            # L2 / MALL; the device's own node bytes (64 B quantised nodes)
            # give the lower device_frac
            "basis": "effective: canonical 32 + 16 + 128 N_node + 48 N_prim bytes per ray",
            "device_frac": round(N_RAYS * dev_b_ray / (kern_c * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            # what the SQ counters show bounds the launch: the VALU issue
            # slots (busy share), with few lanes active per instruction
            "issue": _valu_issue(),
        },
        "cpu_baseline": cpu,
        "detail": {
            "kernel_ms_closest": round(kern_c, 4),
            "bytes_per_ray_closest": round(b_ray, 1),
            "nodes_per_ray_closest": round(nodes_per_ray, 3),
            "prims_per_ray_closest": round(prims_per_ray, 3),
            "device_node_bytes": DEVICE_NODE_BYTES,
            "device_nodes_per_ray_closest": round(cnt[4] / S, 3),
            "device_prims_per_ray_closest": round(cnt[5] / S, 3),
            "device_bytes_per_ray_closest": round(dev_b_ray, 1),
            "device_gbs_closest": round(N_RAYS * dev_b_ray / (kern_c * 1e-3) / 1e9, 1),
            "device_bytes_per_ray_any": round(dev_b_ray_any, 1),
            "any_hit_mrays_s": round(mrays_any, 3),
            "kernel_ms_any": round(kern_a, 4),
            "bytes_per_ray_any": round(b_ray_any, 1),
            "achieved_gbs_any": round(achieved_any, 1),
            "traffic_any": _traffic("any"),
            "frac_any": round(achieved_any / HBM_PEAK_GBS, 4),
            "parity_sample": S,
            "parity_mismatches_closest": mism,
            "parity_mismatches_any": mism_any,
            "tied_hits_closest": n_tied,
            "tied_hit_fraction_closest": round(n_tied / S, 8),
            "tied_hit_mismatches_closest": mism_tied,
            "interval_quirk_rays_closest": n_quirk,
            "fp32_vs_fp64_closest": fp64,
            "measured_copy_gbs": round(copy_gbs, 1),
            "device_frac_of_measured_copy": round(N_RAYS * dev_b_ray / (kern_c * 1e-3) / 1e9 / copy_gbs, 4),
            "compulsory_bytes_closest": int(compulsory),
            "compulsory_ms_closest": round(compulsory / (HBM_PEAK_GBS * 1e9) * 1e3, 4),
            "hit_rate_closest": round(float((ref[0] != oracle.MISS).mean()), 4),
            "qbvh_build_s": round(info["build_seconds"], 3),
            "autotune_closest": tuned_c,
            "autotune_any": tuned_a,
        },
        "render": render,
        "direct": direct,
        "film": film,
        "instanced": instanced,
        "terrain": terrain,
        "guided": guided,
        "batch": batch,
        "train": train,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
