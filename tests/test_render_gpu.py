"""GPU parity of the spectral wavefront path tracer.

1. Same seed, same sample streams: the GPU render (FP32) and the CPU oracle
   render (FP64) follow the same paths, so their images agree far below the
   Monte Carlo noise; only paths whose FP32/FP64 decisions differ diverge.
2. White furnace (test/white_furnace/ashikhmin_shirley.py geometry): a
   convex unit sphere under a constant sky of radiance 1 gives Y = rho for
   Lambert and Mirror, and Y = 1 for a lossless dielectric.
3. Determinism: two renders with the same seed are bit-identical.
"""
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle
from liar_amd import engine as E
from liar_amd import scenery, shaders, spectra, textures
from liar_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_render(ctx, cs, cam, w, h, spp, nl=1, max_depth=0, rr_depth=8, seed=1, lam=(380.0, 780.0)):
    import torch

    gs = E.GpuScene(ctx, cs)
    film = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    stats = gs.render(E.camera_struct(cam), E.render_params(w, h, spp, nl, max_depth, rr_depth, 0.5, lam, seed),
                      film, wgt)
    torch.cuda.synchronize()
    out = film.cpu().numpy().astype(np.float64), wgt.cpu().numpy().astype(np.float64), stats
    gs.close()
    return out


def compare(gx, gw, ox, ow, label):
    assert np.array_equal(gw, ow), "sample counts differ"
    g = gx / gw
    o = ox / ow
    tot_g, tot_o = g[1].sum(), o[1].sum()
    rel_total = abs(tot_g - tot_o) / tot_o
    scale = max(o[1].max(), 1e-12)
    diff = np.abs(g - o) / scale
    frac_bad = float((diff.max(0) > 1e-4).mean())
    say(f"[gpu] {label}: total Y gpu {tot_g:.6f} oracle {tot_o:.6f} rel {rel_total:.2e}; "
        f"max |diff|/max {diff.max():.2e}; pixels > 1e-4: {frac_bad:.4f}")
    return rel_total, frac_bad, diff


def test_cornell_same_seed_matches_oracle(gpu_ctx):
    say("[gpu] cornell 64x64 16spp x 4 lambda vs oracle")
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    gx, gw, st = gpu_render(gpu_ctx, cs, cam, 64, 64, 16, nl=4, max_depth=8, rr_depth=3, seed=1)
    ox, ow = oracle.render(cs, cam, 64, 64, 16, wavelengths=4, max_depth=8, rr_depth=3, seed=1)
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "cornell")
    assert st["paths"] == 64 * 64 * 16 * 4 and st["rays"] >= st["paths"]
    assert rel_total < 2e-3
    assert frac_bad < 0.02


def _mixed_scene():
    au_w = np.array([281.7, 359.3, 413.2, 476.8, 516.6, 563.5, 619.9, 826.5])
    au_n = np.array([1.848, 1.967, 1.950, 1.541, 0.878, 0.613, 0.487, 0.418])
    au_k = np.array([1.87, 1.84, 1.86, 1.71, 2.01, 2.64, 3.31, 5.13])
    floor = scenery.Parallelogram((-4, -4, 0), (8, 0, 0), (0, 8, 0))
    floor.shader = shaders.Lambert(textures.Constant(spectra.Binned(380, 780, np.linspace(0.2, 0.8, 10))))
    objs = [floor]
    mats = [shaders.Lambert(textures.Constant(0.7)), shaders.Mirror(textures.Constant(0.9)),
            shaders.Dielectric(textures.Constant(1.5)),
            shaders.Conductor(textures.Constant(spectra.Sampled(au_w, au_n)),
                              textures.Constant(spectra.Sampled(au_w, au_k)))]
    ct = shaders.CookTorrance(textures.Constant(spectra.Sampled(au_w, au_n)),
                              textures.Constant(spectra.Sampled(au_w, au_k)))
    ct.mdf = shaders.MicrofacetTrowbridgeReitz()
    ct.roughnessU = ct.roughnessV = textures.Constant(0.3)
    mats.append(ct)
    ctb = shaders.CookTorrance(textures.Constant(1.2), textures.Constant(2.0))
    ctb.roughnessU = ctb.roughnessV = textures.Constant(0.2)
    mats.append(ctb)
    for i, m in enumerate(mats):
        s = scenery.Sphere((-2.5 + i, 0.0, 0.5), 0.45)
        s.shader = m
        objs.append(s)
    light = scenery.LightArea(scenery.Parallelogram((-1, -1, 4), (2, 0, 0), (0, -2, 0)))
    light.radiance = textures.Constant(spectra.BlackBody(4500, 8.0))
    sky = scenery.LightSky(textures.Constant(0.3))
    sun = scenery.LightSun((0.3, -0.4, 0.87), np.radians(0.53), textures.Constant(spectra.BlackBody(5800, 2e3)))
    scene = scenery.List(objs + [light, sky, sun])
    cam = la.cameras.PerspectiveCamera()
    cam.position = (0.0, -7.0, 2.5)
    cam.lookAt((0.0, 0.0, 0.4))
    cam.fovAngle = np.radians(45)
    cam.aspectRatio = 1.5
    return scene, cam


def test_mixed_materials_same_seed_matches_oracle(gpu_ctx):
    say("[gpu] mixed BSDFs / lights vs oracle")
    sc, cam = _mixed_scene()
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 96, 64, 16, nl=2, max_depth=12, rr_depth=4, seed=5)
    ox, ow = oracle.render(cs, cam, 96, 64, 16, wavelengths=2, max_depth=12, rr_depth=4, seed=5)
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "mixed")
    assert rel_total < 5e-3
    assert frac_bad < 0.05


@pytest.mark.parametrize("name,shader,expect", [
    ("lambert_1", shaders.Lambert(textures.Constant(1.0)), 1.0),
    ("lambert_0.5", shaders.Lambert(textures.Constant(0.5)), 0.5),
    ("mirror_1", shaders.Mirror(textures.Constant(1.0)), 1.0),
    ("dielectric_lossless", shaders.Dielectric(textures.Constant(1.5)), 1.0),
])
def test_white_furnace(gpu_ctx, name, shader, expect):
    say(f"[gpu] white furnace {name}")
    sc, cam = W.furnace_sphere(shader)
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 128, 128, 32768, nl=1, max_depth=0, rr_depth=8, seed=3)
    img = gx / gw
    # the sphere covers a disc of radius ~60 px at the image centre (fov 11
    # degrees at distance sqrt(125)); test pixels well inside it
    yy, xx = np.mgrid[0:128, 0:128]
    centre = (xx - 63.5) ** 2 + (yy - 63.5) ** 2 < 36 ** 2
    Y = img[1][centre]
    three_sigma = 3 * Y.std() / np.sqrt(Y.size)
    say(f"[gpu] furnace {name}: mean Y {Y.mean():.6f} (expect {expect}), 3 sigma {three_sigma:.1e}, n {Y.size}")
    # enough samples that the noise is well inside the 1e-3 energy bound
    # (test/white_furnace/ashikhmin_shirley.py:20-79), which then holds alone
    assert three_sigma < 5e-4
    assert abs(Y.mean() - expect) < 1e-3
    corner = img[1][:8, :8]
    assert (corner == 0).all()  # camera rays that miss see no sky (generation 0)


def test_render_deterministic(gpu_ctx):
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    a = gpu_render(gpu_ctx, cs, cam, 48, 48, 4, nl=4, max_depth=8, rr_depth=3, seed=9)
    b = gpu_render(gpu_ctx, cs, cam, 48, 48, 4, nl=4, max_depth=8, rr_depth=3, seed=9)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = gpu_render(gpu_ctx, cs, cam, 48, 48, 4, nl=4, max_depth=8, rr_depth=3, seed=10)
    assert not np.array_equal(a[0], c[0])


def test_chunking_is_invisible(gpu_ctx):
    """Splitting the render into many wavefront chunks gives the same image."""
    import torch

    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    gs = E.GpuScene(gpu_ctx, cs)
    out = []
    for chunk in (0, 5000):
        film = torch.zeros((3, 40, 40), dtype=torch.float32, device="cuda")
        wgt = torch.zeros((40, 40), dtype=torch.float32, device="cuda")
        gs.render(E.camera_struct(cam), E.render_params(40, 40, 4, 4, 8, 3, 0.5, (380, 780), 2, chunk), film, wgt)
        torch.cuda.synchronize()
        out.append(film.cpu().numpy())
    gs.close()
    assert np.array_equal(out[0], out[1])


def test_terrain_small_same_seed_matches_oracle(gpu_ctx):
    """The terrain workload's generator at 101^2 vertices and 300 spheres
    (all three shaders, sun + sky) against the oracle, same seed."""
    say("[gpu] small terrain vs oracle")
    cs, cam = W.terrain(n_side=101, n_spheres=300, seed=7, tile=4)
    gx, gw, st = gpu_render(gpu_ctx, cs, cam, 48, 27, 8, nl=1, max_depth=16, rr_depth=16, seed=11)
    ox, ow = oracle.render(cs, cam, 48, 27, 8, wavelengths=1, max_depth=16, rr_depth=16, seed=11)
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "terrain-small")
    assert rel_total < 5e-3
    assert frac_bad < 0.05


def test_terrain_full_size_same_seed_matches_oracle(gpu_ctx):
    """The terrain workload at full size (4,992,800 triangles, 10,000
    spheres, all three shaders, sun + sky) through its own camera at 32x18
    pixels x 2 spp, depth 16, against the oracle's FP64 brute-force render
    of the same paths (same seed).  The GPU traverses its FP32 QBVH, the
    oracle tests every primitive in FP64, so a path whose nearest hit ties
    to FP32 precision may continue differently: the bar is the share of
    pixels that differ (and the image total), not bit equality."""
    import time

    say("[gpu] full terrain 32x18 x 2 spp vs oracle (the oracle takes ~12 s on 16 threads)")
    cs, cam = W.terrain()
    gx, gw, st = gpu_render(gpu_ctx, cs, cam, 32, 18, 2, nl=1, max_depth=16, rr_depth=16, seed=11)
    t0 = time.time()
    ox, ow = oracle.render(cs, cam, 32, 18, 2, wavelengths=1, max_depth=16, rr_depth=16, seed=11)
    say(f"[gpu] full terrain: oracle took {time.time() - t0:.1f} s; {st['rays']} GPU rays for {st['paths']} paths")
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "terrain-full")
    assert frac_bad < 0.02
    assert rel_total < 1e-2


def test_terrain_full_size_renders(gpu_ctx):
    """Full terrain workload (4,992,800 triangles + 10,000 spheres) at reduced
    resolution: finite, non-negative, deterministic image."""
    say("[gpu] full terrain 480x270 x 4 spp")
    cs, cam = W.terrain()
    a = gpu_render(gpu_ctx, cs, cam, 480, 270, 4, nl=1, max_depth=16, rr_depth=16, seed=1)
    img = a[0] / a[1]
    say(f"[gpu] terrain: mean Y {img[1].mean():.4f}, paths {a[2]['paths']}, rays {a[2]['rays']}, "
        f"{a[2]['seconds']:.3f} s")
    assert np.isfinite(img).all() and (img[1] >= 0).all() and img[1].mean() > 0


def _ref_y():
    from pathlib import Path

    d = np.load(Path(__file__).resolve().parent / "golden" / "cornell_ref_4096spp.npz")
    return d["y"].astype(np.float64)


def _mrae(a, b, block=1):
    """mean |a - b| / mean(b) over block x block pixel means."""
    n = a.shape[0] // block
    a = a.reshape(n, block, -1, block).mean((1, 3))
    b = b.reshape(n, block, -1, block).mean((1, 3))
    return float(np.abs(a - b).mean() / b.mean())


def test_cornell_vs_4096spp_cpu_reference(gpu_ctx):
    """Cornell workload (256x256, 4 wavelengths, depth <= 8, RR from 3) against
    the committed 4096-spp FP64 oracle render (seed 1000).  Measures:
    relative mean absolute error  MRAE = mean|Y - Yref| / mean(Yref)
    bias                          = (mean Y - mean Yref) / mean Yref.
    At the workload's 16 spp: MRAE < 0.35 (noise) and |bias| < 1e-2.
    At 1024 spp: |bias| < 2e-3, and MRAE (per pixel and over 16x16-pixel
    blocks, where noise falls ~16x and a structured error stays) below the
    MRAE between two 1024-spp GPU renders of other seeds.  For noise alone
    the first is ~0.8 of the second (a 1024-spp render's noise plus the
    4096-spp reference's, against two 1024-spp renders' noise); a per-pixel
    error of the size of one render's noise pushes it above."""
    ref = _ref_y()
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    ys = {}
    for spp, seed in ((16, 7), (1024, 7), (1024, 8)):
        gx, gw, st = gpu_render(gpu_ctx, cs, cam, 256, 256, spp, nl=4, max_depth=8, rr_depth=3, seed=seed)
        ys[spp, seed] = y = (gx / gw)[1]
        bias = (y.mean() - ref.mean()) / ref.mean()
        say(f"[gpu] cornell {spp} spp seed {seed} vs 4096-spp reference: MRAE {_mrae(y, ref):.4f}, "
            f"16x16 blocks {_mrae(y, ref, 16):.4f}, bias {bias:+.2e}, {st['seconds'] * 1e3:.1f} ms")
        assert abs(bias) < (1e-2 if spp == 16 else 2e-3)
    assert _mrae(ys[16, 7], ref) < 0.35
    a, b = ys[1024, 7], ys[1024, 8]
    for block in (1, 16):
        noise = _mrae(a, b, block)
        say(f"[gpu] block {block}: vs reference {_mrae(a, ref, block):.4f}, between GPU seeds {noise:.4f}")
        assert _mrae(a, ref, block) < noise
