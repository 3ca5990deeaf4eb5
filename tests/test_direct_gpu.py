"""GPU parity of the DirectLighting integrator (lr_render_direct).

1. Same seed, same sample streams: the GPU render (FP32, wavefront over the
   ray tree) and the CPU oracle (FP64 recursion, brute-force shadow rays)
   draw the same random numbers at every tree node, so their images agree
   far below the Monte Carlo noise.
2. Closed forms (pinned on the oracle first, tests/test_oracle_direct.py):
   Lambert floor under a sky or a parallelogram light, Mirror floor, the
   secondary (picked-light) estimate seen through a glossy mirror, the
   emission and generation rules.
3. Chunking: the film does not depend on chunk_paths (each path's random
   numbers are keyed by its index in the image), including chunks small
   enough that the branch and shadow queues must grow and repeat a chunk.
"""
import math
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle
from direct_scenes import AREA, SIDE_LIGHTS, floor_expect, glossy_expect, glossy_view, glow_light, mirror_view
from liar_amd import engine as E
from liar_amd import scenery, shaders, textures
from liar_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_direct(ctx, cs, cam, w, h, spp, nl=1, seed=1, chunk=0, lam=(380.0, 780.0), **kw):
    import torch

    gs = E.GpuScene(ctx, cs)
    film = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    st = gs.render_direct(E.camera_struct(cam), E.render_params(w, h, spp, nl, 0, 8, 0.5, lam, seed, chunk),
                          film, wgt, **kw)
    torch.cuda.synchronize()
    out = film.cpu().numpy().astype(np.float64), wgt.cpu().numpy().astype(np.float64), st
    gs.close()
    return out


def compare(gx, gw, ox, ow, label):
    assert np.array_equal(gw, ow), "sample counts differ"
    g, o = gx / gw, ox / ow
    tot_g, tot_o = g[1].sum(), o[1].sum()
    rel_total = abs(tot_g - tot_o) / tot_o
    diff = np.abs(g - o) / max(o[1].max(), 1e-12)
    frac_bad = float((diff.max(0) > 1e-4).mean())
    say(f"[gpu] {label}: total Y gpu {tot_g:.6f} oracle {tot_o:.6f} rel {rel_total:.2e}; "
        f"max |diff|/max {diff.max():.2e}; pixels > 1e-4: {frac_bad:.4f}")
    return rel_total, frac_bad


def _glass_scene():
    """The Cornell box with a glass, a mirror and a glossy sphere: every
    branch kind (reflection and transmission, high quality and not)."""
    sc, cam = W.cornell_box()
    objs = list(sc.children)
    specs = [(shaders.Dielectric(textures.Constant(1.5)), (1.86, 2.35, 1.7), 0.7),  # on the short box
             (shaders.Mirror(textures.Constant(0.9)), (4.3, 0.6, 1.2), 0.6)]
    ct = shaders.CookTorrance(textures.Constant(1.2), textures.Constant(2.0))
    ct.roughnessU = ct.roughnessV = textures.Constant(0.15)
    specs.append((ct, (1.0, 0.5, 3.8), 0.5))
    for shader, c, r in specs:
        s = scenery.Sphere(c, r)
        s.shader = shader
        objs.append(s)
    return scenery.List(objs), cam


def test_cornell_same_seed_matches_oracle(gpu_ctx):
    say("[gpu] direct cornell 64x64 4spp x 2 lambda vs oracle")
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    gx, gw, st = gpu_direct(gpu_ctx, cs, cam, 64, 64, 4, nl=2, seed=1)
    ox, ow = oracle.render_direct(cs, cam, 64, 64, 4, wavelengths=2, seed=1)
    rel, bad = compare(gx, gw, ox, ow, "direct cornell")
    assert st["paths"] == 64 * 64 * 8 and st["rays"] > st["paths"]
    assert rel < 2e-3 and bad < 0.02


def test_branching_scene_same_seed_matches_oracle(gpu_ctx):
    say("[gpu] direct glass/mirror/glossy spheres 64x64 4spp x 2 lambda vs oracle")
    sc, cam = _glass_scene()
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_direct(gpu_ctx, cs, cam, 64, 64, 4, nl=2, seed=3)
    ox, ow = oracle.render_direct(cs, cam, 64, 64, 4, wavelengths=2, seed=3)
    rel, bad = compare(gx, gw, ox, ow, "direct branching")
    assert rel < 5e-3 and bad < 0.05


def test_terrain_sun_sky_same_seed_matches_oracle(gpu_ctx):
    """Small terrain: Lambert, glossy Cook-Torrance gold and dielectric
    materials under a sun disc and a sky -- glossy continuations (secondary,
    picked-light estimates), dielectric branches, sun and sky shadow rays."""
    say("[gpu] direct small terrain 48x27 4spp vs oracle")
    cs, cam = W.terrain(n_side=101, n_spheres=300, seed=7, tile=4)
    gx, gw, st = gpu_direct(gpu_ctx, cs, cam, 48, 27, 4, nl=1, seed=11)
    ox, ow = oracle.render_direct(cs, cam, 48, 27, 4, wavelengths=1, seed=11)
    rel, bad = compare(gx, gw, ox, ow, "direct terrain")
    assert rel < 5e-3 and bad < 0.05


def test_chunking_and_queue_growth(gpu_ctx):
    """Chunks of 512 paths (queues grown and chunks repeated on overflow) give
    the film of one big chunk bit for bit (a path's radiance is summed in
    fixed point, so the order of its terms does not matter), and a re-render
    is bit-identical."""
    say("[gpu] direct chunk independence (512-path chunks vs one chunk)")
    sc, cam = _glass_scene()
    cs = la.compile_scene(sc)
    a, wa, sa = gpu_direct(gpu_ctx, cs, cam, 48, 48, 4, nl=2, seed=9, chunk=512)
    b, wb, sb = gpu_direct(gpu_ctx, cs, cam, 48, 48, 4, nl=2, seed=9, chunk=1 << 20)
    say(f"[gpu] chunks {sa['chunks']} vs {sb['chunks']}, rays {sa['rays']} vs {sb['rays']}, "
        f"max rel diff {np.abs(a - b).max() / np.abs(b).max():.2e}")
    assert sa["chunks"] == 48 * 48 * 8 // 512 and sb["chunks"] == 1
    assert np.array_equal(wa, wb)
    assert sa["rays"] == sb["rays"]
    assert np.array_equal(a, b)
    c, _, _ = gpu_direct(gpu_ctx, cs, cam, 48, 48, 4, nl=2, seed=9, chunk=1 << 20)
    assert np.array_equal(b, c)


def _mean(ctx, scene, cam, spp=64, size=16, **kw):
    cs = la.compile_scene(scene)
    gx, gw, _ = gpu_direct(ctx, cs, cam, size, size, spp, seed=5, **kw)
    Y = (gx[1] / gw).ravel()
    return Y.mean(), Y.std() / math.sqrt(Y.size), Y


def test_closed_forms(gpu_ctx):
    say("[gpu] direct closed forms")
    sc, cam = W.lit_floor(shaders.Lambert(textures.Constant(0.7)), sky=2.0)
    m, se, _ = _mean(gpu_ctx, sc, cam)
    say(f"[gpu] lambert sky: {m:.5f} (expect 1.4, se {se:.1e})")
    assert abs(m - 1.4) < 4 * se + 1e-4
    for samples in (9, 1):
        area = dict(AREA, samples=samples)
        sc, cam = W.lit_floor(shaders.Lambert(textures.Constant(0.7)), area=area)
        m, se, _ = _mean(gpu_ctx, sc, cam, spp=256)
        want = floor_expect(area, 0.7)
        say(f"[gpu] lambert area light ({samples} samples): {m:.5f} (expect {want:.5f}, se {se:.1e})")
        assert abs(m - want) < 4 * se + 1e-4 * want
    sc, cam = W.lit_floor(shaders.Mirror(textures.Constant(0.8)), area=AREA)
    m, se, _ = _mean(gpu_ctx, sc, cam)
    say(f"[gpu] mirror under light: {m:.5f} (expect 4.0, se {se:.1e})")
    assert abs(m - 4.0) < 4 * se + 1e-4


@pytest.mark.parametrize("secondary", [1, 3])
def test_secondary_estimate_through_glossy_mirror(gpu_ctx, secondary):
    sc, cam = glossy_view(SIDE_LIGHTS)
    m, se, _ = _mean(gpu_ctx, sc, cam, spp=256, secondary_light_samples=secondary)
    want = glossy_expect(SIDE_LIGHTS)
    say(f"[gpu] glossy mirror, {secondary} secondary samples: {m:.5f} (expect {want:.5f}, se {se:.1e})")
    assert abs(m - want) < 4 * se + 2e-4 * want


def test_emission_and_generation_rules(gpu_ctx):
    say("[gpu] direct emission / generation rules")
    parts, cam = mirror_view(shaders.Unshaded(textures.Constant(3.0)))
    m, se, _ = _mean(gpu_ctx, scenery.List(parts), cam)
    assert abs(m - 1.5) < 4 * se + 1e-4
    m, se, _ = _mean(gpu_ctx, scenery.List([glow_light(3.0), parts[1]]), cam)
    assert abs(m - 1.5) < 4 * se + 1e-4
    _, _, Y = _mean(gpu_ctx, scenery.List(parts), cam, max_generation=0)
    assert (Y == 0).all()


def test_invalid_arguments(gpu_ctx):
    sc, cam = W.lit_floor(shaders.Lambert(textures.Constant(0.7)), sky=2.0)
    with pytest.raises(la.LiarError, match="max_generation"):
        _mean(gpu_ctx, sc, cam, max_generation=24)
    with pytest.raises(la.LiarError, match="secondary_light_samples"):
        _mean(gpu_ctx, sc, cam, secondary_light_samples=9)
