"""LightArea over a TriangleMesh on the GPU (light_area.cpp:189-252,
triangle_mesh.cpp:414-444): the hexagon light's floor radiance equals
Lambert's polygon formula for the path tracer and for DirectLighting,
renders match the oracle with the same seed, and the back face of a
single-sided light emits when a ray hits it, as in LiAR."""
import math
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle
from liar_amd import engine as E
from liar_amd import shaders, textures

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from mesh_light_scenes import HEX_L, hexagon_floor_expect, hexagon_light_floor  # noqa: E402
from test_render_gpu import compare, gpu_render, say  # noqa: E402

pytestmark = pytest.mark.gpu


def gpu_direct(ctx, cs, cam, w, h, spp, seed):
    import torch

    gs = E.GpuScene(ctx, cs)
    film = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    gs.render_direct(E.camera_struct(cam), E.render_params(w, h, spp, 1, 0, 8, 0.5, (380.0, 780.0), seed), film, wgt)
    torch.cuda.synchronize()
    out = film.cpu().numpy().astype(np.float64), wgt.cpu().numpy().astype(np.float64)
    gs.close()
    return out


def test_hexagon_light_closed_form(gpu_ctx):
    sc, cam, ring = hexagon_light_floor(shaders.Lambert(textures.Constant(0.7)))
    cs = la.compile_scene(sc)
    want = hexagon_floor_expect(ring, 0.7)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 32, 32, 8192, max_depth=2, seed=3)
    Y = (gx / gw)[1]
    se = Y.std() / math.sqrt(Y.size)
    say(f"[gpu] hexagon light, path tracer: {Y.mean():.5f} (polygon formula {want:.5f}, se {se:.1e})")
    assert abs(Y.mean() - want) < 4 * se + 1e-3 * want
    dx, dw = gpu_direct(gpu_ctx, cs, cam, 32, 32, 1024, seed=4)
    Yd = (dx / dw)[1]
    se = Yd.std() / math.sqrt(Yd.size)
    say(f"[gpu] hexagon light, DirectLighting: {Yd.mean():.5f} (polygon formula {want:.5f}, se {se:.1e})")
    assert abs(Yd.mean() - want) < 4 * se + 1e-3 * want


def test_mesh_light_same_seed_matches_oracle(gpu_ctx):
    sc, cam, _ = hexagon_light_floor(shaders.Lambert(textures.Constant(0.7)), samples=4, double_sided=True)
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 32, 32, 64, max_depth=4, seed=7)
    ox, ow = oracle.render(cs, cam, 32, 32, 64, max_depth=4, seed=7)
    rel, bad, _ = compare(gx, gw, ox, ow, "mesh light apt")
    assert rel < 1e-4 and bad < 0.01
    dx, dw = gpu_direct(gpu_ctx, cs, cam, 32, 32, 16, seed=8)
    ox, ow = oracle.render_direct(cs, cam, 32, 32, 16, seed=8)
    rel, bad, _ = compare(dx, dw, ox, ow, "mesh light direct")
    assert rel < 1e-4 and bad < 0.01


def test_back_face_of_single_sided_light_emits_when_hit(gpu_ctx):
    sc, cam, _ = hexagon_light_floor(shaders.Mirror(textures.Constant(0.8)), facing=+1.0)
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 8, 8, 64, max_depth=3, seed=5)
    dx, dw = gpu_direct(gpu_ctx, cs, cam, 8, 8, 64, seed=5)
    for label, (x, w) in (("apt", (gx, gw)), ("direct", (dx, dw))):
        Y = (x / w)[1]
        say(f"[gpu] back face via mirror ({label}): {Y.mean():.4f} (expect {0.8 * HEX_L})")
        assert abs(Y.mean() - 0.8 * HEX_L) < 0.05
