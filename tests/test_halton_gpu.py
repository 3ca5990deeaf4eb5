"""GPU parity of the Halton sampler (lr_render with LR_SAMPLER_HALTON).

1. Same sequence: the GPU render (FP32) and the oracle render (FP64) take the
   same Halton points and sub-sequence numbers, so their films agree far
   below the Monte Carlo noise; the per-pixel sample counts are equal.
2. Continuation: two renders at sample offsets 0 and n sum to one render of
   2n samples (up to the FP32 order of the film's atomic sums).
3. The generation limit and the renderers that take the Stratifier only
   (DirectLighting, guided, batches) fail with an error.
4. RenderEngine with samplers.Halton renders and advances the sequence.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle
from liar_amd import _native as N
from liar_amd import engine as E
from liar_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def halton_render(gs, cam, w, h, spp, nl=1, max_depth=8, rr_depth=3, seed=1, gens=8, offset=0):
    import torch

    film = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    prm = E.render_params(w, h, spp, nl, max_depth, rr_depth, 0.5, (380.0, 780.0), seed,
                          sampler=N.LR_SAMPLER_HALTON, sampler_generations=gens, sample_offset=offset)
    st = gs.render(E.camera_struct(cam), prm, film, wgt)
    torch.cuda.synchronize()
    return film.cpu().numpy().astype(np.float64), wgt.cpu().numpy().astype(np.float64), st


@pytest.fixture(scope="module")
def cornell(gpu_ctx):
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    gs = E.GpuScene(gpu_ctx, cs)
    yield cs, cam, gs
    gs.close()


def test_halton_same_sequence_matches_oracle(cornell):
    cs, cam, gs = cornell
    say("[gpu] halton cornell 64x64 16 spp x 2 lambda vs oracle")
    gx, gw, st = halton_render(gs, cam, 64, 64, 16, nl=2, seed=4)
    ox, ow = oracle.render(cs, cam, 64, 64, 16, wavelengths=2, max_depth=8, rr_depth=3, seed=4, halton=True,
                           sampler_generations=8)
    assert st["paths"] == 64 * 64 * 16 * 2
    # counts: FP32 radical inverses on both sides, the same pixel rule
    assert np.array_equal(gw, ow)
    tot_g, tot_o = gx[1].sum(), ox[1].sum()
    rel = abs(tot_g - tot_o) / tot_o
    scale = max((ox[1] / np.maximum(ow, 1)).max(), 1e-12)
    diff = np.abs(gx / np.maximum(gw, 1) - ox / np.maximum(ow, 1)) / scale
    bad = float((diff.max(0) > 1e-4).mean())
    say(f"[gpu] halton: total Y gpu {tot_g:.6f} oracle {tot_o:.6f} rel {rel:.2e}; pixels > 1e-4: {bad:.4f}")
    assert rel < 2e-3
    assert bad < 0.02


def test_halton_offset_continues_the_sequence(cornell):
    _, cam, gs = cornell
    n = 48 * 48 * 4
    a = halton_render(gs, cam, 48, 48, 4, seed=2)
    b = halton_render(gs, cam, 48, 48, 4, seed=2, offset=n)
    c = halton_render(gs, cam, 48, 48, 8, seed=2)
    assert np.array_equal(a[1] + b[1], c[1])
    np.testing.assert_allclose(a[0] + b[0], c[0], rtol=1e-4, atol=1e-6 * c[0].max())
    assert not np.array_equal(a[0], b[0])


def test_halton_limits_and_unsupported(cornell, gpu_ctx):
    import torch

    _, cam, gs = cornell
    film = torch.zeros((3, 8, 8), dtype=torch.float32, device="cuda")
    wgt = torch.zeros((8, 8), dtype=torch.float32, device="cuda")
    cam_s = E.camera_struct(cam)
    too_many = E.render_params(8, 8, 1, 1, 8, 3, 0.5, (380.0, 780.0), 1, sampler=N.LR_SAMPLER_HALTON,
                               sampler_generations=29)
    with pytest.raises(RuntimeError, match="28"):
        gs.render(cam_s, too_many, film, wgt)
    prm = E.render_params(8, 8, 1, 1, 8, 3, 0.5, (380.0, 780.0), 1, sampler=N.LR_SAMPLER_HALTON,
                          sampler_generations=8)
    dp = N.DirectParams(8, 1, 1)
    st = N.RenderStats()
    rc = N.lib().lr_render_direct(gs._h, C.byref(cam_s), C.byref(prm), C.byref(dp), C.c_void_p(film.data_ptr()),
                                  C.c_void_p(wgt.data_ptr()), C.byref(st), None)
    assert rc != 0
    bad = E.render_params(8, 8, 1, 1, 8, 3, 0.5, (380.0, 780.0), 1, sampler=7)
    with pytest.raises(RuntimeError, match="sampler"):
        gs.render(cam_s, bad, film, wgt)
    torch.cuda.synchronize()


def test_render_engine_with_halton(gpu_ctx):
    from liar_amd import output, samplers, tracers

    sc, cam = W.cornell_box()
    eng = la.RenderEngine()
    eng.scene, eng.camera = sc, cam
    eng.tracer = tracers.AdjointPhotonTracer()
    eng.sampler = samplers.Halton((32, 32), 4)
    eng.target = output.Raster((32, 32))
    eng.render()
    assert eng.sampler.nextSample == 32 * 32 * 4
    assert eng.target.weight.sum() == 32 * 32 * 4 and np.isfinite(eng.target.xyz_sum).all()
    eng.render()
    assert eng.sampler.nextSample == 2 * 32 * 32 * 4
    eng.tracer.maxRayGeneration = 29
    with pytest.raises(ValueError, match="Too many subsequences"):
        eng.render()
