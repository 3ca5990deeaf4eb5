"""Render-level traversal autotuning (liar_amd.autotune.tune_render): the
fastest configuration is installed in the context, and the film does not
depend on the traversal configuration (bit-identical under each)."""
import numpy as np
import pytest

import liar_amd as la
from liar_amd import engine as E
from liar_amd import workloads as W
from liar_amd.autotune import tune_render

pytestmark = pytest.mark.gpu


def test_tune_render_installs_fastest_and_keeps_film(gpu_ctx):
    import torch

    print("[gpu] render autotune", flush=True)
    sc, cam = W.cornell_box()
    gs = E.GpuScene(gpu_ctx, la.compile_scene(sc))
    cs = E.camera_struct(cam)
    prm = E.render_params(64, 64, 16, 2, 8, 3, 0.5, (380, 780), 3)
    old = gpu_ctx.trace_config()
    configs = [dict(block=b, lds_stack=l, blocks_per_cu=4, refill_min=32, schedule=sc)
               for b in (128, 256) for l in (8, 24) for sc in (1, 2)]
    films = []
    try:
        res = tune_render(gs, cs, prm, repeats=1, configs=configs)
        assert len(res["table"]) == 8 and res["best_ms"] == min(ms for _, ms in res["table"])
        assert gpu_ctx.trace_config() == res["best"]
        for cfg in configs:
            gpu_ctx.set_trace_config(**cfg)
            fx = torch.zeros((3, 64, 64), device="cuda")
            fw = torch.zeros((64, 64), device="cuda")
            gs.render(cs, prm, fx, fw)
            torch.cuda.synchronize()
            films.append(fx.cpu().numpy())
    finally:
        gpu_ctx.set_trace_config(**old)
    for f in films[1:]:
        assert np.array_equal(films[0], f)
    gs.close()


def test_tune_direct_lighting_keeps_film(gpu_ctx):
    """tune_render(direct=True) times lr_render_direct; its film (float
    atomics: equal up to the rounding of the sums) does not depend on the
    traversal schedule."""
    import torch

    print("[gpu] DirectLighting autotune", flush=True)
    sc, cam = W.cornell_box()
    gs = E.GpuScene(gpu_ctx, la.compile_scene(sc))
    cs = E.camera_struct(cam)
    prm = E.render_params(64, 64, 4, 2, 0, 8, 0.5, (380, 780), 5)
    old = gpu_ctx.trace_config()
    configs = [dict(block=256, lds_stack=8, blocks_per_cu=8, refill_min=32, schedule=sc) for sc in (1, 2)]
    films = []
    try:
        res = tune_render(gs, cs, prm, repeats=1, configs=configs, direct=True)
        assert len(res["table"]) == 2 and gpu_ctx.trace_config() == res["best"]
        for cfg in configs:
            gpu_ctx.set_trace_config(**cfg)
            fx = torch.zeros((3, 64, 64), device="cuda")
            fw = torch.zeros((64, 64), device="cuda")
            gs.render_direct(cs, prm, fx, fw)
            torch.cuda.synchronize()
            films.append(fx.cpu().numpy().astype(np.float64))
    finally:
        gpu_ctx.set_trace_config(**old)
    assert films[0].sum() > 0
    np.testing.assert_allclose(films[1], films[0], rtol=1e-4, atol=1e-6 * float(films[0].max()))
    gs.close()
