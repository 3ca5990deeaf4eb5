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
    configs = [dict(block=b, lds_stack=l, blocks_per_cu=4, refill_min=32) for b in (128, 256) for l in (8, 24)]
    films = []
    try:
        res = tune_render(gs, cs, prm, repeats=1, configs=configs)
        assert len(res["table"]) == 4 and res["best_ms"] == min(ms for _, ms in res["table"])
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
