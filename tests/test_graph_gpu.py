"""Graph-captured rendering (lr_render_params.graph): every launch of a render
call (all chunks and bounces, guide inference and training) captured into a
HIP graph kept per scene, updated in place on later calls and launched as one.

Bar: films bit-identical to the eager path, call after call, with eager work
on the same stream between the graph launches (the cached executable is
updated with new seeds, not re-instantiated), on the legacy null stream and
on a side stream.
"""
import numpy as np
import pytest

import liar_amd as la
from liar_amd import engine as E
from liar_amd import workloads as W
from liar_amd.guide import Guide

pytestmark = pytest.mark.gpu


def _film(w, h):
    import torch

    return (torch.zeros((3, h, w), dtype=torch.float32, device="cuda"),
            torch.zeros((h, w), dtype=torch.float32, device="cuda"))


def _np(fx, fw):
    import torch

    torch.cuda.synchronize()
    return fx.cpu().numpy(), fw.cpu().numpy()


@pytest.mark.parametrize("side_stream", [False, True])
def test_graph_render_matches_eager(gpu_ctx, side_stream):
    import torch

    print(f"[gpu] graph vs eager render (side stream {side_stream})", flush=True)
    sc, cam = W.cornell_box()
    gs = E.GpuScene(gpu_ctx, la.compile_scene(sc))
    cs = E.camera_struct(cam)
    stream = torch.cuda.Stream() if side_stream else torch.cuda.current_stream()
    with torch.cuda.stream(stream):
        for call in range(4):
            # chunk_paths forces several chunks (and so several camera / splat groups) per graph
            prm = E.render_params(96, 64, 16, 4, 8, 3, 0.5, (380, 780), 7 + call, 1 << 16)
            fe = _film(96, 64)
            gs.render(cs, prm, *fe)
            eager = _np(*fe)
            prm.graph = 1
            fg = _film(96, 64)
            st = gs.render(cs, prm, *fg)
            # eager work queued behind the graph on the same stream
            junk = torch.ones(1 << 20, device="cuda").cumsum(0)
            got = _np(*fg)
            assert st["chunks"] > 1
            assert np.array_equal(eager[0], got[0]) and np.array_equal(eager[1], got[1]), f"call {call}"
            assert float(junk[-1]) == float(1 << 20)
    gs.close()


def test_graph_guided_inference_matches_eager_and_training_runs(gpu_ctx):
    import torch

    print("[gpu] graph vs eager guided render", flush=True)
    sc, cam = W.cornell_box()
    gs = E.GpuScene(gpu_ctx, la.compile_scene(sc))
    cs = E.camera_struct(cam)
    g = Guide(gpu_ctx, seed=4, learning_rate=2e-3, radiance_weight=1.0)
    # train a little (eager), then compare inference-only renders (with the
    # radiance-driven roulette, whose per-path state is reset by the kernels)
    E.guided_render(gs, g, cs, E.render_params(64, 64, 16, 1, 8, 3, 0.5, (380, 780), 1), *_film(64, 64),
                    train=True, train_steps=8, train_batch=8192)
    for call in range(3):
        prm = E.render_params(64, 64, 16, 2, 8, 3, 0.5, (380, 780), 20 + call, 1 << 15)
        fe = _film(64, 64)
        E.guided_render(gs, g, cs, prm, *fe, learned_selection=True, balance_heuristic=True, learned_rr=True)
        eager = _np(*fe)
        prm.graph = 1
        fg = _film(64, 64)
        E.guided_render(gs, g, cs, prm, *fg, learned_selection=True, balance_heuristic=True, learned_rr=True)
        got = _np(*fg)
        assert np.array_equal(eager[0], got[0]) and np.array_equal(eager[1], got[1]), f"call {call}"
    # training inside the graph (records, the clamp and radiance sums, Adam):
    # the guide's parameters move and stay finite
    p0 = g.params()
    for call in range(3):
        prm = E.render_params(64, 64, 16, 1, 8, 3, 0.5, (380, 780), 40 + call, 1 << 15, graph=True)
        E.guided_render(gs, g, cs, prm, *_film(64, 64), train=True, train_steps=4, train_batch=4096,
                        weight_clamp=20.0, learned_rr=True)
    torch.cuda.synchronize()
    p1 = g.params()
    assert np.isfinite(p1).all() and not np.array_equal(p0, p1)
    g.close()
    gs.close()
