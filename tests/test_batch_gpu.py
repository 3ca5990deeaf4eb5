"""Batched rendering (lr_batch_*): several scenes of uneven size, spp and
depth in one wavefront whose queues mix their paths.

Bar: every scene's film is bit-identical to its own lr_render with the same
parameters (the batch draws the same random numbers per path, traverses a
concatenated tree from the scene's root and splats whole pixels), eagerly and
graph-captured, with chunks that span scene boundaries; per-scene completion
times follow the shortest-first order.
"""
import numpy as np
import pytest

import liar_amd as la
from liar_amd import engine as E
from liar_amd import workloads as W

pytestmark = pytest.mark.gpu


def _film(w, h):
    import torch

    return (torch.zeros((3, h, w), dtype=torch.float32, device="cuda"),
            torch.zeros((h, w), dtype=torch.float32, device="cuda"))


def _scenes():
    cb, cam_cb = W.cornell_box()
    cube, cam_cube = W.random_cube(20000, seed=3)
    ter, cam_ter = W.terrain(65, n_spheres=40, seed=5)
    small, cam_small = W.random_cube(3000, seed=4)
    return [
        (la.compile_scene(cb), cam_cb, dict(width=48, height=40, spp=16, wl=4, depth=8)),
        (cube, cam_cube, dict(width=64, height=64, spp=4, wl=1, depth=5)),
        (ter, cam_ter, dict(width=80, height=45, spp=9, wl=2, depth=12)),
        (small, cam_small, dict(width=16, height=16, spp=64, wl=1, depth=4)),
    ]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_films_equal_solo_renders(gpu_ctx, graph):
    import torch

    print(f"[gpu] batch vs solo renders (graph {graph})", flush=True)
    specs = _scenes()
    gss = [E.GpuScene(gpu_ctx, cs) for cs, _, _ in specs]
    batch = E.Batch(gpu_ctx, gss)
    for call in range(2):
        items, solo = [], []
        for k, ((cs, cam, p), gs) in enumerate(zip(specs, gss)):
            # chunk_paths 20000: chunks end inside scenes and span scene boundaries
            prm = E.render_params(p["width"], p["height"], p["spp"], p["wl"], p["depth"], 3, 0.5, (380, 780),
                                  7 + 13 * k + call, 20000)
            fx, fw = _film(p["width"], p["height"])
            items.append((E.camera_struct(cam), prm, fx, fw))
            sx, sw = _film(p["width"], p["height"])
            gs.render(E.camera_struct(cam), prm, sx, sw)
            solo.append((sx, sw))
        res = batch.render(items, graph=graph)
        torch.cuda.synchronize()
        assert res["chunks"] > len(specs)
        for k, ((_, _, fx, fw), (sx, sw)) in enumerate(zip(items, solo)):
            assert np.array_equal(fx.cpu().numpy(), sx.cpu().numpy()), f"scene {k} xyz, call {call}"
            assert np.array_equal(fw.cpu().numpy(), sw.cpu().numpy()), f"scene {k} weight, call {call}"
        ms = np.array(res["scene_ms"])
        paths = np.array([p["width"] * p["height"] * p["spp"] * p["wl"] for _, _, p in specs])
        order = np.argsort(paths, kind="stable")
        print(f"[gpu] batch call {call}: {res['chunks']} chunks, {res['rays']} rays, scene ms {np.round(ms, 3)}",
              flush=True)
        assert (ms > 0).all() and (np.diff(ms[order]) >= 0).all()
    batch.close()
    for gs in gss:
        gs.close()
