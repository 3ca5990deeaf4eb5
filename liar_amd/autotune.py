"""Render-level traversal autotuning.

lr_autotune_trace (``Qbvh.autotune``) tunes the traversal kernel on one batch
of rays.  A path tracer's bounces trace other rays (incoherent, grazing,
short) and each bounce launch lasts as long as its slowest rays, so the best
configuration for a whole render can differ from the best for camera rays:
deep LDS stacks keep long traversals from spilling, fewer resident blocks
trade occupancy for LDS.  ``tune_render`` times complete lr_render calls of
the caller's workload under each configuration (HIP events on the current
stream) and installs the fastest in the scene's context.
"""
from __future__ import annotations

import itertools

from . import engine as E

BLOCKS = (128, 256)
LDS_STACKS = (8, 16, 24)
BLOCKS_PER_CU = (4, 8, 16)
REFILL = (16, 32, 48)
SCHEDULES = (1, 2)  # LR_TRACE_SCHEDULE_STEP, LR_TRACE_SCHEDULE_PHASED


def tune_render(gs: "E.GpuScene", camera, params, repeats: int = 2, configs=None, direct: bool = False) -> dict:
    """Times lr_render(gs, camera, params) -- lr_render_direct with
    ``direct`` -- under every (block, lds_stack, blocks_per_cu, refill_min,
    schedule) in ``configs`` (default: the product of the module's value
    lists, 108 configurations) and sets the fastest as the context's traversal
    configuration.  Returns {"best", "best_ms", "default_ms", "table"}; the
    film written meanwhile is scratch."""
    import torch

    ctx = gs.ctx
    start = ctx.trace_config()
    if configs is None:
        configs = [dict(block=b, lds_stack=l, blocks_per_cu=c, refill_min=r, schedule=sc)
                   for b, l, c, r, sc in itertools.product(BLOCKS, LDS_STACKS, BLOCKS_PER_CU, REFILL, SCHEDULES)]
    w, h = params.width, params.height
    fx = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    fw = torch.zeros((h, w), dtype=torch.float32, device="cuda")

    render = gs.render_direct if direct else gs.render

    def timed() -> float:
        render(camera, params, fx, fw)  # warm-up (scratch sized for this configuration)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(repeats):
            render(camera, params, fx, fw)
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / repeats

    table = []
    try:
        for cfg in configs:
            ctx.set_trace_config(**cfg)
            table.append((ctx.trace_config(), timed()))
    except Exception:
        ctx.set_trace_config(**start)
        raise
    best, best_ms = min(table, key=lambda x: x[1])
    ctx.set_trace_config(**best)
    default_ms = next((ms for c, ms in table if c == start), None)
    return {"best": best, "best_ms": best_ms, "default_ms": default_ms, "table": table}


CHUNKS = (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)


def tune_chunk(gs: "E.GpuScene", camera, params, chunks=CHUNKS, repeats: int = 1) -> dict:
    """The render's queue tile: times lr_render(gs, camera, params) with each
    ``chunk_paths`` in ``chunks`` (paths traced together per wavefront, 4M ..
    64M by default; whole pixels per chunk) and returns {"best", "best_ms",
    "table"} with ``params.chunk_paths`` left at the fastest.  A chunk pays
    every bounce launch's tail once, so larger chunks amortise it; their
    queues take 104 bytes of HBM per path."""
    import torch

    w, h = params.width, params.height
    fx = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    fw = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    table = []
    for c in chunks:
        params.chunk_paths = int(c)
        gs.render(camera, params, fx, fw)  # warm-up: workspace for this chunk
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(repeats):
            gs.render(camera, params, fx, fw)
        b.record()
        b.synchronize()
        table.append((int(c), a.elapsed_time(b) / repeats))
    best, best_ms = min(table, key=lambda x: x[1])
    params.chunk_paths = best
    return {"best": best, "best_ms": best_ms, "table": table}
