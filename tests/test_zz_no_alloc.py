"""lr_trace_closest_noalloc: the closest-hit entry point that allocates
nothing.  Its hits must equal lr_trace_closest's bit for bit, on two calls in
a row (the second smaller, so the fixed scratch holds the first call's
contents when the second starts), and it must leave the device's free memory
unchanged on a stream the context has never traced on."""
import ctypes as C

import numpy as np
import pytest

import liar_amd as la
from liar_amd import _native as N

pytestmark = pytest.mark.gpu


def _trace(entry, bvh, rays, stream):
    import torch

    n = rays.shape[1]
    prim = torch.full((n,), -7, dtype=torch.int32, device=rays.device)
    tuv = torch.full((3, n), -1.0, dtype=torch.float32, device=rays.device)
    rs = bvh._rays_struct(rays)
    hs = N.HitsSoa(prim.data_ptr(), tuv.data_ptr(), tuv.data_ptr() + 4 * n, tuv.data_ptr() + 8 * n)
    N.check(getattr(N.lib(), entry)(bvh.handle, C.byref(rs), n, C.byref(hs), C.c_void_p(stream)), entry)
    return prim, tuv


def _scene(ctx, n_tris, seed):
    v, t = la.random_triangles(n_tris, seed=seed)
    rng = np.random.default_rng(seed)
    sph = np.concatenate([rng.random((512, 3)), rng.uniform(0.002, 0.02, (512, 1))], 1).astype(np.float32)
    return la.Qbvh(ctx, v, t, sph)


def test_noalloc_matches_closest_on_two_calls(gpu_ctx):
    import torch

    bvh = _scene(gpu_ctx, 300_000, seed=21)
    info = bvh.info()
    print(f"noalloc: 300k tris + 512 spheres, max_stack {info['max_stack']}", flush=True)
    s = torch.cuda.current_stream().cuda_stream
    for n, seed in ((1 << 21, 22), (100_003, 23)):
        rays = torch.from_numpy(la.random_rays(n, seed=seed)).cuda()
        p0, h0 = _trace("lr_trace_closest", bvh, rays, s)
        p1, h1 = _trace("lr_trace_closest_noalloc", bvh, rays, s)
        torch.cuda.synchronize()
        assert int((p0 >= 0).sum()) > n // 10  # LR_MISS is -1 as int32
        assert torch.equal(p0, p1), f"prim ids differ at n={n}"
        assert torch.equal(h0.view(torch.int32), h1.view(torch.int32)), f"t/u/v differ at n={n}"
    bvh.close()


def test_noalloc_allocates_nothing_on_a_new_stream(gpu_ctx):
    import torch

    bvh = _scene(gpu_ctx, 200_000, seed=31)
    rays = torch.from_numpy(la.random_rays(1 << 20, seed=32)).cuda()
    n = rays.shape[1]
    prim = torch.empty(n, dtype=torch.int32, device="cuda")
    tuv = torch.empty((3, n), dtype=torch.float32, device="cuda")
    rs = bvh._rays_struct(rays)
    hs = N.HitsSoa(prim.data_ptr(), tuv.data_ptr(), tuv.data_ptr() + 4 * n, tuv.data_ptr() + 8 * n)
    side = torch.cuda.Stream()
    # the runtime sets up a stream's own queue memory on its first launch:
    # launch something on it before measuring
    with torch.cuda.stream(side):
        tuv.zero_()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    N.check(N.lib().lr_trace_closest_noalloc(bvh.handle, C.byref(rs), n, C.byref(hs), C.c_void_p(side.cuda_stream)),
            "lr_trace_closest_noalloc")
    side.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < (1 << 20), f"device free memory fell by {free0 - free1} bytes"
    # the allocating entry point on that stream grows the context's scratch
    # for it (printed for comparison; its size depends on the occupancy)
    ref = torch.empty_like(prim)
    ref_tuv = torch.empty_like(tuv)
    bvh.trace_closest(rays, ref, ref_tuv, stream=side.cuda_stream)
    side.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    print(f"free memory change: noalloc {free1 - free0} B, lr_trace_closest {free2 - free1} B", flush=True)
    assert torch.equal(ref, prim)
    bvh.close()
