"""GPU parity of the two-level (instanced) traversal (lr_tlas_*,
lr_trace_*_instanced) with the instanced oracle (pinned in
tests/test_oracle_instances.py).

Both sides compute the local ray in FP32 in the same operation order and
test primitives with the same FP32 arithmetic, so primitive ids, instance
ids, distances and barycentrics must agree bit for bit."""
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle
from instance_scenes import geometry, instances, rays_towards
from liar_amd.instancing import Tlas

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(ctx, geoms, inst):
    qs = [la.Qbvh(ctx, v, t, s) for v, t, s in geoms]
    return qs, Tlas(ctx, qs, inst)


def gpu_closest(tl, rays):
    import torch

    r = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    prim, tuv, inst = tl.trace_closest(r)
    torch.cuda.synchronize()
    return (prim.cpu().numpy().view(np.uint32), tuv.cpu().numpy(), inst.cpu().numpy().view(np.uint32))


def gpu_any(tl, rays):
    import torch

    r = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    occ = tl.trace_any(r)
    torch.cuda.synchronize()
    return occ.cpu().numpy()


@pytest.mark.parametrize("n_inst,uniform", [(1, True), (37, False), (300, False)])
def test_closest_and_any_bit_exact(gpu_ctx, n_inst, uniform):
    geoms = [geometry(11), geometry(12, n_tris=150, n_spheres=40), geometry(13, n_tris=600, n_spheres=0)]
    inst = instances(n_inst, len(geoms), seed=n_inst, uniform=uniform)
    qs, tl = build(gpu_ctx, geoms, inst)
    info = tl.info()
    say(f"[gpu] {n_inst} instances: tlas {info['n_nodes']} nodes depth {info['max_depth']} stack {info['max_stack']}")
    center = np.mean([np.asarray(m)[:, 3] for _, m in inst], 0)
    rays = rays_towards(20000, seed=3, box=0.6 if n_inst == 1 else 4.5, center=center)
    gp, gtuv, gi = gpu_closest(tl, rays)
    op, ot, ou, ov, oi = oracle.inst_brute(geoms, [g for g, _ in inst], [m for _, m in inst], rays)
    mism = int((gp != op).sum() + (gi != oi).sum())
    mt = int((gtuv[0].view(np.uint32) != ot.view(np.uint32)).sum())
    muv = int((gtuv[1].view(np.uint32) != ou.view(np.uint32)).sum() + (gtuv[2].view(np.uint32) != ov.view(np.uint32)).sum())
    hits = int((op != oracle.MISS).sum())
    say(f"[gpu] closest: {hits} hits / 20000; mismatches prim/inst {mism}, t {mt}, uv {muv}")
    assert hits > 1000
    assert mism == 0 and mt == 0 and muv == 0
    rays_any = rays.copy()
    rays_any[7] = np.random.default_rng(4).uniform(0.5, 20.0, rays.shape[1]).astype(np.float32)
    ga = gpu_any(tl, rays_any)
    oa = oracle.inst_brute(geoms, [g for g, _ in inst], [m for _, m in inst], rays_any, any_hit=True)
    say(f"[gpu] any: {int(oa.sum())} occluded; mismatches {int((ga != oa).sum())}")
    assert np.array_equal(ga, oa)
    for q in qs:
        q.close()
    tl.close()


def test_shared_geometry_many_instances(gpu_ctx):
    """A 20k-triangle object placed 2000 times (40M instanced triangles):
    any-hit with tfar = inf agrees with closest-hit finding a hit, and hit
    points land on the transformed triangles (FP64 check on a sample)."""
    import torch

    g = geometry(21, n_tris=20000, n_spheres=0, scale=0.01)
    inst = instances(2000, 1, seed=22, spread=30.0)
    q = la.Qbvh(gpu_ctx, *g)
    tl = Tlas(gpu_ctx, [q], inst)
    rays = rays_towards(1 << 20, seed=23, box=30.0)
    gp, gtuv, gi = gpu_closest(tl, rays)
    ga = gpu_any(tl, rays)
    hit = gp != oracle.MISS
    say(f"[gpu] 2000 instances x 20k tris, 1M rays: hit rate {hit.mean():.3f}")
    assert hit.mean() > 0.05
    assert np.array_equal(ga.astype(bool), hit)
    v, tr, _ = g
    o, d = rays[0:3].T.astype(np.float64), rays[3:6].T.astype(np.float64)
    idx = np.nonzero(hit)[0][:2000]
    err = 0.0
    for r in idx:
        m = np.asarray(inst[gi[r]][1], np.float64)
        a, b, c = v[tr[gp[r]]].astype(np.float64)
        qp = m[:, :3] @ (a + float(gtuv[1, r]) * (b - a) + float(gtuv[2, r]) * (c - a)) + m[:, 3]
        err = max(err, float(np.linalg.norm(o[r] + float(gtuv[0, r]) * d[r] - qp)))
    say(f"[gpu] max hit-point distance {err:.2e}")
    assert err < 1e-3
    # timing
    r = torch.from_numpy(rays).cuda()
    prim = torch.empty(r.shape[1], dtype=torch.int32, device="cuda")
    tuv = torch.empty((3, r.shape[1]), dtype=torch.float32, device="cuda")
    ii = torch.empty(r.shape[1], dtype=torch.int32, device="cuda")
    tl.trace_closest(r, prim, tuv, ii)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        tl.trace_closest(r, prim, tuv, ii)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    say(f"[gpu] instanced closest-hit: {ms:.3f} ms for 1M rays ({r.shape[1] / ms / 1e3:.0f} Mrays/s)")
    tl.close()
    q.close()


def test_empty_and_invalid(gpu_ctx):
    import torch

    g = geometry(31)
    q = la.Qbvh(gpu_ctx, *g)
    empty = la.Qbvh(gpu_ctx, np.zeros((0, 3)), np.zeros((0, 3)))
    rays = rays_towards(1000, seed=1)
    tl = Tlas(gpu_ctx, [q, empty], [])
    gp, _, gi = gpu_closest(tl, rays)
    assert (gp == oracle.MISS).all() and (gi == oracle.MISS).all()
    assert not gpu_any(tl, rays).any()
    tl.close()
    eye = np.concatenate([np.eye(3), np.zeros((3, 1))], 1)
    tl = Tlas(gpu_ctx, [q, empty], [(1, eye), (0, eye), (1, eye)])
    assert tl.info()["n_instances"] == 1
    gp, _, gi = gpu_closest(tl, rays_towards(1000, seed=2, box=0.6))
    assert set(np.unique(gi[gp != oracle.MISS]).tolist()) == {1}  # the caller's index of the one instance
    tl.close()
    with pytest.raises(la.LiarError, match="singular"):
        Tlas(gpu_ctx, [q], [(0, np.zeros((3, 4)))])
    with pytest.raises(la.LiarError, match="geometry"):
        Tlas(gpu_ctx, [q], [(5, eye)])
    q.close()
    empty.close()
    del torch
