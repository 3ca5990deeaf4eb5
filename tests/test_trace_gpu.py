"""GPU parity of the QBVH traversal kernels against the CPU oracle.

Bar: primitive ids, t, u, v bit-exact against the oracle (brute force on small
scenes, oracle QBVH traversal at full size; the two agree bit-exactly with
each other, tests/test_oracle.py), any-hit flags bit-exact.  The FP32 error
against the FP64 oracle is reported and bounded.
"""
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle

pytestmark = pytest.mark.gpu
MISS = oracle.MISS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(r):
    import torch

    return torch.from_numpy(np.ascontiguousarray(r)).cuda()


def gpu_closest(b, r):
    import torch

    prim, tuv = b.trace_closest(to_dev(r))
    torch.cuda.synchronize()
    prim = prim.cpu().numpy().view(np.uint32)
    tuv = tuv.cpu().numpy()
    return prim, tuv[0], tuv[1], tuv[2]


def gpu_any(b, r):
    import torch

    occ = b.trace_any(to_dev(r))
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def assert_hits_equal(ref, got, label):
    names = ("prim", "t", "u", "v")
    for name, a, c in zip(names, ref, got):
        bad = np.flatnonzero(a.view(np.uint32) != np.asarray(c).view(np.uint32))
        assert bad.size == 0, f"{label}: {name} differs on {bad.size} rays, first {bad[:5]}: {a[bad[:5]]} vs {np.asarray(c)[bad[:5]]}"


def scene(nt, ns, seed):
    rng = np.random.default_rng(seed)
    v, t = la.random_triangles(nt, seed=seed)
    sph = np.concatenate([rng.random((ns, 3)), rng.uniform(0.002, 0.05, (ns, 1))], 1).astype(np.float32)
    return v, t, (sph if ns else None)


@pytest.mark.parametrize("nt,ns", [(0, 0), (1, 0), (4, 0), (5, 0), (0, 9), (3000, 0), (3000, 400)])
def test_closest_and_any_vs_brute_force(gpu_ctx, nt, ns):
    say(f"[gpu] brute-force parity nt={nt} ns={ns}")
    v, t, s = scene(nt, ns, 100 + nt + ns)
    b = la.Qbvh(gpu_ctx, v, t, s)
    r = la.random_rays(20000, seed=7)
    assert_hits_equal(oracle.brute_closest(v, t, s, r), gpu_closest(b, r), f"closest nt={nt} ns={ns}")
    r2 = la.random_rays(20000, seed=8, tfar_lo=0.01, tfar_hi=2.0)
    np.testing.assert_array_equal(oracle.brute_any(v, t, s, r2), gpu_any(b, r2))


def test_interval_quirk_and_ties_on_gpu(gpu_ctx):
    say("[gpu] interval quirk / ties")
    zA = np.float32(1.0) + np.float32(2.5e-7)
    quad = lambda z: np.array([[0, 0, z], [1, 0, z], [1, 1, z], [0, 1, z]], np.float32)
    verts = np.concatenate([quad(zA), quad(2.0), quad(2.0)])
    tris = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7], [8, 9, 10], [8, 10, 11]], np.uint32)
    sph = np.array([[0.3, 0.2, 3.5, 0.5]], np.float32)
    r = np.zeros((8, 3), np.float32)
    r[0:3] = np.array([[0.3, 0.2, 0.0], [0.3, 0.2, 0.0], [0.7, 0.2, 1.5]]).T
    r[5] = 1.0
    r[6] = [1.0, 0.0, 0.0]
    r[7] = np.inf
    for s in (None, sph):
        b = la.Qbvh(gpu_ctx, verts, tris, s)
        ref = oracle.brute_closest(verts, tris, s, r)
        got = gpu_closest(b, r)
        assert_hits_equal(ref, got, "quirk")
    assert ref[0][0] == 6  # the sphere, the mesh reported a miss
    assert ref[0][1] == 0  # tnear = 0: triangle A
    assert ref[0][2] == 2  # duplicated quads (ids 2 and 4) tie at t = 0.5: smaller id


@pytest.mark.parametrize("schedule", [1, 2])
@pytest.mark.parametrize("block,lds", [(64, 8), (128, 16), (256, 24), (256, 8)])
def test_launch_configs_agree(gpu_ctx, block, lds, schedule):
    """Every (block, LDS stack, schedule) instance, including ones whose LDS
    stack is shorter than the tree's bound (global spill path), gives the
    same hits."""
    say(f"[gpu] launch config block={block} lds={lds} schedule={schedule}")
    v, t, s = scene(200000, 2000, 3)
    b = la.Qbvh(gpu_ctx, v, t, s)
    assert b.info()["max_stack"] > 8
    r = la.random_rays(1 << 16, seed=5)
    nodes, prims = b.export()
    ref = oracle.qbvh_closest(nodes, prims, r)[:4]
    old = gpu_ctx.trace_config()
    try:
        gpu_ctx.set_trace_config(block=block, lds_stack=lds, blocks_per_cu=2, schedule=schedule)
        assert_hits_equal(ref, gpu_closest(b, r), f"block={block} lds={lds} schedule={schedule}")
        r2 = la.random_rays(1 << 16, seed=6, tfar_lo=0.01, tfar_hi=2.0)
        np.testing.assert_array_equal(oracle.qbvh_any(nodes, prims, r2)[0], gpu_any(b, r2))
    finally:
        gpu_ctx.set_trace_config(**old)


@pytest.mark.parametrize("refill,schedule", [(8, 1), (32, 1), (64, 1), (32, 2)])
def test_every_ray_traced_once_across_claim_boundaries(gpu_ctx, refill, schedule):
    """A 512-wave grid (block 64, 2 blocks per CU): ray counts below one ray
    per lane, at the edge of the static share, and with a dynamic remainder
    handed out in shrinking chunks.  Hit buffers start as a sentinel, so a
    ray that no lane traced shows up as a mismatch."""
    import torch

    v, t, s = scene(3000, 100, 12)
    b = la.Qbvh(gpu_ctx, v, t, s)
    old = gpu_ctx.trace_config()
    try:
        gpu_ctx.set_trace_config(block=64, lds_stack=8, blocks_per_cu=2, refill_min=refill, schedule=schedule)
        for n in (1, 65, 32767, 32769, 3 * 32768 + 17, 200003):
            say(f"[gpu] claim boundaries: refill {refill}, {n} rays")
            r = la.random_rays(n, seed=n)
            prim = torch.full((n,), -2, dtype=torch.int32, device="cuda")
            tuv = torch.full((3, n), float("nan"), dtype=torch.float32, device="cuda")
            b.trace_closest(to_dev(r), prim, tuv)
            occ = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
            b.trace_any(to_dev(r), occ)
            torch.cuda.synchronize()
            tuv = tuv.cpu().numpy()
            got = (prim.cpu().numpy().view(np.uint32), tuv[0], tuv[1], tuv[2])
            assert_hits_equal(oracle.brute_closest(v, t, s, r), got, f"n={n}")
            np.testing.assert_array_equal(oracle.brute_any(v, t, s, r), occ.cpu().numpy())
    finally:
        gpu_ctx.set_trace_config(**old)


def test_autotuner_installs_fastest_config_and_keeps_hits(gpu_ctx):
    """lr_autotune_trace times all 9 kernels x 4 blocks_per_cu x 4 refill
    thresholds x 2 schedules, installs the fastest, and traversal under the tuned
    configuration still matches the oracle bit-exactly (closest and any)."""
    say("[gpu] autotuner")
    v, t, s = scene(100000, 1000, 4)
    b = la.Qbvh(gpu_ctx, v, t, s)
    r = la.random_rays(1 << 17, seed=8)
    nodes, prims = b.export()
    old = gpu_ctx.trace_config()
    try:
        res = b.autotune(to_dev(r), repeats=2)
        assert len(res["table"]) == 9 * 4 * 4 * 2
        assert {c["schedule"] for c, _ in res["table"]} == {1, 2}
        assert all(ms > 0 for _, ms in res["table"])
        assert res["best_ms"] == min(ms for _, ms in res["table"])
        assert gpu_ctx.trace_config() == res["best"]
        say(f"[gpu] tuned closest {res['best']} {res['best_ms']:.3f} ms; "
            f"slowest {max(ms for _, ms in res['table']):.3f} ms")
        assert_hits_equal(oracle.qbvh_closest(nodes, prims, r)[:4], gpu_closest(b, r), "tuned")
        r2 = la.random_rays(1 << 17, seed=9, tfar_lo=0.01, tfar_hi=2.0)
        res_any = b.autotune(to_dev(r2), any_hit=True, repeats=2)
        assert gpu_ctx.trace_config() == res_any["best"]
        np.testing.assert_array_equal(oracle.qbvh_any(nodes, prims, r2)[0], gpu_any(b, r2))
    finally:
        gpu_ctx.set_trace_config(**old)


def test_full_size_microbenchmark_parity(gpu_ctx):
    """1M random triangles, 4,194,304 rays: closest and any-hit bit-exact
    against the oracle QBVH traversal."""
    say("[gpu] full-size 1M tris x 4M rays")
    v, t = la.random_triangles(1_000_000, seed=2)
    b = la.Qbvh(gpu_ctx, v, t)
    nodes, prims = b.export()
    ok, msg, _ = oracle.qbvh_check(nodes, prims, v, t)
    assert ok, msg
    n = 1 << 22
    r = la.random_rays(n, seed=3)
    ref = oracle.qbvh_closest(nodes, prims, r)
    got = gpu_closest(b, r)
    assert_hits_equal(ref[:4], got, "C2 closest")
    hit_rate = float((ref[0] != MISS).mean())
    say(f"[gpu] closest hit rate {hit_rate:.4f}, nodes/ray {ref[4] / n:.2f}, prims/ray {ref[5] / n:.2f}")
    r2 = la.random_rays(n, seed=4, tfar_lo=0.01, tfar_hi=2.0)
    occ_ref = oracle.qbvh_any(nodes, prims, r2)[0]
    np.testing.assert_array_equal(occ_ref, gpu_any(b, r2))
    # size-independent property: any-hit with tfar = +inf equals "closest hit found"
    occ_inf = gpu_any(b, r)
    np.testing.assert_array_equal(occ_inf, (got[0] != MISS).astype(np.uint8))


def test_fp32_error_vs_fp64(gpu_ctx):
    """Reports and bounds the FP32 kernel's error against the FP64 oracle."""
    v, t, s = scene(20000, 500, 21)
    b = la.Qbvh(gpu_ctx, v, t, s)
    r = la.random_rays(20000, seed=22)
    p, tt, uu, vv = gpu_closest(b, r)
    fp, ft, fu, fv = oracle.brute_closest_f64(v, t, s, r)
    same = (p == fp) & (p != MISS)
    id_mismatch = int(((p != fp)).sum())
    rel_t = np.abs(tt[same] - ft[same]) / np.abs(ft[same])
    tri = same & (p < 20000)
    err_uv = max(np.abs(uu[tri] - fu[tri]).max(), np.abs(vv[tri] - fv[tri]).max())
    say(f"[gpu] FP32 vs FP64: id mismatches {id_mismatch}/{len(p)}, max rel t err {rel_t.max():.3e}, "
        f"median {np.median(rel_t):.3e}, max |uv| err {err_uv:.3e}")
    assert id_mismatch <= 2
    assert np.median(rel_t) < 1e-6 and rel_t.max() < 1e-3 and err_uv < 1e-3


def test_non_default_stream(gpu_ctx):
    import torch

    v, t, s = scene(5000, 50, 31)
    b = la.Qbvh(gpu_ctx, v, t, s)
    r = la.random_rays(10000, seed=32)
    st = torch.cuda.Stream()
    rd = to_dev(r)
    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        prim, tuv = b.trace_closest(rd)
    st.synchronize()
    ref = oracle.brute_closest(v, t, s, r)
    tuv = tuv.cpu().numpy()
    assert_hits_equal(ref, (prim.cpu().numpy().view(np.uint32), tuv[0], tuv[1], tuv[2]), "stream")
