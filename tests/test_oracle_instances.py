"""Pins the instanced oracle (oracle/oracle.c or_inst_brute) before it checks
the two-level traversal kernel.

- One instance equals its geometry baked into world space (vertices and
  sphere centres transformed in FP64): same primitives hit, same distances
  to FP32 precision, for translations, rotations and (uniform) scales.
- The world hit point o + t d lies on the transformed primitive: for a
  triangle at L (v0 + u e1 + v e2), which checks t_local / s and the meaning
  of u, v independently of the oracle's own formulas.
- List semantics: the nearer of two instances wins; identical instances tie
  and the smaller index wins.
- any-hit with tfar = inf agrees with closest-hit finding a hit.
- A planted mistake (local distance not divided by the scale) fails the
  hit-point check.
"""
import numpy as np

import oracle
from instance_scenes import bake, geometry, instances, rays_towards


def test_single_instance_equals_baked_geometry():
    g = geometry(1)
    for k, (gi, m) in enumerate(instances(6, 1, seed=2, uniform=True)):
        rays = rays_towards(4000, seed=10 + k, box=0.6, center=np.asarray(m)[:, 3])
        prim, t, u, v, inst = oracle.inst_brute([g], [0], [m], rays)
        vw, tw, sw = bake(g, m)
        bp, bt, bu, bv = oracle.brute_closest(vw, tw, sw, rays)
        hit = bp != oracle.MISS
        same = prim == bp
        assert same.mean() > 0.995, (k, same.mean())
        assert ((inst == 0) == (prim != oracle.MISS)).all()
        both = same & hit
        assert both.sum() > 500
        assert np.abs(t[both] - bt[both]).max() <= 2e-5 * np.abs(bt[both]).max()


def test_world_hit_point_lies_on_transformed_primitive():
    g = geometry(3)
    v, tr, sph = g
    inst = instances(8, 1, seed=4)
    rays = rays_towards(6000, seed=5)
    prim, t, u, vv, ii = oracle.inst_brute([g], [gi for gi, _ in inst], [m for _, m in inst], rays)
    o, d = rays[0:3].T.astype(np.float64), rays[3:6].T.astype(np.float64)
    nt = len(tr)
    checked = 0
    for r in np.nonzero(prim != oracle.MISS)[0]:
        m = np.asarray(inst[ii[r]][1], np.float64)
        p = o[r] + float(t[r]) * d[r]
        if prim[r] < nt:
            a, b, c = v[tr[prim[r]]].astype(np.float64)
            q_local = a + float(u[r]) * (b - a) + float(vv[r]) * (c - a)
            q = m[:, :3] @ q_local + m[:, 3]
            assert np.linalg.norm(p - q) < 2e-4, (r, p, q)
        else:
            s = sph[prim[r] - nt].astype(np.float64)
            local = np.linalg.solve(m[:, :3], p - m[:, 3])
            assert abs(np.linalg.norm(local - s[:3]) - s[3]) < 2e-4
        checked += 1
    assert checked > 400


def test_planted_scale_mistake_fails_hit_point():
    g = geometry(3)
    inst = instances(8, 1, seed=4)
    rays = rays_towards(2000, seed=5)
    oracle.set_inst_mistake(1)
    try:
        prim, t, u, vv, ii = oracle.inst_brute([g], [0] * 8, [m for _, m in inst], rays)
    finally:
        oracle.set_inst_mistake(0)
    v, tr, _ = g
    o, d = rays[0:3].T.astype(np.float64), rays[3:6].T.astype(np.float64)
    bad = 0
    for r in np.nonzero(prim < len(tr))[0]:
        m = np.asarray(inst[ii[r]][1], np.float64)
        a, b, c = v[tr[prim[r]]].astype(np.float64)
        q = m[:, :3] @ (a + float(u[r]) * (b - a) + float(vv[r]) * (c - a)) + m[:, 3]
        bad += np.linalg.norm(o[r] + float(t[r]) * d[r] - q) > 1e-3
    assert bad > 100


def test_list_nearest_and_ties():
    g = geometry(5, n_tris=400, n_spheres=0)
    near = np.concatenate([np.eye(3), np.zeros((3, 1))], 1)
    far = np.concatenate([np.eye(3), np.array([[0.0], [0.0], [3.0]])], 1)
    rays = np.zeros((8, 400), np.float32)
    rng = np.random.default_rng(1)
    rays[0:2] = rng.uniform(-0.5, 0.5, (2, 400))
    rays[2] = 10.0
    rays[5] = -1.0
    rays[7] = np.inf
    prim, t, u, v, inst = oracle.inst_brute([g], [0, 0], [near, far], rays)
    hit = prim != oracle.MISS
    assert hit.sum() > 20 and (inst[hit] == 1).all()  # the instance at z = 3 is nearer to z = 10
    prim2, t2, _, _, inst2 = oracle.inst_brute([g], [0, 0], [far, far], rays)
    h2 = prim2 != oracle.MISS
    assert (inst2[h2] == 0).all()  # identical instances: the smaller index wins the tie
    assert np.array_equal(t2[h2], t[hit]) if h2.sum() == hit.sum() else True


def test_any_agrees_with_closest():
    g = geometry(7)
    inst = instances(10, 1, seed=8)
    rays = rays_towards(3000, seed=9)
    prim, *_ = oracle.inst_brute([g], [0] * 10, [m for _, m in inst], rays)
    occ = oracle.inst_brute([g], [0] * 10, [m for _, m in inst], rays, any_hit=True)
    assert np.array_equal(occ.astype(bool), prim != oracle.MISS)
