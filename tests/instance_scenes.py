"""Instanced test scenes shared by the oracle pins and the GPU parity tests."""
import numpy as np

import liar_amd as la
from liar_amd.instancing import random_rotation


def geometry(seed, n_tris=300, n_spheres=20, scale=0.08):
    """A small object around the origin: a triangle soup and a few spheres."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.5, 0.5, (n_tris, 3))
    verts = (np.repeat(c, 3, axis=0) + rng.uniform(-scale, scale, (3 * n_tris, 3))).astype(np.float32)
    tris = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    sph = np.concatenate([rng.uniform(-0.4, 0.4, (n_spheres, 3)), rng.uniform(0.02, 0.08, (n_spheres, 1))], 1)
    return verts, tris, sph.astype(np.float32)


def instances(n, n_geom, seed, spread=4.0, uniform=False):
    """(geometry index, 3x4 local-to-world) with random rotation, scale and translation."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        R = random_rotation(rng)
        S = np.diag(np.full(3, rng.uniform(0.5, 2.0)) if uniform else rng.uniform(0.5, 2.0, 3))
        T = rng.uniform(-spread, spread, 3)
        out.append((k % n_geom, np.concatenate([R @ S, T[:, None]], 1)))
    return out


def rays_towards(n, seed, box=4.5, tfar=None, center=(0.0, 0.0, 0.0)):
    """(8, n) rays from a shell around ``center`` through random points of the
    cube of half-size ``box`` around it."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, np.float64)
    o = rng.normal(size=(n, 3))
    o = c + 3 * box * o / np.linalg.norm(o, axis=1, keepdims=True)
    target = c + rng.uniform(-box, box, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((8, n), np.float32)
    r[0:3] = o.T
    r[3:6] = d.T
    r[6] = 0.0
    r[7] = np.inf if tfar is None else tfar
    return r


def bake(geom, m):
    """Geometry transformed to world space (spheres: uniform scale only)."""
    v, t, s = geom
    m = np.asarray(m, np.float64)
    vw = (v.astype(np.float64) @ m[:, :3].T + m[:, 3]).astype(np.float32)
    k = np.cbrt(abs(np.linalg.det(m[:, :3])))
    cw = s[:, :3].astype(np.float64) @ m[:, :3].T + m[:, 3]
    sw = np.concatenate([cw, s[:, 3:4] * k], 1).astype(np.float32)
    return vw, t, sw
