"""Two-level ray queries: a top-level tree over instances of shared
geometry (mirrors LiAR's ``scenery.Transformation`` of an object inside a
``List`` / ``QbvhTree``, src/liar/scenery/transformation.cpp:117-131).

Each instance places one geometry (a :class:`liar_amd.Qbvh`, built once)
with a 3x4 (or 4x4 affine) local-to-world matrix.  Traces return the
primitive id within the instance's geometry, the world distance and the
instance index.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .qbvh import Context, Qbvh


def _affine(m) -> np.ndarray:
    m = np.asarray(m, np.float64)
    if m.shape == (4, 4):
        if not np.allclose(m[3], [0, 0, 0, 1]):
            raise ValueError("only affine transforms are supported (last row 0 0 0 1)")
        m = m[:3]
    if m.shape != (3, 4):
        raise ValueError(f"transform must be 3x4 or 4x4, got {m.shape}")
    return m


class Tlas:
    def __init__(self, ctx: Context, geometries, instances):
        """geometries: list of Qbvh; instances: list of (geometry index, matrix)."""
        self.ctx = ctx
        self.geometries = list(geometries)
        gs = (C.c_void_p * max(1, len(self.geometries)))(*[g.handle.value for g in self.geometries])
        arr = (N.Instance * max(1, len(instances)))()
        for k, (g, m) in enumerate(instances):
            arr[k].geometry = int(g)
            arr[k].local_to_world[:] = [float(x) for x in _affine(m).ravel()]
        h = C.c_void_p()
        N.check(N.lib().lr_tlas_create(ctx.handle, gs, len(self.geometries), arr, len(instances), C.byref(h)),
                "lr_tlas_create")
        self._h = h
        self.n_instances = len(instances)

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise N.LiarError("tlas destroyed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            N.lib().lr_tlas_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = N.TlasInfo()
        N.check(N.lib().lr_tlas_get_info(self.handle, C.byref(i)), "lr_tlas_get_info")
        return {"n_instances": i.n_instances, "n_nodes": i.n_nodes, "max_depth": i.max_depth,
                "max_stack": i.max_stack, "n_geometry_nodes": i.n_geometry_nodes,
                "n_geometry_prims": i.n_geometry_prims, "build_seconds": i.build_seconds,
                "scene_lo": list(i.scene_lo), "scene_hi": list(i.scene_hi)}

    def trace_closest(self, rays, prim=None, tuv=None, inst=None, stream=None):
        """(8, n) CUDA rays -> (prim int32 (n,), tuv float32 (3, n), instance int32 (n,))."""
        import torch

        n = rays.shape[1]
        prim = torch.empty(n, dtype=torch.int32, device=rays.device) if prim is None else prim
        tuv = torch.empty((3, n), dtype=torch.float32, device=rays.device) if tuv is None else tuv
        inst = torch.empty(n, dtype=torch.int32, device=rays.device) if inst is None else inst
        rs = Qbvh._rays_struct(rays)
        hs = N.HitsSoa(prim.data_ptr(), tuv.data_ptr(), tuv.data_ptr() + 4 * n, tuv.data_ptr() + 8 * n)
        s = torch.cuda.current_stream(rays.device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_trace_closest_instanced(self.handle, C.byref(rs), n, C.byref(hs),
                                                   C.c_void_p(inst.data_ptr()), C.c_void_p(s)),
                "lr_trace_closest_instanced")
        return prim, tuv, inst

    def trace_any(self, rays, occluded=None, stream=None):
        import torch

        n = rays.shape[1]
        occluded = torch.empty(n, dtype=torch.uint8, device=rays.device) if occluded is None else occluded
        rs = Qbvh._rays_struct(rays)
        s = torch.cuda.current_stream(rays.device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_trace_any_instanced(self.handle, C.byref(rs), n, C.c_void_p(occluded.data_ptr()),
                                               C.c_void_p(s)), "lr_trace_any_instanced")
        return occluded


def random_rotation(rng) -> np.ndarray:
    """Uniform random rotation (QR of a Gaussian matrix, sign-fixed)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q
