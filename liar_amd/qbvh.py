"""Host-side handles for the ray-query path: device context, QBVH, ray batches.

Mirrors the batched form of LiAR's ``SceneObject::intersect`` /
``isIntersecting`` (src/liar/kernel/scene_object.h:110-113): a whole batch of
rays goes to the GPU in one call instead of one virtual call per ray.

Ray batches live in HBM as one float32 tensor of shape (8, n) whose rows are
(ox, oy, oz, dx, dy, dz, tnear, tfar); hits as an int32 tensor (n,) of
primitive ids (LR_MISS = -1 as int32) plus a float32 tensor (3, n) of
(t, u, v).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

RAY_ROWS = ("ox", "oy", "oz", "dx", "dy", "dz", "tnear", "tfar")


class Context:
    """One HIP device (LiAR has one process-global render set-up instead)."""

    def __init__(self, device: int = 0):
        L = N.lib()
        h = C.c_void_p()
        N.check(L.lr_context_create(int(device), C.byref(h)), "lr_context_create")
        self._h = h
        self.device = int(device)

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise N.LiarError("context destroyed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            N.lib().lr_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def trace_config(self) -> dict:
        cfg = N.TraceConfig()
        N.check(N.lib().lr_get_trace_config(self.handle, C.byref(cfg)), "lr_get_trace_config")
        return {"block": cfg.block, "blocks_per_cu": cfg.blocks_per_cu, "lds_stack": cfg.lds_stack,
                "refill_min": cfg.refill_min, "schedule": cfg.schedule}

    def set_trace_config(self, block: int = 0, blocks_per_cu: int = 0, lds_stack: int = 0,
                         refill_min: int = 0, schedule: int = 0) -> None:
        """0 keeps a field; schedule 1 = one node or leaf per lane and
        iteration, 2 = phased node / leaf iterations (LR_TRACE_SCHEDULE_*)."""
        cfg = N.TraceConfig(block, blocks_per_cu, lds_stack, refill_min, schedule)
        N.check(N.lib().lr_set_trace_config(self.handle, C.byref(cfg)), "lr_set_trace_config")


def _f32(a, cols):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if a.size == 0:
        return a.reshape(0, cols)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected an array of shape (n, {cols}), got {a.shape}")
    return a


class Qbvh:
    """SAH 4-wide QBVH over triangles and spheres (one object-level tree).

    ``ctx=None`` builds host-only (for inspection; cannot be traced).
    """

    def __init__(self, ctx: Context | None, verts, tris, spheres=None):
        verts = _f32(verts, 3)
        tris = np.ascontiguousarray(np.asarray(tris, dtype=np.uint32))
        tris = tris.reshape(0, 3) if tris.size == 0 else tris
        if tris.ndim != 2 or tris.shape[1] != 3:
            raise ValueError(f"tris must have shape (n, 3), got {tris.shape}")
        spheres = _f32(np.zeros((0, 4)) if spheres is None else spheres, 4)
        h = C.c_void_p()
        N.check(
            N.lib().lr_qbvh_create(ctx.handle if ctx is not None else None, N.ptr(verts), verts.shape[0],
                                   N.ptr(tris), tris.shape[0], N.ptr(spheres) if len(spheres) else None,
                                   spheres.shape[0], C.byref(h)),
            "lr_qbvh_create")
        self._h = h
        self.ctx = ctx
        self.n_triangles = int(tris.shape[0])
        self.n_spheres = int(spheres.shape[0])

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise N.LiarError("qbvh destroyed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            N.lib().lr_qbvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = N.QbvhInfo()
        N.check(N.lib().lr_qbvh_get_info(self.handle, C.byref(i)), "lr_qbvh_get_info")
        return {
            "n_nodes": i.n_nodes, "n_prims": i.n_prims, "n_triangles": i.n_triangles,
            "n_spheres": i.n_spheres, "max_depth": i.max_depth, "max_stack": i.max_stack,
            "sah_cost": i.sah_cost, "build_seconds": i.build_seconds,
            "scene_lo": list(i.scene_lo), "scene_hi": list(i.scene_hi),
        }

    def export(self):
        """(nodes uint8 (n_nodes, 128), prims uint8 (n_prims, 48)) host copies."""
        inf = self.info()
        nodes = np.zeros((inf["n_nodes"], N.NODE_BYTES), dtype=np.uint8)
        prims = np.zeros((inf["n_prims"], N.PRIM_BYTES), dtype=np.uint8)
        N.check(N.lib().lr_qbvh_export(self.handle, N.ptr(nodes), N.ptr(prims)), "lr_qbvh_export")
        return nodes, prims

    def export_q8(self):
        """uint8 (n_nodes, 64): the quantised nodes the GPU traverses (lr_node_q8)."""
        q = np.zeros((self.info()["n_nodes"], N.NODE_Q8_BYTES), dtype=np.uint8)
        N.check(N.lib().lr_qbvh_export_q8(self.handle, N.ptr(q)), "lr_qbvh_export_q8")
        return q

    # -- device traces (torch CUDA tensors) ---------------------------------
    @staticmethod
    def _rays_struct(rays) -> N.RaysSoa:
        if rays.dim() != 2 or rays.shape[0] != 8 or not rays.is_contiguous():
            raise ValueError("rays must be a contiguous (8, n) float32 tensor")
        if str(rays.dtype) != "torch.float32":
            raise ValueError("rays must be float32")
        n = rays.shape[1]
        base = rays.data_ptr()
        return N.RaysSoa(*[base + 4 * n * k for k in range(8)])

    def trace_closest(self, rays, prim=None, tuv=None, stream=None):
        """Closest hit of an (8, n) CUDA ray tensor -> (prim int32 (n,), tuv float32 (3, n))."""
        import torch

        n = rays.shape[1]
        if prim is None:
            prim = torch.empty(n, dtype=torch.int32, device=rays.device)
        if tuv is None:
            tuv = torch.empty((3, n), dtype=torch.float32, device=rays.device)
        if prim.numel() != n or tuv.shape != (3, n) or not tuv.is_contiguous():
            raise ValueError("hit buffers must be (n,) int32 and contiguous (3, n) float32")
        rs = self._rays_struct(rays)
        hs = N.HitsSoa(prim.data_ptr(), tuv.data_ptr(), tuv.data_ptr() + 4 * n, tuv.data_ptr() + 8 * n)
        s = torch.cuda.current_stream(rays.device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_trace_closest(self.handle, C.byref(rs), n, C.byref(hs), C.c_void_p(s)),
                "lr_trace_closest")
        return prim, tuv

    def trace_any(self, rays, occluded=None, stream=None):
        """Any hit of an (8, n) CUDA ray tensor -> uint8 (n,)."""
        import torch

        n = rays.shape[1]
        if occluded is None:
            occluded = torch.empty(n, dtype=torch.uint8, device=rays.device)
        rs = self._rays_struct(rays)
        s = torch.cuda.current_stream(rays.device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_trace_any(self.handle, C.byref(rs), n, C.c_void_p(occluded.data_ptr()),
                                     C.c_void_p(s)), "lr_trace_any")
        return occluded

    def autotune(self, rays, any_hit=False, repeats=3, stream=None) -> dict:
        """lr_autotune_trace on an (8, n) CUDA ray tensor: times every traversal
        configuration on these rays, installs the fastest in the context and
        returns {"best": config, "best_ms": ms, "table": [(config, ms), ...]}."""
        import torch

        n = rays.shape[1]
        rs = self._rays_struct(rays)
        s = torch.cuda.current_stream(rays.device).cuda_stream if stream is None else stream
        cap = 1024
        entries = (N.TuneEntry * cap)()
        count = C.c_uint32(0)
        best = N.TraceConfig()
        if any_hit:
            occ = torch.empty(n, dtype=torch.uint8, device=rays.device)
            hs_p, occ_p = None, C.c_void_p(occ.data_ptr())
        else:
            prim = torch.empty(n, dtype=torch.int32, device=rays.device)
            tuv = torch.empty((3, n), dtype=torch.float32, device=rays.device)
            hs = N.HitsSoa(prim.data_ptr(), tuv.data_ptr(), tuv.data_ptr() + 4 * n, tuv.data_ptr() + 8 * n)
            hs_p, occ_p = C.byref(hs), None
        N.check(N.lib().lr_autotune_trace(self.handle, C.byref(rs), n, hs_p, occ_p, int(repeats), entries, cap,
                                          C.byref(count), C.byref(best), C.c_void_p(s)), "lr_autotune_trace")

        def cfg(c):
            return {"block": c.block, "blocks_per_cu": c.blocks_per_cu, "lds_stack": c.lds_stack,
                    "refill_min": c.refill_min, "schedule": c.schedule}

        table = [(cfg(entries[k].cfg), float(entries[k].ms)) for k in range(count.value)]
        return {"best": cfg(best), "best_ms": min(ms for _, ms in table), "table": table}


# -- deterministic synthetic inputs -------------------------------------------

def random_triangles(n: int, seed: int = 2, scale: float = 0.01):
    """Triangle soup: centre ~ U[0,1)^3, vertices centre + scale * U[-1,1)^3."""
    verts = np.empty((3 * n, 3), dtype=np.float32)
    tris = np.empty((n, 3), dtype=np.uint32)
    N.check(N.lib().lr_gen_random_triangles(seed, n, scale, N.ptr(verts), N.ptr(tris)),
            "lr_gen_random_triangles")
    return verts, tris


def random_rays(n: int, seed: int = 3, tfar_lo: float = 0.0, tfar_hi: float = 0.0) -> np.ndarray:
    """(8, n) float32: origins ~ U[0,1)^3, directions uniform on S^2,
    tnear = 0, tfar = inf (or ~ U[tfar_lo, tfar_hi) when tfar_hi > 0)."""
    r = np.empty((8, n), dtype=np.float32)
    N.check(N.lib().lr_gen_random_rays(seed, n, tfar_lo, tfar_hi, *[N.ptr(r[k]) for k in range(8)]),
            "lr_gen_random_rays")
    return r


def philox4x32(key: int, c0: int, c1: int = 0, c2: int = 0, c3: int = 0) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    N.lib().lr_philox4x32(key, c0, c1, c2, c3, N.ptr(out))
    return out
