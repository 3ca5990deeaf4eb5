"""Batched BSDF queries on the GPU (lr_bsdf_eval / lr_bsdf_sample): LiAR's
``Bsdf::evaluate`` and ``Bsdf::sample`` (src/liar/kernel/bsdf.h:163-164) for
one BSDF whose inputs are looked up at one wavelength, on many directions at
once, evaluated by the same device functions as the renderer's shading
kernels."""
from __future__ import annotations

import ctypes as C

from . import _native as N

LAMBERT, MIRROR, DIELECTRIC, CONDUCTOR, COOK_TORRANCE_GGX, COOK_TORRANCE_BECKMANN = 1, 2, 3, 4, 5, 6


class Bsdf:
    """One BSDF in its local frame (z = shading normal towards the viewer).
    ``ior`` is the dielectric's n_incident / n_transmitted; ``roughness_u``
    / ``roughness_v`` the CookTorrance roughness (alpha = max(r, 1e-3)^2)."""

    def __init__(self, ctx, kind, reflectance=1.0, transmittance=1.0, eta=1.0, kappa=0.0, ior=1.0,
                 roughness_u=0.1, roughness_v=None, fuzz=0.0):
        self.ctx = ctx
        rv = roughness_u if roughness_v is None else roughness_v
        self._d = N.BsdfDesc(int(kind), float(reflectance), float(transmittance), float(eta), float(kappa),
                             float(ior), float(roughness_u), float(rv), float(fuzz))

    def evaluate(self, wi, wo, stream=None):
        """wi, wo: CUDA float32 (n, 3) -> (value, pdf) float32 (n,); 0, 0 where
        the BSDF has no value."""
        import torch

        wi, wo = wi.contiguous(), wo.contiguous()
        n = wo.shape[0]
        value = torch.empty(n, dtype=torch.float32, device=wo.device)
        pdf = torch.empty_like(value)
        s = torch.cuda.current_stream(wo.device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_bsdf_eval(self.ctx.handle, C.byref(self._d), C.c_void_p(wi.data_ptr()),
                                     C.c_void_p(wo.data_ptr()), n, C.c_void_p(value.data_ptr()),
                                     C.c_void_p(pdf.data_ptr()), C.c_void_p(s)), "lr_bsdf_eval")
        return value, pdf

    def sample(self, wi, u, allow=3, stream=None):
        """wi (n, 3), u (n, 3) CUDA float32 -> (wo (n, 3), value, pdf); pdf 0
        where no direction was sampled."""
        import torch

        wi, u = wi.contiguous(), u.contiguous()
        n = u.shape[0]
        wo = torch.empty((n, 3), dtype=torch.float32, device=u.device)
        value = torch.empty(n, dtype=torch.float32, device=u.device)
        pdf = torch.empty_like(value)
        s = torch.cuda.current_stream(u.device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_bsdf_sample(self.ctx.handle, C.byref(self._d), C.c_void_p(wi.data_ptr()),
                                       C.c_void_p(u.data_ptr()), n, int(allow), C.c_void_p(wo.data_ptr()),
                                       C.c_void_p(value.data_ptr()), C.c_void_p(pdf.data_ptr()), C.c_void_p(s)),
                "lr_bsdf_sample")
        return wo, value, pdf
