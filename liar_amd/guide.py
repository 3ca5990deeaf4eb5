"""Path-guiding network handle (C-ABI lr_guide_*): inference and Adam
training on the GPU.  Tensors are torch CUDA tensors (device memory)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class Guide:
    def __init__(self, ctx, seed: int = 1, learning_rate: float = 1e-3, radiance_weight: float = 0.0):
        h = C.c_void_p()
        N.check(N.lib().lr_guide_create(ctx.handle, int(seed), float(learning_rate), C.byref(h)), "lr_guide_create")
        self._h = h
        self.ctx = ctx
        self.n_params = N.lib().lr_guide_param_count()
        if radiance_weight:
            self.set_radiance_weight(radiance_weight)

    def set_radiance_weight(self, weight: float) -> None:
        """Weight of the reflected-radiance loss in training (0: output 42,
        the radiance estimate that radiance-driven roulette reads, is not
        trained)."""
        N.check(N.lib().lr_guide_set_radiance_weight(self._h, float(weight)), "lr_guide_set_radiance_weight")

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.lib().lr_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        N.check(N.lib().lr_guide_get_params(self._h, N.ptr(p)), "lr_guide_get_params")
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        N.check(N.lib().lr_guide_set_params(self._h, N.ptr(p)), "lr_guide_set_params")

    @staticmethod
    def _stream(t, stream):
        import torch

        return torch.cuda.current_stream(t.device).cuda_stream if stream is None else stream

    def infer(self, inputs, mixture=None, stream=None):
        """inputs: (n, 8) float32 CUDA -> (n, 48): 8 lobes x (weight, mean
        xyz, kappa), then the strategy probabilities (light, guide, BSDF)
        and the reflected-radiance estimate."""
        import torch

        n = inputs.shape[0]
        assert inputs.shape[1] == 8 and inputs.is_contiguous()
        if mixture is None:
            mixture = torch.empty((n, 48), dtype=torch.float32, device=inputs.device)
        N.check(N.lib().lr_guide_infer(self._h, C.c_void_p(inputs.data_ptr()), n, None,
                                       C.c_void_p(mixture.data_ptr()), C.c_void_p(self._stream(inputs, stream))),
                "lr_guide_infer")
        return mixture

    def train_step(self, inputs, dirs, weights, loss=None, grad=None, stream=None):
        n = inputs.shape[0]
        assert inputs.shape == (n, 8) and dirs.shape == (n, 3) and weights.shape == (n,)
        N.check(N.lib().lr_guide_train_step(
            self._h, C.c_void_p(inputs.data_ptr()), C.c_void_p(dirs.data_ptr()), C.c_void_p(weights.data_ptr()), n,
            C.c_void_p(loss.data_ptr()) if loss is not None else None,
            C.c_void_p(grad.data_ptr()) if grad is not None else None,
            C.c_void_p(self._stream(inputs, stream))), "lr_guide_train_step")
