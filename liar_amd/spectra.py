"""Spectra (mirrors ``liar.spectra``: src/liar/spectra/).

Wavelengths are accepted in metres (LiAR's unit) or nanometres (any value > 1,
the same rule as LiAR's observer loader, src/liar/__init__.py:51-63) and kept
in nanometres.
"""
from __future__ import annotations

import numpy as np

from . import _native as N


def _to_nm(ws):
    ws = np.asarray(ws, dtype=np.float64)
    return ws if (ws > 1).any() else ws * 1e9


class Spectrum:
    """Base class; subclasses describe themselves as an lr_spectrum."""

    def _describe(self, keep: list) -> N.Spectrum:
        raise NotImplementedError

    def evaluate(self, ctx, wavelengths_nm, reflectant=False):
        """The values the GPU renderer uses at these wavelengths [nm]
        (lr_spectrum_eval on ``ctx``'s device)."""
        return _evaluate(self, ctx, wavelengths_nm, reflectant)


class Sampled(Spectrum):
    """Piecewise-linear tabulated spectrum, zero outside the table
    (src/liar/spectra/sampled.cpp:102-131, src/liar/kernel/spectral.cpp:185-210)."""

    def __init__(self, wavelengths, values):
        w = _to_nm(wavelengths).astype(np.float32)
        v = np.asarray(values, dtype=np.float32)
        if w.ndim != 1 or w.shape != v.shape or len(w) < 2:
            raise ValueError("Sampled needs matching 1-D wavelength/value tables of length >= 2")
        if not (np.diff(w) > 0).all():
            raise ValueError("Sampled wavelengths must be strictly increasing")
        self.wavelengths = w
        self.values = v

    def __call__(self, wavelength):
        """``Sampled.__call__`` (sampled.cpp:102-123): linear interpolation,
        the last value at the last wavelength, 0 outside the table.  (The
        renderer's per-sample evaluation, Spectral::fromSampled, is 0 at the
        last wavelength too: ``lr_spectrum_eval``.)"""
        w = float(_to_nm([wavelength])[0])
        ws, vs = self.wavelengths.astype(np.float64), self.values.astype(np.float64)
        if w < ws[0] or w > ws[-1]:
            return 0.0
        k = int(np.searchsorted(ws, w, side="right"))  # upper_bound
        if k >= len(ws):
            return float(vs[-1])
        t = (w - ws[k - 1]) / (ws[k] - ws[k - 1])
        return float(vs[k - 1] + t * (vs[k] - vs[k - 1]))

    def resample(self, wavelengths):
        """``Sampled.resample(wavelengths)`` (sampled.cpp:63-93): a Sampled
        on the given ascending wavelengths, each value interpolated in the
        table interval [w_k-1, w_k) that holds it, 0 below the first and from
        the last wavelength on.  Wavelengths in metres or nanometres (kept
        in nanometres, like every Sampled)."""
        ws = _to_nm(np.asarray(wavelengths, np.float64))
        tw, tv = self.wavelengths.astype(np.float64), self.values.astype(np.float64)
        out = np.zeros(len(ws), np.float64)
        i, n = 0, len(ws)
        while i < n and ws[i] < tw[0]:
            i += 1
        for k in range(1, len(tw)):
            while i < n and ws[i] < tw[k]:
                out[i] = tv[k - 1] + (ws[i] - tw[k - 1]) / (tw[k] - tw[k - 1]) * (tv[k] - tv[k - 1])
                i += 1
        return Sampled(ws, out)

    def _describe(self, keep):
        keep += [self.wavelengths, self.values]
        return N.Spectrum(N.SPEC_SAMPLED, len(self.values), self.wavelengths.ctypes.data,
                          self.values.ctypes.data, 0.0, 0.0, 0.0, 0.0, 0.0)


def _evaluate(sp, ctx, wavelengths_nm, reflectant):
    import ctypes as C

    import torch

    w = torch.as_tensor(np.asarray(wavelengths_nm, np.float32), device=f"cuda:{ctx.device}")
    out = torch.empty_like(w)
    keep = []
    desc = sp._describe(keep)
    s = torch.cuda.current_stream(w.device).cuda_stream
    N.check(N.lib().lr_spectrum_eval(ctx.handle, C.byref(desc), C.c_void_p(w.data_ptr()), len(w),
                                     int(bool(reflectant)), C.c_void_p(out.data_ptr()), C.c_void_p(s)),
            "lr_spectrum_eval")
    return out.cpu().numpy()


class Binned(Spectrum):
    """Piecewise-constant spectrum: ``len(values)`` equal bins over [lo, hi)
    nanometres, zero outside."""

    def __init__(self, lo, hi, values):
        self.lo, self.hi = float(lo), float(hi)
        self.values = np.asarray(values, dtype=np.float32)
        if self.values.ndim != 1 or len(self.values) < 1 or not self.hi > self.lo:
            raise ValueError("Binned needs lo < hi and at least one value")

    def _describe(self, keep):
        keep.append(self.values)
        return N.Spectrum(N.SPEC_BINNED, len(self.values), None, self.values.ctypes.data, 0.0, self.lo,
                          self.hi, 0.0, 0.0)


class BlackBody(Spectrum):
    """Planck emitter normalised to ``scale`` at its peak wavelength
    (src/liar/spectra/black_body.cpp:47-118)."""

    def __init__(self, temperature, scale=1.0):
        self.temperature = float(temperature)
        self.scale = float(scale)
        if not self.temperature > 0:
            raise ValueError("temperature must be > 0")

    @property
    def peakWavelength(self):
        return 2.897771955e-3 / self.temperature

    def _describe(self, keep):
        return N.Spectrum(N.SPEC_BLACKBODY, 0, None, None, 0.0, 0.0, 0.0, self.temperature, self.scale)


class Constant(Spectrum):
    def __init__(self, value):
        self.value = float(value)

    def _describe(self, keep):
        return N.Spectrum(N.SPEC_CONSTANT, 0, None, None, self.value, 0.0, 0.0, 0.0, 0.0)
