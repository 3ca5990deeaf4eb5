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
        w = float(_to_nm([wavelength])[0])
        if w < self.wavelengths[0] or w >= self.wavelengths[-1]:
            return 0.0
        return float(np.interp(w, self.wavelengths, self.values))

    def _describe(self, keep):
        keep += [self.wavelengths, self.values]
        return N.Spectrum(N.SPEC_SAMPLED, len(self.values), self.wavelengths.ctypes.data,
                          self.values.ctypes.data, 0.0, 0.0, 0.0, 0.0, 0.0)


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
