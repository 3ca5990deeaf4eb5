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


class RecoveryMengSimon:
    """``spectra.RecoveryMengSimon(wavelengths, samples)``
    (src/liar/spectra/recovery_meng_simon.cpp:83-320): spectral recovery from
    basis spectra at chromaticities (Meng, Simon, Hanika, Dachsbacher 2015).
    The chromaticity (x, y) of a tristimulus value is located in the Delaunay
    triangulation of the basis points and the corners of the box
    [-0.3, 1] x [-0.3, 1.2]; the spectrum is the barycentric blend of the
    triangle's basis spectra (a corner contributes nothing and the other
    weights are renormalised), scaled by X + Y + Z.  ``samples``: a mapping
    (x, y) -> values on ``wavelengths`` (metres or nanometres)."""

    BOX = ((-0.3, -0.3), (1.0, 1.2))

    def __init__(self, wavelengths, samples):
        from scipy.spatial import Delaunay

        self.wavelengths = _to_nm(wavelengths)
        n = len(self.wavelengths)
        if n < 2:
            raise ValueError("Requires at least two wavelengths.")
        if not (np.diff(self.wavelengths) > 0).all():
            raise ValueError("Requires strict increasing wavelengths.")
        (x0, y0), (x1, y1) = self.BOX
        xy, values = [], []
        for p, v in dict(samples).items():
            p = (float(p[0]), float(p[1]))
            if not (x0 <= p[0] <= x1 and y0 <= p[1] <= y1):
                raise ValueError("chromaticities must be in box [0, 1]x[0, 1].")
            v = np.asarray(v, np.float64)
            if v.shape != (n,):
                raise ValueError("Requires as many values as wavelengths.")
            xy.append(p)
            values.append(v)
        self.points = np.array([(x0, y0), (x1, y0), (x1, y1), (x0, y1)] + xy, np.float64)
        # rows 0..3: the box corners, which carry no spectrum
        self.values = np.vstack([np.zeros((4, n))] + [v[None] for v in values])
        self.maxima = self.values.max(axis=1)
        self.mesh = Delaunay(self.points)

    def weights(self, xyz):
        """(three point rows, three weights already scaled by X + Y + Z), or
        None when the colour is black or its chromaticity is outside the box."""
        X, Y, Z = (float(c) for c in xyz)
        s = X + Y + Z
        if s == 0:
            return None
        p = np.array([X / s, Y / s])
        (x0, y0), (x1, y1) = self.BOX
        if not (x0 <= p[0] <= x1 and y0 <= p[1] <= y1):
            return None
        t = int(self.mesh.find_simplex(p))
        if t < 0:
            return None
        rows = self.mesh.simplices[t]
        T = self.mesh.transform[t]
        b = T[:2] @ (p - T[2])
        w = np.array([b[0], b[1], 1.0 - b[0] - b[1]])
        w[rows < 4] = 0.0
        tot = w.sum()
        if tot == 0:
            return None
        return rows, w * (s / tot)

    def recover(self, xyz):
        """The recovered spectrum's values on ``wavelengths``."""
        r = self.weights(xyz)
        if r is None:
            return np.zeros(len(self.wavelengths))
        rows, w = r
        return w @ self.values[rows]

    def recover_at(self, xyz, wavelength):
        """Its value at one wavelength (metres or nanometres): linear in the
        table interval that holds it, 0 before the first wavelength and from
        the last on (recovery_meng_simon.cpp:158-210)."""
        w = float(_to_nm([wavelength])[0])
        ws = self.wavelengths
        k = int(np.searchsorted(ws, w, side="right"))
        if k == 0 or k >= len(ws):
            return 0.0
        v = self.recover(xyz)
        t = (w - ws[k - 1]) / (ws[k] - ws[k - 1])
        return float(v[k - 1] + t * (v[k] - v[k - 1]))

    def reflectant_maximum(self, xyz):
        """The divisor of the recovered spectrum as a reflectance when above 1:
        the blend of the basis spectra's maxima (an upper bound of the blend's
        maximum, recovery_meng_simon.cpp:169-178)."""
        r = self.weights(xyz)
        if r is None:
            return 0.0
        rows, w = r
        return float(w @ self.maxima[rows])


class Recovery:
    """``liar.Recovery``: the standard spectral recovery that XYZ spectra use
    (src/liar/kernel/recovery.cpp:62-98).  LiAR's package installs the
    Meng-Simon model of its data file at import (src/liar/__init__.py:72-79);
    here the same data (``data/recovery_meng_simon.json``) is loaded on first
    use."""

    _standard = None

    @staticmethod
    def standard():
        if Recovery._standard is None:
            import json
            from pathlib import Path

            d = json.loads((Path(__file__).resolve().parent / "data" / "recovery_meng_simon.json").read_text())
            Recovery._standard = RecoveryMengSimon(
                d["wavelengths"], {tuple(s["xy"]): s["spectrum"] for s in d["spectra"]})
        return Recovery._standard

    @staticmethod
    def setStandard(recovery):
        if not isinstance(recovery, RecoveryMengSimon):
            raise TypeError("the standard recovery must be a spectra.RecoveryMengSimon")
        Recovery._standard = recovery


class XYZ(Spectrum):
    """``liar.Spectrum.XYZ`` / ``textures.Constant((X, Y, Z))`` / ``liar.rgb``:
    a tristimulus value (illuminant E) rendered as the standard recovery's
    spectrum of it (spectrum.cpp:231-253, spectral.cpp:180-183).  On the GPU
    it is the recovered table as a Sampled spectrum whose Reflectant divisor
    is the recovery's blended maximum."""

    def __init__(self, *xyz):
        v = xyz[0] if len(xyz) == 1 else xyz
        v = tuple(float(c) for c in v)
        if len(v) != 3:
            raise ValueError("XYZ needs three tristimulus values")
        self.value = v

    def __call__(self, wavelength):
        return Recovery.standard().recover_at(self.value, wavelength)

    def recovery_table(self):
        """(wavelengths [nm], recovered values, Reflectant divisor) of the
        standard recovery."""
        rec = Recovery.standard()
        return rec.wavelengths, rec.recover(self.value), rec.reflectant_maximum(self.value)

    def _describe(self, keep):
        ws, vs, mx = self.recovery_table()
        w, v = ws.astype(np.float32), vs.astype(np.float32)
        keep += [w, v]
        return N.Spectrum(N.SPEC_SAMPLED, len(v), w.ctypes.data, v.ctypes.data, float(mx), 0.0, 0.0, 0.0, 0.0)


def rgb(*args):
    """``liar.rgb(r, g, b[, rgbSpace])`` or ``liar.rgb((r, g, b)[, rgbSpace])``
    (src/liar/kernel/rgb_space.cpp:402-420): the XYZ spectrum of a colour in
    ``rgbSpace`` (default sRGB; gamma-encoded values)."""
    from .rgb_space import RgbSpace, sRGB

    space = sRGB
    if args and isinstance(args[-1], RgbSpace):
        space, args = args[-1], args[:-1]
    if len(args) == 1:
        args = tuple(args[0])
    if len(args) not in (3, 4):
        raise TypeError("rgb needs (r, g, b[, a]) and an optional RgbSpace")
    return XYZ(space.toXYZ(np.asarray(args[:3], np.float64)))

