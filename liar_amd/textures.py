"""Textures (mirrors ``liar.textures``; only Constant, src/liar/textures/constant.cpp)."""
from __future__ import annotations

from . import spectra


class Texture:
    pass


class Constant(Texture):
    """A constant scalar or spectrum."""

    def __init__(self, value=1.0):
        if isinstance(value, spectra.Spectrum):
            self.spectrum = value
            self.value = None
        elif isinstance(value, (tuple, list)):
            # Constant(const XYZ&) (constant.cpp:50-55): a tristimulus triple
            self.spectrum = spectra.XYZ(value)
            self.value = None
        else:
            self.value = float(value)
            self.spectrum = spectra.Constant(self.value)

    def scalar(self) -> float:
        if self.value is None:
            raise TypeError("a spectral Constant has no scalar value")
        return self.value


def as_texture(x) -> Texture:
    if isinstance(x, Texture):
        return x
    return Constant(x)
