"""RGB colour spaces (mirrors ``liar.RgbSpace``, src/liar/kernel/rgb_space.cpp).

A space is four CIE xy chromaticities (red, green, blue primaries and the
white point) and a gamma.  Linear RGB maps to CIE XYZ relative to
illuminant E -- the renderer's white is a flat spectrum -- through the
primaries' matrix and a Bradford adaptation from the white point to E; the
matrices come from the native library (``lr_rgb_space_matrices``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class RgbSpace:
    def __init__(self, red, green, blue, white, gamma=1.0):
        self.red = (float(red[0]), float(red[1]))
        self.green = (float(green[0]), float(green[1]))
        self.blue = (float(blue[0]), float(blue[1]))
        self.white = (float(white[0]), float(white[1]))
        self.gamma = float(gamma)
        self._m = None  # (rgb_to_xyz, xyz_to_rgb), from the native library on first use
        if not (self.gamma > 0):
            raise ValueError(f"Invalid gamma: {self.gamma}. Must be strictly positive.")

    def _matrices(self):
        if self._m is None:
            rgb_to_xyz = np.zeros(9, np.float32)
            xyz_to_rgb = np.zeros(9, np.float32)
            N.check(N.lib().lr_rgb_space_matrices(C.byref(self.struct()), N.ptr(rgb_to_xyz), N.ptr(xyz_to_rgb)),
                    "lr_rgb_space_matrices")
            self._m = (rgb_to_xyz.reshape(3, 3).astype(np.float64), xyz_to_rgb.reshape(3, 3).astype(np.float64))
        return self._m

    @property
    def _rgb_to_xyz(self):
        return self._matrices()[0]

    @property
    def _xyz_to_rgb(self):
        return self._matrices()[1]

    def validate(self) -> "RgbSpace":
        """Checks the chromaticities (RgbSpace::init, rgb_space.cpp:324-334); raises LiarError."""
        self._matrices()
        return self

    def struct(self) -> N.RgbSpace:
        return N.RgbSpace(self.red, self.green, self.blue, self.white, self.gamma)

    def linearSpace(self) -> "RgbSpace":
        """Same chromaticities, gamma 1 (rgb_space.cpp:240-247)."""
        return self.withGamma(1.0)

    def withGamma(self, gamma) -> "RgbSpace":
        return RgbSpace(self.red, self.green, self.blue, self.white, gamma)

    def _curve(self, x, g):
        x = np.asarray(x, np.float64)
        return np.sign(x) * np.power(np.abs(x), g)

    def toXYZlinear(self, rgb):
        return self._rgb_to_xyz @ np.asarray(rgb, np.float64)[:3]

    def toRGBAlinear(self, xyz):
        return self._xyz_to_rgb @ np.asarray(xyz, np.float64)[:3]

    def toXYZ(self, rgb):
        """XYZ (illuminant E) of a gamma-encoded RGB triple."""
        return self.toXYZlinear(self._curve(np.asarray(rgb, np.float64)[:3], self.gamma))

    def toRGBA(self, xyz):
        """Gamma-encoded RGB of an XYZ (illuminant E) triple."""
        return self._curve(self.toRGBAlinear(xyz), 1.0 / self.gamma)

    def __eq__(self, other):
        return isinstance(other, RgbSpace) and (self.red, self.green, self.blue, self.white, self.gamma) == (
            other.red, other.green, other.blue, other.white, other.gamma)

    def __repr__(self):
        return f"liar_amd.RgbSpace({self.red}, {self.green}, {self.blue}, {self.white}, {self.gamma})"


# built-in spaces (src/liar/kernel/rgb_space.h:124-140)
CIEXYZ = RgbSpace((1.0, 0.0), (0.0, 1.0), (0.0, 0.0), (1.0 / 3, 1.0 / 3), 1.0)
sRGB = RgbSpace((0.64, 0.33), (0.30, 0.60), (0.15, 0.06), (0.3127, 0.3290), 2.2)
