"""Render targets (mirrors ``liar.output``): an accumulating XYZ raster with
LiAR's box-filter weights (src/liar/output/raster.cpp:547-651)."""
from __future__ import annotations

import numpy as np


class Raster:
    def __init__(self, resolution=(320, 240)):
        self.resolution = (int(resolution[0]), int(resolution[1]))
        w, h = self.resolution
        self.xyz_sum = np.zeros((3, h, w), np.float64)
        self.weight = np.zeros((h, w), np.float64)

    def accumulate(self, xyz_sum, weight):
        self.xyz_sum += xyz_sum
        self.weight += weight

    def xyz(self):
        """Weighted mean XYZ per pixel (Raster::weighted), shape (h, w, 3)."""
        w = np.where(self.weight > 0, self.weight, 1.0)
        return np.moveaxis(self.xyz_sum / w, 0, -1)

    def rgb(self):
        """Linear sRGB (D65) of xyz()."""
        m = np.array([[3.2404542, -1.5371385, -0.4985314],
                      [-0.9692660, 1.8760108, 0.0415560],
                      [0.0556434, -0.2040259, 1.0572252]])
        return self.xyz() @ m.T


class Image(Raster):
    """Raster saved to ``filename`` at the end of a render: ``.npy`` (XYZ,
    float32) or ``.pfm`` (linear sRGB)."""

    def __init__(self, filename, resolution=(320, 240)):
        super().__init__(resolution)
        self.filename = str(filename)

    def save(self):
        if self.filename.endswith(".pfm"):
            rgb = self.rgb().astype("<f4")[::-1]
            with open(self.filename, "wb") as f:
                f.write(b"PF\n%d %d\n-1.0\n" % (self.resolution[0], self.resolution[1]))
                f.write(rgb.tobytes())
        else:
            np.save(self.filename, self.xyz().astype(np.float32))
