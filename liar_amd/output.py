"""Render targets (mirrors ``liar.output``): an accumulating XYZ raster with
LiAR's box-filter weights and tone mapping (src/liar/output/raster.cpp), and
an Image that saves it through the image codecs (src/liar/output/image.cpp).

Tone mapping runs on the GPU (``lr_tonemap``); the codecs are the native
library's (``lr_image_write``): .png and .tga are written in ``rgbSpace``
(gamma encoded), .exr, .hdr / .rgbe / .pic and .pfm in its linear space, as
LiAR's codecs choose (src/liar/codecs/openexr.cpp:71-72,
src/liar/kernel/image_codec.cpp:200-209).  ``.npy`` stores the raw weighted
XYZ (float32, h x w x 3).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as N
from .rgb_space import RgbSpace, sRGB

# Raster::ToneMapping names (raster.cpp:30-40), in the C-ABI's order
TONE_MAPPINGS = ("Linear", "CompressY", "CompressRGB", "Reinhard2002Y", "Reinhard2002RGB", "ExponentialY",
                 "ExponentialRGB", "DuikerY", "DuikerRGB")
LINEAR_CODECS = (".exr", ".hdr", ".rgbe", ".pic", ".pfm")


class _ToneMappingNames:
    """``Raster.ToneMapping.<name>``: the operator names, usable wherever a
    toneMapping string is (raster.cpp:30-40)."""


for _name in TONE_MAPPINGS:
    setattr(_ToneMappingNames, _name, _name)


class Raster:
    ToneMapping = _ToneMappingNames

    def __init__(self, resolution=(320, 240)):
        self.resolution = (int(resolution[0]), int(resolution[1]))
        w, h = self.resolution
        self.xyz_sum = np.zeros((3, h, w), np.float64)
        self.weight = np.zeros((h, w), np.float64)
        # Raster's defaults (raster.cpp:277-289)
        self.rgbSpace = sRGB
        self.toneMapping = "Linear"
        self._exposure_stops = 0.0
        self.exposureCorrectionStops = 0.0
        self.autoExposure = True
        self.middleGrey = 0.184
        self.maxSampleLuminance = math.inf
        self.device = 0
        self.lastTonemap = None

    # Raster::setExposureStops turns automatic exposure off (raster.cpp:203-214)
    @property
    def exposureStops(self):
        return self._exposure_stops

    @exposureStops.setter
    def exposureStops(self, stops):
        self._exposure_stops = float(stops)
        self.autoExposure = False

    def accumulate(self, xyz_sum, weight):
        self.xyz_sum += xyz_sum
        self.weight += weight

    def xyz(self):
        """Weighted mean XYZ per pixel (Raster::weighted, gain 1), shape (h, w, 3)."""
        w = np.where(self.weight > 0, self.weight, 1.0)
        return np.moveaxis(self.xyz_sum / w, 0, -1)

    def tonemap_params(self, space: RgbSpace | None = None) -> N.TonemapParams:
        if self.toneMapping not in TONE_MAPPINGS:
            raise ValueError(f"unknown toneMapping {self.toneMapping!r}; one of {TONE_MAPPINGS}")
        sp = self.rgbSpace if space is None else space
        return N.TonemapParams(sp.struct(), TONE_MAPPINGS.index(self.toneMapping), int(bool(self.autoExposure)),
                               float(self._exposure_stops), float(self.exposureCorrectionStops),
                               float(self.middleGrey))

    def tonemap(self, space: RgbSpace | None = None):
        """Raster::tonemap on the GPU: display RGBA (h, w, 4) float32 in
        ``space`` (default ``rgbSpace``).  The exposure used is left in
        ``lastTonemap``."""
        import torch

        from .qbvh import Context

        w, h = self.resolution
        dev = torch.device("cuda", self.device)
        xyz = torch.from_numpy(self.xyz_sum.astype(np.float32)).to(dev)
        wgt = torch.from_numpy(self.weight.astype(np.float32)).to(dev)
        rgba = torch.empty((h, w, 4), dtype=torch.float32, device=dev)
        info = N.TonemapInfo()
        ctx = Context(self.device)
        try:
            s = torch.cuda.current_stream(dev).cuda_stream
            prm = self.tonemap_params(space)
            N.check(N.lib().lr_tonemap(ctx.handle, C.c_void_p(xyz.data_ptr()), C.c_void_p(wgt.data_ptr()), None, w,
                                       h, C.byref(prm), C.c_void_p(rgba.data_ptr()), None, C.byref(info),
                                       C.c_void_p(s)), "lr_tonemap")
            out = rgba.cpu().numpy()
        finally:
            ctx.close()
        self.lastTonemap = {"exposure_stops": info.exposure_stops, "gain": info.gain,
                            "average_luminance": info.average_luminance, "max_luminance": info.max_luminance}
        return out


class Image(Raster):
    """Raster saved to ``filename`` at the end of a render (Image::save,
    image.cpp:75-84): tone mapped into the codec's space and written by the
    codec of the file's extension."""

    def __init__(self, filename, resolution=(320, 240)):
        super().__init__(resolution)
        self.filename = str(filename)

    def save(self):
        ext = self.filename[self.filename.rfind("."):].lower() if "." in self.filename else ""
        if ext == ".npy":
            np.save(self.filename, self.xyz().astype(np.float32))
            return
        space = self.rgbSpace.linearSpace() if ext in LINEAR_CODECS else self.rgbSpace
        rgba = np.ascontiguousarray(self.tonemap(space), np.float32)
        write_image(self.filename, rgba, space)


def write_image(path, rgba, space: RgbSpace | None = None) -> None:
    """lr_image_write of a host (h, w, 4) float32 RGBA array (rows top to bottom)."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError("rgba must have shape (h, w, 4)")
    h, w, _ = rgba.shape
    sp = None if space is None else C.byref(space.struct())
    N.check(N.lib().lr_image_write(str(path).encode(), N.ptr(rgba), w, h, sp), "lr_image_write")
