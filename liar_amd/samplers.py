"""Samplers (mirrors ``liar.samplers``).  Sample values come from the
counter-based Philox4x32-10 stream keyed by (seed, path, generation); the
Stratifier keeps LiAR's per-pixel screen strata (src/liar/samplers/stratifier.cpp)."""
from __future__ import annotations


class Stratifier:
    def __init__(self, resolution=(320, 240), samplesPerPixel=1):
        self.resolution = (int(resolution[0]), int(resolution[1]))
        self.samplesPerPixel = int(samplesPerPixel)
        self.jittered = True


class Halton(Stratifier):
    """Accepted for API compatibility; sampled like the Stratifier."""

    def __init__(self, resolution=(320, 240), samplesPerPixel=1):
        super().__init__(resolution, samplesPerPixel)
