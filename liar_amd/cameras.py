"""Cameras (mirrors ``liar.cameras.PerspectiveCamera``,
src/liar/cameras/perspective_camera.cpp; pinhole only)."""
from __future__ import annotations

import math

import numpy as np


class PerspectiveCamera:
    def __init__(self):
        self.position = (0.0, 0.0, 0.0)
        self.direction = (0.0, 1.0, 0.0)
        self.sky = (0.0, 0.0, 1.0)
        self.fovAngle = math.radians(60.0)
        self.aspectRatio = 4.0 / 3.0
        self.nearLimit = 0.0
        self.lensRadius = 0.0

    def lookAt(self, target):
        d = np.asarray(target, dtype=np.float64) - np.asarray(self.position, dtype=np.float64)
        n = np.linalg.norm(d)
        if n == 0:
            raise ValueError("lookAt target equals the camera position")
        self.direction = tuple(d / n)
