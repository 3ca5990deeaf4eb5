"""Synthetic workloads of the benchmark configurations, built with the
LiAR-style scene API."""
from __future__ import annotations

import math

import numpy as np

from . import cameras, scenery, shaders, spectra, textures

# 10 bins of 40 nm over [380, 780) nm; a bin takes the value of its centre
# (400, 440, ..., 760 nm).
_CENTRES = 400.0 + 40.0 * np.arange(10)
WHITE = spectra.Binned(380.0, 780.0, np.full(10, 0.75))
RED = spectra.Binned(380.0, 780.0, np.where(_CENTRES > 600.0, 0.6, 0.05))
GREEN = spectra.Binned(380.0, 780.0, np.where((_CENTRES >= 500.0) & (_CENTRES <= 570.0), 0.6, 0.05))


def _quad(o, a, b, shader):
    p = scenery.Parallelogram(o, a, b)
    p.shader = shader
    return p


def _block(corners, heights, shader):
    """Rotated box: 4 floor corners (x, z) extruded between two heights."""
    ps = [(x, y, z) for y in heights for (x, z) in corners]
    faces = [(0, 2, 1), (0, 3, 2), (4, 6, 7), (4, 5, 6), (0, 5, 4), (0, 1, 5),
             (1, 6, 5), (1, 2, 6), (2, 7, 6), (2, 3, 7), (3, 4, 7), (3, 0, 4)]
    m = scenery.TriangleMesh(ps, [], [], faces)
    m.shader = shader
    return m


def cornell_box(light_scale: float = 50.0):
    """Cornell-box-style scene: 5 axis-aligned walls (10 triangles), two rotated
    boxes (24 triangles), one quad light (2 triangles) emitting a 3000 K
    blackbody (peak radiance ``light_scale``).  All surfaces Lambert with the
    10-bin reflectances above.  Returns (scene, camera)."""
    s = 0.01
    W, H, D = 5.56, 5.488, 5.592
    white = shaders.Lambert(textures.Constant(WHITE))
    red = shaders.Lambert(textures.Constant(RED))
    green = shaders.Lambert(textures.Constant(GREEN))
    walls = [
        _quad((0, 0, 0), (W, 0, 0), (0, 0, D), white),      # floor
        _quad((0, H, 0), (0, 0, D), (W, 0, 0), white),      # ceiling
        _quad((0, 0, D), (W, 0, 0), (0, H, 0), white),      # back wall
        _quad((W, 0, 0), (0, 0, D), (0, H, 0), red),        # x = W wall
        _quad((0, 0, 0), (0, H, 0), (0, 0, D), green),      # x = 0 wall
    ]
    short = _block([(1.30, 0.65), (0.82, 2.25), (2.40, 2.72), (2.90, 1.14)], [0.0, 1.65], white)
    tall = _block([(4.23, 2.47), (2.65, 2.96), (3.14, 4.56), (4.72, 4.06)], [0.0, 3.30], white)
    # light quad just below the ceiling, facing down (normal = dx x dy = -y)
    surf = scenery.Parallelogram((2.13, H - 0.001, 2.27), (1.30, 0, 0), (0, 0, 1.05))
    light = scenery.LightArea(surf)
    bb = spectra.BlackBody(3000.0, light_scale)
    light.radiance = textures.Constant(bb)
    light.shader = shaders.Sum([shaders.Unshaded(textures.Constant(bb)), shaders.Lambert(textures.Constant(WHITE))])
    scene = scenery.List([light] + walls + [short, tall])
    cam = cameras.PerspectiveCamera()
    cam.position = (2.78, 2.73, -8.0)
    cam.direction = (0.0, 0.0, 1.0)
    cam.sky = (0.0, 1.0, 0.0)
    cam.fovAngle = math.radians(40.0)
    cam.aspectRatio = 1.0
    del s
    return scene, cam


def furnace_sphere(shader, sky=1.0):
    """White-furnace set-up (test/white_furnace/ashikhmin_shirley.py:20-79):
    unit sphere under a constant sky, camera at (0, -10, 5) looking at the
    origin with an 11 degree field of view."""
    sph = scenery.Sphere((0.0, 0.0, 0.0), 1.0)
    sph.shader = shader
    scene = scenery.List([sph, scenery.LightSky(textures.Constant(sky))])
    cam = cameras.PerspectiveCamera()
    cam.position = (0.0, -10.0, 5.0)
    cam.lookAt((0.0, 0.0, 0.0))
    cam.sky = (0.0, 0.0, 1.0)
    cam.fovAngle = math.radians(11.0)
    cam.aspectRatio = 1.0
    return scene, cam
