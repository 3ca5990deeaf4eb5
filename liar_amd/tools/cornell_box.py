"""The Cornell box (mirrors ``liar.tools.cornell_box``,
src/liar/tools/cornell_box.py:19-213): the measured scene of the Program of
Computer Graphics (http://www.graphics.cornell.edu/online/box/), its
millimetres scaled by 0.01 as the reference tool does, with its measured
spectra (``data/cornell_box.spd``).

``scene()`` = ``lights() + walls() + blocks()`` as a List; ``camera()`` the
box's pinhole camera.  The light is a 16-sample LightArea whose surface also
reflects (Unshaded + Lambert 0.78); walls are Parallelograms, the two blocks
TriangleMeshes whose corners carry their faces' normals."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

from .. import cameras, scenery, shaders, textures
from . import spd

_DATA = Path(__file__).resolve().parent.parent / "data" / "cornell_box.spd"
with open(_DATA) as _fp:
    WHITE, RED, GREEN, LIGHT = spd.load(_fp)

# the box's measurements in millimetres (x to the left wall, y up, z away
# from the camera)
_LIGHT_QUAD = ((213.0, 548.79, 227.0), (130.0, 0.0, 0.0), (0.0, 0.0, 105.0))
_WALLS = {  # name: (support, edge u, edge v, colour)
    "floor": ((0.0, 0.0, 559.2), (552.8, 0.0, 0.0), (0.0, 0.0, -559.2), "white"),
    "ceiling": ((0.0, 548.8, 0.0), (556.0, 0.0, 0.0), (0.0, 0.0, 559.2), "white"),
    "back": ((556.0, 0.0, 559.2), (-556.0, 0.0, 0.0), (0.0, 548.8, 0.0), "white"),
    "right": ((0.0, 0.0, 559.2), (0.0, 0.0, -559.2), (0.0, 548.8, 0.0), "green"),
    "left": ((552.8, 0.0, 0.0), (0.0, 0.0, 559.2), (0.0, 548.8, 0.0), "red"),
}
# each block: footprint corners (x, z) in order and its (bottom, top) heights
_SHORT = (((130.0, 65.0), (82.0, 225.0), (240.0, 272.0), (290.0, 114.0)), (0.05, 165.0))
_TALL = (((423.0, 247.0), (265.0, 296.0), (314.0, 456.0), (472.0, 406.0)), (0.05, 330.0))


def _m(p):
    return tuple(0.01 * float(x) for x in p)


def _lambert(name):
    return shaders.Lambert(textures.Constant({"white": WHITE, "red": RED, "green": GREEN}[name]))


def scene():
    return scenery.List(lights() + walls() + blocks())


def camera():
    cam = cameras.PerspectiveCamera()
    cam.aspectRatio = 1.0
    cam.fovAngle = 2.0 * math.atan(0.5 * 0.025 / 0.035)  # 25 mm film at 35 mm focal length
    cam.position = _m((278.0, 273.0, -800.0))
    cam.sky = (0.0, 1.0, 0.0)
    cam.direction = (0.0, 0.0, 1.0)
    return cam


def lights():
    support, du, dv = (_m(p) for p in _LIGHT_QUAD)
    light = scenery.LightArea(scenery.Parallelogram(support, du, dv))
    light.radiance = textures.Constant(LIGHT)
    light.numberOfEmissionSamples = 16
    light.shader = shaders.Sum([shaders.Unshaded(textures.Constant(LIGHT)), shaders.Lambert(textures.Constant(0.78))])
    return [light]


def walls():
    out = []
    for support, du, dv, colour in _WALLS.values():
        wall = scenery.Parallelogram(_m(support), _m(du), _m(dv))
        wall.shader = _lambert(colour)
        out.append(wall)
    return out


# a block's triangles as (corner vertices, face), vertices 0-3 the bottom ring
# and 4-7 the top ring of its footprint; each face's normal is the unit
# cross product of the edges of its corner triple (the same tables as the
# reference tool, so the meshes are the same triangles in the same order)
_FACE_NORMAL_CORNERS = ((0, 3, 1), (4, 5, 7), (0, 1, 4), (1, 2, 5), (2, 3, 6), (3, 0, 7))
_BLOCK_TRIANGLES = (((0, 2, 1), 0), ((0, 3, 2), 0), ((4, 6, 7), 1), ((4, 5, 6), 1), ((0, 5, 4), 2), ((0, 1, 5), 2),
                    ((1, 6, 5), 3), ((1, 2, 6), 3), ((2, 7, 6), 4), ((2, 3, 7), 4), ((3, 4, 7), 5), ((3, 0, 4), 5))


def _block(footprint, heights):
    verts = [_m((x, h, z)) for h in heights for x, z in footprint]
    normals = []
    for a, b, c in _FACE_NORMAL_CORNERS:
        pa, pb, pc = (np.asarray(verts[i], np.float64) for i in (a, b, c))
        n = np.cross(pb - pa, pc - pa)
        normals.append(tuple(n / np.linalg.norm(n)))
    tris = [tuple((v, face) for v in corners) for corners, face in _BLOCK_TRIANGLES]
    return scenery.TriangleMesh(verts, normals, [], tris)


def blocks():
    out = []
    for footprint, heights in (_SHORT, _TALL):
        b = _block(footprint, heights)
        b.shader = _lambert("white")
        out.append(b)
    return out
