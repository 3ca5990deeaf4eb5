"""Scenes and closed forms shared by the DirectLighting tests."""
import math

import numpy as np

import liar_amd as la
from liar_amd import scenery, shaders, textures
from liar_amd import workloads as W


def polygon_irradiance(p, n, verts, radiance):
    """Lambert's formula for the irradiance at point p (normal n) from a
    uniformly emitting convex polygon (fully visible, above the tangent plane)."""
    total = 0.0
    m = len(verts)
    for i in range(m):
        a = np.asarray(verts[i], np.float64) - p
        b = np.asarray(verts[(i + 1) % m], np.float64) - p
        theta = math.acos(np.clip(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)), -1, 1))
        g = np.cross(a, b)
        total += theta * np.dot(n, g / np.linalg.norm(g))
    return 0.5 * radiance * abs(total)


def footprint_mean(fn, half, k=24):
    """Mean of fn(x, y) over the square |x|, |y| <= half (midpoint rule)."""
    c = (np.arange(k) + 0.5) / k * 2 * half - half
    return float(np.mean([fn(x, y) for x in c for y in c]))


def light_verts(area):
    o, a, b = (np.asarray(area[k], np.float64) for k in ("origin", "dx", "dy"))
    return [o, o + a, o + a + b, o + b]


AREA = dict(origin=(-0.5, 0.5, 2.0), dx=(1.0, 0.0, 0.0), dy=(0.0, -1.0, 0.0), radiance=5.0)


def floor_expect(area, rho, cam_height=1.0, fov_deg=4.0):
    half = cam_height * math.tan(math.radians(fov_deg) / 2)
    n = np.array([0.0, 0.0, 1.0])
    vs = light_verts(area)
    return footprint_mean(lambda x, y: rho / math.pi * polygon_irradiance(np.array([x, y, 0.0]), n, vs,
                                                                         area["radiance"]), half)


def glossy_view(lights, secondary=1):
    """Camera at z = 1.8 looking up at a near-perfect glossy mirror ceiling
    (z = 2, facing down); the floor (z = 0, Lambert 0.7) is seen through it."""
    floor = W._quad((-50.0, -50.0, 0.0), (100.0, 0.0, 0.0), (0.0, 100.0, 0.0),
                    shaders.Lambert(textures.Constant(0.7)))
    ct = shaders.CookTorrance()
    ct.reflectance = textures.Constant(1.0)
    ct.refractionIndex = textures.Constant(1.0)
    ct.absorptionCoefficient = textures.Constant(1000.0)
    ct.roughnessU = ct.roughnessV = textures.Constant(1e-3)
    ct.mdf = shaders.MicrofacetTrowbridgeReitz()
    ceiling = W._quad((-50.0, -50.0, 2.0), (100.0, 0.0, 0.0), (0.0, 100.0, 0.0), ct)
    parts = [floor, ceiling]
    for L in lights:
        surf = scenery.Parallelogram(L["origin"], L["dx"], L["dy"])
        surf.shader = shaders.Unshaded(textures.Constant(L["radiance"]))
        la_ = scenery.LightArea(surf)
        la_.radiance = textures.Constant(L["radiance"])
        parts.append(la_)
    return scenery.List(parts), _up_camera()


# two lights of different power beside the footprint, both facing the floor
SIDE_LIGHTS = [dict(origin=(2.5, 0.5, 1.5), dx=(1.0, 0.0, 0.0), dy=(0.0, -1.0, 0.0), radiance=20.0),
               dict(origin=(-3.5, -1.5, 1.0), dx=(0.0, 3.0, 0.0), dy=(1.0, 0.0, 0.0), radiance=4.0)]


def glossy_expect(lights):
    """Radiance of the Lambert 0.7 floor seen through glossy_view's ceiling:
    rho E / pi averaged over the reflected footprint (the camera at 0.2 below
    the ceiling, the floor 2 below it)."""
    half = (0.2 + 2.0) * math.tan(math.radians(2.0))
    n = np.array([0.0, 0.0, 1.0])
    vss = [light_verts(L) for L in lights]
    return footprint_mean(lambda x, y: sum(0.7 / math.pi * polygon_irradiance(np.array([x, y, 0.0]), n, vs,
                                                                              L["radiance"])
                                           for vs, L in zip(vss, lights)), half)


def glow_light(radiance):
    """A double-sided parallelogram light covering the floor (z = 0)."""
    surf = scenery.Parallelogram((-50.0, 50.0, 0.0), (100.0, 0.0, 0.0), (0.0, -100.0, 0.0))
    surf.shader = shaders.Unshaded(textures.Constant(radiance))
    light = scenery.LightArea(surf)
    light.radiance = textures.Constant(radiance)
    light.isDoubleSided = True
    return light


def mirror_view(floor_shader, mirror_rho=0.5):
    """Camera at z = 1.8 looking up at a Mirror ceiling (z = 2); the floor
    (z = 0) with ``floor_shader`` is seen through it."""
    floor = W._quad((-50.0, -50.0, 0.0), (100.0, 0.0, 0.0), (0.0, 100.0, 0.0), floor_shader)
    mirror = W._quad((-50.0, -50.0, 2.0), (100.0, 0.0, 0.0), (0.0, 100.0, 0.0),
                     shaders.Mirror(textures.Constant(mirror_rho)))
    return [floor, mirror], _up_camera()


def _up_camera():
    cam = la.cameras.PerspectiveCamera()
    cam.position = (0.0, 0.0, 1.8)
    cam.lookAt((0.0, 0.0, 2.0))
    cam.sky = (0.0, 1.0, 0.0)
    cam.fovAngle = math.radians(4.0)
    cam.aspectRatio = 1.0
    return cam
