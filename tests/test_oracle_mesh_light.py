"""Pins the oracle's LightArea over a TriangleMesh (light_area.cpp:189-252,
triangle_mesh.cpp:69-85,414-444) and LiAR's emission sidedness:

- a Lambert floor lit by a hexagonal mesh light (six triangles of unequal
  area) has the radiance rho E / pi of Lambert's polygon formula, with the
  AdjointPhotonTracer (light picked by area table + BSDF sampling, one-sample
  MIS) and with DirectLighting (9 and 1 emission samples);
- a mirror floor sees the back face of a single-sided light: LightArea's
  emission along a ray re-intersects the surface, whose normal the
  IntersectionContext turns towards the ray (scene_object.cpp:252-266), so
  the single-sided test never rejects it -- the mirror shows rho_m L, as in
  LiAR -- while sampling (sampleEmission) never picks the back face.
"""
import math

import pytest

import liar_amd as la
import oracle
from liar_amd import shaders, textures
from mesh_light_scenes import HEX_L, hexagon_floor_expect, hexagon_light_floor


def mean_and_se(xyz, w):
    Y = (xyz[1] / w).ravel()
    return Y.mean(), Y.std() / math.sqrt(Y.size)


def test_hexagon_light_path_tracer():
    sc, cam, ring = hexagon_light_floor(shaders.Lambert(textures.Constant(0.7)))
    xyz, w = oracle.render(la.compile_scene(sc), cam, 16, 16, 512, max_depth=2, seed=3)
    m, se = mean_and_se(xyz, w)
    want = hexagon_floor_expect(ring, 0.7)
    assert abs(m - want) < 4 * se + 1e-3 * want, (m, want, se)


@pytest.mark.parametrize("samples", [9, 1])
def test_hexagon_light_direct_lighting(samples):
    sc, cam, ring = hexagon_light_floor(shaders.Lambert(textures.Constant(0.7)), samples=samples)
    xyz, w = oracle.render_direct(la.compile_scene(sc), cam, 16, 16, 64 if samples == 9 else 256, seed=4)
    m, se = mean_and_se(xyz, w)
    want = hexagon_floor_expect(ring, 0.7)
    assert abs(m - want) < 4 * se + 1e-3 * want, (m, want, se)


@pytest.mark.parametrize("direct", [False, True])
def test_single_sided_light_back_face_emits_when_hit(direct):
    # the light faces up (away from the mirror floor below it)
    sc, cam, _ = hexagon_light_floor(shaders.Mirror(textures.Constant(0.8)), facing=+1.0)
    cs = la.compile_scene(sc)
    if direct:
        xyz, w = oracle.render_direct(cs, cam, 8, 8, 16, seed=5)
    else:
        xyz, w = oracle.render(cs, cam, 8, 8, 16, max_depth=3, seed=5)
    m, se = mean_and_se(xyz, w)
    assert abs(m - 0.8 * HEX_L) < 4 * se + 1e-6, m
