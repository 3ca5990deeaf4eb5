"""Pins the CPU DirectLighting oracle (oracle/render.c, or_render_direct) by
closed forms before it is used to check the GPU integrator.

- A Lambert floor under a constant sky of radiance s: the light estimate
  integrates the whole upper hemisphere, radiance rho s.
- A Lambert floor under a parallelogram light: Lambert's polygon formula
  (irradiance E = L/2 sum_i Theta_i (n . Gamma_i), Lambert 1760) gives
  rho E / pi, for 9 and 1 emission samples alike (the power-heuristic
  weights with sample counts, ray_tracer.cpp:188-268).
- A Mirror floor: its BSDF sample reaches the sky or the light with MIS
  weight 1 (specular), radiance rho_m s; the continuation adds nothing.
- A floor seen through a near-perfect glossy mirror: the glossy continuation
  is not high quality, so the floor's estimate picks one light by power and
  divides by its pick pdf (direct_lighting.cpp:279-307); with two lights of
  different power the result is still rho E / pi.
- Emission of a non-light object seen by a continuation counts; that of a
  light does not (direct_lighting.cpp:248-255).
- A mistake planted in the oracle (MIS weights dropped) is caught.
"""
import math

import numpy as np
import pytest

import liar_amd as la
import oracle
from direct_scenes import (AREA, SIDE_LIGHTS, floor_expect, glossy_expect, glossy_view, glow_light, mirror_view,
                           polygon_irradiance)
from liar_amd import scenery, shaders, textures
from liar_amd import workloads as W


def render_mean(scene, cam, size=16, spp=64, seed=5, **kw):
    cs = la.compile_scene(scene)
    xyz, w = oracle.render_direct(cs, cam, size, size, spp, seed=seed, **kw)
    Y = (xyz[1] / w).ravel()
    return Y.mean(), Y.std() / math.sqrt(Y.size), Y


def test_polygon_irradiance_closed_form():
    # a huge square light covers the hemisphere: E -> pi L
    big = [(-1e5, -1e5, 1.0), (1e5, -1e5, 1.0), (1e5, 1e5, 1.0), (-1e5, 1e5, 1.0)]
    assert abs(polygon_irradiance(np.zeros(3), np.array([0, 0, 1.0]), big, 2.0) - 2 * math.pi) < 1e-4
    # a small light of area A at height h straight above: E ~ L A / h^2
    h, s = 10.0, 0.01
    small = [(-s, -s, h), (s, -s, h), (s, s, h), (-s, s, h)]
    e = polygon_irradiance(np.zeros(3), np.array([0, 0, 1.0]), small, 1.0)
    assert abs(e - (2 * s) ** 2 / h ** 2) / e < 1e-5


def test_direct_lambert_sky():
    sc, cam = W.lit_floor(shaders.Lambert(textures.Constant(0.7)), sky=2.0)
    m, se, _ = render_mean(sc, cam)
    assert abs(m - 1.4) < 4 * se + 1e-9, (m, se)


@pytest.mark.parametrize("samples", [9, 1])
def test_direct_lambert_area_light(samples):
    area = dict(AREA, samples=samples)
    sc, cam = W.lit_floor(shaders.Lambert(textures.Constant(0.7)), area=area)
    m, se, _ = render_mean(sc, cam, spp=64 if samples == 9 else 256)
    want = floor_expect(area, 0.7)
    assert abs(m - want) < 4 * se + 1e-9, (m, want, se)
    assert se / want < 0.02


def test_direct_mirror_sky_and_light():
    sc, cam = W.lit_floor(shaders.Mirror(textures.Constant(0.8)), sky=2.0)
    m, se, _ = render_mean(sc, cam)
    assert abs(m - 1.6) < 4 * se + 1e-9, (m, se)
    # the reflected camera rays all reach the light: 0.8 x 5
    sc, cam = W.lit_floor(shaders.Mirror(textures.Constant(0.8)), area=AREA)
    m, se, _ = render_mean(sc, cam)
    assert abs(m - 4.0) < 4 * se + 1e-9, (m, se)


@pytest.mark.parametrize("secondary", [1, 3])
def test_direct_secondary_estimate_through_glossy_mirror(secondary):
    sc, cam = glossy_view(SIDE_LIGHTS)
    m, se, _ = render_mean(sc, cam, spp=256, secondary_light_samples=secondary)
    want = glossy_expect(SIDE_LIGHTS)
    # the near-mirror's reflectance (Fresnel of eta 1, kappa 1000) is 1 - O(1e-5)
    assert abs(m - want) < 4 * se + 1e-4 * want, (m, want, se)
    assert se / want < 0.03


def test_direct_emission_rules():
    """An Unshaded (non-light) floor seen through a mirror counts; the light's
    own surface seen through a mirror does not (only its estimate does)."""
    parts, cam = mirror_view(shaders.Unshaded(textures.Constant(3.0)))
    m, se, _ = render_mean(scenery.List(parts), cam)
    assert abs(m - 1.5) < 4 * se + 1e-9
    # the same floor as a light: the mirror's estimate sees it (0.5 x 3),
    # its surface hit by the continuation adds nothing more
    m, se, _ = render_mean(scenery.List([glow_light(3.0), parts[1]]), cam)
    assert abs(m - 1.5) < 4 * se + 1e-9, (m, se)


def test_direct_generation_limit():
    """max_generation 0 stops at the camera hit: a mirror shows its own
    estimate only (none here: no lights); 1 lets it show the glowing floor."""
    parts, cam = mirror_view(shaders.Unshaded(textures.Constant(3.0)))
    _, _, Y = render_mean(scenery.List(parts), cam, max_generation=0)
    assert (Y == 0).all()
    m, _, _ = render_mean(scenery.List(parts), cam, max_generation=1)
    assert m > 1.0


def test_planted_mis_mistake_fails_area_light():
    area = dict(AREA, samples=9)
    sc, cam = W.lit_floor(shaders.Lambert(textures.Constant(0.7)), area=area)
    want = floor_expect(area, 0.7)
    oracle.set_planted_mistake(2)
    try:
        m, se, _ = render_mean(sc, cam)
    finally:
        oracle.set_planted_mistake(0)
    assert abs(m - want) > 10 * se, (m, want, se)
