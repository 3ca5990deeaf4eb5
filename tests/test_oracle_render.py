"""Pins the CPU path-tracing oracle (oracle/render.c) by closed forms before
it is used to check the GPU renderer.

- Observer: the wavelength pdf integrates to one and a unit spectrum has
  Y = 1 (src/liar/kernel/observer.cpp:105-123); a constant-1 sky seen after
  one Lambert(1) bounce therefore gives Y = 1 (white furnace).
- Closed forms of the furnace sphere: Lambert rho gives rho, Mirror rho gives
  rho, a lossless dielectric gives 1; a black sphere
  gives 0; a Lambert sphere under a sky of radiance s gives rho * s.
- BlackBody: normalised so that its peak value equals `scale`
  (src/liar/spectra/black_body.cpp:76,96-99).
- A mistake planted in the oracle (MIS weights dropped) is caught by the
  furnace test.
"""
import math

import numpy as np
import pytest

import liar_amd as la
import oracle
from liar_amd import shaders, spectra, textures
from liar_amd import workloads as W


def furnace(shader, sky=1.0, spp=64, size=32, seed=3, lam=(380.0, 780.0)):
    sc, cam = W.furnace_sphere(shader, sky)
    cs = la.compile_scene(sc)
    xyz, w = oracle.render(cs, cam, size, size, spp, wavelengths=1, max_depth=0, rr_depth=8, seed=seed,
                           lambda_range=lam)
    img = xyz / w
    yy, xx = np.mgrid[0:size, 0:size]
    c = (size - 1) / 2
    centre = (xx - c) ** 2 + (yy - c) ** 2 < (0.28 * size) ** 2
    return img[1][centre], img


@pytest.mark.parametrize("rho", [1.0, 0.5])
def test_oracle_lambert_furnace(rho):
    Y, _ = furnace(shaders.Lambert(textures.Constant(rho)), spp=256)
    assert abs(Y.mean() - rho) < 4 * Y.std() / math.sqrt(Y.size) + 1e-6


def test_oracle_mirror_furnace():
    # each path returns exactly 0.8 x sky; the only noise is the wavelength
    # estimator ybar(lambda) / pdf(lambda)
    Y, img = furnace(shaders.Mirror(textures.Constant(0.8)), spp=256)
    assert abs(Y.mean() - 0.8) < 4 * Y.std() / math.sqrt(Y.size) + 1e-6
    assert (img[1][:3, :3] == 0).all()


def test_oracle_dielectric_furnace():
    Y, _ = furnace(shaders.Dielectric(textures.Constant(1.5)), spp=256)
    assert abs(Y.mean() - 1.0) < 4 * Y.std() / math.sqrt(Y.size) + 1e-6


def test_oracle_sky_scale_and_black():
    Y, _ = furnace(shaders.Lambert(textures.Constant(0.5)), sky=3.0, spp=256)
    assert abs(Y.mean() - 1.5) < 4 * Y.std() / math.sqrt(Y.size) + 1e-6
    Y0, _ = furnace(shaders.Lambert(textures.Constant(0.0)))
    assert (Y0 == 0).all()


def test_oracle_spectral_sky_matches_luminance():
    """Sky with a BlackBody spectrum over a Mirror(1): each path returns the
    sky value at its wavelength, so the mean Y equals the luminance of the
    spectrum on the sampling domain (trapezoid integral of S * ybar)."""
    bb = spectra.BlackBody(4000, 2.0)
    Y, _ = furnace(shaders.Mirror(textures.Constant(1.0)), sky=bb, spp=512, size=24)
    w, xyz = la.observer.standard_observer()
    m = (w >= 380) & (w <= 780)
    ww, yb = w[m].astype(np.float64), xyz[m, 1].astype(np.float64)
    h, c, k = 6.626070040e-34, 299792458.0, 1.3806488e-23
    pl = lambda x: 2 * h * c * c / (x ** 5 * np.expm1(h * c / (k * x * 4000)))
    s = 2.0 * pl(ww * 1e-9) / pl(2.897771955e-3 / 4000)
    fine = np.linspace(380, 780, 40001)
    lum = np.trapezoid(np.interp(fine, ww, s * 1.0) * np.interp(fine, ww, yb), fine) / np.trapezoid(
        np.interp(fine, ww, yb), fine)
    assert abs(Y.mean() - lum) / lum < 5e-3, (Y.mean(), lum)


def test_blackbody_peak_equals_scale():
    bb = spectra.BlackBody(3000, 7.0)
    lam = bb.peakWavelength * 1e9
    # evaluate through the oracle's spectrum path: a sky of this spectrum over
    # a mirror restricted to a 0.01 nm window around the peak
    sc, cam = W.furnace_sphere(shaders.Mirror(textures.Constant(1.0)), bb)
    cs = la.compile_scene(sc)
    xyz, w = oracle.render(cs, cam, 8, 8, 4, lambda_range=(lam - 0.005, lam + 0.005), seed=1,
                           observer=(np.array([lam - 1, lam + 1], np.float32),
                                     np.array([[1, 1, 1], [1, 1, 1]], np.float32)))
    img = xyz / w
    assert abs(img[1][4, 4] - 7.0) < 1e-6 * 7


def test_planted_mis_mistake_fails_furnace():
    """A mistake planted in the oracle (both one-sample MIS weights dropped)
    doubles the Lambert furnace, so the furnace test catches it."""
    oracle.set_planted_mistake(1)
    try:
        Y, _ = furnace(shaders.Lambert(textures.Constant(1.0)), spp=128)
    finally:
        oracle.set_planted_mistake(0)
    assert abs(Y.mean() - 2.0) < 0.05
    Y, _ = furnace(shaders.Lambert(textures.Constant(1.0)), spp=128)
    assert abs(Y.mean() - 1.0) < 0.02


def test_compile_scene_cornell():
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    assert cs.tris.shape == (36, 3) and cs.spheres.shape == (0, 4)
    assert len(cs.lights) == 1 and cs.lights[0]["type"] == 0
    # the two light triangles carry the light's material, whose light index is 0
    m = cs.materials[cs.prim_material[0]]
    assert m["light"] == 0 and m["emission"] != 0xFFFFFFFF and m["type"] == 1
    # walls: 5 quads = 10 triangles; boxes 24
    assert (cs.prim_material[2:] != cs.prim_material[0]).all()


def test_binned_spectra_of_cornell():
    red = W.RED.values
    green = W.GREEN.values
    np.testing.assert_array_equal(red, np.float32([0.05] * 6 + [0.6] * 4))
    np.testing.assert_array_equal(green, np.float32([0.05] * 3 + [0.6] * 2 + [0.05] * 5))


def test_oracle_cornell_consistent_with_committed_reference():
    """The oracle at 64 spp on the Cornell workload agrees with the committed
    4096-spp oracle reference within Monte Carlo noise (bias < 1 %)."""
    from pathlib import Path

    ref = np.load(Path(__file__).resolve().parent / "golden" / "cornell_ref_4096spp.npz")["y"].astype(np.float64)
    sc, cam = W.cornell_box()
    cs = la.compile_scene(sc)
    xyz, w = oracle.render(cs, cam, 256, 256, 16, wavelengths=4, max_depth=8, rr_depth=3, seed=5)
    y = (xyz / w)[1]
    bias = (y.mean() - ref.mean()) / ref.mean()
    mrae = np.abs(y - ref).mean() / ref.mean()
    assert abs(bias) < 1e-2 and mrae < 0.35, (bias, mrae)
