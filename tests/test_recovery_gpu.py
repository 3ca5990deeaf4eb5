"""Tristimulus colours on the GPU (recovered spectra, LiAR's Meng-Simon
recovery, src/liar/spectra/recovery_meng_simon.cpp).

1. The GPU's lookup of an XYZ spectrum (a Sampled table with the recovery's
   Reflectant divisor) equals the oracle's.
2. Closed form: a Lambert sphere of reflectance Constant((X, Y, Z)) under a
   constant sky of radiance 1 renders to (X, Y, Z): the one-bounce estimate's
   expectation is the tristimulus value of the reflectance spectrum, which
   the recovery makes equal to the colour (for a colour inside the basis
   points' hull and a blended maximum below 1).
3. Same seed: an rgb-coloured scene matches the oracle render.
"""
import math
import sys

import numpy as np
import pytest

import liar_amd as la
import oracle as O
from liar_amd import shaders, spectra, textures
from liar_amd import workloads as W

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_render_gpu import compare, gpu_render, say  # noqa: E402

pytestmark = pytest.mark.gpu


def test_xyz_spectrum_lookup_matches_oracle(gpu_ctx):
    w = np.linspace(360.0, 830.0, 941)
    for xyz in ((0.2, 0.3, 0.1), (0.9, 1.0, 0.6), (1.5, 1.4, 0.2)):
        sp = spectra.XYZ(xyz)
        for refl in (False, True):
            got = sp.evaluate(gpu_ctx, w, reflectant=refl)
            want = O.spectrum_eval(sp, w.astype(np.float32).astype(np.float64), reflectant=refl)
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7, err_msg=f"{xyz} reflectant {refl}")


@pytest.mark.parametrize("xyz", [(0.3, 0.35, 0.25), (0.4, 0.35, 0.2)])
def test_lambert_xyz_sphere_furnace_renders_its_colour(gpu_ctx, xyz):
    rec = la.Recovery.standard()
    rows, _ = rec.weights(xyz)
    assert (rows >= 4).all() and rec.reflectant_maximum(xyz) < 1.0  # the closed form's conditions
    sc, cam = W.furnace_sphere(shaders.Lambert(textures.Constant(xyz)))
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 128, 128, 16384, nl=1, max_depth=0, rr_depth=8, seed=3,
                           lam=(360.0, 830.0))
    img = gx / gw
    yy, xx = np.mgrid[0:128, 0:128]
    centre = (xx - 63.5) ** 2 + (yy - 63.5) ** 2 < 36 ** 2
    got = np.array([img[c][centre].mean() for c in range(3)])
    sigma = np.array([img[c][centre].std() / math.sqrt(centre.sum()) for c in range(3)])
    say(f"[gpu] XYZ furnace {xyz}: rendered {np.round(got, 5)}, 3 sigma {np.round(3 * sigma, 6)}")
    assert (np.abs(got - np.array(xyz)) < 3 * sigma + 1e-3).all()


def test_rgb_scene_same_seed_matches_oracle(gpu_ctx):
    sc, cam = W.furnace_sphere(shaders.Lambert(textures.Constant(la.rgb(0.8, 0.3, 0.1))))
    cs = la.compile_scene(sc)
    gx, gw, _ = gpu_render(gpu_ctx, cs, cam, 48, 48, 16, nl=2, max_depth=4, rr_depth=8, seed=7)
    ox, ow = O.render(cs, cam, 48, 48, 16, wavelengths=2, max_depth=4, rr_depth=8, seed=7)
    rel, bad, _ = compare(gx, gw, ox, ow, "rgb sphere")
    assert rel < 2e-3 and bad < 0.02
