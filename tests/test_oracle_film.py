"""Pins the film-output oracle (oracle/film.py) before it checks lr_tonemap.

- The reference's own colour table (test/test_codecs.cpp:35-76, committed as
  tests/golden/srgb_colors.json): sRGB 8-bit colours and their XYZ under
  illuminant E.  Decoded with the IEC 61966-2-1 sRGB curve (as the table's
  external tool did) they match the oracle's RGB -> XYZ(E) matrix.
- Bradford: the oracle's adaptation matrices equal Lindbloom's published
  D65 -> D50 and D50 -> D65 Bradford matrices.
- A planted mistake (no adaptation to E) fails the colour table.
- Tone-mapping operators and automatic exposure against closed forms.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle.film as F

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "srgb_colors.json").read_text())


def srgb_decode(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def srgb_matrix(adapt=True):
    m, _ = F.rgb_space_matrices(F.SRGB["red"], F.SRGB["green"], F.SRGB["blue"], F.SRGB["white"])
    if not adapt:
        W = np.array([0.3127 / 0.3290, 1.0, (1 - 0.3127 - 0.3290) / 0.3290])
        m = np.linalg.inv(F.bradford(W, np.ones(3))) @ m
    return m


def table_error(m):
    err = 0.0
    for rgb_row, xyz_row in zip(GOLD["rgb8"], GOLD["xyz_e"]):
        for rgb, xyz in zip(rgb_row, xyz_row):
            got = m @ srgb_decode(np.array(rgb) / 255.0)
            err = max(err, float(np.abs(got - np.array(xyz)).max()))
    return err


def test_srgb_matrix_matches_reference_colour_table():
    # the table has 4 decimals
    assert table_error(srgb_matrix()) < 1.0e-4
    m, mi = F.rgb_space_matrices(F.SRGB["red"], F.SRGB["green"], F.SRGB["blue"], F.SRGB["white"])
    assert np.allclose(m @ np.ones(3), 1.0, atol=1e-12)  # RGB white -> E white
    assert np.allclose(m @ mi, np.eye(3), atol=1e-12)


def test_planted_missing_adaptation_fails_colour_table():
    assert table_error(srgb_matrix(adapt=False)) > 1e-2


def test_bradford_matches_published_matrices():
    # Lindbloom, "Chromatic Adaptation": Bradford D65 -> D50 and D50 -> D65
    d65, d50 = [0.95047, 1.0, 1.08883], [0.96422, 1.0, 0.82521]
    want = np.array([[1.0478112, 0.0228866, -0.0501270],
                     [0.0295424, 0.9904844, -0.0170491],
                     [-0.0092345, 0.0150436, 0.7521316]])
    assert np.abs(F.bradford(d65, d50) - want).max() < 2e-7
    back = np.array([[0.9555766, -0.0230393, 0.0631636],
                     [-0.0282895, 1.0099416, 0.0210077],
                     [0.0122982, -0.0204830, 1.3299098]])
    assert np.abs(F.bradford(d50, d65) - back).max() < 2e-7


def test_filmic_inverse_and_known_points():
    for g in (0.05, 0.184, 0.5, 0.9):
        assert abs(float(F.filmic(F.inv_filmic(g))) - g) < 1e-9
    assert float(F.filmic(0.004)) == 0.0
    assert float(F.filmic(-1.0)) == 0.0
    assert float(F.filmic(np.nan)) == 0.0  # std::max(0.f, NaN) returns 0


def film(values, w=None):
    """A film whose pixels have XYZ = values * (1, 1, 1) (E white) and weight w."""
    values = np.asarray(values, np.float64)
    w = np.ones_like(values) * 2.0 if w is None else np.asarray(w, np.float64)
    return np.stack([values * w] * 3), w


LIN = dict(F.SRGB, gamma=1.0)


@pytest.mark.parametrize("op,expect", [
    (F.LINEAR, lambda y: y),
    (F.COMPRESS_Y, lambda y: y / (1 + y)),
    (F.COMPRESS_RGB, lambda y: y / (1 + y)),
    (F.EXP_Y, lambda y: -math.expm1(-y)),
    (F.EXP_RGB, lambda y: -math.expm1(-y)),
    (F.DUIKER_Y, lambda y: float(F.filmic(y))),
    (F.DUIKER_RGB, lambda y: float(F.filmic(y))),
])
def test_operators_on_grey(op, expect):
    # E-white greys map to R = G = B = Y in a gamma-1 space, so every
    # operator reduces to its scalar curve
    vals = [0.0, 0.01, 0.3, 1.0, 4.0]
    xyz, w = film(vals)
    out, info = F.tonemap(xyz, w, op, space=LIN, auto=False, stops=0.0)
    assert info["gain"] == 1.0
    for k, y in enumerate(vals):
        assert np.allclose(out[k, :3], expect(y), rtol=1e-12, atol=1e-12), (op, y, out[k])
        assert out[k, 3] == 1.0


def test_reinhard_and_gamma():
    vals = [0.25, 1.0, 3.0]
    xyz, w = film(vals)
    out, info = F.tonemap(xyz, w, F.REINHARD_Y, space=LIN, auto=False, stops=1.0)
    lw = 2.0 * 3.0  # gain 2^1 times the largest Y
    for k, y in enumerate(vals):
        L = 2.0 * y
        assert np.allclose(out[k, :3], L * (1 + L / lw ** 2) / (1 + L), rtol=1e-12)
    out_g, _ = F.tonemap(xyz, w, F.LINEAR, space=F.SRGB, auto=False, stops=0.0)
    assert np.allclose(out_g[:, 0], np.array(vals) ** (1 / 2.2), rtol=1e-12)


def test_auto_exposure_uniform_film():
    xyz, w = film([0.02] * 16)
    for op, gain in ((F.LINEAR, 0.184 / 0.02), (F.COMPRESS_Y, 0.184 / (0.02 * (1 - 0.184))),
                     (F.EXP_RGB, -math.log1p(-0.184) / 0.02), (F.DUIKER_Y, F.inv_filmic(0.184) / 0.02)):
        g, stops, avg, mx = F.exposure(xyz, w, op)
        assert abs(avg - 0.02) < 1e-15 and abs(mx - 0.02) < 1e-15
        v = 3 * math.log2(gain)
        assert stops == math.copysign(math.floor(abs(v) + 0.5), v) / 3
        assert g == 2.0 ** stops
    # uncovered pixels and black pixels do not count; exposure correction adds
    xyz, w = film([0.02, 0.0, 0.08], w=[1.0, 1.0, 0.0])
    g, stops, avg, mx = F.exposure(xyz, w, F.LINEAR, correction=0.5)
    assert abs(avg - 0.02) < 1e-15 and g == 2.0 ** (stops + 0.5)


def test_nan_is_pink_and_uncovered_is_clear():
    xyz, w = film([np.nan, 1.0], w=[1.0, 0.0])
    out, _ = F.tonemap(xyz, w, F.LINEAR, auto=False)
    assert tuple(out[0]) == (1.0, 0.0, 1.0, 1.0)
    assert tuple(out[1]) == (0.0, 0.0, 0.0, 0.0)


def test_quantize8():
    q = F.quantize8([-0.1, 0.0, 0.5, 254.9 / 255, 1.0, 2.0, np.nan])
    assert list(q) == [0, 0, 127, 254, 255, 255, 0]
