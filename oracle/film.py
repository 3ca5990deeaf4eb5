"""FP64 numpy oracle of the film output path -- TEST INFRASTRUCTURE ONLY.

Restates, for checking liar_amd's lr_rgb_space_matrices / lr_tonemap:
  - RgbSpace::init (src/liar/kernel/rgb_space.cpp:335-386): RGB -> XYZ under
    the white point, then the Bradford adaptation white -> E
    (rgb_space.cpp:424-440);
  - RgbSpace::toGamma's tone curve (rgb_space.cpp:156-175);
  - Raster::exposureStops / sceneGain / averageSceneLuminance / weighted
    (src/liar/output/raster.cpp:99-150, 613-651) and the nine operators of
    Raster::tonemap (raster.cpp:306-494), with filmic / invFilmic
    (raster.cpp:60-80).
Pinned by tests/test_oracle_film.py: the reference's own sRGB -> XYZ(E)
colour table (tests/golden/srgb_colors.json), Lindbloom's published Bradford
D65 -> D50 matrix, and closed forms of the operators.
"""
from __future__ import annotations

import math

import numpy as np

LINEAR, COMPRESS_Y, COMPRESS_RGB, REINHARD_Y, REINHARD_RGB, EXP_Y, EXP_RGB, DUIKER_Y, DUIKER_RGB = range(9)
Y_OPS = (LINEAR, COMPRESS_Y, REINHARD_Y, EXP_Y, DUIKER_Y)

SRGB = dict(red=(0.64, 0.33), green=(0.30, 0.60), blue=(0.15, 0.06), white=(0.3127, 0.3290), gamma=2.2)

BRADFORD = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])


def bradford(src_white_xyz, dst_white_xyz):
    """Von Kries adaptation in the Bradford cone space: M^-1 diag(M d / M s) M."""
    s = BRADFORD @ np.asarray(src_white_xyz, np.float64)
    d = BRADFORD @ np.asarray(dst_white_xyz, np.float64)
    return np.linalg.inv(BRADFORD) @ np.diag(d / s) @ BRADFORD


def rgb_space_matrices(red, green, blue, white, gamma=1.0):
    """(rgb_to_xyz, xyz_to_rgb): linear RGB <-> XYZ relative to illuminant E."""
    prim = np.array([[red[0], green[0], blue[0]], [red[1], green[1], blue[1]],
                     [1 - red[0] - red[1], 1 - green[0] - green[1], 1 - blue[0] - blue[1]]], np.float64)
    W = np.array([white[0] / white[1], 1.0, (1 - white[0] - white[1]) / white[1]])
    S = np.linalg.solve(prim, W)
    m_xyz = prim * S[None, :]
    m = bradford(W, np.ones(3)) @ m_xyz
    return m, np.linalg.inv(m)


def tonecurve(x, g):
    return np.where(x < 0, -np.power(np.abs(x), g), np.power(np.abs(x), g))


def filmic(x):
    # std::max(0.f, x - 0.004f) is (0 < v) ? v : 0, so NaN becomes 0
    v = np.asarray(x, np.float64) - 0.004
    with np.errstate(invalid="ignore"):
        x = np.where(v > 0, v, 0.0)
    y = (x * (6.2 * x + 0.5)) / (x * (6.2 * x + 1.7) + 0.06)
    return np.power(y, 2.2)


def inv_filmic(y):
    y = min(max(y, 0.0), 0.99999)
    y = y ** (1 / 2.2)
    a, b, c = 6.2 * y - 6.2, 1.7 * y - 0.5, 0.06 * y
    return (-b - np.sqrt(b * b - 4 * a * c)) / (2 * a) + 0.004


def exposure(xyz, w, op, auto=True, stops=0.0, correction=0.0, middle_grey=0.184):
    """(gain, stops, log-average Y, max Y) of a film (xyz (3, n), w (n,))."""
    y = xyz[1]
    cov = (w > 0) & (y > 0)
    l = y[cov] / w[cov]
    avg = float(np.exp(np.mean(np.log(l)))) if l.size else 0.0
    mx = float(l.max()) if l.size else 0.0
    if auto:
        g = middle_grey
        gain = 1.0
        if avg > 0:
            if op == LINEAR:
                gain = g / avg
            elif op in (COMPRESS_Y, COMPRESS_RGB, REINHARD_Y, REINHARD_RGB):
                gain = g / (avg * (1 - g))
            elif op in (EXP_Y, EXP_RGB):
                gain = -np.log1p(-g) / avg
            else:
                gain = inv_filmic(g) / avg
        v = 3 * float(np.log2(gain))
        stops = math.copysign(math.floor(abs(v) + 0.5), v) / 3  # std::round: halves away from zero
    return 2.0 ** (stops + correction), stops, avg, mx


def tonemap(xyz, w, op, space=SRGB, alpha=None, **exp_kw):
    """Display RGBA (n, 4) of a film: xyz (3, n) sums, w (n,) weights,
    alpha (n,) sums or None (alpha 1 where w > 0)."""
    xyz = np.asarray(xyz, np.float64)
    w = np.asarray(w, np.float64)
    gain, stops, avg, mx = exposure(xyz, w, op, **exp_kw)
    _, to_rgb = rgb_space_matrices(space["red"], space["green"], space["blue"], space["white"])
    pos = w > 0
    s = np.where(pos, gain / np.where(pos, w, 1.0), 0.0)
    X, Y, Z = xyz[0] * s, xyz[1] * s, xyz[2] * s
    a = np.where(pos, (alpha / np.where(pos, w, 1.0)) if alpha is not None else 1.0, 0.0)
    lw = gain * mx
    inv_lw2 = 1.0 / (lw * lw) if lw > 0 else np.inf
    with np.errstate(all="ignore"):
        if op in Y_OPS:
            if op == LINEAR:
                sc = np.ones_like(Y)
            elif op == COMPRESS_Y:
                sc = 1 / (1 + Y)
            elif op == REINHARD_Y:
                sc = (1 + Y * inv_lw2) / (1 + Y)
            elif op == EXP_Y:
                sc = np.where(Y < 1e-8, 1 - Y / 2, -np.expm1(-Y) / np.where(Y == 0, 1, Y))
            else:
                sc = filmic(Y) / np.maximum(Y, 0.004)
            X, Y, Z = X * sc, Y * sc, Z * sc
            nan = np.isnan(X) | np.isnan(Y) | np.isnan(Z) | np.isnan(a)
            rgb = to_rgb @ np.stack([X, Y, Z])
        else:
            rgb = to_rgb @ np.stack([X, Y, Z])
            if op == COMPRESS_RGB:
                rgb = rgb / (1 + rgb)
            elif op == REINHARD_RGB:
                rgb = rgb * (1 + rgb * inv_lw2) / (1 + rgb)
            elif op == EXP_RGB:
                rgb = -np.expm1(-rgb)
            else:
                rgb = filmic(rgb)
            nan = np.isnan(rgb).any(0) | np.isnan(a)
        if space["gamma"] != 1:
            rgb = tonecurve(rgb, 1.0 / space["gamma"])
    out = np.concatenate([rgb, a[None]], 0).T.copy()
    out[nan] = (1.0, 0.0, 1.0, 1.0)
    return out, {"gain": gain, "stops": stops, "average_luminance": avg, "max_luminance": mx}


def quantize8(rgba):
    """LiAR's PNG quantisation: clamp(255 x, 0, 255), truncated."""
    return np.clip(np.nan_to_num(255 * np.asarray(rgba, np.float64), nan=0.0), 0, 255).astype(np.uint8)
