"""GPU parity of the film output path (lr_tonemap) with the FP64 oracle
(oracle/film.py, pinned in tests/test_oracle_film.py): all nine tone-mapping
operators, automatic and manual exposure, with and without an alpha plane,
gamma-encoded and linear spaces; the 8-bit output against the oracle's
quantisation; Image.save through the codecs."""
import ctypes as C
import sys
import zlib

import numpy as np
import pytest

import liar_amd as la
import oracle.film as F
from liar_amd import _native as N
from liar_amd.output import Image

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_film(h, w, seed=1, nan=True):
    rng = np.random.default_rng(seed)
    wgt = rng.integers(0, 9, (h, w)).astype(np.float32)  # some pixels uncovered
    lum = np.exp(rng.normal(-1.0, 2.0, (h, w)))  # wide dynamic range
    # colours inside the sRGB gamut: positive linear RGB, so that the RGB
    # operators stay away from their pole at r = -1
    rgb = rng.uniform(0.05, 1.0, (3, h * w))
    m, _ = F.rgb_space_matrices(F.SRGB["red"], F.SRGB["green"], F.SRGB["blue"], F.SRGB["white"])
    chroma = (m @ rgb).reshape(3, h, w)
    xyz = (chroma * lum * wgt).astype(np.float32)
    if nan:
        xyz[1, 3, 4] = np.nan
        wgt[3, 4] = 2.0
    alpha = (wgt * rng.uniform(0, 1, (h, w))).astype(np.float32)
    return xyz, wgt, alpha


def gpu_tonemap(ctx, xyz, wgt, alpha, op, space, auto, stops=0.0, corr=0.0, grey=0.184):
    import torch

    h, w = wgt.shape
    dx = torch.from_numpy(xyz).cuda()
    dw = torch.from_numpy(wgt).cuda()
    da = None if alpha is None else torch.from_numpy(alpha).cuda()
    rgba = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    rgba8 = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    prm = N.TonemapParams(space.struct(), op, int(auto), stops, corr, grey)
    info = N.TonemapInfo()
    s = torch.cuda.current_stream().cuda_stream
    N.check(N.lib().lr_tonemap(ctx.handle, C.c_void_p(dx.data_ptr()), C.c_void_p(dw.data_ptr()),
                               None if da is None else C.c_void_p(da.data_ptr()), w, h, C.byref(prm),
                               C.c_void_p(rgba.data_ptr()), C.c_void_p(rgba8.data_ptr()), C.byref(info),
                               C.c_void_p(s)), "lr_tonemap")
    torch.cuda.synchronize()
    return rgba.cpu().numpy(), rgba8.cpu().numpy(), info


def space_dict(sp):
    return dict(red=sp.red, green=sp.green, blue=sp.blue, white=sp.white, gamma=sp.gamma)


@pytest.mark.parametrize("op", range(9))
def test_tonemap_matches_oracle(gpu_ctx, op):
    xyz, wgt, alpha = random_film(37, 53, seed=op)
    h, w = wgt.shape
    for space in (la.sRGB, la.sRGB.linearSpace()):
        for auto, stops, corr in ((True, 0.0, 0.0), (False, -1.5, 0.25)):
            for a in (None, alpha):
                got, got8, info = gpu_tonemap(gpu_ctx, xyz, wgt, a, op, space, auto, stops, corr)
                ref, rinfo = F.tonemap(xyz.reshape(3, -1).astype(np.float64), wgt.ravel(), op,
                                       space=space_dict(space),
                                       alpha=None if a is None else a.ravel().astype(np.float64),
                                       auto=auto, stops=stops, correction=corr)
                ref = ref.reshape(h, w, 4)
                assert info.exposure_stops == pytest.approx(rinfo["stops"], abs=1e-6)
                assert info.gain == pytest.approx(rinfo["gain"], rel=1e-6)
                assert info.average_luminance == pytest.approx(rinfo["average_luminance"], rel=1e-5)
                assert info.max_luminance == pytest.approx(rinfo["max_luminance"], rel=1e-6)
                err = np.abs(got - ref)
                tol = 2e-5 + 1e-4 * np.abs(ref)
                bad = err > tol
                assert not bad.any(), (op, space.gamma, auto, a is None, err.max(), np.argwhere(bad)[:5])
                if op != F.DUIKER_RGB:  # LiAR's filmic curve turns NaN into 0 (std::max)
                    assert tuple(got[3, 4]) == (1.0, 0.0, 1.0, 1.0)  # NaN pixel is pink
                r8 = F.quantize8(ref)
                near = np.abs(255 * ref - np.round(255 * ref)) < 2e-3 * np.maximum(1, 255 * np.abs(ref))
                mism = (got8 != r8) & ~near
                assert not mism.any(), (op, np.argwhere(mism)[:5])
    say(f"[gpu] tonemap op {op}: matches the oracle (4 configurations x 2 spaces)")


def test_tonemap_large_film_reduction(gpu_ctx):
    """1920x1080: the exposure reduction spans the full partial-sum grid."""
    xyz, wgt, _ = random_film(1080, 1920, seed=11, nan=False)
    got, _, info = gpu_tonemap(gpu_ctx, xyz, wgt, None, F.REINHARD_RGB, la.sRGB, True)
    ref, rinfo = F.tonemap(xyz.reshape(3, -1).astype(np.float64), wgt.ravel(), F.REINHARD_RGB,
                           space=space_dict(la.sRGB))
    say(f"[gpu] 1080p tonemap: stops {info.exposure_stops:.4f} (oracle {rinfo['stops']:.4f}), "
        f"avg {info.average_luminance:.6g} ({rinfo['average_luminance']:.6g})")
    assert info.average_luminance == pytest.approx(rinfo["average_luminance"], rel=1e-6)
    assert info.exposure_stops == pytest.approx(rinfo["stops"], abs=1e-6)
    assert np.allclose(got.reshape(-1, 4), ref, rtol=1e-4, atol=2e-5)


def test_image_save_png(gpu_ctx, tmp_path):
    xyz, wgt, _ = random_film(20, 30, seed=5, nan=False)
    img = Image(str(tmp_path / "t.png"), (30, 20))
    img.accumulate(xyz.astype(np.float64), wgt.astype(np.float64))
    img.toneMapping = "CompressY"
    img.save()
    b = (tmp_path / "t.png").read_bytes()
    # IHDR, IDAT (single), IEND
    n = int.from_bytes(b[33:37], "big")
    raw = np.frombuffer(zlib.decompress(b[41:41 + n]), np.uint8).reshape(20, 1 + 120)[:, 1:].reshape(20, 30, 4)
    ref, _ = F.tonemap(xyz.reshape(3, -1).astype(np.float64), wgt.ravel().astype(np.float64), F.COMPRESS_Y,
                       space=space_dict(la.sRGB))
    r8 = F.quantize8(ref).reshape(20, 30, 4)
    assert (np.abs(raw.astype(int) - r8.astype(int)) <= 1).all()
    assert (raw == r8).mean() > 0.99
    assert img.lastTonemap["gain"] > 0
