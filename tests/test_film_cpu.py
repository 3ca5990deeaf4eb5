"""Film output on the host side of the C-ABI (no GPU needed):
lr_rgb_space_matrices against the FP64 oracle, and every image codec of
lr_image_write decoded back from its public file format (PNG chunks and
zlib stream with their CRC / Adler checks, the EXR header and half
scanlines, Radiance RGBE, PFM, TGA)."""
import struct
import zlib

import numpy as np
import pytest

import liar_amd as la
import oracle.film as F
from liar_amd.output import write_image


def test_rgb_space_matrices_match_oracle():
    for sp in (la.sRGB, la.CIEXYZ, la.RgbSpace((0.64, 0.33), (0.21, 0.71), (0.15, 0.06), (0.3127, 0.329), 2.2)):
        m, mi = F.rgb_space_matrices(sp.red, sp.green, sp.blue, sp.white)
        assert np.abs(sp._rgb_to_xyz - m).max() < 1e-6
        assert np.abs(sp._xyz_to_rgb - mi).max() < 5e-6


def test_rgb_space_rejects_bad_chromaticities():
    with pytest.raises(la.LiarError, match="chromaticity"):
        la.RgbSpace((0.9, 0.3), (0.3, 0.6), (0.15, 0.06), (0.3127, 0.329)).validate()
    with pytest.raises(ValueError, match="gamma"):
        la.RgbSpace((0.64, 0.33), (0.3, 0.6), (0.15, 0.06), (0.3127, 0.329), 0.0)
    bad = la.RgbSpace((0.64, 0.33), (0.3, 0.6), (0.15, 0.06), (0.3127, 0.329))
    bad.gamma = 0.0
    with pytest.raises(la.LiarError, match="gamma"):
        bad.validate()


def sample_rgba(h=5, w=7, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.2, 1.3, (h, w, 4)).astype(np.float32)
    a[0, 0] = (0.0, 0.5, 1.0, 1.0)
    a[1, 2] = (254.5 / 255, 1.0 / 255, 0.99999, 0.0)
    return a


def test_png_roundtrip(tmp_path):
    rgba = sample_rgba()
    p = tmp_path / "x.png"
    write_image(p, rgba)
    b = p.read_bytes()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(b):
        n = struct.unpack(">I", b[pos:pos + 4])[0]
        typ, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", b[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(typ + data) & 0xFFFFFFFF
        chunks.append((typ, data))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (w, h, depth, ctype, comp, filt, inter) == (7, 5, 8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])  # checks the Adler-32 too
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    assert np.array_equal(rows[:, 1:].reshape(h, w, 4), F.quantize8(rgba))


def half_bits(x):
    with np.errstate(over="ignore"):
        return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def test_exr_layout_and_values(tmp_path):
    rgba = sample_rgba()
    rgba[2, 3] = (7e4, -3e-6, 1e-9, 65504.0)  # overflow, subnormal, underflow, max half
    p = tmp_path / "x.exr"
    write_image(p, rgba, la.sRGB.linearSpace())
    b = p.read_bytes()
    assert struct.unpack("<II", b[:8]) == (20000630, 2)
    pos, attrs = 8, {}
    while b[pos] != 0:
        e = b.index(0, pos)
        name = b[pos:e].decode()
        e2 = b.index(0, e + 1)
        typ = b[e + 1:e2].decode()
        n = struct.unpack("<I", b[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    pos += 1
    for req in ("channels", "compression", "dataWindow", "displayWindow", "lineOrder", "pixelAspectRatio",
                "screenWindowCenter", "screenWindowWidth"):
        assert req in attrs
    assert attrs["compression"][1] == b"\x00"
    assert struct.unpack("<4i", attrs["dataWindow"][1]) == (0, 0, 6, 4)
    chrom = struct.unpack("<8f", attrs["chromaticities"][1])
    assert np.allclose(chrom, [0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.3127, 0.3290], atol=1e-7)
    h, w = 5, 7
    offsets = struct.unpack(f"<{h}Q", b[pos:pos + 8 * h])
    for y in range(h):
        o = offsets[y]
        yy, size = struct.unpack("<ii", b[o:o + 8])
        assert yy == y and size == 8 * w
        px = np.frombuffer(b[o + 8:o + 8 + size], np.uint16).reshape(4, w)  # A, B, G, R
        for c, row in zip((3, 2, 1, 0), px):
            assert np.array_equal(row, half_bits(rgba[y, :, c])), (y, c)


def test_rgbe_pfm_tga(tmp_path):
    rgba = np.abs(sample_rgba()) + 0.01
    h, w = rgba.shape[:2]
    p = tmp_path / "x.hdr"
    write_image(p, rgba)
    b = p.read_bytes()
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 5 +X 7\n"
    assert b.startswith(head)
    e = np.frombuffer(b[len(head):], np.uint8).reshape(h, w, 4).astype(np.float64)
    unit = np.ldexp(1.0, (e[..., 3] - 136).astype(int))[..., None]  # value of one mantissa step
    c = rgba[..., :3].astype(np.float64)
    # each mantissa is the component truncated to the shared exponent's steps
    assert (e[..., :3] * unit <= c * (1 + 1e-6)).all() and (c < (e[..., :3] + 1) * unit * (1 + 1e-6)).all()
    assert (e[..., :3].max(-1) >= 128).all()  # the largest component is normalised
    p = tmp_path / "x.pfm"
    write_image(p, rgba)
    b = p.read_bytes()
    head = b"PF\n7 5\n-1.0\n"
    assert b.startswith(head)
    f = np.frombuffer(b[len(head):], "<f4").reshape(h, w, 3)[::-1]
    assert np.array_equal(f, rgba[..., :3])
    p = tmp_path / "x.tga"
    write_image(p, rgba)
    b = p.read_bytes()
    assert b[2] == 2 and struct.unpack("<HH", b[12:16]) == (7, 5) and b[16] == 32 and b[17] == 0x28
    px = np.frombuffer(b[18:], np.uint8).reshape(h, w, 4)
    assert np.array_equal(px[..., [2, 1, 0, 3]], F.quantize8(rgba))


def test_unknown_extension(tmp_path):
    with pytest.raises(la.LiarError, match="no image codec"):
        write_image(tmp_path / "x.bmp", sample_rgba())
