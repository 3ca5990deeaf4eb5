"""Scene import (CPU): the PLY reader (lr_load_ply, for TriangleMesh.loadPly,
src/liar/scenery/triangle_mesh.cpp:183-281) and the OBJ loader
(liar_amd.tools.wavefront_obj, for src/liar/tools/wavefront_obj.py).

Pinned by:
- PLY: meshes written here in ascii, binary little- and big-endian with
  mixed property types decode to the same arrays; LiAR's rule that a polygon
  keeps only its last fan triangle (the reference's example cube,
  tests/golden/cube_quads.ply from examples/tools/ply.ply, gives one
  triangle per quad); malformed files raise.
- OBJ: the polygon triangulation equals the index triples of LiAR's own
  geometry.triangulate_3d on 21 polygons (tests/golden/triangulation_vectors.json,
  written by tests/golden/make_triangulation_vectors.py), and triangles cover
  each polygon's area; the reference's example OBJ/MTL
  (tests/golden/cube_groups.obj, wavefront_obj.mtl from examples/tools/) loads
  as 6 groups with their materials.
"""
import json
import struct
from pathlib import Path

import numpy as np
import pytest

import liar_amd as la
from liar_amd import scenery
from liar_amd.tools import wavefront_obj

GOLDEN = Path(__file__).resolve().parent / "golden"


def _write_ply(path, fmt, verts, faces, normals=None, uvs=None, extra_element=True):
    """verts (n,3) float, faces: lists of indices."""
    props = [("float", "x"), ("float", "y"), ("double", "z")]
    if normals is not None:
        props += [("float", "nx"), ("float", "ny"), ("float", "nz")]
    if uvs is not None:
        props += [("float", "u"), ("float", "v")]
    lines = ["ply", f"format {fmt} 1.0", "comment written by test_import"]
    if extra_element:
        lines += ["element camera 1", "property float fov", "property list uchar int tags"]
    lines += [f"element vertex {len(verts)}"] + [f"property {t} {n}" for t, n in props]
    lines += ["property uchar flags", f"element face {len(faces)}", "property list uchar int vertex_indices",
              "property ushort material", "end_header"]
    head = ("\n".join(lines) + "\n").encode()
    rows = []
    for k, v in enumerate(verts):
        r = list(v)
        if normals is not None:
            r += list(normals[k])
        if uvs is not None:
            r += list(uvs[k])
        rows.append(r)
    if fmt == "ascii":
        body = []
        if extra_element:
            body.append("0.5 2 7 9")
        body += [" ".join(repr(float(x)) for x in r) + " 1" for r in rows]
        body += [" ".join(str(x) for x in [len(f)] + list(f)) + " 3" for f in faces]
        path.write_bytes(head + ("\n".join(body) + "\n").encode())
        return
    e = "<" if fmt == "binary_little_endian" else ">"
    buf = bytearray(head)
    if extra_element:
        buf += struct.pack(e + "fBii", 0.5, 2, 7, 9)
    for r in rows:
        x, y, z, *rest = r
        buf += struct.pack(e + "ffd", x, y, z) + struct.pack(e + "f" * len(rest), *rest) + struct.pack(e + "B", 1)
    for f in faces:
        buf += struct.pack(e + "B" + "i" * len(f), len(f), *f) + struct.pack(e + "H", 3)
    path.write_bytes(bytes(buf))


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_formats_decode_identically(tmp_path, fmt):
    rng = np.random.default_rng(3)
    v = rng.normal(size=(7, 3)).astype(np.float32)
    n = rng.normal(size=(7, 3)).astype(np.float32)
    uv = rng.random((7, 2)).astype(np.float32)
    faces = [[0, 1, 2], [2, 3, 4, 5], [6, 0, 3, 1, 5]]
    p = tmp_path / f"m_{fmt}.ply"
    _write_ply(p, fmt, v, faces, n, uv)
    m = scenery.TriangleMesh.loadPly(p)
    np.testing.assert_array_equal(m.vertices(), v)
    np.testing.assert_array_equal(m.normals(), n)
    np.testing.assert_array_equal(m.uvs(), uv)
    # LiAR keeps the last fan triangle of a polygon: (i0, i_{n-2}, i_{n-1})
    assert m.triangles().tolist() == [[0, 1, 2], [2, 4, 5], [6, 1, 5]]
    full = scenery.TriangleMesh.loadPly(p, fan_all=True)
    assert full.triangles().tolist() == [[0, 1, 2], [2, 3, 4], [2, 4, 5], [6, 0, 3], [6, 3, 1], [6, 1, 5]]


def test_ply_reference_example_cube():
    m = scenery.TriangleMesh.loadPly(GOLDEN / "cube_quads.ply")
    assert m.vertices().shape == (8, 3) and len(m.normals()) == 0
    # quads "4 a b c d" -> (a, c, d)
    assert m.triangles().tolist() == [[0, 6, 4], [1, 7, 3], [4, 7, 5], [6, 3, 7], [2, 1, 3], [0, 5, 1]]
    full = scenery.TriangleMesh.loadPly(GOLDEN / "cube_quads.ply", fan_all=True)
    assert len(full.triangles()) == 12


@pytest.mark.parametrize("text, msg", [
    ("not a ply\n", "not a PLY"),
    ("ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nend_header\n0 0\n", "x, y, z"),
    ("ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\nproperty float z\n"
     "element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 1 1\n3 0 1 2\n", "out of range"),
    ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
     "end_header\n0 0 0\n1 1 1\n", "truncated"),
    ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
     "element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 1 1\n2 2 2\n2 0 1\n",
     "at least 3"),
])
def test_ply_malformed_raise(tmp_path, text, msg):
    p = tmp_path / "bad.ply"
    p.write_text(text)
    with pytest.raises(la.LiarError, match=msg):
        scenery.TriangleMesh.loadPly(p)


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian"])
def test_ply_overstated_counts_raise_without_allocating(tmp_path, fmt):
    """A header whose counts exceed what the body can hold is rejected before
    any buffer is sized by it (a few-byte file must not ask for 50 GB)."""
    head = (f"ply\nformat {fmt} 1.0\nelement vertex 4294967295\nproperty float x\nproperty float y\n"
            "property float z\nend_header\n").encode()
    p = tmp_path / "huge_v.ply"
    p.write_bytes(head + (b"0 0 0\n" if fmt == "ascii" else bytes(12)))
    with pytest.raises(la.LiarError, match="more records than the file"):
        scenery.TriangleMesh.loadPly(p)
    # one face whose list count claims 2^32 - 1 indices
    head = (f"ply\nformat {fmt} 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
            "element face 1\nproperty list uint int vertex_indices\nend_header\n").encode()
    if fmt == "ascii":
        body = b"0 0 0\n1 0 0\n0 1 0\n4294967295 0 1 2\n"
    else:
        body = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], "<f4").tobytes() + np.array([2**32 - 1, 0, 1, 2], "<u4").tobytes()
    p = tmp_path / "huge_f.ply"
    p.write_bytes(head + body)
    with pytest.raises(la.LiarError, match="longer than the file"):
        scenery.TriangleMesh.loadPly(p)


def test_triangulation_matches_liar_vectors():
    cases = json.loads((GOLDEN / "triangulation_vectors.json").read_text())
    assert len(cases) == 21
    for c in cases:
        poly = [tuple(p) for p in c["polygon"]]
        got = [list(t) for t in wavefront_obj.triangulate_3d(*poly)]
        assert got == c["triangles"], c["name"]


def test_triangulation_covers_the_polygon():
    # independent of the reference: n-2 triangles whose areas sum to the
    # polygon's (shoelace), for convex and concave planar polygons
    for poly in ([(0, 0, 0), (2, 0, 0), (2, 1, 0), (1, 1, 0), (1, 2, 0), (0, 2, 0)],
                 [(0, 0, 0), (3, 0, 0), (3, 3, 0), (2, 1, 0), (1, 3, 0), (0, 3, 0)],
                 [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)]):
        tris = wavefront_obj.triangulate_3d(*poly)
        assert len(tris) == len(poly) - 2
        P = np.array(poly, float)[:, :2]
        shoelace = 0.5 * abs(np.dot(P[:, 0], np.roll(P[:, 1], -1)) - np.dot(P[:, 1], np.roll(P[:, 0], -1)))
        def tri_area(a, b, c):
            u, w = P[b] - P[a], P[c] - P[a]
            return 0.5 * abs(u[0] * w[1] - u[1] * w[0])

        area = sum(tri_area(a, b, c) for a, b, c in tris)
        assert abs(area - shoelace) < 1e-12


def test_obj_reference_example_groups():
    groups, mats = wavefront_obj.load(GOLDEN / "cube_groups.obj")
    assert len(groups) == 6
    assert sorted(mats) == ["", "black", "blue", "green", "red", "white", "yellow"]
    for g in groups:
        assert g.vertices().shape == (4, 3) and g.triangles().shape == (2, 3)
    # materials follow usemtl: black, white, red, yellow, green, blue, each
    # Lambert(Constant(rgb(Kd))) in sRGB: the XYZ of the Kd colour
    from liar_amd import sRGB

    kds = [(0, 0, 0), (1, 1, 1), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1)]
    for g, kd in zip(groups, kds):
        np.testing.assert_allclose(g.shader.diffuse.spectrum.value, sRGB.toXYZ(kd), atol=1e-12)
    # every group is a square of area 4 (side 2)
    for g in groups:
        v = g.vertices().astype(float)
        area = sum(0.5 * np.linalg.norm(np.cross(v[b] - v[a], v[c] - v[a])) for a, b, c in g.triangles())
        assert abs(area - 4.0) < 1e-12
