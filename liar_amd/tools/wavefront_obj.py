"""Wavefront OBJ import, mirroring ``liar.tools.wavefront_obj``
(src/liar/tools/wavefront_obj.py:26-318).

``load(filename)`` / ``decode(lines)`` return ``(groups, materials)``: one
``scenery.TriangleMesh`` per ``g`` group (vertices compressed to those its
faces use, in order of first use), each with the shader of the ``usemtl`` in
force when the group closed, and the dict of materials by name.

Faces are triangulated like LiAR's ``geometry.triangulate_3d``
(src/liar/tools/geometry.py:4-93): the face is projected along the major axis
of its fan-summed normal and ear-clipped, each step cutting the ear of
largest area (ties: the later vertex); a degenerate (zero-area) face becomes
the n cyclic triples.  When the normal's major component is negative LiAR
clips the reversed polygon and applies the indices found there to the face in
its original order; this module does the same, so the triangles match.

Materials: a library's ``Kd`` becomes a Lambert of ``rgb(Kd)`` (the colour
in the default sRGB space, recovered as a spectrum, as LiAR's loader does,
wavefront_obj.py:100-110); ``Ks``, ``Ns`` and texture maps are ignored.  ``mtllib`` paths are tried as given, then next
to the OBJ file.  Negative ``vt`` / ``vn`` indices count back from the uv /
normal lists (LiAR counts ``vt`` back from the normals and ``vn`` from the uvs).
"""
from __future__ import annotations

import math
import os

from .. import scenery, shaders, spectra, textures


def _default_material(kd=(1.0, 1.0, 1.0)):
    return shaders.Lambert(textures.Constant(spectra.rgb(kd)))


# -- polygon triangulation -------------------------------------------------------

def _sub(a, b):
    return tuple(x - y for x, y in zip(a, b))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _area2(pts):
    """Signed area of a 2D polygon (fan from its first point)."""
    if len(pts) < 3:
        return 0.0
    ax, ay = pts[0]
    s = 0.0
    for (bx, by), (cx, cy) in zip(pts[1:-1], pts[2:]):
        s += (bx - ax) * (cy - ay) - (cx - ax) * (by - ay)
    return 0.5 * s


def _inside(a, b, c, d):
    """d in the closed triangle abc (edges included)."""
    if _area2([a, b, c]) < 0:
        b, c = c, b
    return all(_area2([p, q, d]) >= 0 for p, q in ((a, b), (b, c), (c, a)))


def _clip_ears(pts):
    """Ear clipping of a counter-clockwise 2D polygon -> index triples."""
    n = len(pts)
    ring = list(range(n))
    area = _area2(pts)
    if area < 0:
        raise ValueError(f"polygon {pts!r} has negative area {area!r}")
    if area == 0:
        return [(ring[k], ring[(k + 1) % n], ring[(k + 2) % n]) for k in range(n)]
    out = []
    while len(ring) > 3:
        m = len(ring)
        best = None
        for pos in range(m):
            tri = (ring[pos - 1], ring[pos], ring[(pos + 1) % m])
            a, b, c = (pts[k] for k in tri)
            ear = _area2([a, b, c])
            if ear <= 0:
                continue
            if any(_inside(a, b, c, pts[k]) for k in ring if k not in tri):
                continue
            key = (ear, pos, tri)
            if best is None or key > best:
                best = key
        if best is None:
            raise ValueError(f"no ear to clip in polygon {pts!r}")
        out.append(best[2])
        del ring[best[1]]
    out.append(tuple(ring))
    return out


def triangulate_3d(*verts):
    """Index triples covering the 3D polygon ``verts`` (see the module text)."""
    if len(verts) < 3:
        raise ValueError("a polygon needs 3 vertices")
    a = verts[0]
    nrm = (0.0, 0.0, 0.0)
    for b, c in zip(verts[1:-1], verts[2:]):
        nrm = tuple(x + y for x, y in zip(nrm, _cross(_sub(b, a), _sub(c, a))))
    k = max((abs(nrm[q]), q) for q in range(3))[1]
    i, j = (k + 1) % 3, (k + 2) % 3
    flat = [(v[i], v[j]) for v in verts]
    if nrm[k] < 0:
        return list(reversed(_clip_ears(flat[::-1])))
    return _clip_ears(flat)


# -- OBJ / MTL -------------------------------------------------------------------

def _load_materials(path, materials, material):
    current = None
    for line in open(path):
        f = line.split()
        if not f:
            continue
        cmd, args = f[0], f[1:]
        if cmd == "newmtl":
            name = args[0] if args else ""
            if name not in materials:
                current = {"kd": (1.0, 1.0, 1.0)}
                materials[name] = current
            else:
                current = None  # an existing material is kept
        elif cmd == "Kd" and current is not None:
            if len(args) != 3:
                raise ValueError(f"material library {path}: bad Kd line {line!r}")
            current["kd"] = tuple(float(x) for x in args)
    for name, m in list(materials.items()):
        if isinstance(m, dict):
            materials[name] = material(m["kd"]) if material is _default_material else material()


def load(filename, *args, **kwargs):
    with open(filename) as f:
        return decode(f, filename, *args, **kwargs)


def decode(lines, filename="", material=None, initial_materials=None):
    material = material or _default_material
    base = os.path.dirname(os.path.abspath(filename)) if filename else os.getcwd()
    verts, normals, uvs = [], [], []
    faces, groups = [], []
    materials = dict(initial_materials or {})
    if "" not in materials:
        materials[""] = _default_material()
    current = materials[""]

    def index(tok, count, what):
        if tok is None:
            return None
        q = int(tok)
        if q > 0:
            return q - 1
        if q < 0:
            return count + q
        raise ValueError(f"obj file {filename}: 0 {what} index")

    def close_group():
        used = {}
        for face in faces:
            for v in face:
                used.setdefault(v, len(used))
        order = sorted(used, key=used.get)
        mesh = scenery.TriangleMesh([verts[v] for v in order], (), (),
                                    [tuple(used[v] for v in face) for face in faces])
        mesh.shader = current
        groups.append(mesh)

    for line in lines:
        if not isinstance(line, str):
            line = line.decode()
        f = line.split()
        if not f:
            continue
        cmd, args = f[0], f[1:]
        if cmd == "mtllib":
            for m in args:
                path = m if os.path.exists(m) else os.path.join(base, m)
                _load_materials(path, materials, material)
        elif cmd == "usemtl":
            name = args[0] if args else ""
            current = materials.get(name, materials[""])
        elif cmd == "g":
            if faces:
                close_group()
                faces = []
        elif cmd == "v":
            if len(args) != 3:
                raise ValueError(f"obj file {filename}: bad vertex line {line!r}")
            verts.append(tuple(float(x) for x in args))
        elif cmd == "vn":
            n = tuple(float(x) for x in args)
            s = math.sqrt(sum(x * x for x in n))
            normals.append(tuple(x / s for x in n) if s else n)
        elif cmd == "vt":
            uvs.append(tuple(float(x) for x in args)[:2])
        elif cmd == "f":
            if len(args) < 3:
                raise ValueError(f"obj file {filename}: face with fewer than 3 vertices: {line!r}")
            corners = []
            for tok in args:
                parts = (tok.split("/") + [None, None])[:3]
                v = index(parts[0], len(verts), "vertex")
                if v is None or not 0 <= v < len(verts):
                    raise ValueError(f"obj file {filename}: vertex index out of range: {line!r}")
                corners.append(v)
            for a, b, c in triangulate_3d(*[verts[v] for v in corners]):
                faces.append((corners[a], corners[b], corners[c]))
    if faces:
        close_group()
    return groups, materials
