"""Scene objects (mirrors ``liar.scenery``: src/liar/scenery/).

Geometry is flattened by the RenderEngine into one QBVH over triangles and
spheres; aggregates (List, QbvhTree) only group children.
"""
from __future__ import annotations

import numpy as np

from .textures import as_texture


class SceneObject:
    def __init__(self):
        self.shader = None


class TriangleMesh(SceneObject):
    """``TriangleMesh(vertices, normals, uvs, triangles)``; each triangle is a
    3-tuple of vertex indices or of ``(vertex[, normal[, uv]])`` index tuples
    (src/liar/scenery/triangle_mesh.cpp:39-85), a None index meaning none.
    A triangle whose three corners name normals shades with those normals
    interpolated at the hit (triangle_mesh.cpp:334-358); otherwise, and for a
    plain vertex index, with its geometric normal.  Texture coordinates are
    kept but not used (the engine has constant textures only)."""

    def __init__(self, vertices, normals=(), uvs=(), triangles=()):
        super().__init__()
        self._vertices = np.asarray(vertices, dtype=np.float32).reshape(-1, 3)
        self._normals = normals if isinstance(normals, np.ndarray) else list(normals)
        self._uvs = uvs if isinstance(uvs, np.ndarray) else list(uvs)
        self._normal_idx = None  # (nt, 3) int64, -1 = no normal
        if isinstance(triangles, np.ndarray) and triangles.ndim == 2 and triangles.shape[1] == 3:
            self._triangles = np.ascontiguousarray(triangles, dtype=np.uint32)
        else:
            idx, nidx = [], []
            for tri in triangles:
                if len(tri) != 3:
                    raise ValueError("a triangle needs three corners")
                idx.append([c[0] if isinstance(c, (tuple, list)) else c for c in tri])
                nidx.append([c[1] if isinstance(c, (tuple, list)) and len(c) > 1 and c[1] is not None else -1
                             for c in tri])
            self._triangles = np.asarray(idx, dtype=np.uint32).reshape(-1, 3)
            nidx = np.asarray(nidx, dtype=np.int64).reshape(-1, 3)
            if (nidx >= 0).any():
                if nidx.max() >= len(self._normals):
                    raise ValueError("triangle references a missing normal")
                self._normal_idx = nidx
        if len(self._triangles) and self._triangles.max() >= len(self._vertices):
            raise ValueError("triangle references a missing vertex")

    def corner_normals(self, M=None):
        """(nt, 9) float32: each triangle's three corner normals, zero for a
        triangle that does not name all three; None when no triangle has
        normals.  M: a local-to-world 4x4 matrix; normals transform by its
        inverse transpose (LiAR's normalTransform)."""
        if self._normal_idx is None:
            return None
        ns = np.asarray(self._normals, np.float64).reshape(-1, 3)
        if M is not None:
            ns = ns @ np.linalg.inv(np.asarray(M, np.float64)[:3, :3])  # rows: (M^-T n)^T
        ok = (self._normal_idx >= 0).all(axis=1)
        out = np.zeros((len(self._triangles), 9), np.float64)
        out[ok] = ns[self._normal_idx[ok]].reshape(-1, 9)
        return out.astype(np.float32)

    @staticmethod
    def loadPly(path, fan_all=False):
        """``TriangleMesh.loadPly(path)`` (src/liar/scenery/triangle_mesh.cpp:183-281)
        through the library's PLY reader (lr_load_ply).  Like LiAR, a polygon
        face keeps only its last fan triangle; ``fan_all=True`` keeps all."""
        import ctypes as C

        from . import _native as N

        m = N.MeshData()
        N.check(N.lib().lr_load_ply(str(path).encode(), N.LR_PLY_FAN_ALL if fan_all else 0, C.byref(m)),
                "lr_load_ply")
        try:
            nv, nt = m.n_verts, m.n_tris

            def arr(ptr, n, dtype):
                if not ptr or n == 0:
                    return np.zeros((0,), dtype)
                return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

            verts = arr(m.verts, 3 * nv, np.float32).reshape(-1, 3)
            normals = arr(m.normals, 3 * nv, np.float32).reshape(-1, 3)
            uvs = arr(m.uvs, 2 * nv, np.float32).reshape(-1, 2)
            tris = arr(m.tris, 3 * nt, np.uint32).reshape(-1, 3)
        finally:
            N.lib().lr_mesh_data_free(C.byref(m))
        return TriangleMesh(verts, normals, uvs, tris)

    def vertices(self):
        return self._vertices

    def triangles(self):
        return self._triangles

    def normals(self):
        return self._normals

    def uvs(self):
        return self._uvs


class TriangleMeshComposite(SceneObject):
    """``TriangleMeshComposite(children)``
    (src/liar/scenery/triangle_mesh_composite.cpp:94-183): the triangles of
    several TriangleMesh children in one mesh, each child keeping its own
    shader.  An intersection applies the mesh rule to all of them at once --
    the nearest triangle with t > near, then LiAR's interval check on it --
    and reports the child and the triangle's index in that child (:186-216).

    The engine merges every mesh of a scene into one triangle set, which is
    this rule; ``mesh()`` gives the merged arrays for ray queries
    (``liar_amd.Qbvh``) and ``child_of(k)`` maps a hit's triangle index back."""

    def __init__(self, children=()):
        super().__init__()
        self._children = []
        self.add(children)

    def add(self, child):
        """Adds a TriangleMesh or a sequence of them (``add(child)`` /
        ``add(children)``, :54-64)."""
        items = child if isinstance(child, (list, tuple)) else [child]
        for c in items:
            if not isinstance(c, TriangleMesh):
                raise TypeError("TriangleMeshComposite children are TriangleMesh objects")
            self._children.append(c)

    def children(self):
        return list(self._children)

    def offsets(self):
        """First merged triangle index of each child (the back links, :172)."""
        return np.cumsum([0] + [len(c.triangles()) for c in self._children[:-1]]).astype(np.int64)

    def mesh(self):
        """(vertices float32 (nv, 3), triangles uint32 (nt, 3)): children's
        arrays concatenated in order, vertex indices rebased (:127-173)."""
        verts, tris, base = [], [], 0
        for c in self._children:
            verts.append(np.asarray(c.vertices(), np.float32).reshape(-1, 3))
            tris.append(np.asarray(c.triangles(), np.uint32).reshape(-1, 3) + np.uint32(base))
            base += len(verts[-1])
        if not verts:
            return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
        return np.concatenate(verts), np.concatenate(tris)

    def child_of(self, k):
        """Merged triangle index -> (child index, triangle index in the child)."""
        off = self.offsets()
        c = int(np.searchsorted(off, k, side="right") - 1)
        if k < 0 or c < 0 or k - off[c] >= len(self._children[c].triangles()):
            raise IndexError("triangle index out of range")
        return c, int(k - off[c])


class Sphere(SceneObject):
    def __init__(self, center=(0.0, 0.0, 0.0), radius=1.0):
        super().__init__()
        self.center = tuple(float(c) for c in center)
        self.radius = float(radius)


class Parallelogram(SceneObject):
    """support + u * sizeU + v * sizeV, u, v in [0, 1]; normal = sizeU x sizeV
    (src/liar/scenery/parallelogram.cpp)."""

    def __init__(self, support, sizeU, sizeV):
        super().__init__()
        self.support = np.asarray(support, dtype=np.float32)
        self.sizeU = np.asarray(sizeU, dtype=np.float32)
        self.sizeV = np.asarray(sizeV, dtype=np.float32)

    def _mesh(self):
        o, a, b = self.support, self.sizeU, self.sizeV
        v = np.stack([o, o + a, o + a + b, o + b]).astype(np.float32)
        return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


class List(SceneObject):
    def __init__(self, children=()):
        super().__init__()
        self.children = list(children)

    def add(self, child):
        self.children.append(child)


class Transformation(SceneObject):
    """``Transformation(child, localToWorld)`` (src/liar/scenery/transformation.cpp):
    places ``child`` with an affine local-to-world matrix (4x4, or 3x4).

    For rendering, the engine bakes the transform into the child's geometry
    (vertices, parallelograms and, for a uniform scale, sphere centres and
    radii).  Ray queries over many placements of shared geometry use the
    two-level tree of ``liar_amd.instancing.Tlas`` instead."""

    def __init__(self, child, localToWorld):
        super().__init__()
        self.child = child
        m = np.asarray(localToWorld, dtype=np.float64)
        if m.shape == (3, 4):
            m = np.vstack([m, [0.0, 0.0, 0.0, 1.0]])
        if m.shape != (4, 4) or not np.allclose(m[3], [0.0, 0.0, 0.0, 1.0]):
            raise ValueError("localToWorld must be an affine 4x4 (or 3x4) matrix")
        if abs(np.linalg.det(m[:3, :3])) == 0.0:
            raise ValueError("localToWorld is singular")
        self.localToWorld = m

    def worldToLocal(self):
        return np.linalg.inv(self.localToWorld)


class QbvhTree(List):
    """Object-level QBVH (src/liar/scenery/object_trees.h); here the engine
    builds one QBVH over all primitives, so this is a grouping node."""


class AabbTree(List):
    pass


class LightArea(SceneObject):
    """Area light over a surface (src/liar/scenery/light_area.cpp): a
    Parallelogram, or a TriangleMesh of at most 4096 triangles (sampled by
    area, triangle_mesh.cpp:414-444).  As in LiAR, a single-sided light is
    sampled on its front face only, but emits from both faces when a ray hits
    it."""

    def __init__(self, surface):
        super().__init__()
        if not isinstance(surface, (Parallelogram, TriangleMesh)):
            raise TypeError("LightArea surfaces are limited to Parallelogram and TriangleMesh")
        self.surface = surface
        self.radiance = as_texture(1.0)
        self.isDoubleSided = False
        self.numberOfEmissionSamples = 9  # light_area.cpp:50 (DirectLighting)


class LightSky(SceneObject):
    """Constant environment light (src/liar/scenery/light_sky.cpp)."""

    def __init__(self, radiance=None):
        super().__init__()
        self.radiance = as_texture(1.0 if radiance is None else radiance)
        self.numberOfEmissionSamples = 1  # light_sky.cpp:395 (DirectLighting)


class LightSun(SceneObject):
    """Distant disc light: radiance within ``angle`` (apparent diameter, rad)
    of ``direction`` (towards the sun).  Not a LiAR class; models the sun disc
    of the terrain workload."""

    def __init__(self, direction=(0.0, 0.0, 1.0), angle=np.radians(0.5), radiance=None):
        super().__init__()
        self.direction = tuple(float(x) for x in direction)
        self.angle = float(angle)
        self.radiance = as_texture(1.0 if radiance is None else radiance)
        self.numberOfEmissionSamples = 1
