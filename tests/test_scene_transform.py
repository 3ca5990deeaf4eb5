"""scenery.Transformation in render scenes: the engine bakes the
transform into the child's geometry (src/liar/scenery/transformation.cpp
semantics for a single placement: a point p of the child is at L p)."""
import numpy as np
import pytest

import liar_amd as la
from liar_amd import scenery, shaders, textures


def rot_z(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def affine(A, t):
    m = np.eye(4)
    m[:3, :3] = A
    m[:3, 3] = t
    return m


def test_mesh_parallelogram_sphere_and_light_are_baked():
    mesh = scenery.TriangleMesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [], [], [(0, 1, 2)])
    quad = scenery.Parallelogram((0, 0, 1), (1, 0, 0), (0, 1, 0))
    ball = scenery.Sphere((1, 2, 3), 0.5)
    surf = scenery.Parallelogram((0, 0, 5), (1, 0, 0), (0, -1, 0))
    light = scenery.LightArea(surf)
    for o in (mesh, quad, ball, surf):
        o.shader = shaders.Lambert(textures.Constant(0.5))
    M = affine(2.0 * rot_z(0.3), (1.0, -2.0, 0.5))
    sc = scenery.Transformation(scenery.List([mesh, quad, ball, light]), M)
    cs = la.compile_scene(sc)
    ref = la.compile_scene(scenery.List([mesh, quad, ball, light]))
    want = ref.verts.astype(np.float64) @ M[:3, :3].T + M[:3, 3]
    assert np.allclose(cs.verts, want, atol=1e-5)
    assert np.array_equal(cs.tris, ref.tris)
    assert np.allclose(cs.spheres[0, :3], M[:3, :3] @ [1, 2, 3] + M[:3, 3], atol=1e-5)
    assert cs.spheres[0, 3] == pytest.approx(1.0)  # radius 0.5 x uniform scale 2
    L = cs.lights[0]
    assert np.allclose(L["origin"], M[:3, :3] @ [0, 0, 5] + M[:3, 3])
    assert np.allclose(L["dx"], M[:3, :3] @ [1, 0, 0]) and np.allclose(L["dy"], M[:3, :3] @ [0, -1, 0])


def test_nested_transformations_compose():
    mesh = scenery.TriangleMesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [], [], [(0, 1, 2)])
    mesh.shader = shaders.Lambert(textures.Constant(0.5))
    A, B = affine(rot_z(0.4), (1, 0, 0)), affine(np.diag([1.0, 2.0, 3.0]), (0, 0, 1))
    cs = la.compile_scene(scenery.Transformation(scenery.Transformation(mesh, B), A))
    C = A @ B
    want = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0)], np.float64) @ C[:3, :3].T + C[:3, 3]
    assert np.allclose(cs.verts, want, atol=1e-6)


def test_sphere_under_non_uniform_scale_is_rejected():
    ball = scenery.Sphere((0, 0, 0), 1.0)
    ball.shader = shaders.Lambert(textures.Constant(0.5))
    with pytest.raises(NotImplementedError):
        la.compile_scene(scenery.Transformation(ball, affine(np.diag([1.0, 2.0, 1.0]), (0, 0, 0))))


def test_singular_or_projective_transforms_are_rejected():
    with pytest.raises(ValueError):
        scenery.Transformation(scenery.List(), np.zeros((4, 4)))
    m = np.eye(4)
    m[3, 0] = 1.0
    with pytest.raises(ValueError):
        scenery.Transformation(scenery.List(), m)
