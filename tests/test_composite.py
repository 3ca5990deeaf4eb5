"""TriangleMeshComposite (src/liar/scenery/triangle_mesh_composite.cpp): the
children's triangles as one mesh, hits reported as (child, local index).

Pinned on the CPU:
- mesh() concatenates the children with rebased vertex indices and
  child_of() inverts the back links;
- the rule is the one-mesh rule: a composite whose nearest triangle lies
  inside LiAR's near tolerance reports a miss although another child has a
  valid hit further on (where a List of the same meshes would report that
  hit) -- the oracle's brute force over mesh() shows both answers;
- compile_scene gives a composite the arrays and per-child shaders of the
  same meshes listed one by one, a child without a shader taking the
  composite's.
"""
import numpy as np
import pytest

import liar_amd as la
import oracle
from liar_amd import scenery, shaders


def quad(z, x0=0.0):
    v = np.array([[x0, 0, z], [x0 + 1, 0, z], [x0 + 1, 1, z], [x0, 1, z]], np.float32)
    return scenery.TriangleMesh(v, (), (), [(0, 1, 2), (0, 2, 3)])


def test_mesh_and_child_of_round_trip():
    a, b, c = quad(1.0), quad(2.0), quad(3.0, x0=5.0)
    comp = scenery.TriangleMeshComposite([a, b])
    comp.add(c)
    v, t = comp.mesh()
    assert v.shape == (12, 3) and t.shape == (6, 3)
    assert np.array_equal(t[2:4], b.triangles() + 4) and np.array_equal(t[4:], c.triangles() + 8)
    assert list(comp.offsets()) == [0, 2, 4]
    assert [comp.child_of(k) for k in range(6)] == [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)]
    with pytest.raises(IndexError):
        comp.child_of(6)
    with pytest.raises(TypeError):
        comp.add(scenery.Sphere())


def test_one_mesh_rule_versus_a_list_of_meshes():
    near = np.float32(1.0)
    zA = near + np.float32(2.5e-7)  # inside near * (1 + 1e-6): rejected by the interval check
    a, b = quad(zA), quad(2.0)
    comp = scenery.TriangleMeshComposite([a, b])
    r = np.zeros((8, 1), np.float32)
    r[0:3, 0] = [0.3, 0.2, 0.0]
    r[5, 0] = 1.0
    r[6, 0] = near
    r[7, 0] = np.inf
    v, t = comp.mesh()
    prim = oracle.brute_closest(v, t, None, r)[0]
    assert prim[0] == oracle.MISS  # the composite: one mesh, its nearest triangle is doomed
    # the same meshes as separate objects of a List: mesh A misses, mesh B is hit at t = 2
    va, ta = a.vertices(), a.triangles()
    vb, tb = b.vertices(), b.triangles()
    assert oracle.brute_closest(va, ta, None, r)[0][0] == oracle.MISS
    pb, tt = oracle.brute_closest(vb, tb, None, r)[:2]
    assert pb[0] != oracle.MISS and tt[0] == 2.0
    # an ordinary ray reports the child and its local triangle
    r2 = r.copy()
    r2[6, 0] = 0.0
    k = int(oracle.brute_closest(v, t, None, r2)[0][0])
    assert comp.child_of(k)[0] == 0


def test_compile_scene_matches_listed_meshes_and_shaders():
    red = shaders.Lambert(0.6)
    grey = shaders.Lambert(0.3)
    a, b = quad(1.0), quad(2.0)
    a.shader = red
    comp = scenery.TriangleMeshComposite([a, b])
    comp.shader = grey  # b has no shader of its own
    listed_b = quad(2.0)
    listed_b.shader = grey
    c1 = la.compile_scene(scenery.List([comp]))
    c2 = la.compile_scene(scenery.List([a, listed_b]))
    assert np.array_equal(c1.verts, c2.verts) and np.array_equal(c1.tris, c2.tris)
    assert np.array_equal(c1.prim_material, c2.prim_material)
    assert len(c1.materials) == 2 and c1.prim_material[0] != c1.prim_material[2]


@pytest.mark.gpu
def test_composite_ray_queries_on_gpu(gpu_ctx):
    """A composite of three random meshes traced through lr_qbvh over its
    merged arrays: ids bit-exact against the oracle, and every hit maps back
    to a child whose own triangle, transformed by nothing, is the one hit."""
    import torch

    print("[gpu] composite queries", flush=True)
    meshes = []
    for k in range(3):
        v, t = la.random_triangles(2000 + 500 * k, seed=40 + k)
        meshes.append(scenery.TriangleMesh(v, (), (), t))
    comp = scenery.TriangleMeshComposite(meshes)
    v, t = comp.mesh()
    b = la.Qbvh(gpu_ctx, v, t)
    r = la.random_rays(30000, seed=44)
    prim, tuv = b.trace_closest(torch.from_numpy(r).cuda())
    torch.cuda.synchronize()
    prim = prim.cpu().numpy().view(np.uint32)
    ref = oracle.brute_closest(v, t, None, r)
    assert np.array_equal(prim, ref[0]) and np.array_equal(tuv.cpu().numpy()[0].view(np.uint32), ref[1].view(np.uint32))
    hit = np.nonzero(prim != oracle.MISS)[0]
    assert hit.size > 1000
    for k in hit[:200]:
        c, j = comp.child_of(int(prim[k]))
        assert np.array_equal(meshes[c].vertices()[meshes[c].triangles()[j]], v[t[prim[k]]])
    b.close()
