"""The pbrt-v3 importer (liar_amd.tools.pbrt, mirroring liar.tools.pbrt).

The reference ships no pbrt scenes or tests, so the importer is pinned by
hand-written files whose scene graph is fixed by pbrt's rules: transform
order (each statement post-multiplies the current transform), LookAt's
camera frame, pbrt's fov on the shorter image side, attribute/transform
blocks, object instancing, named materials, area lights and the
combination of a file's meshes into one TriangleMeshComposite.  The GPU test
renders a white furnace (a Kd = 1 sphere under a unit sky: every pixel 1) and
checks the image's orientation against pbrt's left-handed camera frame.
"""
import math

import numpy as np
import pytest

import liar_amd as la
from liar_amd import scenery, shaders, tracers
from liar_amd.tools import pbrt

QUAD = '"point P" [-1 0 -1  1 0 -1  1 0 1  -1 0 1] "integer indices" [0 1 2  0 2 3]'


def scene_of(text, **kw):
    s = pbrt.PbrtScene(**kw)
    s.parse_string(text)
    return s


def test_tokens_comments_lists_and_types():
    toks = list(pbrt._tokens("t", 'Shape "sphere" # a comment\n "float radius" [ 2.5 ]\n "bool x" "true" [1 -2 .5e1]'))
    kinds = [t.kind for t in toks]
    assert kinds == ["ident", "string", "param", "lbr", "number", "rbr", "param", "bool", "lbr", "number", "number",
                     "number", "rbr"]
    assert toks[2].value == ("float", "radius") and toks[4].value == 2.5 and toks[7].value is True
    assert [t.value for t in toks[9:12]] == [1, -2, 5.0] and toks[9].line == 3


def test_transform_order_blocks_and_coordinate_systems():
    s = scene_of(f"""
        WorldBegin
        AttributeBegin
          Translate 1 0 0
          Rotate 90 0 0 1
          Shape "sphere" "float radius" 0.5
          CoordinateSystem "here"
        AttributeEnd
        TransformBegin
          Scale 2 2 2
          Shape "sphere"
        TransformEnd
        CoordSysTransform "here"
        Translate 1 0 0
        Shape "sphere"
        WorldEnd""")
    cs = la.compile_scene(s.engine.scene)
    # Translate then Rotate: the rotation acts first, the translation last
    c = cs.spheres
    assert np.allclose(c[0], [1, 0, 0, 0.5], atol=1e-6)
    assert np.allclose(c[1], [0, 0, 0, 2], atol=1e-6)
    # in "here" (T R), a further x translation moves along the rotated x axis (+y)
    assert np.allclose(c[2], [1, 1, 0, 1], atol=1e-6)


def test_transform_and_concat_transform_are_column_major():
    s = scene_of("""
        WorldBegin
        Transform [1 0 0 0  0 1 0 0  0 0 1 0  3 4 5 1]
        ConcatTransform [2 0 0 0  0 2 0 0  0 0 2 0  0 0 0 1]
        Shape "sphere" "float radius" 1
        WorldEnd""")
    assert np.allclose(la.compile_scene(s.engine.scene).spheres[0], [3, 4, 5, 2])


def test_camera_film_sampler_integrator():
    s = scene_of("""
        LookAt 0 1 -5  0 1 0  0 1 0
        Camera "perspective" "float fov" [40] "float hither" 0.05
        Film "image" "integer xresolution" [64] "integer yresolution" [48] "string filename" "out.png"
        Sampler "stratified" "integer xsamples" 2 "integer ysamples" 3
        Integrator "directlighting" "integer maxdepth" 3
        WorldBegin
        Shape "sphere"
        WorldEnd""")
    e = s.engine
    assert np.allclose(e.camera.position, (0, 1, -5)) and np.allclose(e.camera.direction, (0, 0, 1))
    assert np.allclose(e.camera.sky, (0, 1, 0)) and e.camera.nearLimit == 0.05
    aspect = 64 / 48
    assert e.camera.aspectRatio == pytest.approx(aspect)
    # pbrt's fov spans the shorter (vertical) side; the camera's spans the width
    assert e.camera.fovAngle == pytest.approx(2 * math.atan(math.tan(math.radians(20)) * aspect))
    assert e.sampler.resolution == (64, 48) and e.sampler.samplesPerPixel == 6
    assert isinstance(e.tracer, tracers.DirectLighting) and e.tracer.maxRayGeneration == 3
    assert e.target.filename == "out.png" and e.target.resolution == (64, 48)


def test_materials_named_materials_and_fallback():
    s = scene_of(f"""
        WorldBegin
        MakeNamedMaterial "gold" "string type" "metal" "spectrum eta" [400 1.5 700 0.2]
            "spectrum k" [400 1.9 700 3.8] "float roughness" 0.1
        Material "glass" "float index" 1.33
        Shape "trianglemesh" {QUAD}
        NamedMaterial "gold"
        Shape "sphere"
        Material "plastic" "rgb Kd" [0.2 0.4 0.6]
        Shape "trianglemesh" "point P" [-1 3 -1  1 3 -1  1 3 1  -1 3 1] "integer indices" [0 1 2  0 2 3]
        Material "mirror"
        Shape "sphere" "float radius" 0.25
        WorldEnd""")
    objs = s.engine.scene.children
    comp = [o for o in objs if isinstance(o, scenery.TriangleMeshComposite)]
    assert len(comp) == 1 and len(comp[0].children()) == 2  # the two meshes combined
    glass = comp[0].children()[0].shader
    assert isinstance(glass, shaders.Dielectric) and glass.innerRefractionIndex.scalar() == pytest.approx(1.33)
    plastic = comp[0].children()[1]
    assert isinstance(plastic, scenery.TriangleMesh)
    sph = [o for o in objs if isinstance(o, scenery.Sphere)]
    gold = sph[0].shader
    assert isinstance(gold, shaders.CookTorrance) and isinstance(gold.mdf, shaders.MicrofacetTrowbridgeReitz)
    assert gold.roughnessU.scalar() == pytest.approx(pbrt.remap_roughness(0.1))
    cs = la.compile_scene(s.engine.scene)
    assert len(cs.tris) == 4 and len(cs.spheres) == 2
    # the plastic quad (3 up) falls back to matte with its Kd
    assert np.allclose(cs.verts[4:, 1], 3.0)
    kinds = [cs.materials[m]["type"] for m in cs.prim_material]
    assert len(set(kinds)) == 4  # dielectric, lambert, cook-torrance, mirror


def test_area_light_and_infinite_light():
    s = scene_of(f"""
        WorldBegin
        LightSource "infinite" "rgb L" [0.5 0.5 0.5] "integer nsamples" 4
        AttributeBegin
          Translate 0 2 0
          AreaLightSource "diffuse" "blackbody L" [3000 2] "integer nsamples" 9
          Shape "trianglemesh" {QUAD}
        AttributeEnd
        Shape "sphere"
        WorldEnd""")
    objs = s.engine.scene.children
    sky = [o for o in objs if isinstance(o, scenery.LightSky)][0]
    # "rgb L" [0.5 0.5 0.5]: rgb() of the linear-sRGB grey, the XYZ of E white at Y = 0.5
    np.testing.assert_allclose(sky.radiance.spectrum.value, (0.5, 0.5, 0.5), atol=1e-7)
    assert sky.numberOfEmissionSamples == 4
    placed = [o for o in objs if isinstance(o, scenery.Transformation)][0]
    light = placed.child
    assert isinstance(light, scenery.LightArea) and light.numberOfEmissionSamples == 9
    assert isinstance(light.surface.shader, shaders.Unshaded)
    cs = la.compile_scene(s.engine.scene)
    (L,) = [li for li in cs.lights if li["type"] == la._native.LIGHT_PARALLELOGRAM]
    o, a, b = (np.asarray(L[k], np.float64) for k in ("origin", "dx", "dy"))
    corners = {tuple(np.round(p, 6)) for p in (o, o + a, o + b, o + a + b)}
    assert corners == {(-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)}
    # the mesh's winding (0 1 2: normal -y) is kept
    assert np.dot(np.cross(a, b), [0, -1, 0]) > 0
    # the mesh stays visible as an emitter in the scene geometry
    assert len(cs.tris) == 2 and cs.materials[cs.prim_material[0]]["light"] != la._native.LR_NONE


def test_object_instancing_and_include(tmp_path):
    inc = tmp_path / "shapes.pbrt"
    inc.write_text('ObjectBegin "ball"\n  Shape "sphere" "float radius" 0.5\nObjectEnd\n')
    main = tmp_path / "main.pbrt"
    main.write_text('WorldBegin\nInclude "shapes.pbrt"\nTranslate 2 0 0\nObjectInstance "ball"\n'
                    'Translate 0 0 3\nObjectInstance "ball"\nWorldEnd\n')
    s = pbrt.parse(str(main))
    cs = la.compile_scene(s.engine.scene)
    assert np.allclose(cs.spheres, [[2, 0, 0, 0.5], [2, 0, 3, 0.5]])


def test_remap_roughness_fits_pbrt_alpha():
    """LiAR's polynomial is a fit of sqrt(alpha) with pbrt-v3's published
    RoughnessToAlpha; the two agree to 1% over the useful range, and without
    remapping roughness is alpha."""
    for r in np.geomspace(0.005, 1.0, 25):
        x = math.log(r)
        alpha = 1.62142 + 0.819955 * x + 0.1734 * x * x + 0.0171201 * x ** 3 + 0.000640711 * x ** 4
        assert pbrt.remap_roughness(r) == pytest.approx(math.sqrt(alpha), rel=1e-2)
    assert pbrt.remap_roughness(0.04, remaproughness=False) == pytest.approx(0.2)
    assert pbrt.remap_roughness(1e-9) == pbrt.remap_roughness(1e-3)


def test_errors():
    with pytest.raises(NotImplementedError, match="Shape 'cone'"):
        scene_of('WorldBegin\nShape "cone"\nWorldEnd')
    with pytest.raises(NotImplementedError, match="Texture"):
        scene_of('WorldBegin\nTexture "t" "spectrum" "imagemap"\nWorldEnd')
    with pytest.raises(pbrt.ParseError, match="unterminated"):
        scene_of('WorldBegin\nShape "sphere" "float radius" [1\n')
    with pytest.raises(pbrt.ParseError, match="only between WorldBegin"):
        scene_of('Shape "sphere"')
    with pytest.raises(NotImplementedError, match="parallelogram"):
        scene_of('WorldBegin\nAreaLightSource "diffuse"\nShape "sphere"\nWorldEnd')


def test_mirrored_target_reverses_columns():
    class T:
        resolution = (3, 2)

        def accumulate(self, x, w):
            self.x, self.w = x, w

    t = T()
    m = pbrt._MirroredTarget(t)
    x = np.arange(18, dtype=np.float64).reshape(3, 2, 3)
    m.accumulate(x, x[0])
    assert np.array_equal(t.x[:, :, 0], x[:, :, 2]) and np.array_equal(t.w, x[0][:, ::-1])
    assert m.resolution == (3, 2)


@pytest.mark.gpu
def test_pbrt_render_furnace_and_orientation(gpu_ctx, tmp_path):
    """A Kd = 1 sphere under a unit sky renders 1 inside its disc (one bounce
    off a convex sphere escapes to the sky; camera rays that miss see no sky,
    as in LiAR's AdjointPhotonTracer); with a white sphere at -x and a black
    one at +x, pbrt's frame for this LookAt (right = up x forward = +x) puts
    the white one in the left half of the image."""
    print("[gpu] pbrt furnace + orientation", flush=True)
    head = f"""LookAt 0 0 -5  0 0 0  0 1 0
        Camera "perspective" "float fov" 50
        Film "image" "integer xresolution" 64 "integer yresolution" 48 "string filename" "{tmp_path / 'f.npy'}"
        Sampler "stratified" "integer xsamples" 16 "integer ysamples" 16
        Integrator "path" "integer maxdepth" 8
        WorldBegin
        LightSource "infinite" "rgb L" [1 1 1]"""
    s = scene_of(head + """
        Material "matte" "rgb Kd" [1 1 1]
        Shape "sphere" "float radius" 1.5
        WorldEnd""")
    s.render()
    y = s.image.xyz()[..., 1]
    yy, xx = np.mgrid[0:48, 0:64]
    disc = (xx - 31.5) ** 2 + (yy - 23.5) ** 2 < 9 ** 2
    Y = y[disc]
    print(f"[gpu] furnace Y in the disc: mean {Y.mean():.5f}, std {Y.std():.4f}, n {Y.size}; "
          f"corner {y[:3, :3].max():.3g}", flush=True)
    assert abs(Y.mean() - 1.0) < 2e-3 + 3 * Y.std() / np.sqrt(Y.size)
    assert (y[:3, :3] == 0).all()
    s = scene_of(head + """
        AttributeBegin
          Material "matte" "rgb Kd" [1 1 1]
          Translate -1.5 0 0
          Shape "sphere" "float radius" 0.8
        AttributeEnd
        Material "matte" "rgb Kd" [0 0 0]
        Translate 1.5 0 0
        Shape "sphere" "float radius" 0.8
        WorldEnd""")
    s.render()
    y = s.image.xyz()[..., 1]
    left, right = y[:, :32].mean(), y[:, 32:].mean()
    print(f"[gpu] white sphere at -x, black at +x: left half {left:.3f}, right half {right:.3f}", flush=True)
    assert left > 0.05 and right < 1e-6
