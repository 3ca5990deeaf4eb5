"""The reference's scripts run against liar_amd with only their import
changed: the Cornell box tool (src/liar/tools/cornell_box.py:19-213), the
spectral data tool (src/liar/tools/spd.py:27-75), Observer.setStandard
(src/liar/__init__.py:51-69) and the script helpers used by
examples/cornell_box/direct_lighting.py."""
import io
import math

import numpy as np
import pytest

import liar_amd
import oracle
from liar_amd.observer import standard_observer
from liar_amd.tools import cornell_box, scripting, spd


def test_direct_lighting_example_builds(tmp_path):
    # examples/cornell_box/direct_lighting.py with `liar` -> `liar_amd`
    ns = {}
    exec("from liar_amd import *\nfrom liar_amd.tools import cornell_box, scripting\n", ns)
    options = scripting.renderOptions([], width=40, height=30, super_sampling=9)
    engine = ns["RenderEngine"]()
    engine.tracer = ns["tracers"].DirectLighting()
    engine.sampler = ns["samplers"].Stratifier((options.width, options.height), options.super_sampling)
    engine.scene = cornell_box.scene()
    engine.camera = cornell_box.camera()
    engine.target = scripting.makeRenderTarget(options.width, options.height, str(tmp_path / "direct.hdr"),
                                               "Cornell Box with Direct Lighting only")
    assert engine.target.resolution == (40, 30) and engine.target.toneMapping == "Linear"
    assert engine.sampler.samplesPerPixel == 9
    cs = liar_amd.compile_scene(engine.scene)
    assert cs.tris.shape == (2 + 10 + 24, 3)
    assert len(cs.lights) == 1 and cs.lights[0]["samples"] == 16
    # the blocks' corners carry their faces' normals: the geometric normals
    V, T, N = cs.verts, cs.tris, cs.tri_normals
    has = np.abs(N).sum(1) > 0
    assert has.sum() == 24
    g = np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]])
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    np.testing.assert_allclose(np.einsum("ij,ij->i", g[has], N[has, :3]), 1.0, atol=1e-6)
    cam = cornell_box.camera()
    assert cam.position == pytest.approx((2.78, 2.73, -8.0))
    assert cam.fovAngle == pytest.approx(2 * math.atan(0.025 / 0.035 / 2))


def test_render_options_parse_switches():
    o = scripting.renderOptions(["--width", "64", "--super-sampling", "4"], width=800, height=800, super_sampling=9)
    assert (o.width, o.height, o.super_sampling) == (64, 800, 4)


def test_spd_load_dump_roundtrip():
    text = "# comment\n; also\n400 0.1 1.5\n500 0.2 1.6\n600 0.3 1.7\n"
    a, b = spd.loads(text)
    np.testing.assert_allclose(a.wavelengths, [400, 500, 600])
    np.testing.assert_allclose(b.values, [1.5, 1.6, 1.7], rtol=1e-6)
    buf = io.StringIO()
    spd.dump([a, b], buf)
    a2, b2 = spd.loads(buf.getvalue())
    np.testing.assert_array_equal(a2.values, a.values)
    np.testing.assert_array_equal(b2.values, b.values)
    # metres are recognised
    (m,) = spd.loads("4e-7 1\n5e-7 2\n")
    np.testing.assert_allclose(m.wavelengths, [400, 500])


@pytest.mark.parametrize("metal", ["Ag", "Al", "Au", "Cu"])
def test_builtin_refractive_index(metal):
    n, k = spd.load_builtin_refractive_index(metal)
    assert len(n.values) == len(k.values) >= 10
    assert (np.diff(n.wavelengths) != 0).all() and (k.values > 0).all()


def test_builtin_refractive_index_unknown():
    with pytest.raises(ValueError, match="not found"):
        spd.load_builtin_refractive_index("Unobtainium")


def test_observer_set_standard():
    bundled = standard_observer()
    flat = liar_amd.Observer([400e-9, 700e-9], [(1.0, 1.0, 1.0), (1.0, 1.0, 1.0)])
    try:
        liar_amd.Observer.setStandard(flat)
        w, xyz = standard_observer()
        np.testing.assert_allclose(w, [400, 700])
        # the oracle renders with the current standard observer: a Mirror(1)
        # under a sky of 2 on [400, 500) nm and 0 from 520 nm on gives, with a
        # flat observer on [400, 700], Y = (2 * 100 + 20) / 300
        from liar_amd import shaders, spectra, textures
        from liar_amd import workloads as W

        sky = spectra.Sampled([400, 500, 520, 700], [2, 2, 0, 0])
        sc, cam = W.furnace_sphere(shaders.Mirror(textures.Constant(1.0)), sky)
        xyz_, wt = oracle.render(liar_amd.compile_scene(sc), cam, 8, 8, 4096, lambda_range=(400, 700), seed=2)
        Y = (xyz_ / wt)[1][3:5, 3:5]
        assert abs(Y.mean() - 220 / 300) < 0.03
    finally:
        liar_amd.observer._standard = None
    assert standard_observer() is bundled


def test_oracle_renders_tool_cornell_box():
    cs = liar_amd.compile_scene(cornell_box.scene())
    xyz, w = oracle.render_direct(cs, cornell_box.camera(), 16, 16, 4, seed=3, secondary_light_samples=1)
    y = (xyz / w)[1]
    assert np.isfinite(y).all() and (y >= 0).all() and y.mean() > 0
