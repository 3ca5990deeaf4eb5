"""The reference's Cornell-box scripts rendered on the GPU through the drop-in
API (RenderEngine + liar_amd.tools.cornell_box, src/liar/tools/
cornell_box.py:19-213, examples/cornell_box/direct_lighting.py) match the CPU
oracle with the same seed, for DirectLighting and the AdjointPhotonTracer;
Observer.setStandard (src/liar/__init__.py:51-69) changes what the GPU
renders as it does the oracle."""
import math
import sys

import numpy as np
import pytest

import liar_amd
import oracle
from liar_amd import output, samplers, tracers
from liar_amd.tools import cornell_box

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_render_gpu import compare, say  # noqa: E402

pytestmark = pytest.mark.gpu


def engine_render(tracer, w, h, spp, seed):
    engine = liar_amd.RenderEngine()
    engine.tracer = tracer
    engine.sampler = samplers.Stratifier((w, h), spp)
    engine.scene = cornell_box.scene()
    engine.camera = cornell_box.camera()
    engine.target = output.Raster((w, h))
    engine.seed = seed
    engine.render()
    return engine.target.xyz_sum, engine.target.weight


def test_tool_cornell_direct_lighting_matches_oracle():
    say("[gpu] tool Cornell box, DirectLighting through RenderEngine vs oracle")
    gx, gw = engine_render(tracers.DirectLighting(), 48, 48, 4, seed=7)
    lo, hi = (float(x) for x in liar_amd.observer.standard_observer()[0][[0, -1]])
    cs = liar_amd.compile_scene(cornell_box.scene())
    ox, ow = oracle.render_direct(cs, cornell_box.camera(), 48, 48, 4, max_generation=8, secondary_light_samples=1,
                                  trace_glossy=True, lambda_range=(lo, hi), seed=7)
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "tool cornell direct")
    assert rel_total < 1e-4 and frac_bad < 0.01


def test_tool_cornell_path_tracer_matches_oracle():
    say("[gpu] tool Cornell box, AdjointPhotonTracer through RenderEngine vs oracle")
    gx, gw = engine_render(tracers.AdjointPhotonTracer(), 48, 48, 16, seed=8)
    lo, hi = (float(x) for x in liar_amd.observer.standard_observer()[0][[0, -1]])
    cs = liar_amd.compile_scene(cornell_box.scene())
    ox, ow = oracle.render(cs, cornell_box.camera(), 48, 48, 16, wavelengths=1, max_depth=0, rr_depth=8,
                           rr_survival=0.5, lambda_range=(lo, hi), seed=8)
    rel_total, frac_bad, _ = compare(gx, gw, ox, ow, "tool cornell apt")
    assert rel_total < 2e-3 and frac_bad < 0.02


def test_observer_set_standard_on_gpu():
    from liar_amd import scenery, shaders, spectra, textures
    from liar_amd import workloads as W

    flat = liar_amd.Observer([400e-9, 700e-9], [(1.0, 1.0, 1.0), (1.0, 1.0, 1.0)])
    try:
        liar_amd.Observer.setStandard(flat)
        sky = spectra.Sampled([400, 500, 520, 700], [2, 2, 0, 0])
        sc, cam = W.furnace_sphere(shaders.Mirror(textures.Constant(1.0)), sky)
        engine = liar_amd.RenderEngine()
        engine.tracer = tracers.AdjointPhotonTracer()
        engine.sampler = samplers.Stratifier((16, 16), 16384)
        engine.scene, engine.camera = sc, cam
        engine.target = output.Raster((16, 16))
        engine.render()
        Y = (engine.target.xyz_sum / engine.target.weight)[1][6:10, 6:10]
        say(f"[gpu] flat observer: Y {Y.mean():.4f} (expect {220 / 300:.4f})")
        assert abs(Y.mean() - 220 / 300) < 0.01
    finally:
        liar_amd.observer._standard = None
