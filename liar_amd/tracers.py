"""Ray tracers (mirrors ``liar.tracers.AdjointPhotonTracer``,
src/liar/tracers/adjoint_photon_tracer.cpp).

LiAR plays Russian roulette with survival 0.5 from ``maxRayGeneration`` on
and has no hard depth cap.  ``russianRouletteStart``, ``survivalProbability``
and ``maxDepth`` (extensions) express other rules, e.g. "depth <= 8, roulette
from 3"; ``maxDepth = None`` caps paths at 64 generations.
"""
from __future__ import annotations


class AdjointPhotonTracer:
    def __init__(self):
        self.maxRayGeneration = 8
        self.russianRouletteStart = None
        self.survivalProbability = 0.5
        self.maxDepth = None


class DirectLighting:
    """Direct lighting with recursive specular / glossy rays (mirrors
    ``liar.tracers.DirectLighting``, src/liar/tracers/direct_lighting.cpp:42-80):
    camera rays and the specular continuations of camera rays estimate every
    light with its ``numberOfEmissionSamples``; other hits use
    ``numSecondaryLightSamples`` lights picked by power.  Rendered by
    lr_render_direct."""

    def __init__(self):
        self.maxRayGeneration = 8
        self._secondary = 1
        self.traceGlossy = True

    @property
    def numSecondaryLightSamples(self):
        return self._secondary

    @numSecondaryLightSamples.setter
    def numSecondaryLightSamples(self, n):
        self._secondary = max(int(n), 1)  # direct_lighting.cpp:60-67
