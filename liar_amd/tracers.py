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
