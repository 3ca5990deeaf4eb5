"""Samplers (mirrors ``liar.samplers``).

Stratifier: LiAR's per-pixel screen strata (src/liar/samplers/stratifier.cpp)
with the counter-based Philox4x32-10 stream keyed by (seed, path, generation).

Halton: LiAR's global Halton sequence over the film
(src/liar/samplers/halton.cpp:180-230): sample j of the sequence puts its
screen point at radical inverses of j in bases 2 and 3, its wavelength in base
13, and draws the AdjointPhotonTracer's per-generation numbers from digit-
scrambled sub-sequences on the primes from 17 on (rendered by lr_render with
LR_SAMPLER_HALTON)."""
from __future__ import annotations


class Stratifier:
    def __init__(self, resolution=(320, 240), samplesPerPixel=1):
        self.resolution = (int(resolution[0]), int(resolution[1]))
        self.samplesPerPixel = int(samplesPerPixel)
        self.jittered = True


class Halton:
    """``Halton()``: a progressive sampler.  LiAR's renders until cancelled;
    here one ``RenderEngine.render()`` takes ``samplesPerPixel`` x width x
    height samples (an extension: the film's sample budget), and the next
    render continues the sequence where the last one stopped (``nextSample``,
    the state LiAR keeps as its next task id).  ``samplesPerTask`` is kept for
    LiAR's state (halton.cpp:46-56); the sequence does not depend on it.

    The tracer's maxRayGeneration generations each take 4 one-dimensional and
    3 two-dimensional sub-sequences; at most 284 primes are available, so more
    than 28 generations raise like LiAR's "Too many subsequences requested"
    (halton.cpp:147-151)."""

    PRIMES = 284

    def __init__(self, resolution=(320, 240), samplesPerPixel=1):
        self.resolution = (int(resolution[0]), int(resolution[1]))
        self.samplesPerPixel = int(samplesPerPixel)
        self._samples_per_task = 1024
        self.nextSample = 0

    @property
    def samplesPerTask(self):
        return self._samples_per_task

    @samplesPerTask.setter
    def samplesPerTask(self, n):
        self._samples_per_task = max(int(n), 1)

    def getState(self):
        """(next task id, samplesPerTask), LiAR's state (halton.cpp:81-85);
        a partly used task counts as used."""
        return (-(-self.nextSample // self._samples_per_task), self._samples_per_task)

    def setState(self, state):
        task, self.samplesPerTask = state
        self.nextSample = int(task) * self._samples_per_task

    def seed(self, randomSeed):
        """A deterministic sequence: seeding changes nothing (halton.cpp:69-72)."""

    def _check_generations(self, generations):
        if 10 * int(generations) > self.PRIMES:
            raise ValueError("Too many subsequences requested.")
