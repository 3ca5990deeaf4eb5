"""liar_amd -- MI355X (gfx950) native implementation of LiAR's ray-tracing hot
path: QBVH closest/any-hit traversal and the spectral path-tracing integrand.

The compute path is hand-written HIP in ``liar_amd/csrc`` behind the C-ABI of
``include/liar_amd.h``; this package is the host-side mirror of LiAR's
interfaces on top of that C-ABI (``scenery``, ``shaders``, ``textures``,
``spectra``, ``cameras``, ``samplers``, ``tracers``, ``output``,
``RenderEngine``).
"""
from . import cameras, output, samplers, scenery, shaders, shading, spectra, textures, tracers  # noqa: F401
from ._native import LIB_PATH, LiarError  # noqa: F401
from .engine import RenderEngine, compile_scene  # noqa: F401
from .observer import Observer  # noqa: F401
from .qbvh import Context, Qbvh, philox4x32, random_rays, random_triangles  # noqa: F401
from .rgb_space import CIEXYZ, RgbSpace, sRGB  # noqa: F401
from .spectra import Recovery, rgb  # noqa: F401

# ``from liar_amd import *`` gives a script LiAR's top-level names
__all__ = ["RenderEngine", "Observer", "cameras", "output", "samplers", "scenery", "shaders", "spectra", "textures",
           "tracers", "sRGB", "CIEXYZ", "RgbSpace", "Recovery", "rgb"]

__version__ = "0.1.0"
