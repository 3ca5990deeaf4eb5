"""Shaders (mirrors ``liar.shaders``: src/liar/shaders/).

Each shader is flattened into one lr_material; see include/liar_amd.h for the
BSDF each type evaluates on the GPU.
"""
from __future__ import annotations

from .textures import Constant, as_texture


class Shader:
    pass


class Lambert(Shader):
    def __init__(self, diffuse=None):
        self.diffuse = as_texture(1.0 if diffuse is None else diffuse)


class Mirror(Shader):
    def __init__(self, reflectance=None):
        self.reflectance = as_texture(1.0 if reflectance is None else reflectance)
        self.fuzz = None


class Dielectric(Shader):
    def __init__(self, innerRefractionIndex=None, outerRefractionIndex=None):
        self.innerRefractionIndex = as_texture(1.5 if innerRefractionIndex is None else innerRefractionIndex)
        self.outerRefractionIndex = as_texture(1.0 if outerRefractionIndex is None else outerRefractionIndex)
        self.reflectance = Constant(1.0)
        self.transmittance = Constant(1.0)


class Conductor(Shader):
    def __init__(self, refractionIndex=None, absorptionCoefficient=None):
        self.refractionIndex = as_texture(1.0 if refractionIndex is None else refractionIndex)
        self.absorptionCoefficient = as_texture(0.0 if absorptionCoefficient is None else absorptionCoefficient)
        self.reflectance = Constant(1.0)


class MicrofacetTrowbridgeReitz:
    """GGX distribution (src/liar/shaders/microfacet_trowbridge_reitz.cpp)."""


class MicrofacetBeckmann:
    """Beckmann distribution; LiAR's CookTorrance default (cook_torrance.cpp:73)."""


class CookTorrance(Shader):
    def __init__(self, refractionIndex=None, absorptionCoefficient=None):
        self.refractionIndex = as_texture(1.0 if refractionIndex is None else refractionIndex)
        self.absorptionCoefficient = as_texture(0.0 if absorptionCoefficient is None else absorptionCoefficient)
        self.reflectance = Constant(1.0)
        self.roughnessU = Constant(0.1)
        self.roughnessV = Constant(0.1)
        self.mdf = MicrofacetBeckmann()
        self.numberOfSamples = 1


class Unshaded(Shader):
    """Pure emitter seen by camera rays (src/liar/shaders/unshaded.cpp)."""

    def __init__(self, colour=None):
        self.colour = as_texture(1.0 if colour is None else colour)


class Sum(Shader):
    """Sum of shaders (src/liar/shaders/sum.cpp); here at most one BSDF plus
    Unshaded emitters."""

    def __init__(self, children):
        self.children = list(children)
