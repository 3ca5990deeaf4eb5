"""pbrt-v3 scene files into a RenderEngine (mirrors ``liar.tools.pbrt``,
src/liar/tools/pbrt.py:43-1163).

``parse(path)`` reads a scene, sets up ``scene.engine`` (camera, sampler,
tracer, target, scene graph) and renders it when ``render_immediately``.
The subset of pbrt this engine can render is supported; the rest raises
``NotImplementedError`` naming the statement:

* transforms: Identity, Translate, Scale, Rotate, LookAt, Transform,
  ConcatTransform, CoordinateSystem, CoordSysTransform, ReverseOrientation;
  TransformBegin/End, AttributeBegin/End, ObjectBegin/End/ObjectInstance;
  Include;
* Camera "perspective" (fov on the shorter image side, frameaspectratio,
  hither); Film "image" (resolution, filename, scale, maxsampleluminance);
  Sampler (pixelsamples or xsamples * ysamples); PixelFilter "box";
  Integrator "path" / "directlighting" (maxdepth); Accelerator is ignored;
* Shape "trianglemesh", "plymesh", "sphere"; a file's meshes are combined
  into one TriangleMeshComposite at WorldEnd (and per ObjectBegin block);
* Material / MakeNamedMaterial / NamedMaterial "matte" (Kd), "mirror" (Kr),
  "glass" (eta or index, Kr, Kt), "metal" (eta, k, roughness with pbrt's
  roughness remapping); other materials fall back to matte with their Kd;
* AreaLightSource "diffuse" (L, nsamples) on a planar parallelogram (a
  two-triangle trianglemesh); LightSource "infinite" (L, nsamples, no map);
* parameter types integer, float, point, vector, normal, bool, string, rgb /
  color (``rgb`` of the linear-sRGB value, as LiAR's importer makes it), xyz,
  blackbody, spectrum (inline pairs).

pbrt's camera frame is left handed (right = up x forward); this engine's
camera takes right = forward x sky, so the importer renders with sky = up and
mirrors the film horizontally, which gives pbrt's orientation.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .. import cameras, engine, output, samplers, scenery, shaders, spectra, textures, tracers
from ..rgb_space import sRGB

# pbrt colours are linear sRGB values (LiAR's importer, pbrt.py:1395)
_RGB_LINEAR = sRGB.linearSpace()

__all__ = ["parse", "PbrtScene", "ParseError", "remap_roughness"]


class ParseError(Exception):
    def __init__(self, path: str, line_number: int, message: str):
        super().__init__(f"{path}:{line_number}: error: {message}")
        self.path = path
        self.line_number = line_number
        self.message = message


def parse(path, *args, **kwargs) -> "PbrtScene":
    """``liar.tools.pbrt.parse(path)`` (pbrt.py:43-46)."""
    scene = PbrtScene(*args, **kwargs)
    scene.parse(path)
    return scene


# -- tokens ---------------------------------------------------------------------

_TYPES = ("integer", "float", "point", "vector", "normal", "bool", "string", "texture", "color", "rgb", "xyz",
          "blackbody", "spectrum", "point3", "vector3", "normal3", "point2", "vector2")
_TOKEN = re.compile(r"""
    (?P<comment>\#[^\n]*)
  | (?P<space>\s+)
  | (?P<param>"(?:%s)\s+[A-Za-z_][A-Za-z0-9_]*")
  | (?P<string>"[^"]*")
  | (?P<lbr>\[) | (?P<rbr>\])
  | (?P<number>[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)
  | (?P<ident>[A-Za-z_][A-Za-z0-9_]*)
""" % "|".join(_TYPES), re.VERBOSE)


@dataclass
class _Tok:
    kind: str
    value: Any
    line: int


def _tokens(path: str, text: str):
    pos, line = 0, 1
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ParseError(path, line, f"unexpected input {text[pos:pos + 20]!r}")
        kind = m.lastgroup
        s = m.group()
        if kind == "param":
            t, name = s[1:-1].split()
            yield _Tok("param", (t, name), line)
        elif kind == "string":
            v = s[1:-1]
            yield _Tok("bool" if v in ("true", "false") else "string", (v == "true") if v in ("true", "false") else v,
                       line)
        elif kind == "number":
            yield _Tok("number", float(s) if any(c in s for c in ".eE") else int(s), line)
        elif kind in ("ident", "lbr", "rbr"):
            yield _Tok(kind, s, line)
        line += s.count("\n")
        pos = m.end()


@dataclass
class _Statement:
    line: int
    keyword: str
    args: list = field(default_factory=list)
    kwargs: dict = field(default_factory=dict)


# -- values ---------------------------------------------------------------------

# LiAR's fit of sqrt(pbrt alpha) in x = log(max(roughness, 1e-3))
# (src/liar/textures/remap_pbrt_roughness.cpp:60-72), Horner order
_REMAP = (1.2730328, 0.31971025, 0.02405594, -0.0018453393, -0.00026452233)


def remap_roughness(r: float, remaproughness: bool = True) -> float:
    """pbrt's roughness parameter as LiAR's roughness, the square root of the
    microfacet alpha (RemapPbrtRoughness::remap, used by pbrt.py:1404-1420):
    with ``remaproughness`` pbrt maps roughness to alpha by a fit in log
    roughness, else roughness is alpha."""
    r = float(r)
    if not remaproughness:
        return math.sqrt(r)
    x = math.log(max(r, 1e-3))
    c0, c1, c2, c3, c4 = _REMAP
    return c0 + (c1 + (c2 + (c3 + c4 * x) * x) * x) * x


def _spectrum(v):
    """A pbrt colour value (number, rgb triple or spectrum object) as a LiAR
    texture input."""
    if isinstance(v, spectra.Spectrum):
        return v
    if isinstance(v, (int, float)):
        return float(v)
    r, g, b = (float(x) for x in v)
    return spectra.rgb(r, g, b, _RGB_LINEAR)  # LiAR's importer: pbrt.py:1355-1362


def _rotation(angle_deg, axis):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(math.radians(angle_deg)), math.sin(math.radians(angle_deg))
    x, y, z = a
    m = np.eye(4)
    m[:3, :3] = [[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                 [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                 [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]]
    return m


class _MirroredTarget:
    """A film target whose columns are reversed on the way in: renders pbrt's
    left-handed camera frame with this engine's right-handed one."""

    def __init__(self, target):
        self.target = target

    def __getattr__(self, name):
        return getattr(self.target, name)

    def accumulate(self, xyz_sum, weight):
        self.target.accumulate(np.ascontiguousarray(xyz_sum[:, :, ::-1]), np.ascontiguousarray(weight[:, ::-1]))


# -- the scene ------------------------------------------------------------------

class PbrtScene:
    """State machine over pbrt statements (pbrt.py:49-1100): options block
    (camera, film, sampler, integrator), then the world block."""

    def __init__(self, render_immediately: bool = False, device: int = 0):
        self.engine = engine.RenderEngine()
        self.engine.device = device
        self._render_immediately = render_immediately
        self._ctm = np.eye(4)
        self._named_cs: dict[str, np.ndarray] = {}
        self._world = False
        self._named_materials: dict[str, Any] = {}
        self._dir = "."
        self.resolution = (640, 480)
        self._film_kwargs: dict = {}
        self.Camera()
        self.Film()
        self.Sampler()
        self.Integrator()

    # transforms (pbrt.py:78-133): each statement post-multiplies the CTM
    def Identity(self):
        self._ctm = np.eye(4)

    def Translate(self, x, y, z):
        m = np.eye(4)
        m[:3, 3] = (x, y, z)
        self._ctm = self._ctm @ m

    def Scale(self, x, y, z):
        self._ctm = self._ctm @ np.diag([x, y, z, 1.0])

    def Rotate(self, angle, x, y, z):
        self._ctm = self._ctm @ _rotation(angle, (x, y, z))

    def LookAt(self, ex, ey, ez, lx, ly, lz, ux, uy, uz):
        eye, look, up = (np.array(v, np.float64) for v in ((ex, ey, ez), (lx, ly, lz), (ux, uy, uz)))
        d = look - eye
        d /= np.linalg.norm(d)
        right = np.cross(up / np.linalg.norm(up), d)
        if np.linalg.norm(right) == 0:
            raise ValueError("LookAt: up is parallel to the view direction")
        right /= np.linalg.norm(right)
        new_up = np.cross(d, right)
        cam_to_world = np.eye(4)
        cam_to_world[:3, 0], cam_to_world[:3, 1], cam_to_world[:3, 2], cam_to_world[:3, 3] = right, new_up, d, eye
        self._ctm = self._ctm @ np.linalg.inv(cam_to_world)

    def CoordinateSystem(self, name):
        self._named_cs[name] = self._ctm.copy()

    def CoordSysTransform(self, name):
        self._ctm = self._named_cs[name].copy()

    def Transform(self, *m):
        self._ctm = np.asarray(_flat(m), np.float64).reshape(4, 4).T.copy()

    def ConcatTransform(self, *m):
        self._ctm = self._ctm @ np.asarray(_flat(m), np.float64).reshape(4, 4).T

    def ReverseOrientation(self):
        # as pbrt.py:131-133 does it: a point reflection of what follows
        self.Scale(-1, -1, -1)

    # options block
    def Camera(self, name="perspective", fov=90, frameaspectratio=None, hither=1e-2, **kwargs):
        self._verify_options()
        if name != "perspective":
            raise NotImplementedError(f"Camera {name!r}: only 'perspective' is supported")
        if kwargs.get("lensradius", 0):
            raise NotImplementedError("Camera lensradius: the camera is a pinhole")
        self._named_cs["camera"] = np.linalg.inv(self._ctm)
        c2w = self._named_cs["camera"]
        cam = cameras.PerspectiveCamera()
        cam.position = tuple(c2w[:3, 3])
        cam.direction = tuple(c2w[:3, 2] / np.linalg.norm(c2w[:3, 2]))
        cam.sky = tuple(c2w[:3, 1] / np.linalg.norm(c2w[:3, 1]))
        cam.nearLimit = float(hither)
        self.engine.camera = cam
        self._fov, self._frameaspect = float(fov), frameaspectratio

    def Film(self, name="image", xresolution=640, yresolution=480, cropwindow=(0, 1, 0, 1), **kwargs):
        self._verify_options()
        if name != "image":
            raise NotImplementedError(f"Film {name!r}: only 'image' is supported")
        if tuple(cropwindow) != (0, 1, 0, 1):
            raise NotImplementedError("Film cropwindow")
        self.resolution = (int(xresolution), int(yresolution))
        self._film_kwargs = kwargs

    def Sampler(self, name="halton", pixelsamples=None, xsamples=2, ysamples=2, **kwargs):
        self._verify_options()
        spp = int(pixelsamples) if pixelsamples else int(xsamples) * int(ysamples)
        if name == "halton" and pixelsamples is None:
            spp = 16  # pbrt-v3's default pixelsamples for halton
        self.engine.sampler = samplers.Stratifier(self.resolution, spp)

    def PixelFilter(self, name="box", **kwargs):
        self._verify_options()
        if name != "box":
            raise NotImplementedError(f"PixelFilter {name!r}: the film filter is a box")

    def Integrator(self, name="path", maxdepth=5, **kwargs):
        self._verify_options()
        if name == "path":
            t = tracers.AdjointPhotonTracer()
        elif name == "directlighting":
            t = tracers.DirectLighting()
        else:
            raise NotImplementedError(f"Integrator {name!r}: 'path' and 'directlighting' are supported")
        t.maxRayGeneration = int(maxdepth)
        self.engine.tracer = t

    def Accelerator(self, *args, **kwargs):
        self._verify_options()

    # world block (pbrt.py:135-281)
    def WorldBegin(self):
        self._verify_options()
        self.Identity()
        self.CoordinateSystem("world")
        self._world = True
        self._stack = []
        self._material = shaders.Lambert(0.5)
        self._area_light = None
        self._objects: list = []
        self._instances: dict[str, list] = {}
        self._cur_instance = None

    def WorldEnd(self):
        self._verify_world()
        w, h = self.resolution
        cam = self.engine.camera
        cam.aspectRatio = float(self._frameaspect) if self._frameaspect else w / h
        # pbrt's fov spans the shorter image side; the camera's spans the width
        if cam.aspectRatio > 1:
            cam.fovAngle = 2 * math.atan(math.tan(0.5 * math.radians(self._fov)) * cam.aspectRatio)
        else:
            cam.fovAngle = math.radians(self._fov)
        fk = self._film_kwargs
        img = output.Image(fk.get("filename", "pbrt.exr"), self.resolution)
        if "scale" in fk:
            img.exposureStops = math.log2(float(fk["scale"]))
        if "maxsampleluminance" in fk:
            img.maxSampleLuminance = min(float(fk["maxsampleluminance"]), 1e38)
        self.image = img
        self.engine.target = _MirroredTarget(img)
        self.engine.sampler.resolution = self.resolution
        self.engine.scene = scenery.QbvhTree(_combine_meshes(self._objects))
        self._world = False
        if self._render_immediately:
            self.render()

    def render(self):
        self.engine.render()

    def TransformBegin(self):
        self._verify_world()
        self._stack.append(("T", self._ctm.copy()))

    def TransformEnd(self):
        self._verify_world()
        kind, ctm = self._stack.pop()
        self._ctm = ctm

    def AttributeBegin(self):
        self._verify_world()
        self._stack.append(("A", (self._ctm.copy(), self._material, self._area_light)))

    def AttributeEnd(self):
        self._verify_world()
        kind, state = self._stack.pop()
        if kind != "A":
            raise ValueError("AttributeEnd without AttributeBegin")
        self._ctm, self._material, self._area_light = state

    def ObjectBegin(self, name):
        self.AttributeBegin()
        if self._cur_instance is not None:
            raise ValueError("ObjectBegin inside an object definition")
        self._cur_instance = self._instances[name] = []

    def ObjectEnd(self):
        self.AttributeEnd()
        if self._cur_instance is None:
            raise ValueError("ObjectEnd without ObjectBegin")
        self._cur_instance[:] = _combine_meshes(self._cur_instance)
        self._cur_instance = None

    def ObjectInstance(self, name):
        self._verify_world()
        for shape in self._instances[name]:
            self._add(shape)

    def _placed(self, obj):
        if np.allclose(self._ctm, np.eye(4)):
            return obj
        return scenery.Transformation(obj, self._ctm.copy())

    def _add(self, obj):
        obj = self._placed(obj)
        (self._cur_instance if self._cur_instance is not None else self._objects).append(obj)

    # shapes (pbrt.py:390-443)
    def Shape(self, name, **kwargs):
        self._verify_world()
        maker = getattr(self, "_shape_" + name, None)
        if maker is None:
            raise NotImplementedError(f"Shape {name!r}")
        shape = maker(**kwargs)
        shape.shader = self._material
        if self._area_light:
            light_name, lk = self._area_light
            shape = self._area(shape, light_name, **lk)
        self._add(shape)

    def _shape_sphere(self, radius=1.0, **kwargs):
        return scenery.Sphere((0.0, 0.0, 0.0), float(radius))

    def _shape_trianglemesh(self, indices, P, N=None, uv=None, st=None, **kwargs):
        v = np.asarray(_flat(P), np.float64).reshape(-1, 3)
        t = np.asarray(_flat(indices), np.int64).reshape(-1, 3)
        return scenery.TriangleMesh(v.astype(np.float32), (), (), t.astype(np.uint32))

    def _shape_plymesh(self, filename, **kwargs):
        if "alpha" in kwargs or "shadowalpha" in kwargs:
            raise NotImplementedError("plymesh alpha masks")
        return scenery.TriangleMesh.loadPly(os.path.join(self._dir, filename))

    # materials (pbrt.py:449-788)
    def Material(self, name="matte", **kwargs):
        self._verify_world()
        self._material = self._make_material(name, **kwargs)

    def MakeNamedMaterial(self, name, **kwargs):
        self._verify_world()
        kind = kwargs.pop("type")
        self._named_materials[name] = self._make_material(kind, **kwargs)

    def NamedMaterial(self, name):
        self._verify_world()
        self._material = self._named_materials[name]

    def _make_material(self, name, **kwargs):
        kwargs.pop("bumpmap", None)
        if name == "matte":
            if kwargs.get("sigma", 0):
                raise NotImplementedError("matte sigma (Oren-Nayar)")
            return shaders.Lambert(_spectrum(kwargs.get("Kd", 0.5)))
        if name == "mirror":
            return shaders.Mirror(_spectrum(kwargs.get("Kr", 1.0)))
        if name == "glass":
            if kwargs.get("uroughness", 0) or kwargs.get("vroughness", 0):
                raise NotImplementedError("rough glass")
            eta = kwargs.get("index", kwargs.get("eta", 1.5))
            g = shaders.Dielectric(_spectrum(eta))
            g.reflectance = textures.as_texture(_spectrum(kwargs.get("Kr", 1.0)))
            g.transmittance = textures.as_texture(_spectrum(kwargs.get("Kt", 1.0)))
            return g
        if name == "metal":
            if "eta" not in kwargs or "k" not in kwargs:
                raise NotImplementedError("metal needs eta and k (no built-in copper table)")
            remap = bool(kwargs.get("remaproughness", True))
            r = float(kwargs.get("uroughness", kwargs.get("roughness", 0.01)))
            m = shaders.CookTorrance(_spectrum(kwargs["eta"]), _spectrum(kwargs["k"]))
            m.mdf = shaders.MicrofacetTrowbridgeReitz()
            m.roughnessU = m.roughnessV = textures.as_texture(remap_roughness(r, remap))
            return m
        # pbrt.py:462-476: an unknown material falls back to matte with its Kd
        return shaders.Lambert(_spectrum(kwargs.get("Kd", 0.5)))

    # lights (pbrt.py:953-1042)
    def LightSource(self, name, **kwargs):
        self._verify_world()
        if name != "infinite":
            raise NotImplementedError(f"LightSource {name!r}: only 'infinite' is supported")
        if kwargs.get("mapname"):
            raise NotImplementedError("infinite light maps")
        light = scenery.LightSky(_spectrum(kwargs.get("L", 1.0)))
        light.numberOfEmissionSamples = int(kwargs.get("nsamples", 1))
        self._objects.append(light)

    def AreaLightSource(self, name, **kwargs):
        self._verify_world()
        if name not in ("diffuse", "area"):
            raise NotImplementedError(f"AreaLightSource {name!r}")
        self._area_light = (name, kwargs)

    def _area(self, shape, name, L=1.0, nsamples=1, twosided=False, **kwargs):
        # pbrt.py:393-401, 962-967: the shape becomes a LightArea (placed by
        # the CTM like any shape) whose surface shows its radiance
        light = scenery.LightArea(_as_parallelogram(shape))
        light.radiance = textures.as_texture(_spectrum(L))
        light.numberOfEmissionSamples = int(nsamples)
        light.isDoubleSided = bool(twosided)
        light.surface.shader = shaders.Unshaded(light.radiance)
        return light

    # statements
    def parse(self, path: str):
        with open(path) as f:
            text = f.read()
        old = self._dir
        self._dir = os.path.dirname(os.path.abspath(path))
        try:
            self.parse_string(text, path)
        finally:
            self._dir = old

    def parse_string(self, text: str, path: str = "<string>"):
        st: _Statement | None = None
        pending = None  # (type, name) of a typed parameter awaiting its value
        toks = _tokens(path, text)
        for tok in toks:
            if tok.kind == "ident":
                if st is not None:
                    self._execute(st, path)
                    st = None
                if tok.value == "Include":
                    nxt = next(toks, None)
                    if nxt is None or nxt.kind != "string":
                        raise ParseError(path, tok.line, "Include must be followed by a string")
                    self.parse(os.path.join(self._dir, nxt.value))
                else:
                    st = _Statement(tok.line, tok.value)
            elif tok.kind == "param":
                if st is None:
                    raise ParseError(path, tok.line, "a parameter outside a statement")
                pending = tok.value
            else:
                if st is None:
                    raise ParseError(path, tok.line, "expected a statement name")
                if tok.kind == "lbr":
                    vals = []
                    for t in toks:
                        if t.kind == "rbr":
                            break
                        if t.kind not in ("number", "string", "bool"):
                            raise ParseError(path, t.line, "lists hold numbers and strings only")
                        vals.append(t.value)
                    else:
                        raise ParseError(path, tok.line, "unterminated list")
                elif tok.kind == "rbr":
                    raise ParseError(path, tok.line, "unbalanced ]")
                else:
                    vals = [tok.value]
                if pending is not None:
                    st.kwargs[_pyname(pending[1])] = _convert(pending[0], vals, path, tok.line)
                    pending = None
                else:
                    st.args.append(vals[0] if len(vals) == 1 else vals)
        if st is not None:
            self._execute(st, path)

    def _execute(self, st: _Statement, path: str):
        fn = getattr(self, st.keyword, None)
        if fn is None or st.keyword.startswith("_"):
            raise NotImplementedError(f"{path}:{st.line}: pbrt statement {st.keyword!r} is not supported")
        try:
            fn(*st.args, **st.kwargs)
        except (NotImplementedError, ParseError):
            raise
        except (TypeError, ValueError, KeyError, IndexError) as e:
            raise ParseError(path, st.line, f"{st.keyword}: {e}") from e

    def _verify_options(self):
        if self._world:
            raise ValueError("statement allowed only before WorldBegin")

    def _verify_world(self):
        if not self._world:
            raise ValueError("statement allowed only between WorldBegin and WorldEnd")


def _pyname(name):
    return name + "_" if name in ("from", "in", "lambda") else name


def _flat(xs):
    out = []
    for x in xs:
        if isinstance(x, (list, tuple)):
            out.extend(_flat(x))
        else:
            out.append(x)
    return out


def _convert(kind, vals, path, line):
    if kind == "integer":
        v = [int(x) for x in vals]
    elif kind in ("float",):
        v = [float(x) for x in vals]
    elif kind in ("point", "vector", "normal", "point3", "vector3", "normal3"):
        if len(vals) % 3:
            raise ParseError(path, line, f"{kind} values come in threes")
        return [float(x) for x in vals] if len(vals) > 3 else tuple(float(x) for x in vals)
    elif kind in ("point2", "vector2"):
        return [float(x) for x in vals]
    elif kind == "bool":
        v = [bool(x) for x in vals]
    elif kind == "string":
        v = [str(x) for x in vals]
    elif kind in ("rgb", "color"):
        if len(vals) != 3:
            raise ParseError(path, line, "an rgb value has three components")
        return tuple(float(x) for x in vals)
    elif kind == "xyz":
        if len(vals) != 3:
            raise ParseError(path, line, "an xyz value has three components")
        return spectra.XYZ(tuple(float(x) for x in vals))
    elif kind == "blackbody":
        t = float(vals[0])
        scale = float(vals[1]) if len(vals) > 1 else 1.0
        return spectra.BlackBody(t, scale)
    elif kind == "spectrum":
        if len(vals) == 1 and isinstance(vals[0], str):
            raise NotImplementedError("spectrum files")
        ws, xs = vals[0::2], vals[1::2]
        return spectra.Sampled([float(w) for w in ws], [float(x) for x in xs])
    else:
        raise NotImplementedError(f"parameter type {kind!r}")
    return v[0] if len(v) == 1 else v


def _as_parallelogram(shape):
    """A two-triangle planar parallelogram mesh (or a Parallelogram) as a
    Parallelogram for LightArea, with the mesh's winding (normal)."""
    if isinstance(shape, scenery.Parallelogram):
        o, a, b = shape.support, shape.sizeU, shape.sizeV
        v = np.stack([o, o + a, o + a + b, o + b])
    elif isinstance(shape, scenery.TriangleMesh) and len(shape.triangles()) == 2:
        v = np.asarray(shape.vertices(), np.float64)
        if len(v) != 4:
            raise NotImplementedError("area lights: the mesh must be one quad (4 vertices, 2 triangles)")
    else:
        raise NotImplementedError("area lights are limited to parallelogram quads")
    w = np.asarray(v, np.float64)
    tri = np.asarray(shape.triangles()) if isinstance(shape, scenery.TriangleMesh) else np.array([[0, 1, 2], [0, 2, 3]])
    # the quad's corners in order around it: the two triangles share a diagonal
    shared = sorted(set(tri[0]) & set(tri[1]))
    if len(shared) != 2:
        raise NotImplementedError("area lights: the two triangles must share an edge")
    other0 = [i for i in tri[0] if i not in shared][0]
    other1 = [i for i in tri[1] if i not in shared][0]
    # keep the first triangle's winding: other0 -> shared in the order tri[0] visits them
    k = list(tri[0]).index(other0)
    s_next, s_prev = tri[0][(k + 1) % 3], tri[0][(k + 2) % 3]
    o = w[s_next]
    a = w[other1] - o  # s_next -> other1 -> s_prev is the second half
    b = w[other0] - o
    if not np.allclose(w[s_prev], o + a + b, rtol=1e-5, atol=1e-6 * max(1.0, float(np.abs(w).max()))):
        raise NotImplementedError("area lights: the quad is not a parallelogram")
    # orient like the mesh: normal = (b - o) x ... of the first triangle's winding
    n_mesh = np.cross(w[tri[0][1]] - w[tri[0][0]], w[tri[0][2]] - w[tri[0][0]])
    if np.dot(np.cross(a, b), n_mesh) < 0:
        a, b = b, a
    pg = scenery.Parallelogram(o.astype(np.float32), a.astype(np.float32), b.astype(np.float32))
    pg.shader = shape.shader
    return pg


def _combine_meshes(objs):
    """pbrt.py:182-192 / 241-254: the meshes of a block as one composite."""
    out, meshes = [], []
    for o in objs:
        (meshes if isinstance(o, scenery.TriangleMesh) else out).append(o)
    if len(meshes) > 1:
        out.append(scenery.TriangleMeshComposite(meshes))
    else:
        out.extend(meshes)
    return out
