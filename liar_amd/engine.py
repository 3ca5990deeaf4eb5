"""RenderEngine (mirrors ``liar.RenderEngine``, src/liar/kernel/render_engine.cpp:35-247)
and the flattening of a scene graph into the C-ABI scene description.

``compile_scene`` turns the object graph into plain arrays (triangles,
spheres, one material per primitive, spectra, lights); ``RenderEngine.render``
builds the QBVH and scene on the GPU and runs ``lr_render`` into its target.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import cameras, samplers, scenery, shaders, spectra, textures, tracers
from .observer import ensure_standard_observer


@dataclass
class CompiledScene:
    verts: np.ndarray          # (nv, 3) float32
    tris: np.ndarray           # (nt, 3) uint32
    spheres: np.ndarray        # (ns, 4) float32
    prim_material: np.ndarray  # (nt + ns,) uint32
    materials: list = field(default_factory=list)   # dicts (type, slots, roughness)
    spectra: list = field(default_factory=list)     # spectra.Spectrum objects
    lights: list = field(default_factory=list)      # dicts
    tri_normals: np.ndarray = None  # (nt, 9) float32 corner normals, or None (geometric normals only)


def _spectrum_of(tex, table, index):
    if tex is None:
        return N.LR_NONE
    tex = textures.as_texture(tex)
    sp = tex.spectrum
    key = id(sp)
    if key not in index:
        index[key] = len(table)
        table.append(sp)
    return index[key]


def _scalar_of(tex, default):
    if tex is None:
        return default
    return textures.as_texture(tex).scalar()


def _material_of(shader, table, index):
    m = {"type": N.MAT_NONE, "reflectance": N.LR_NONE, "transmittance": N.LR_NONE, "eta": N.LR_NONE,
         "kappa": N.LR_NONE, "emission": N.LR_NONE, "light": N.LR_NONE, "roughness": 0.0, "roughness_v": 0.0,
         "fuzz": 0.0}
    parts = shader.children if isinstance(shader, shaders.Sum) else ([shader] if shader is not None else [])
    bsdfs = [p for p in parts if not isinstance(p, shaders.Unshaded)]
    emit = [p for p in parts if isinstance(p, shaders.Unshaded)]
    if len(bsdfs) > 1:
        raise NotImplementedError("Sum of more than one BSDF shader is not supported")
    if len(emit) > 1:
        raise NotImplementedError("Sum of more than one Unshaded shader is not supported")
    if emit:
        m["emission"] = _spectrum_of(emit[0].colour, table, index)
    if bsdfs:
        s = bsdfs[0]
        if isinstance(s, shaders.Lambert):
            m.update(type=N.MAT_LAMBERT, reflectance=_spectrum_of(s.diffuse, table, index))
        elif isinstance(s, shaders.Mirror):
            fuzz = _scalar_of(s.fuzz, 0.0) if s.fuzz is not None else 0.0
            if not (fuzz >= 0.0 and math.isfinite(fuzz)):
                raise ValueError("Mirror.fuzz must be finite and >= 0")
            m.update(type=N.MAT_MIRROR, reflectance=_spectrum_of(s.reflectance, table, index), fuzz=float(fuzz))
        elif isinstance(s, shaders.Dielectric):
            m.update(type=N.MAT_DIELECTRIC, reflectance=_spectrum_of(s.reflectance, table, index),
                     transmittance=_spectrum_of(s.transmittance, table, index),
                     eta=_spectrum_of(s.innerRefractionIndex, table, index),
                     kappa=_spectrum_of(s.outerRefractionIndex, table, index))
        elif isinstance(s, shaders.Conductor):
            m.update(type=N.MAT_CONDUCTOR, reflectance=_spectrum_of(s.reflectance, table, index),
                     eta=_spectrum_of(s.refractionIndex, table, index),
                     kappa=_spectrum_of(s.absorptionCoefficient, table, index))
        elif isinstance(s, shaders.CookTorrance):
            ru, rv = _scalar_of(s.roughnessU, 0.1), _scalar_of(s.roughnessV, 0.1)
            t = N.MAT_COOK_TORRANCE if isinstance(s.mdf, shaders.MicrofacetTrowbridgeReitz) \
                else N.MAT_COOK_TORRANCE_BECKMANN
            m.update(type=t, reflectance=_spectrum_of(s.reflectance, table, index),
                     eta=_spectrum_of(s.refractionIndex, table, index),
                     kappa=_spectrum_of(s.absorptionCoefficient, table, index), roughness=float(ru),
                     roughness_v=float(rv))
        else:
            raise NotImplementedError(f"shader {type(s).__name__} is not supported")
    return m


def compile_scene(scene) -> CompiledScene:
    """Flattens a scene graph into arrays (LiAR's object graph -> one QBVH)."""
    verts, tris, tri_mat, tri_nrm = [], [], [], []
    sph, sph_mat = [], []
    materials, mat_index = [], {}
    table, spec_index = [], {}
    lights = []
    nverts = [0]

    def material(shader, light=N.LR_NONE):
        key = (id(shader), light)
        if key not in mat_index:
            m = _material_of(shader, table, spec_index)
            m["light"] = light
            mat_index[key] = len(materials)
            materials.append(m)
        return mat_index[key]

    def add_mesh(v, t, mat, cn=None):
        verts.append(np.asarray(v, np.float32))
        tris.append(np.asarray(t, np.uint32) + nverts[0])
        tri_mat.append(np.full(len(t), mat, np.uint32))
        tri_nrm.append(cn if cn is not None else np.zeros((len(t), 9), np.float32))
        tri_nrm_any[0] = tri_nrm_any[0] or cn is not None
        nverts[0] += len(v)

    tri_nrm_any = [False]

    def point(M, p):
        p = np.asarray(p, np.float64)
        return p if M is None else (M[:3, :3] @ p + M[:3, 3])

    def vector(M, v):
        v = np.asarray(v, np.float64)
        return v if M is None else M[:3, :3] @ v

    def walk(obj, inherited, M=None):
        shader = obj.shader if obj.shader is not None else inherited
        if isinstance(obj, scenery.Transformation):
            L = obj.localToWorld if M is None else M @ obj.localToWorld
            walk(obj.child, shader, L)
        elif isinstance(obj, scenery.List):
            for c in obj.children:
                walk(c, shader, M)
        elif isinstance(obj, scenery.TriangleMesh):
            v = obj.vertices()
            if M is not None:
                v = (np.asarray(v, np.float64) @ M[:3, :3].T + M[:3, 3]).astype(np.float32)
            add_mesh(v, obj.triangles(), material(shader), obj.corner_normals(M))
        elif isinstance(obj, scenery.TriangleMeshComposite):
            # one triangle set is this object's rule; each child keeps its shader
            for c in obj.children():
                cs = c.shader if c.shader is not None else shader
                v = c.vertices()
                if M is not None:
                    v = (np.asarray(v, np.float64) @ M[:3, :3].T + M[:3, 3]).astype(np.float32)
                add_mesh(v, c.triangles(), material(cs), c.corner_normals(M))
        elif isinstance(obj, scenery.Parallelogram):
            v, t = obj._mesh()
            if M is not None:
                v = (np.asarray(v, np.float64) @ M[:3, :3].T + M[:3, 3]).astype(np.float32)
            add_mesh(v, t, material(shader))
        elif isinstance(obj, scenery.Sphere):
            r = obj.radius
            if M is not None:
                A = M[:3, :3]
                k = np.cbrt(abs(np.linalg.det(A)))
                if not np.allclose(A.T @ A, k * k * np.eye(3), rtol=1e-6, atol=1e-9 * k * k):
                    raise NotImplementedError("a Sphere under a non-uniform scale or shear is not supported")
                r = obj.radius * k
            sph.append([*point(M, obj.center), r])
            sph_mat.append(material(shader))
        elif isinstance(obj, scenery.LightArea) and isinstance(obj.surface, scenery.TriangleMesh):
            surf = obj.surface
            li = len(lights)
            first = sum(len(t) for t in tris)
            v = surf.vertices()
            if M is not None:
                v = (np.asarray(v, np.float64) @ M[:3, :3].T + M[:3, 3]).astype(np.float32)
            lights.append({"type": N.LIGHT_MESH, "radiance": _spectrum_of(obj.radiance, table, spec_index),
                           "double_sided": bool(obj.isDoubleSided), "samples": int(obj.numberOfEmissionSamples),
                           "first_triangle": first, "n_triangles": len(surf.triangles())})
            s = obj.shader if obj.shader is not None else surf.shader
            add_mesh(v, surf.triangles(), material(s, li), surf.corner_normals(M))
        elif isinstance(obj, scenery.LightArea):
            surf = obj.surface
            li = len(lights)
            o, a, b = point(M, surf.support), vector(M, surf.sizeU), vector(M, surf.sizeV)
            lights.append({"type": N.LIGHT_PARALLELOGRAM,
                           "radiance": _spectrum_of(obj.radiance, table, spec_index),
                           "double_sided": bool(obj.isDoubleSided),
                           "origin": o, "dx": a, "dy": b,
                           "samples": int(obj.numberOfEmissionSamples)})
            s = obj.shader if obj.shader is not None else surf.shader
            v = np.stack([o, o + a, o + a + b, o + b]).astype(np.float32)
            add_mesh(v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), material(s, li))
        elif isinstance(obj, scenery.LightSky):
            lights.append({"type": N.LIGHT_SKY, "radiance": _spectrum_of(obj.radiance, table, spec_index),
                           "samples": int(obj.numberOfEmissionSamples)})
        elif isinstance(obj, scenery.LightSun):
            dsun = vector(M, obj.direction)
            dsun = dsun / np.linalg.norm(dsun)
            lights.append({"type": N.LIGHT_SUN, "radiance": _spectrum_of(obj.radiance, table, spec_index),
                           "direction": tuple(float(x) for x in dsun), "angle": obj.angle,
                           "samples": int(obj.numberOfEmissionSamples)})
        else:
            raise NotImplementedError(f"scene object {type(obj).__name__} is not supported")

    walk(scene, None)
    V = np.concatenate(verts) if verts else np.zeros((0, 3), np.float32)
    T = np.concatenate(tris) if tris else np.zeros((0, 3), np.uint32)
    TM = np.concatenate(tri_mat) if tri_mat else np.zeros(0, np.uint32)
    S = np.asarray(sph, np.float32).reshape(-1, 4)
    PM = np.concatenate([TM, np.asarray(sph_mat, np.uint32)]).astype(np.uint32)
    CN = np.concatenate(tri_nrm).astype(np.float32) if tri_nrm_any[0] else None
    return CompiledScene(V, T, S, PM, materials, table, lights, CN)


def camera_struct(cam: cameras.PerspectiveCamera) -> N.Camera:
    c = N.Camera()
    c.position[:] = [float(x) for x in cam.position]
    c.direction[:] = [float(x) for x in cam.direction]
    c.sky[:] = [float(x) for x in cam.sky]
    c.fov = float(cam.fovAngle)
    c.aspect = float(cam.aspectRatio)
    c.near_limit = float(cam.nearLimit)
    return c


class GpuScene:
    """A compiled scene resident on one GPU (QBVH + lr_scene)."""

    def __init__(self, ctx, cs: CompiledScene):
        from .qbvh import Qbvh

        ensure_standard_observer()
        self.ctx = ctx
        self.compiled = cs
        self.qbvh = Qbvh(ctx, cs.verts, cs.tris, cs.spheres)
        keep = []
        specs = (N.Spectrum * max(1, len(cs.spectra)))(*[s._describe(keep) for s in cs.spectra])
        mats = (N.Material * max(1, len(cs.materials)))(*[N.material_struct(m) for m in cs.materials])
        ls = (N.Light * max(1, len(cs.lights)))()
        for i, L in enumerate(cs.lights):
            ls[i].type = L["type"]
            ls[i].radiance = L["radiance"]
            ls[i].double_sided = int(L.get("double_sided", False))
            for name in ("origin", "dx", "dy", "direction"):
                if name in L:
                    getattr(ls[i], name)[:] = [float(x) for x in L[name]]
            ls[i].angle = float(L.get("angle", 0.0))
            ls[i].samples = int(L.get("samples", 0))
            ls[i].first_triangle = int(L.get("first_triangle", 0))
            ls[i].n_triangles = int(L.get("n_triangles", 0))
        pm = np.ascontiguousarray(cs.prim_material, np.uint32)
        h = C.c_void_p()
        N.check(N.lib().lr_scene_create(ctx.handle, self.qbvh.handle, N.ptr(pm), mats, len(cs.materials), specs,
                                        len(cs.spectra), ls, len(cs.lights), C.byref(h)), "lr_scene_create")
        self._h = h
        if cs.tri_normals is not None:
            cn = np.ascontiguousarray(cs.tri_normals, np.float32)
            N.check(N.lib().lr_scene_set_normals(h, N.ptr(cn), len(cs.tris)), "lr_scene_set_normals")

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.lib().lr_scene_destroy(self._h)
            self._h = None
        if getattr(self, "qbvh", None) is not None:
            self.qbvh.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, camera: N.Camera, params: N.RenderParams, film_xyz, film_w, stream=None) -> dict:
        import torch

        s = torch.cuda.current_stream(film_xyz.device).cuda_stream if stream is None else stream
        st = N.RenderStats()
        N.check(N.lib().lr_render(self._h, C.byref(camera), C.byref(params), C.c_void_p(film_xyz.data_ptr()),
                                  C.c_void_p(film_w.data_ptr()), C.byref(st), C.c_void_p(s)), "lr_render")
        return {"paths": st.paths, "rays": st.rays, "chunks": st.chunks, "seconds": st.seconds}


    def render_direct(self, camera: N.Camera, params: N.RenderParams, film_xyz, film_w, max_generation=8,
                      secondary_light_samples=1, trace_glossy=True, stream=None) -> dict:
        """lr_render_direct: LiAR's DirectLighting integrand (``params``'
        depth and roulette fields are ignored)."""
        import torch

        s = torch.cuda.current_stream(film_xyz.device).cuda_stream if stream is None else stream
        dp = N.DirectParams(int(max_generation), int(secondary_light_samples), int(bool(trace_glossy)))
        st = N.RenderStats()
        N.check(N.lib().lr_render_direct(self._h, C.byref(camera), C.byref(params), C.byref(dp),
                                         C.c_void_p(film_xyz.data_ptr()), C.c_void_p(film_w.data_ptr()),
                                         C.byref(st), C.c_void_p(s)), "lr_render_direct")
        return {"paths": st.paths, "rays": st.rays, "chunks": st.chunks, "seconds": st.seconds}


class Batch:
    """Many GpuScenes rendered by one wavefront (lr_batch_*): every bounce's
    queue mixes the paths of all scenes in a chunk; each film equals the
    scene's own lr_render with the same parameters."""

    def __init__(self, ctx, scenes):
        self.ctx = ctx
        self.scenes = list(scenes)
        hs = (C.c_void_p * len(self.scenes))(*[g._h.value for g in self.scenes])
        h = C.c_void_p()
        N.check(N.lib().lr_batch_create(ctx.handle, hs, len(self.scenes), C.byref(h)), "lr_batch_create")
        self._h = h

    def render(self, items, graph=False, stream=None) -> dict:
        """items: one (camera, params, film_xyz, film_w) per scene.  Returns the
        stats and per-scene completion times ``scene_ms`` (from the call's
        first launch, device clock)."""
        import torch

        assert len(items) == len(self.scenes)
        arr = (N.BatchItem * len(items))()
        for k, (cam, prm, fx, fw) in enumerate(items):
            arr[k].camera = cam
            arr[k].params = prm
            arr[k].film_xyz = fx.data_ptr()
            arr[k].film_weight = fw.data_ptr()
        ms = (C.c_float * len(items))()
        st = N.RenderStats()
        s = torch.cuda.current_stream(items[0][2].device).cuda_stream if stream is None else stream
        N.check(N.lib().lr_render_batch(self._h, arr, N.LR_BATCH_GRAPH if graph else 0, ms, C.byref(st),
                                        C.c_void_p(s)), "lr_render_batch")
        return {"paths": st.paths, "rays": st.rays, "chunks": st.chunks, "seconds": st.seconds,
                "scene_ms": [float(x) for x in ms]}

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.lib().lr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def guided_render(gs: "GpuScene", guide, camera: N.Camera, params: N.RenderParams, film_xyz, film_w,
                  selection=0.5, train=False, train_steps=0, train_batch=65536, max_records=0, weight_clamp=0.0,
                  light_share=0.5, learned_selection=False, balance_heuristic=False, image_weighting=False,
                  record_depth=0, guide_depth=0, learned_rr=False, rr_scale=1.0, rr_min_survival=0.05,
                  rr_start=0, stream=None) -> dict:
    """lr_render_guided on a GpuScene with a liar_amd.guide.Guide.  With
    ``learned_selection`` the network's per-vertex strategy probabilities
    replace ``selection`` / ``light_share``; ``guide_depth`` > 0 guides the
    first guide_depth bounces only; ``learned_rr`` plays radiance-driven
    Russian roulette on the network's reflected-radiance estimate (train it
    with ``Guide.set_radiance_weight``)."""
    import torch

    s = torch.cuda.current_stream(film_xyz.device).cuda_stream if stream is None else stream
    gp = N.GuideParams(float(selection), int(bool(train)), int(train_steps), int(train_batch), int(max_records),
                       float(weight_clamp), float(light_share), int(bool(learned_selection)),
                       int(bool(balance_heuristic)), int(bool(image_weighting)), int(record_depth),
                       int(guide_depth), int(bool(learned_rr)), float(rr_scale),
                       float(rr_min_survival), int(rr_start))
    st = N.RenderStats()
    N.check(N.lib().lr_render_guided(gs._h, C.byref(camera), C.byref(params), guide.handle, C.byref(gp),
                                     C.c_void_p(film_xyz.data_ptr()), C.c_void_p(film_w.data_ptr()), C.byref(st),
                                     C.c_void_p(s)), "lr_render_guided")
    return {"paths": st.paths, "rays": st.rays, "chunks": st.chunks, "seconds": st.seconds}


def render_params(width, height, spp, wavelengths=1, max_depth=0, rr_depth=8, rr_survival=0.5,
                  lambda_range=(380.0, 780.0), seed=1, chunk_paths=0, max_sample_luminance=0.0,
                  graph=False, sampler=N.LR_SAMPLER_STRATIFIER, sampler_generations=0,
                  sample_offset=0) -> N.RenderParams:
    return N.RenderParams(int(width), int(height), int(spp), int(wavelengths), int(max_depth), int(rr_depth),
                          float(rr_survival), float(lambda_range[0]), float(lambda_range[1]), int(seed),
                          int(chunk_paths), float(max_sample_luminance), int(bool(graph)), int(sampler),
                          int(sampler_generations), int(sample_offset))


class RenderEngine:
    """``engine.render()`` renders ``scene`` through ``camera`` with ``tracer``
    and ``sampler`` into ``target`` on GPU ``device``."""

    def __init__(self):
        self.camera = None
        self.tracer = None
        self.sampler = None
        self.target = None
        self.scene = None
        self.numberOfThreads = 0  # accepted for API compatibility
        self.seed = 1
        self.device = 0
        self.wavelengthsPerSample = 1
        self.wavelengthRange = None  # (lo, hi) nm; None = the observer's full range
        self.lastStats = None

    def render(self, time=0.0, bucket=None):
        import torch

        from .observer import standard_observer
        from .qbvh import Context

        for name in ("camera", "tracer", "sampler", "target", "scene"):
            if getattr(self, name) is None:
                raise RuntimeError(f"RenderEngine.{name} is not set")
        w, h = self.sampler.resolution
        if tuple(self.target.resolution) != (w, h):
            raise RuntimeError("sampler and target resolutions differ")
        tr = self.tracer
        rr_start = getattr(tr, "russianRouletteStart", None)
        rr_start = tr.maxRayGeneration if rr_start is None else rr_start
        lr = self.wavelengthRange or (float(standard_observer()[0][0]), float(standard_observer()[0][-1]))
        msl = float(getattr(self.target, "maxSampleLuminance", math.inf))
        halton = isinstance(self.sampler, samplers.Halton)
        if halton:
            if isinstance(tr, tracers.DirectLighting):
                raise NotImplementedError("DirectLighting renders with the Stratifier only")
            self.sampler._check_generations(tr.maxRayGeneration)
        params = render_params(w, h, self.sampler.samplesPerPixel, self.wavelengthsPerSample,
                               getattr(tr, "maxDepth", None) or 0, rr_start, getattr(tr, "survivalProbability", 0.5),
                               lr, self.seed, max_sample_luminance=msl if math.isfinite(msl) else 0.0,
                               sampler=N.LR_SAMPLER_HALTON if halton else N.LR_SAMPLER_STRATIFIER,
                               sampler_generations=tr.maxRayGeneration if halton else 0,
                               sample_offset=self.sampler.nextSample if halton else 0)
        ctx = Context(self.device)
        gs = GpuScene(ctx, compile_scene(self.scene))
        dev = torch.device("cuda", self.device)
        film = torch.zeros((3, h, w), dtype=torch.float32, device=dev)
        wgt = torch.zeros((h, w), dtype=torch.float32, device=dev)
        if isinstance(tr, tracers.DirectLighting):
            self.lastStats = gs.render_direct(camera_struct(self.camera), params, film, wgt, tr.maxRayGeneration,
                                              tr.numSecondaryLightSamples, tr.traceGlossy)
        else:
            self.lastStats = gs.render(camera_struct(self.camera), params, film, wgt)
        torch.cuda.synchronize(dev)
        if halton:
            self.sampler.nextSample += w * h * params.spp * params.wavelengths
        self.target.accumulate(film.cpu().numpy().astype(np.float64), wgt.cpu().numpy().astype(np.float64))
        if hasattr(self.target, "save"):
            self.target.save()
        gs.close()
        ctx.close()
