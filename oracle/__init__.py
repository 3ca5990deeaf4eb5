"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg
of ``bench.py``, always as the checker (or the timed CPU baseline), never as a
code path of the library.  Wraps ``oracle/liboracle.so`` (built by
``oracle/Makefile``) and the pure-Python/numpy restatements in this package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
MISS = 0xFFFFFFFF

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing; run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        P, U32, I = C.c_void_p, C.c_uint32, C.c_int
        RAYS = P  # const float* const[8]
        L.or_brute_closest.argtypes = [P, P, U32, P, U32, RAYS, U32, P, P, P, P, I]
        L.or_brute_any.argtypes = [P, P, U32, P, U32, RAYS, U32, P, I]
        L.or_brute_closest_f64.argtypes = [P, P, U32, P, U32, RAYS, U32, P, P, P, P, I]
        L.or_qbvh_closest.argtypes = [P, U32, P, RAYS, U32, P, P, P, P, I, P, P]
        L.or_qbvh_any.argtypes = [P, U32, P, RAYS, U32, P, I, P, P]
        L.or_qbvh_check.argtypes = [P, U32, P, U32, P, P, U32, P, U32, P]
        L.or_last_message.restype = C.c_char_p
        F4 = C.POINTER(C.c_float)
        L.or_tri_f32.argtypes = [P, P, P, P, P, P, P, P]
        L.or_sphere_f32.argtypes = [P, P, P, C.c_float, C.c_float, P]
        for name in ("or_brute_closest", "or_brute_any", "or_brute_closest_f64", "or_qbvh_closest",
                     "or_qbvh_any", "or_qbvh_check", "or_tri_f32", "or_sphere_f32"):
            getattr(L, name).restype = C.c_int
        del F4
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _rays_arg(rays: np.ndarray):
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    assert rays.ndim == 2 and rays.shape[0] == 8
    arr = (C.c_void_p * 8)(*[rays[k].ctypes.data for k in range(8)])
    return rays, arr


def _scene(verts, tris, spheres):
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    if spheres is None:
        spheres = np.zeros((0, 4), np.float32)
    spheres = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    return verts, tris, spheres


def default_threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def brute_closest(verts, tris, spheres, rays, threads: int | None = None):
    """LiAR List{TriangleMesh, Spheres...}::intersect by brute force (FP32).
    Returns (prim uint32 (n,), t, u, v float32 (n,))."""
    verts, tris, spheres = _scene(verts, tris, spheres)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    prim = np.empty(n, np.uint32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    rc = lib().or_brute_closest(_p(verts), _p(tris), len(tris), _p(spheres), len(spheres), ra, n,
                                _p(prim), _p(t), _p(u), _p(v), threads or default_threads())
    assert rc == 0
    return prim, t, u, v


def brute_any(verts, tris, spheres, rays, threads: int | None = None):
    verts, tris, spheres = _scene(verts, tris, spheres)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    occ = np.empty(n, np.uint8)
    rc = lib().or_brute_any(_p(verts), _p(tris), len(tris), _p(spheres), len(spheres), ra, n,
                            _p(occ), threads or default_threads())
    assert rc == 0
    return occ


class _Geom(C.Structure):
    _fields_ = [("verts", C.c_void_p), ("tris", C.c_void_p), ("nt", C.c_uint32), ("spheres", C.c_void_p),
                ("ns", C.c_uint32)]


def inst_brute(geoms, inst_geom, local_to_world, rays, any_hit=False, threads: int | None = None):
    """Instances by brute force (LiAR's Transformation of each geometry inside
    a List): geoms = [(verts, tris, spheres)], inst_geom (m,), local_to_world
    (m, 3, 4).  Returns (prim, t, u, v, inst) or occluded (uint8)."""
    L = lib()
    if not hasattr(L, "_inst_sig"):
        P = C.c_void_p
        L.or_inst_brute.argtypes = [P, C.c_uint32, P, P, C.c_uint32, P, C.c_uint32, C.c_int, P, P, P, P, P, P,
                                    C.c_int]
        L.or_inst_brute.restype = C.c_int
        L._inst_sig = True
    keep = []
    arr = (_Geom * max(1, len(geoms)))()
    for k, (v, t, s) in enumerate(geoms):
        v, t, s = _scene(v, t, s)
        keep += [v, t, s]
        arr[k] = _Geom(_p(v), _p(t), len(t), _p(s) if len(s) else None, len(s))
    ig = np.ascontiguousarray(inst_geom, np.uint32).ravel()
    m = np.ascontiguousarray(local_to_world, np.float32).reshape(-1, 12)
    assert len(ig) == len(m)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    prim = np.empty(n, np.uint32)
    inst = np.empty(n, np.uint32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    occ = np.empty(n, np.uint8)
    rc = L.or_inst_brute(arr, len(geoms), _p(ig), _p(m), len(ig), ra, n, int(bool(any_hit)), _p(prim), _p(t), _p(u),
                         _p(v), _p(inst), _p(occ), threads or default_threads())
    assert rc == 0, rc
    return occ if any_hit else (prim, t, u, v, inst)


def set_inst_mistake(k: int) -> None:
    """Test hook of the instanced oracle: 1 forgets t_local / s."""
    lib().or_set_inst_mistake(int(k))


def brute_closest_f64(verts, tris, spheres, rays, threads: int | None = None):
    """Same candidates, all arithmetic in FP64 (error reference)."""
    verts, tris, spheres = _scene(verts, tris, spheres)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    prim = np.empty(n, np.uint32)
    t, u, v = (np.empty(n, np.float64) for _ in range(3))
    rc = lib().or_brute_closest_f64(_p(verts), _p(tris), len(tris), _p(spheres), len(spheres), ra, n,
                                    _p(prim), _p(t), _p(u), _p(v), threads or default_threads())
    assert rc == 0
    return prim, t, u, v


def qbvh_closest(nodes, prims, rays, threads: int | None = None):
    """Closest hit over flattened QBVH arrays (uint8 (n,128), (m,48)).
    Returns (prim, t, u, v, nodes_visited, prims_tested)."""
    nodes = np.ascontiguousarray(nodes)
    prims = np.ascontiguousarray(prims)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    prim = np.empty(n, np.uint32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    nv, npr = C.c_uint64(), C.c_uint64()
    rc = lib().or_qbvh_closest(_p(nodes), len(nodes), _p(prims), ra, n, _p(prim), _p(t), _p(u), _p(v),
                               threads or default_threads(), C.byref(nv), C.byref(npr))
    assert rc == 0, "oracle traversal stack overflow"
    return prim, t, u, v, nv.value, npr.value


def qbvh_closest_ties(nodes, prims, rays, threads: int | None = None):
    """qbvh_closest plus a per-ray flag: 1 where the hit distance is reached
    by two or more primitives (a numerical tie, broken by the smaller id).
    Returns (prim, t, u, v, flags uint8, nodes_visited, prims_tested); flags
    bit 0: the hit distance is shared by two or more primitives (a tie), bit
    1: LiAR's interval quirk rejected the mesh's nearest triangle."""
    L = lib()
    if not getattr(L, "_ties_sig", False):
        P, U32 = C.c_void_p, C.c_uint32
        L.or_qbvh_closest_ties.argtypes = [P, U32, P, P, U32, P, P, P, P, P, C.c_int, P, P]
        L.or_qbvh_closest_ties.restype = C.c_int
        L._ties_sig = True
    nodes = np.ascontiguousarray(nodes)
    prims = np.ascontiguousarray(prims)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    prim = np.empty(n, np.uint32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    tied = np.zeros(n, np.uint8)
    nv, npr = C.c_uint64(), C.c_uint64()
    rc = L.or_qbvh_closest_ties(_p(nodes), len(nodes), _p(prims), ra, n, _p(prim), _p(t), _p(u), _p(v), _p(tied),
                                threads or default_threads(), C.byref(nv), C.byref(npr))
    assert rc == 0, "oracle traversal stack overflow"
    return prim, t, u, v, tied, nv.value, npr.value


def qbvh_any(nodes, prims, rays, threads: int | None = None):
    nodes = np.ascontiguousarray(nodes)
    prims = np.ascontiguousarray(prims)
    rays, ra = _rays_arg(rays)
    n = rays.shape[1]
    occ = np.empty(n, np.uint8)
    nv, npr = C.c_uint64(), C.c_uint64()
    rc = lib().or_qbvh_any(_p(nodes), len(nodes), _p(prims), ra, n, _p(occ), threads or default_threads(),
                           C.byref(nv), C.byref(npr))
    assert rc == 0, "oracle traversal stack overflow"
    return occ, nv.value, npr.value


def decode_q8(q8):
    """lr_node_q8 records (uint8 (n, 64)) -> lr_node records (uint8 (n, 128)).

    Every plane is fmaf(q, scale, origin).  scale is a power of two and
    q < 256, so the FP32 product q * scale is exact and one FP32 addition
    gives the fused value bit for bit; the child words are copied.  The
    oracle's QBVH traversal over the decoded nodes therefore makes exactly the
    box tests of the GPU traversal over the quantised ones.  (The quantised
    format is the library's device format; LiAR's QBVH keeps FP32 boxes.)"""
    q8 = np.ascontiguousarray(q8, dtype=np.uint8)
    n = q8.shape[0]
    hdr = q8[:, :24].copy().view(np.float32)  # origin x, y, z, scale x, y, z
    q = q8[:, 24:48].reshape(n, 6, 4).astype(np.float32)  # qlo_x, qhi_x, qlo_y, qhi_y, qlo_z, qhi_z
    planes = np.empty((n, 6, 4), np.float32)
    for k in range(6):
        ax = k // 2
        planes[:, k, :] = hdr[:, ax:ax + 1] + q[:, k, :] * hdr[:, 3 + ax:4 + ax]
    out = np.zeros((n, 128), np.uint8)
    out[:, :96] = planes.reshape(n, 24).view(np.uint8)
    out[:, 96:112] = q8[:, 48:64]
    return out


def qbvh_check(nodes, prims, verts, tris, spheres=None):
    """Structural invariants of a flattened QBVH; returns (ok, message, max_depth)."""
    verts, tris, spheres = _scene(verts, tris, spheres)
    nodes = np.ascontiguousarray(nodes)
    prims = np.ascontiguousarray(prims)
    md = C.c_uint32()
    rc = lib().or_qbvh_check(_p(nodes), len(nodes), _p(prims), len(prims), _p(verts), _p(tris), len(tris),
                             _p(spheres), len(spheres), C.byref(md))
    return rc == 0, lib().or_last_message().decode(), md.value


def tri_f32(o, d, v0, e1, e2):
    """Single FP32 ray-triangle test -> (hit, t, u, v)."""
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, v0, e1, e2)]
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    hit = lib().or_tri_f32(*[_p(x) for x in a], _p(t), _p(u), _p(v))
    return bool(hit), float(t[0]), float(u[0]), float(v[0])


def sphere_f32(o, d, c, r, tmin=0.0):
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, c)]
    t = np.zeros(1, np.float32)
    r2 = np.float32(r) * np.float32(r)
    hit = lib().or_sphere_f32(*[_p(x) for x in a], C.c_float(r2), C.c_float(tmin), _p(t))
    return bool(hit), float(t[0])


# ---------------------------------------------------------------------------
# path-tracing oracle (oracle/render.c)
# ---------------------------------------------------------------------------

class _Spec(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("count", C.c_uint32), ("w", C.c_void_p), ("v", C.c_void_p),
                ("value", C.c_float), ("lo", C.c_float), ("hi", C.c_float), ("temperature", C.c_float),
                ("scale", C.c_float)]


class _Mat(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in ("type", "refl", "trans", "eta", "kappa", "emission", "light")] + \
               [("roughness", C.c_float), ("roughness_v", C.c_float), ("fuzz", C.c_float)]


class _Light(C.Structure):
    _fields_ = [("type", C.c_uint32), ("radiance", C.c_uint32), ("double_sided", C.c_uint32),
                ("origin", C.c_float * 3), ("dx", C.c_float * 3), ("dy", C.c_float * 3),
                ("direction", C.c_float * 3), ("angle", C.c_float), ("samples", C.c_uint32),
                ("first_triangle", C.c_uint32), ("n_triangles", C.c_uint32)]


class _Cam(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("sky", C.c_float * 3),
                ("fov", C.c_float), ("aspect", C.c_float), ("near_limit", C.c_float)]


class _Params(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("wavelengths", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("rr_survival", C.c_float), ("lambda_min", C.c_float), ("lambda_max", C.c_float),
                ("seed", C.c_uint64), ("chunk_paths", C.c_uint32), ("max_sample_luminance", C.c_float),
                ("sampler", C.c_uint32), ("sampler_generations", C.c_uint32), ("sample_offset", C.c_uint64)]


def _spec_struct(sp, keep):
    """Describes a liar_amd.spectra object with the oracle's own layout."""
    kind = type(sp).__name__
    if kind == "Constant":
        return _Spec(0, 0, None, None, sp.value, 0, 0, 0, 0)
    if kind == "Sampled":
        w = np.ascontiguousarray(sp.wavelengths, np.float32)
        v = np.ascontiguousarray(sp.values, np.float32)
        keep += [w, v]
        return _Spec(1, len(v), w.ctypes.data, v.ctypes.data, 0, 0, 0, 0, 0)
    if kind == "XYZ":
        # the recovered table, with the recovery's Reflectant divisor
        rec = sp.recovery_table()
        w = np.ascontiguousarray(rec[0], np.float32)
        v = np.ascontiguousarray(rec[1], np.float32)
        keep += [w, v]
        return _Spec(1, len(v), w.ctypes.data, v.ctypes.data, rec[2], 0, 0, 0, 0)
    if kind == "Binned":
        v = np.ascontiguousarray(sp.values, np.float32)
        keep.append(v)
        return _Spec(2, len(v), None, v.ctypes.data, 0, sp.lo, sp.hi, 0, 0)
    if kind == "BlackBody":
        return _Spec(3, 0, None, None, 0, 0, 0, sp.temperature, sp.scale)
    raise TypeError(kind)


def render(compiled, camera, width, height, spp, wavelengths=1, max_depth=0, rr_depth=8, rr_survival=0.5,
           lambda_range=(380.0, 780.0), seed=1, observer=None, threads=None, max_sample_luminance=0.0,
           halton=False, sampler_generations=0, sample_offset=0):
    """FP64 CPU render (AdjointPhotonTracer) of a compiled scene -> (xyz sums
    (3, h, w), weight (h, w)).  ``halton``: the Halton sampler over the whole
    film (width * height * spp * wavelengths samples starting at sequence
    index sample_offset + 1), generations below sampler_generations drawing
    from its scrambled sub-sequences; the weight then counts the samples that
    fell in each pixel."""
    return _render(compiled, camera, width, height, spp, wavelengths, max_depth, rr_depth, rr_survival,
                   lambda_range, seed, observer, threads, max_sample_luminance, None,
                   (1 if halton else 0, sampler_generations, sample_offset))


def render_direct(compiled, camera, width, height, spp, wavelengths=1, max_generation=8,
                  secondary_light_samples=1, trace_glossy=True, lambda_range=(380.0, 780.0), seed=1,
                  observer=None, threads=None, max_sample_luminance=0.0):
    """FP64 CPU render (DirectLighting) of a compiled scene -> (xyz sums
    (3, h, w), weight (h, w))."""
    cfg = np.array([max_generation, secondary_light_samples, int(bool(trace_glossy))], np.uint32)
    return _render(compiled, camera, width, height, spp, wavelengths, 0, 8, 0.5, lambda_range, seed, observer,
                   threads, max_sample_luminance, cfg)


def _render(compiled, camera, width, height, spp, wavelengths, max_depth, rr_depth, rr_survival, lambda_range, seed,
            observer, threads, max_sample_luminance, dl_cfg, sampler=(0, 0, 0)):
    L = lib()
    if not hasattr(L, "_render_sig"):
        head = [C.c_void_p] * 2 + [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                   C.c_void_p, C.c_void_p]
        L.or_render.argtypes = head + [C.c_void_p, C.c_void_p, C.c_int]
        L.or_render.restype = C.c_int
        L.or_render_direct.argtypes = head + [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.or_render_direct.restype = C.c_int
        L._render_sig = True
    keep = []
    cs = compiled
    V = np.ascontiguousarray(cs.verts, np.float32)
    T = np.ascontiguousarray(cs.tris, np.uint32)
    S = np.ascontiguousarray(cs.spheres, np.float32)
    PM = np.ascontiguousarray(cs.prim_material, np.uint32)
    specs = (_Spec * max(1, len(cs.spectra)))(*[_spec_struct(s, keep) for s in cs.spectra])
    mats = (_Mat * max(1, len(cs.materials)))(*[
        _Mat(m["type"], m["reflectance"], m["transmittance"], m["eta"], m["kappa"], m["emission"], m["light"],
             m["roughness"], m.get("roughness_v", m["roughness"]), m.get("fuzz", 0.0)) for m in cs.materials])
    CN = getattr(cs, "tri_normals", None)
    CN = None if CN is None else np.ascontiguousarray(CN, np.float32)
    lights = (_Light * max(1, len(cs.lights)))()
    for i, Ld in enumerate(cs.lights):
        lights[i].type = Ld["type"]
        lights[i].radiance = Ld["radiance"]
        lights[i].double_sided = int(Ld.get("double_sided", False))
        for name in ("origin", "dx", "dy", "direction"):
            if name in Ld:
                getattr(lights[i], name)[:] = [float(x) for x in Ld[name]]
        lights[i].angle = float(Ld.get("angle", 0.0))
        lights[i].samples = int(Ld.get("samples", 0))
        lights[i].first_triangle = int(Ld.get("first_triangle", 0))
        lights[i].n_triangles = int(Ld.get("n_triangles", 0))
    cam = _Cam()
    cam.position[:] = [float(x) for x in camera.position]
    cam.direction[:] = [float(x) for x in camera.direction]
    cam.sky[:] = [float(x) for x in camera.sky]
    cam.fov, cam.aspect, cam.near_limit = float(camera.fovAngle), float(camera.aspectRatio), float(camera.nearLimit)
    P = _Params(width, height, spp, wavelengths, max_depth, rr_depth, rr_survival, lambda_range[0],
                lambda_range[1], seed, 0, max_sample_luminance, *sampler)
    if observer is None:
        from liar_amd.observer import standard_observer
        observer = standard_observer()
    ow = np.ascontiguousarray(observer[0], np.float32)
    oxyz = np.ascontiguousarray(observer[1], np.float32)
    xyz = np.zeros((3, height, width), np.float64)
    wgt = np.zeros((height, width), np.float64)
    args = [V.ctypes.data, T.ctypes.data, len(T), S.ctypes.data if len(S) else None, len(S),
            CN.ctypes.data if CN is not None else None, PM.ctypes.data, mats,
            specs, len(cs.spectra), lights, len(cs.lights), ow.ctypes.data, oxyz.ctypes.data, len(ow),
            C.byref(cam), C.byref(P)]
    tail = [xyz.ctypes.data, wgt.ctypes.data, threads or default_threads()]
    if dl_cfg is None:
        rc = L.or_render(*args, *tail)
    else:
        rc = L.or_render_direct(*args, dl_cfg.ctypes.data, *tail)
    assert rc == 0, "oracle render failed"
    return xyz, wgt


def halton_tables():
    """The Halton sampler's sub-sequence digit permutations: (perm uint16,
    off uint32 (285,)); prime k's permutation is perm[off[k]:off[k+1]], of
    length prime k (17, 19, 23, ...)."""
    L = lib()
    L.or_halton_tables.argtypes = [C.c_void_p, C.c_void_p]
    L.or_halton_tables.restype = C.c_uint32
    off = np.zeros(285, np.uint32)
    n = L.or_halton_tables(None, off.ctypes.data)
    perm = np.zeros(n, np.uint16)
    L.or_halton_tables(perm.ctypes.data, off.ctypes.data)
    return perm, off


def halton_sample(j: int, base: int = 2, prime_index: int = -1) -> float:
    """The oracle's FP32 radical inverse of sequence index j: plain in
    ``base``, or (prime_index >= 0) scrambled on that sub-sequence prime."""
    L = lib()
    L.or_halton_sample.argtypes = [C.c_uint64, C.c_uint32, C.c_int]
    L.or_halton_sample.restype = C.c_float
    return float(L.or_halton_sample(int(j), int(base), int(prime_index)))


def set_planted_mistake(k: int) -> None:
    """Test hook of the oracle renderer: 1 drops the path tracer's MIS
    weights, 2 those of the DirectLighting light estimate."""
    lib().or_set_planted_mistake(int(k))


# ---------------------------------------------------------------------------
# single-function probes of the renderer oracle (BSDFs, distributions,
# spectra), for the tests that pin them
# ---------------------------------------------------------------------------

MAT_LAMBERT, MAT_MIRROR, MAT_DIELECTRIC, MAT_CONDUCTOR, MAT_GGX, MAT_BECKMANN = 1, 2, 3, 4, 5, 6


def bsdf_params(kind, refl=1.0, trans=1.0, eta=1.0, kappa=0.0, ior=1.0, roughness_u=0.1, roughness_v=None,
                fuzz=0.0):
    """The oracle's Bsdf record: alpha = max(roughness, 1e-3)^2 per axis and
    eta, kappa floored at 1e-9, as the shaders' doBsdf set them."""
    rv = roughness_u if roughness_v is None else roughness_v
    au, av = max(roughness_u, 1e-3) ** 2, max(rv, 1e-3) ** 2
    return np.array([kind, refl, trans, max(eta, 1e-9), max(kappa, 1e-9), ior, au, av, fuzz], np.float64)


def _probe_sigs(L):
    if getattr(L, "_probe_sig", False):
        return
    P, U32, D = C.c_void_p, C.c_uint32, C.c_double
    L.or_bsdf_eval_n.argtypes = [P, P, P, U32, P, P]
    L.or_bsdf_sample_n.argtypes = [P, P, P, U32, U32, P, P]
    L.or_fresnel_conductor.argtypes = [D, D, D]
    L.or_fresnel_conductor.restype = D
    L.or_mdf_d.argtypes = [U32, D, D, P, P]
    L.or_mdf_d.restype = D
    L.or_mdf_sample_h.argtypes = [U32, D, D, D, D, P]
    L.or_spectrum_eval.argtypes = [P, P, U32, C.c_int, P]
    L._probe_sig = True


def bsdf_eval(prm, wi, wo):
    """Bsdf doEvaluate of the oracle for wi, wo (n, 3) -> (ok bool (n,),
    value (n,), pdf (n,)); no geometric-hemisphere masking (callers pass wo
    on the shading side, as the renderer's bsdf_evaluate lets through)."""
    L = lib()
    _probe_sigs(L)
    wi = np.ascontiguousarray(np.broadcast_to(wi, np.shape(wo)), np.float64)
    wo = np.ascontiguousarray(wo, np.float64)
    n = len(wo)
    out = np.zeros((n, 2), np.float64)
    ok = np.zeros(n, np.uint8)
    L.or_bsdf_eval_n(_p(np.ascontiguousarray(prm, np.float64)), _p(wi), _p(wo), n, _p(out), _p(ok))
    return ok.astype(bool), out[:, 0], out[:, 1]


def bsdf_sample(prm, wi, u, allow=3):
    """Bsdf::sample of the oracle: wi (3,) or (n, 3), u (n, 3) -> (ok, wo (n, 3),
    value, pdf)."""
    L = lib()
    _probe_sigs(L)
    u = np.ascontiguousarray(u, np.float64)
    n = len(u)
    wi = np.ascontiguousarray(np.broadcast_to(wi, (n, 3)), np.float64)
    out = np.zeros((n, 5), np.float64)
    ok = np.zeros(n, np.uint8)
    L.or_bsdf_sample_n(_p(np.ascontiguousarray(prm, np.float64)), _p(wi), _p(u), n, int(allow), _p(out), _p(ok))
    return ok.astype(bool), out[:, :3], out[:, 3], out[:, 4]


def fresnel_conductor(eta, kappa, cos_i):
    L = lib()
    _probe_sigs(L)
    return float(L.or_fresnel_conductor(float(eta), float(kappa), float(cos_i)))


def mdf_d(kind, alpha_u, alpha_v, h):
    """D(h) and pdf_h = D cos(theta_h) of the oracle's distribution (kind
    MAT_GGX or MAT_BECKMANN) at unit vectors h (n, 3)."""
    L = lib()
    _probe_sigs(L)
    h = np.ascontiguousarray(np.atleast_2d(h), np.float64)
    d = np.empty(len(h))
    p = np.empty(len(h))
    pk = np.zeros(1, np.float64)
    for i in range(len(h)):
        d[i] = L.or_mdf_d(int(kind), float(alpha_u), float(alpha_v), _p(h[i]), _p(pk))
        p[i] = pk[0]
    return d, p


def mdf_sample_h(kind, alpha_u, alpha_v, u1, u2):
    L = lib()
    _probe_sigs(L)
    out = np.zeros(3, np.float64)
    L.or_mdf_sample_h(int(kind), float(alpha_u), float(alpha_v), float(u1), float(u2), _p(out))
    return out


def spectrum_eval(spectrum, wavelengths_nm, reflectant=False):
    """A liar_amd.spectra object evaluated by the oracle at wavelengths [nm],
    as its renderer does for one sampled wavelength."""
    L = lib()
    _probe_sigs(L)
    keep = []
    sp = _spec_struct(spectrum, keep)
    w = np.ascontiguousarray(wavelengths_nm, np.float64)
    out = np.zeros(len(w), np.float64)
    L.or_spectrum_eval(C.byref(sp), _p(w), len(w), int(bool(reflectant)), _p(out))
    return out
