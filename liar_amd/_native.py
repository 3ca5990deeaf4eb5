"""ctypes binding of the C-ABI in ``include/liar_amd.h``.

The shared library is built in-tree by ``python -m liar_amd.build``.  Loading
fails loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libliar_amd.so"

LR_OK = 0
LR_MISS = 0xFFFFFFFF
LR_CHILD_EMPTY = 0xFFFFFFFF
NODE_BYTES = 128
NODE_Q8_BYTES = 64  # lr_node_q8, the device traversal's node
PRIM_BYTES = 48


class LiarError(RuntimeError):
    """A non-zero status from the native library (message from lr_last_error)."""


class RaysSoa(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tnear", "tfar")]


class HitsSoa(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("prim", "t", "u", "v")]


class QbvhInfo(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("n_prims", C.c_uint32),
        ("n_triangles", C.c_uint32),
        ("n_spheres", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("max_stack", C.c_uint32),
        ("sah_cost", C.c_double),
        ("build_seconds", C.c_double),
        ("scene_lo", C.c_float * 3),
        ("scene_hi", C.c_float * 3),
    ]


class TraceConfig(C.Structure):
    _fields_ = [("block", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("lds_stack", C.c_uint32),
                ("refill_min", C.c_uint32), ("schedule", C.c_uint32)]


class TuneEntry(C.Structure):
    _fields_ = [("cfg", TraceConfig), ("ms", C.c_float)]


class Spectrum(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("count", C.c_uint32), ("w", C.c_void_p), ("v", C.c_void_p),
                ("value", C.c_float), ("lo", C.c_float), ("hi", C.c_float), ("temperature", C.c_float),
                ("scale", C.c_float)]


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("reflectance", C.c_uint32), ("transmittance", C.c_uint32),
                ("eta", C.c_uint32), ("kappa", C.c_uint32), ("emission", C.c_uint32), ("light", C.c_uint32),
                ("roughness", C.c_float), ("roughness_v", C.c_float), ("fuzz", C.c_float)]


def material_struct(m: dict) -> "Material":
    """lr_material of a compiled material dict (roughness_v defaults to
    roughness: isotropic; fuzz to 0: a perfect mirror)."""
    return Material(m["type"], m["reflectance"], m["transmittance"], m["eta"], m["kappa"], m["emission"],
                    m["light"], float(m["roughness"]), float(m.get("roughness_v", m["roughness"])),
                    float(m.get("fuzz", 0.0)))


class BsdfDesc(C.Structure):
    """lr_bsdf."""
    _fields_ = [("type", C.c_uint32), ("reflectance", C.c_float), ("transmittance", C.c_float),
                ("eta", C.c_float), ("kappa", C.c_float), ("ior", C.c_float), ("roughness_u", C.c_float),
                ("roughness_v", C.c_float), ("fuzz", C.c_float)]


class Light(C.Structure):
    _fields_ = [("type", C.c_uint32), ("radiance", C.c_uint32), ("double_sided", C.c_uint32),
                ("origin", C.c_float * 3), ("dx", C.c_float * 3), ("dy", C.c_float * 3),
                ("direction", C.c_float * 3), ("angle", C.c_float), ("samples", C.c_uint32),
                ("first_triangle", C.c_uint32), ("n_triangles", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("sky", C.c_float * 3),
                ("fov", C.c_float), ("aspect", C.c_float), ("near_limit", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("wavelengths", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("rr_survival", C.c_float), ("lambda_min", C.c_float), ("lambda_max", C.c_float),
                ("seed", C.c_uint64), ("chunk_paths", C.c_uint32), ("max_sample_luminance", C.c_float),
                ("graph", C.c_uint32), ("sampler", C.c_uint32), ("sampler_generations", C.c_uint32),
                ("sample_offset", C.c_uint64)]


LR_SAMPLER_STRATIFIER = 0
LR_SAMPLER_HALTON = 1


class MeshData(C.Structure):
    _fields_ = [("n_verts", C.c_uint32), ("n_tris", C.c_uint32), ("verts", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)), ("tris", C.POINTER(C.c_uint32))]


LR_PLY_FAN_ALL = 1


class BatchItem(C.Structure):
    _fields_ = [("camera", Camera), ("params", RenderParams), ("film_xyz", C.c_void_p), ("film_weight", C.c_void_p)]


LR_BATCH_GRAPH = 1


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays", C.c_uint64), ("chunks", C.c_uint32), ("seconds", C.c_double)]


class DirectParams(C.Structure):
    _fields_ = [("max_generation", C.c_uint32), ("secondary_light_samples", C.c_uint32),
                ("trace_glossy", C.c_uint32)]


class Instance(C.Structure):
    _fields_ = [("geometry", C.c_uint32), ("local_to_world", C.c_float * 12)]


class TlasInfo(C.Structure):
    _fields_ = [("n_instances", C.c_uint32), ("n_nodes", C.c_uint32), ("max_depth", C.c_uint32),
                ("max_stack", C.c_uint32), ("n_geometry_nodes", C.c_uint64), ("n_geometry_prims", C.c_uint64),
                ("build_seconds", C.c_double), ("scene_lo", C.c_float * 3), ("scene_hi", C.c_float * 3)]


class RgbSpace(C.Structure):
    _fields_ = [("red", C.c_float * 2), ("green", C.c_float * 2), ("blue", C.c_float * 2), ("white", C.c_float * 2),
                ("gamma", C.c_float)]


class TonemapParams(C.Structure):
    _fields_ = [("space", RgbSpace), ("tone_mapping", C.c_uint32), ("auto_exposure", C.c_uint32),
                ("exposure_stops", C.c_float), ("exposure_correction_stops", C.c_float), ("middle_grey", C.c_float)]


class TonemapInfo(C.Structure):
    _fields_ = [("exposure_stops", C.c_float), ("gain", C.c_float), ("average_luminance", C.c_float),
                ("max_luminance", C.c_float)]


class GuideParams(C.Structure):
    _fields_ = [("selection", C.c_float), ("train", C.c_uint32), ("train_steps", C.c_uint32),
                ("train_batch", C.c_uint32), ("max_records", C.c_uint32), ("weight_clamp", C.c_float),
                ("light_share", C.c_float), ("learned_selection", C.c_uint32), ("balance_heuristic", C.c_uint32),
                ("image_weighting", C.c_uint32), ("record_depth", C.c_uint32), ("guide_depth", C.c_uint32),
                ("learned_rr", C.c_uint32), ("rr_scale", C.c_float), ("rr_min_survival", C.c_float),
                ("rr_start", C.c_uint32)]


LR_NONE = 0xFFFFFFFF
SPEC_CONSTANT, SPEC_SAMPLED, SPEC_BINNED, SPEC_BLACKBODY = 0, 1, 2, 3
MAT_NONE, MAT_LAMBERT, MAT_MIRROR, MAT_DIELECTRIC, MAT_CONDUCTOR, MAT_COOK_TORRANCE = 0, 1, 2, 3, 4, 5
MAT_COOK_TORRANCE_BECKMANN = 6
LIGHT_PARALLELOGRAM, LIGHT_SKY, LIGHT_SUN, LIGHT_MESH = 0, 1, 2, 3

_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    "lr_last_error": (C.c_char_p, []),
    "lr_version": (C.c_char_p, []),
    "lr_context_create": (_I, [_I, C.POINTER(_P)]),
    "lr_context_destroy": (_I, [_P]),
    "lr_qbvh_create": (_I, [_P, _P, _U32, _P, _U32, _P, _U32, C.POINTER(_P)]),
    "lr_qbvh_destroy": (_I, [_P]),
    "lr_qbvh_get_info": (_I, [_P, C.POINTER(QbvhInfo)]),
    "lr_qbvh_export": (_I, [_P, _P, _P]),
    "lr_qbvh_export_q8": (_I, [_P, _P]),
    "lr_trace_closest": (_I, [_P, C.POINTER(RaysSoa), _U32, C.POINTER(HitsSoa), _P]),
    "lr_trace_closest_noalloc": (_I, [_P, C.POINTER(RaysSoa), _U32, C.POINTER(HitsSoa), _P]),
    "lr_trace_any": (_I, [_P, C.POINTER(RaysSoa), _U32, _P, _P]),
    "lr_set_trace_config": (_I, [_P, C.POINTER(TraceConfig)]),
    "lr_get_trace_config": (_I, [_P, C.POINTER(TraceConfig)]),
    "lr_autotune_trace": (_I, [_P, C.POINTER(RaysSoa), _U32, _P, _P, _U32, _P, _U32, C.POINTER(_U32),
                               C.POINTER(TraceConfig), _P]),
    "lr_gen_random_triangles": (_I, [_U64, _U32, _F, _P, _P]),
    "lr_gen_random_rays": (_I, [_U64, _U32, _F, _F, _P, _P, _P, _P, _P, _P, _P, _P]),
    "lr_philox4x32": (None, [_U64, _U32, _U32, _U32, _U32, _P]),
    "lr_gen_terrain": (_I, [_U64, _U32, _U32, _F, _F, _F, _P, _P]),
    "lr_render_guided": (_I, [_P, C.POINTER(Camera), C.POINTER(RenderParams), _P, C.POINTER(GuideParams), _P, _P,
                              C.POINTER(RenderStats), _P]),
    "lr_load_ply": (_I, [C.c_char_p, _U32, C.POINTER(MeshData)]),
    "lr_mesh_data_free": (None, [C.POINTER(MeshData)]),
    "lr_batch_create": (_I, [_P, _P, _U32, C.POINTER(_P)]),
    "lr_batch_destroy": (_I, [_P]),
    "lr_render_batch": (_I, [_P, _P, _U32, _P, C.POINTER(RenderStats), _P]),
    "lr_guide_param_count": (_I, []),
    "lr_guide_create": (_I, [_P, _U64, _F, C.POINTER(_P)]),
    "lr_guide_destroy": (_I, [_P]),
    "lr_guide_get_params": (_I, [_P, _P]),
    "lr_guide_set_params": (_I, [_P, _P]),
    "lr_guide_infer": (_I, [_P, _P, _U32, _P, _P, _P]),
    "lr_guide_train_step": (_I, [_P, _P, _P, _P, _U32, _P, _P, _P]),
    "lr_guide_set_radiance_weight": (_I, [_P, _F]),
    "lr_set_standard_observer": (_I, [_P, _P, _U32]),
    "lr_scene_create": (_I, [_P, _P, _P, _P, _U32, _P, _U32, _P, _U32, C.POINTER(_P)]),
    "lr_scene_destroy": (_I, [_P]),
    "lr_scene_set_normals": (_I, [_P, _P, _U32]),
    "lr_bsdf_eval": (_I, [_P, _P, _P, _P, _U32, _P, _P, _P]),
    "lr_bsdf_sample": (_I, [_P, _P, _P, _P, _U32, _U32, _P, _P, _P, _P]),
    "lr_spectrum_eval": (_I, [_P, _P, _P, _U32, _U32, _P, _P]),
    "lr_halton_permutation": (_I, [_U32, _P, _P, _U32]),
    "lr_render": (_I, [_P, C.POINTER(Camera), C.POINTER(RenderParams), _P, _P, C.POINTER(RenderStats), _P]),
    "lr_render_direct": (_I, [_P, C.POINTER(Camera), C.POINTER(RenderParams), C.POINTER(DirectParams), _P, _P,
                              C.POINTER(RenderStats), _P]),
    "lr_tlas_create": (_I, [_P, _P, _U32, _P, _U32, C.POINTER(_P)]),
    "lr_tlas_destroy": (_I, [_P]),
    "lr_tlas_get_info": (_I, [_P, C.POINTER(TlasInfo)]),
    "lr_trace_closest_instanced": (_I, [_P, C.POINTER(RaysSoa), _U32, C.POINTER(HitsSoa), _P, _P]),
    "lr_trace_any_instanced": (_I, [_P, C.POINTER(RaysSoa), _U32, _P, _P]),
    "lr_rgb_space_matrices": (_I, [C.POINTER(RgbSpace), _P, _P]),
    "lr_tonemap": (_I, [_P, _P, _P, _P, _U32, _U32, C.POINTER(TonemapParams), _P, _P, C.POINTER(TonemapInfo), _P]),
    "lr_image_write": (_I, [C.c_char_p, _P, _U32, _U32, C.POINTER(RgbSpace)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libliar_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"native library {LIB_PATH} is missing; build it with `python -m liar_amd.build`")
        # torch's HIP runtime must be the process's HIP runtime: device
        # pointers and streams pass between torch and this library, so torch
        # (which ships its own runtime) is loaded first and this library then
        # binds to the same one
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != LR_OK:
        msg = lib().lr_last_error().decode(errors="replace")
        raise LiarError(f"{what}: status {rc}: {msg}" if what else f"status {rc}: {msg}")


def ptr(a) -> int:
    """Address of a numpy array or torch tensor (contiguous)."""
    if a is None:
        return 0
    if hasattr(a, "data_ptr"):
        return int(a.data_ptr())
    return int(a.ctypes.data)
