"""ctypes binding of the luzgpu C ABI (include/luzgpu.h).

The library is the in-tree ``luz_amd/libluzgpu.so``. Loading fails loudly when it is
missing: there is no CPU fallback for the device path.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libluzgpu.so")

OK = 0
E_INVALID = -1
E_UNSUPPORTED = -2
E_OOM = -3
E_DEVICE = -4
E_IO = -5

INTEGRATOR_WAVEFRONT = 0
INTEGRATOR_MEGAKERNEL = 1
INTEGRATOR_AUTO = 2

EXPORTED = [
    "luzgpu_create", "luzgpu_destroy", "luzgpu_last_error", "luzgpu_scene_load",
    "luzgpu_scene_release", "luzgpu_scene_get_info", "luzgpu_render", "luzgpu_primary_hits",
    "luzgpu_sampler_eval", "luzgpu_scene_tlas_topology", "luzgpu_scene_blas", "luzgpu_render_batch",
    "luzgpu_adjoint", "luzgpu_scene_set_materials", "luzgpu_autotune", "luzgpu_render_denoised", "luzgpu_nfor",
    "luzgpu_sample_counts", "luzgpu_scene_post_settings", "luzgpu_postprocess", "luzgpu_render_noalloc",
    "luzgpu_scene_create",
]


class Settings(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint32),
        ("samples", ctypes.c_int32),
        ("max_bounces", ctypes.c_int32),
        ("sample_begin", ctypes.c_uint32),
        ("sample_count", ctypes.c_int32),
        ("integrator", ctypes.c_uint32),
        ("output_on_device", ctypes.c_uint32),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("count_visits", ctypes.c_uint32),
        ("stage_times", ctypes.c_uint32),
        ("traversal_grid", ctypes.c_uint32),
        ("extend_refill", ctypes.c_int32),
        ("shadow_refill", ctypes.c_int32),
        ("fetch_chunk", ctypes.c_uint32),
        ("lds_stack", ctypes.c_uint32),
        ("autotune", ctypes.c_uint32),
        ("adaptive", ctypes.c_uint32),
        ("adaptive_min_samples", ctypes.c_uint32),
        ("adaptive_check_interval", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("adaptive_threshold", ctypes.c_double),
        ("postprocess", ctypes.c_uint32),
        ("reserved1", ctypes.c_uint32),
    ]


class PostSettings(ctypes.Structure):
    _fields_ = [
        ("exposure", ctypes.c_double),
        ("contrast", ctypes.c_double),
        ("bloom", ctypes.c_uint32),
        ("view_transform", ctypes.c_uint32),
    ]


VIEW_STANDARD, VIEW_AGX, VIEW_ACES, VIEW_RAW = 0, 1, 2, 3
VIEW_NAMES = {"standard": VIEW_STANDARD, "agx": VIEW_AGX, "aces": VIEW_ACES, "raw": VIEW_RAW}


ADAPTIVE_SCENE, ADAPTIVE_OFF, ADAPTIVE_ON = 0, 1, 2


class Tuning(ctypes.Structure):
    _fields_ = [
        ("traversal_grid", ctypes.c_uint32),
        ("extend_refill", ctypes.c_int32),
        ("shadow_refill", ctypes.c_int32),
        ("fetch_chunk", ctypes.c_uint32),
        ("lds_stack", ctypes.c_uint32),
        ("trials", ctypes.c_uint32),
        ("calibration_paths", ctypes.c_uint64),
        ("default_ms", ctypes.c_double),
        ("tuned_ms", ctypes.c_double),
        ("tune_ms", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("rendered_samples", ctypes.c_uint64),
        ("closest_hit_rays", ctypes.c_uint64),
        ("shadow_rays", ctypes.c_uint64),
        ("render_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("nodes_visited", ctypes.c_uint64),
        ("prims_tested", ctypes.c_uint64),
        ("kernel_launches", ctypes.c_uint32),
        ("extend_launches", ctypes.c_uint32),
        ("closest_nodes", ctypes.c_uint64),
        ("closest_prims", ctypes.c_uint64),
        ("extend_ms", ctypes.c_double),
        ("shade_ms", ctypes.c_double),
        ("shadow_ms", ctypes.c_double),
        ("tune_ms", ctypes.c_double),
        ("denoise_ms", ctypes.c_double),
    ]


class NforSettings(ctypes.Structure):
    _fields_ = [
        ("feature_prefilter_radius", ctypes.c_uint32),
        ("regression_radius", ctypes.c_uint32),
        ("patch_radius", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("min_variance", ctypes.c_double),
    ]


class NforBuffers(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32)] + [
        (n, ctypes.c_void_p) for n in ("image", "color_a", "color_b", "color_variance", "features_a", "features_b",
                                       "feature_variance")]


class SceneInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("width", "height", "samples")] + [
        ("max_bounces", ctypes.c_int32)] + [
        (n, ctypes.c_uint32) for n in ("hittables", "triangles", "meshes", "instances", "lights",
                                        "tlas_nodes", "blas_nodes", "env_width", "env_height",
                                        "tlas_depth", "blas_depth", "materials")]


class LuzGpuError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"luzgpu error {code}: {message}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run luz_amd/build.py (or __graft_entry__.build())")
    # torch ships its own HIP runtime under this library's HIP soname (libamdhip64.so.7).
    # Importing torch first makes the loader bind this library to that runtime, so the
    # process holds one HIP/HSA runtime; loaded the other way round there are two, and the
    # one initialised second finds no devices.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.luzgpu_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    lib.luzgpu_destroy.argtypes = [vp]
    lib.luzgpu_destroy.restype = None
    lib.luzgpu_last_error.restype = ctypes.c_char_p
    lib.luzgpu_scene_load.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp)]
    lib.luzgpu_scene_release.argtypes = [vp]
    lib.luzgpu_scene_release.restype = None
    lib.luzgpu_scene_get_info.argtypes = [vp, ctypes.POINTER(SceneInfo)]
    lib.luzgpu_render.argtypes = [vp, vp, ctypes.POINTER(Settings), vp, ctypes.POINTER(Stats)]
    lib.luzgpu_render_noalloc.argtypes = [vp, vp, ctypes.POINTER(Settings), vp, ctypes.POINTER(Stats)]
    lib.luzgpu_primary_hits.argtypes = [vp, vp, ctypes.POINTER(Settings), vp, vp]
    lib.luzgpu_adjoint.argtypes = [vp, vp, ctypes.POINTER(Settings), vp, vp, vp, vp]
    lib.luzgpu_scene_set_materials.argtypes = [vp, ctypes.c_uint32, vp, vp, vp]
    lib.luzgpu_autotune.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(vp), ctypes.POINTER(Settings), ctypes.POINTER(Tuning)]
    lib.luzgpu_render_denoised.argtypes = [vp, vp, ctypes.POINTER(Settings), ctypes.POINTER(NforSettings), vp, vp,
                                           ctypes.POINTER(NforBuffers), ctypes.POINTER(Stats)]
    lib.luzgpu_nfor.argtypes = [vp, ctypes.POINTER(NforBuffers), ctypes.POINTER(NforSettings), vp, ctypes.POINTER(ctypes.c_double)]
    lib.luzgpu_render_batch.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(vp), ctypes.POINTER(Settings),
                                        ctypes.POINTER(vp), ctypes.POINTER(Stats)]
    lib.luzgpu_sampler_eval.argtypes = [vp, ctypes.c_size_t, vp, vp, vp]
    lib.luzgpu_sample_counts.argtypes = [vp, vp, ctypes.c_size_t]
    lib.luzgpu_scene_post_settings.argtypes = [vp, ctypes.POINTER(PostSettings)]
    lib.luzgpu_postprocess.argtypes = [vp, ctypes.POINTER(PostSettings), ctypes.c_uint32, ctypes.c_uint32, vp, vp,
                                       ctypes.c_uint32]
    lib.luzgpu_scene_tlas_topology.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.luzgpu_scene_blas.argtypes = [vp, ctypes.c_uint32, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64)]
    for name in ("luzgpu_create", "luzgpu_scene_load", "luzgpu_scene_get_info", "luzgpu_render", "luzgpu_render_noalloc",
                 "luzgpu_render_batch", "luzgpu_adjoint", "luzgpu_scene_set_materials",
                 "luzgpu_primary_hits", "luzgpu_sampler_eval", "luzgpu_scene_tlas_topology", "luzgpu_scene_blas"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != OK:
        msg = load().luzgpu_last_error().decode(errors="replace")
        raise LuzGpuError(code, msg)
