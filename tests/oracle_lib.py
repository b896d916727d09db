"""ctypes access to the FP64 oracle (oracle/liboracle.so) — the checker, never the product."""

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
        lib = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        lib.oracle_sampler.argtypes = [ctypes.c_size_t, vp, vp, vp]
        lib.oracle_sampler.restype = None
        lib.oracle_render.argtypes = [ctypes.c_char_p] + [ctypes.c_uint32] * 7 + [vp, vp]
        lib.oracle_render.restype = ctypes.c_int
        lib.oracle_primary.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp]
        lib.oracle_primary.restype = ctypes.c_int
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_ggx_pdf_integral.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint32]
        lib.oracle_ggx_pdf_integral.restype = ctypes.c_double
        lib.oracle_ggx_ndf_integral.argtypes = [ctypes.c_double, ctypes.c_uint32]
        lib.oracle_ggx_ndf_integral.restype = ctypes.c_double
        lib.oracle_dielectric_fresnel.argtypes = [ctypes.c_double] * 3
        lib.oracle_dielectric_fresnel.restype = ctypes.c_double
        lib.oracle_conductor_fresnel.argtypes = [ctypes.c_double] * 3
        lib.oracle_conductor_fresnel.restype = ctypes.c_double
        lib.oracle_hg_pdf.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.oracle_hg_pdf.restype = ctypes.c_double
        lib.oracle_hg_sample_cos.argtypes = [ctypes.c_double, ctypes.c_float, ctypes.c_float]
        lib.oracle_hg_sample_cos.restype = ctypes.c_double
        lib.oracle_adjoint.argtypes = [ctypes.c_char_p] + [ctypes.c_uint32] * 6 + [vp, vp, ctypes.c_uint32]
        lib.oracle_adjoint.restype = ctypes.c_int
        lib.oracle_ggx_dlog_alpha.argtypes = [ctypes.c_double] * 5 + [vp]
        lib.oracle_ggx_dlog_alpha.restype = ctypes.c_double
        lib.oracle_ggx_fallback.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, vp, vp]
        lib.oracle_ggx_fallback.restype = ctypes.c_double
        lib.oracle_render_adaptive.argtypes = ([ctypes.c_char_p] + [ctypes.c_uint32] * 8 + [ctypes.c_double, vp, vp])
        lib.oracle_render_adaptive.restype = ctypes.c_int
        _lib = lib
    return _lib


def sampler(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    n = keys.shape[0]
    s1 = np.zeros(n, dtype=np.float64)
    s2 = np.zeros((n, 2), dtype=np.float32)
    load().oracle_sampler(n, keys.ctypes.data, s1.ctypes.data, s2.ctypes.data)
    return s1, s2


def render(path, seed, spp, width, height, sample_begin=0, max_bounces=None, threads=None):
    lib = load()
    out = np.zeros((height, width, 3), dtype=np.float64)
    ms = ctypes.c_double()
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.oracle_render(path.encode(), seed, spp, sample_begin, width, height,
                           0 if max_bounces is None else max_bounces + 1, threads, out.ctypes.data, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return out, ms.value


def render_adaptive(path, seed, spp, width, height, min_samples=0, check_interval=0, threshold=0.0,
                    max_bounces=None, threads=None):
    """Adaptive per-pixel stopping (0 / 0.0: the scene file's values): (image, counts), or None
    when the rule cannot stop a pixel early."""
    lib = load()
    out = np.zeros((height, width, 3), dtype=np.float64)
    counts = np.zeros((height, width), dtype=np.uint32)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.oracle_render_adaptive(path.encode(), seed, spp, width, height,
                                    0 if max_bounces is None else max_bounces + 1, threads, min_samples,
                                    check_interval, threshold, out.ctypes.data, counts.ctypes.data)
    if rc < 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return None if rc == 1 else (out, counts)


def primary(path, seed, width, height):
    lib = load()
    ids = np.zeros((height, width, 2), dtype=np.int64)
    t = np.zeros((height, width), dtype=np.float64)
    rc = lib.oracle_primary(path.encode(), seed, width, height, ids.ctypes.data, t.ctypes.data)
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return ids, t


def adjoint(path, seed, spp, width, height, weights, material_count, max_bounces=None, threads=None):
    """Path-replay gradients of sum(weights * image): (material_count, 7) FP64, per material
    albedo rgb, GGX roughness, emission rgb."""
    lib = load()
    w = np.ascontiguousarray(weights, dtype=np.float32)
    assert w.shape == (height, width, 3)
    grads = np.zeros((material_count, 7), dtype=np.float64)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib.oracle_adjoint(path.encode(), seed, spp, width, height, 0 if max_bounces is None else max_bounces + 1,
                            threads, w.ctypes.data, grads.ctypes.data, material_count)
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return grads


def ggx_dlog_alpha(mu_v, mu_l, phi, alpha, h=1e-6):
    fd = ctypes.c_double()
    d = load().oracle_ggx_dlog_alpha(mu_v, mu_l, phi, alpha, h, ctypes.byref(fd))
    return d, fd.value


def ggx_fallback(mu, alpha, nodes=0, grid=0):
    """(P, dP/dalpha, brute-force frequency or None) of the metal sampler's mirror fallback."""
    dp, brute = ctypes.c_double(), ctypes.c_double()
    p = load().oracle_ggx_fallback(mu, alpha, nodes, grid, ctypes.byref(dp), ctypes.byref(brute))
    return p, dp.value, (brute.value if grid else None)
