"""Python mirror of the reference's entry points for the radiance path.

``SceneFile.read(path)`` and ``Renderer.render(scene, ...)`` keep the reference's names
(include/luz/SceneFile/SceneFile.hpp:8, include/luz/Renderer/Renderer.hpp:7) and the same
error behaviour: bad files raise (the reference throws ``std::runtime_error``), a render
returns the raw scene-linear image that ``_manageThreads`` leaves in ``Scene::getImage()``.
Everything runs on the GPU through the C ABI; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


class Context:
    """One HIP device and stream (the reference's worker pool, Threads.cpp:375-548)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        self.device = int(device)
        self._last_resolution = None  # (W, H) of the last single-scene render
        _lib.check(self._lib.luzgpu_create(ctypes.byref(self._h), device))

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            self._lib.luzgpu_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sampler_eval(self, keys: np.ndarray):
        """keys: (n, 7) uint32 rows (seed, x, y, sample, stream, bounce, dim)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        n = keys.shape[0]
        s1 = np.zeros(n, dtype=np.float64)
        s2 = np.zeros((n, 2), dtype=np.float32)
        _lib.check(self._lib.luzgpu_sampler_eval(self._h, n, keys.ctypes.data, s1.ctypes.data, s2.ctypes.data))
        return s1, s2


@dataclass
class RenderStats:
    rendered_samples: int
    closest_hit_rays: int
    shadow_rays: int
    render_ms: float
    kernel_ms: float
    kernel_launches: int
    nodes_visited: int
    prims_tested: int
    extend_launches: int = 0
    closest_nodes: int = 0
    closest_prims: int = 0
    extend_ms: float = 0.0
    shade_ms: float = 0.0
    shadow_ms: float = 0.0
    tune_ms: float = 0.0
    denoise_ms: float = 0.0

    @staticmethod
    def of(st) -> "RenderStats":
        return RenderStats(st.rendered_samples, st.closest_hit_rays, st.shadow_rays, st.render_ms, st.kernel_ms,
                           st.kernel_launches, st.nodes_visited, st.prims_tested, st.extend_launches,
                           st.closest_nodes, st.closest_prims, st.extend_ms, st.shade_ms, st.shadow_ms, st.tune_ms,
                           st.denoise_ms)


@dataclass
class Tuning:
    """Launch shape chosen by the autotuner (luzgpu_autotune) and its calibration."""
    grid: int
    extend_refill: int
    shadow_refill: int
    fetch_chunk: int
    lds_stack: int
    trials: int
    calibration_paths: int
    default_ms: float
    tuned_ms: float
    tune_ms: float

    def as_dict(self) -> dict:
        return {"grid": self.grid, "extend_refill": self.extend_refill, "shadow_refill": self.shadow_refill,
                "fetch_chunk": self.fetch_chunk, "lds_stack": self.lds_stack}


class Scene:
    """A scene uploaded to the device (SceneFile::read + Renderer.cpp:18-20 pre-render sync)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self._lib = _lib.load()
        self._h = handle
        info = _lib.SceneInfo()
        _lib.check(self._lib.luzgpu_scene_get_info(self._h, ctypes.byref(info)))
        self.info = {name: getattr(info, name) for name, _ in _lib.SceneInfo._fields_}

    @property
    def width(self) -> int:
        return self.info["width"]

    @property
    def height(self) -> int:
        return self.info["height"]

    def release(self) -> None:
        if self._h:
            self._lib.luzgpu_scene_release(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def tlas_topology(self) -> str:
        need = ctypes.c_size_t()
        _lib.check(self._lib.luzgpu_scene_tlas_topology(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _lib.check(self._lib.luzgpu_scene_tlas_topology(self._h, buf, need.value, None))
        return buf.value.decode()

    def blas(self, mesh: int):
        nodes, tris = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check(self._lib.luzgpu_scene_blas(self._h, mesh, None, None, None, ctypes.byref(nodes), ctypes.byref(tris)))
        boxes = np.zeros((nodes.value, 6), dtype=np.float64)
        links = np.zeros((nodes.value, 5), dtype=np.int64)
        idx = np.zeros(tris.value, dtype=np.int64)
        _lib.check(self._lib.luzgpu_scene_blas(self._h, mesh, boxes.ctypes.data, links.ctypes.data, idx.ctypes.data,
                                                   None, None))
        return boxes, links, idx

    def post_settings(self):
        """The scene file's post-processing settings (a luzgpu_post_settings)."""
        ps = _lib.PostSettings()
        _lib.check(self._lib.luzgpu_scene_post_settings(self._h, ctypes.byref(ps)))
        return ps

    def set_materials(self, albedo=None, roughness=None, emission=None) -> None:
        """Per-material parameters (scene-file order): albedo (n, 3) for Lambertian, roughness
        (n,) for metal, emission (n, 3) radiance for emitters; None keeps the current values."""
        n = self.info["materials"]

        def arr(x, shape):
            if x is None:
                return None
            if hasattr(x, "detach"):
                x = x.detach().cpu().numpy()
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(shape))
            return a
        a, r, e = arr(albedo, (n, 3)), arr(roughness, (n,)), arr(emission, (n, 3))
        _lib.check(self._lib.luzgpu_scene_set_materials(self._h, n, a.ctypes.data if a is not None else None,
                                                        r.ctypes.data if r is not None else None,
                                                        e.ctypes.data if e is not None else None))

    def adjoint(self, dl_dimage, seed: int, samples=None, max_bounces=None, sample_begin=0, sample_count=None,
                resolution=None, adaptive=None):
        """Path-replay adjoint of Renderer.render with the same arguments. dl_dimage: dLoss/dImage
        (H, W, 3), a numpy array or a device tensor. Returns (d_albedo (n,3), d_roughness (n,),
        d_emission (n,3)) as float64 numpy arrays. The adjoint replays every sample of every
        pixel, so it is refused (LuzGpuError) where adaptive sampling would be in effect: pass
        adaptive=False, to this call and to the render it differentiates."""
        n = self.info["materials"]
        on_device = hasattr(dl_dimage, "data_ptr") and getattr(dl_dimage, "is_cuda", False)
        s = _settings(seed, samples, max_bounces, sample_begin, sample_count, "megakernel", on_device=on_device,
                      resolution=resolution, adaptive=adaptive)
        w, h = resolution if resolution is not None else (self.width, self.height)
        if on_device:
            g = dl_dimage.contiguous().float()
            ptr = _device_buffer(g, self.ctx, w * h * 3, "Scene.adjoint(dl_dimage)")
        else:
            g = np.ascontiguousarray(np.asarray(dl_dimage, dtype=np.float32))
            if g.size != w * h * 3:
                raise ValueError(f"Scene.adjoint(dl_dimage): expected {w * h * 3} elements (H*W*3), got {g.size}")
            ptr = g.ctypes.data
        da = np.zeros((n, 3), dtype=np.float64)
        dr = np.zeros(n, dtype=np.float64)
        de = np.zeros((n, 3), dtype=np.float64)
        _lib.check(self._lib.luzgpu_adjoint(self.ctx.handle, self._h, ctypes.byref(s), ctypes.c_void_p(ptr), da.ctypes.data,
                                            dr.ctypes.data, de.ctypes.data))
        return da, dr, de


class SceneFile:
    @staticmethod
    def read(ctx, path: str) -> Scene:
        """Loads ``path``; with ``ctx=None`` the scene is built on the host only (inspection)."""
        h = ctypes.c_void_p()
        lib = _lib.load()
        _lib.check(lib.luzgpu_scene_load(ctx.handle if ctx is not None else None, path.encode(), ctypes.byref(h)))
        return Scene(ctx, h)


def _device_buffer(t, ctx, count: int, what: str) -> int:
    """The device address of a caller tensor the library writes or reads ``count`` floats
    through, after checking that it is one: a GPU tensor on the context's device, float32,
    contiguous, with exactly ``count`` elements. Anything else raises ValueError here instead
    of reaching a kernel as a bad pointer."""
    if not hasattr(t, "data_ptr") or not getattr(t, "is_cuda", False):
        raise ValueError(f"{what}: expected a GPU tensor (torch, device cuda:{ctx.device})")
    if str(t.dtype) != "torch.float32":
        raise ValueError(f"{what}: expected float32, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous tensor")
    if t.numel() != count:
        raise ValueError(f"{what}: expected {count} elements (H*W*3), got {t.numel()}")
    if t.device.index != ctx.device:
        raise ValueError(f"{what}: tensor on cuda:{t.device.index}, context on cuda:{ctx.device}")
    return t.data_ptr()


def _settings(seed, samples, max_bounces, sample_begin, sample_count, integrator, on_device=False, resolution=None,
              count_visits=False, stage_times=False, tuning=None, adaptive=None, postprocess=False):
    s = _lib.Settings()
    s.postprocess = 1 if postprocess else 0
    # adaptive: None = the scene's adaptive= setting, False / True = --no-adaptive / --adaptive,
    # a dict {"min_samples", "check_interval", "threshold"} = on with those values
    if adaptive is False:
        s.adaptive = _lib.ADAPTIVE_OFF
    elif adaptive is True:
        s.adaptive = _lib.ADAPTIVE_ON
    elif isinstance(adaptive, dict):
        s.adaptive = _lib.ADAPTIVE_ON
        s.adaptive_min_samples = int(adaptive.get("min_samples", 0))
        s.adaptive_check_interval = int(adaptive.get("check_interval", 0))
        s.adaptive_threshold = float(adaptive.get("threshold", 0.0))
    elif adaptive is not None:
        raise ValueError("adaptive: None, a bool, or a dict of min_samples / check_interval / threshold")
    if tuning == "auto":
        s.autotune = 1
    elif tuning:
        s.traversal_grid = int(tuning.get("grid", 0))
        s.extend_refill = int(tuning.get("extend_refill", 0))
        s.shadow_refill = int(tuning.get("shadow_refill", 0))
        s.fetch_chunk = int(tuning.get("fetch_chunk", 0))
        s.lds_stack = int(tuning.get("lds_stack", 0))
    s.count_visits = 1 if count_visits else 0
    s.stage_times = 1 if stage_times else 0
    if resolution is not None:
        s.width, s.height = int(resolution[0]), int(resolution[1])
    s.seed = int(seed) & 0xFFFFFFFF
    s.samples = -1 if samples is None else int(samples)
    s.max_bounces = -1 if max_bounces is None else int(max_bounces)
    s.sample_begin = int(sample_begin)
    s.sample_count = -1 if sample_count is None else int(sample_count)
    s.integrator = {"wavefront": _lib.INTEGRATOR_WAVEFRONT, "megakernel": _lib.INTEGRATOR_MEGAKERNEL,
                    "auto": _lib.INTEGRATOR_AUTO}[integrator]
    s.output_on_device = 1 if on_device else 0
    return s


def _post_settings(post: dict):
    ps = _lib.PostSettings()
    ps.exposure = float(post.get("exposure", 0.0))
    ps.contrast = float(post.get("contrast", 1.0))
    ps.bloom = 1 if post.get("bloom", True) else 0
    ps.view_transform = _lib.VIEW_NAMES[post.get("view_transform", "agx")]
    return ps


class Renderer:
    """``Renderer::render`` for the device path."""

    @staticmethod
    def render(scene: Scene, seed: int, samples=None, max_bounces=None, sample_begin=0, sample_count=None,
               integrator: str = "wavefront", out=None, resolution=None, count_visits=False, stage_times=False,
               tuning=None, adaptive=None, postprocess=False):
        """Returns (image[H, W, 3] float32, RenderStats). ``out`` may be a device tensor
        (torch, float32, contiguous, H*W*3) that receives the image without a host copy.
        ``tuning`` ({"grid", "extend_refill", "shadow_refill", "fetch_chunk", "lds_stack"}) sets
        the wavefront launch shape, or "auto" uses the context's autotuned one (tuning it
        first if needed); it never changes the image. ``adaptive``: None follows the scene
        file, False / True turn per-pixel stopping off / on, a dict {"min_samples",
        "check_interval", "threshold"} turns it on with those values;
        ``Renderer.sample_counts`` then gives each pixel's sample count. ``postprocess``:
        the image is the scene's display post-processing of the raw image (exposure, bloom,
        view transform, firefly suppression), as Renderer::render leaves Scene::getImage()."""
        s = _settings(seed, samples, max_bounces, sample_begin, sample_count, integrator, on_device=out is not None,
                      resolution=resolution, count_visits=count_visits, stage_times=stage_times, tuning=tuning,
                      adaptive=adaptive, postprocess=postprocess)
        st = _lib.Stats()
        w, h = resolution if resolution is not None else (scene.width, scene.height)
        if out is not None:
            ptr = _device_buffer(out, scene.ctx, w * h * 3, "Renderer.render(out=...)")
            img = out
        else:
            img = np.zeros((h, w, 3), dtype=np.float32)
            ptr = img.ctypes.data
        _lib.check(scene.ctx._lib.luzgpu_render(scene.ctx.handle, scene._h, ctypes.byref(s), ctypes.c_void_p(ptr),
                                                ctypes.byref(st)))
        scene.ctx._last_resolution = (w, h)
        return img, RenderStats.of(st)

    @staticmethod
    def render_batch(scenes, seeds, samples=None, max_bounces=None, outs=None, resolutions=None, count_visits=False,
                     stage_times=False, tuning=None, adaptive=None, integrator: str = "wavefront"):
        """Renders several scenes in one multi-scene launch sequence (luzgpu_render_batch).

        ``seeds``, ``samples``, ``max_bounces``, ``resolutions`` and ``adaptive`` are
        per-scene lists (or None for the scene defaults); ``outs`` optional per-scene device
        tensors. ``integrator``: "wavefront" (multi-scene passes) or "megakernel" (every
        scene's paths in one plain-kernel launch, for scenes too small to fill the GPU).
        Returns (images, aggregate RenderStats)."""
        n = len(scenes)
        if n == 0:
            raise ValueError("render_batch needs at least one scene")
        ctx = scenes[0].ctx
        arr = (_lib.Settings * n)()
        images, ptrs = [], (ctypes.c_void_p * n)()
        for i, sc in enumerate(scenes):
            res = resolutions[i] if resolutions else None
            arr[i] = _settings(seeds[i], samples[i] if samples else None, max_bounces[i] if max_bounces else None, 0, None,
                               integrator, on_device=outs is not None, resolution=res,
                               count_visits=count_visits, stage_times=stage_times, tuning=tuning,
                               adaptive=adaptive[i] if adaptive else None)
            w, h = res if res is not None else (sc.width, sc.height)
            if outs is not None:
                images.append(outs[i])
                ptrs[i] = _device_buffer(outs[i], ctx, w * h * 3, f"Renderer.render_batch(outs[{i}])")
            else:
                img = np.zeros((h, w, 3), dtype=np.float32)
                images.append(img)
                ptrs[i] = img.ctypes.data
        handles = (ctypes.c_void_p * n)(*[sc._h.value for sc in scenes])
        st = _lib.Stats()
        _lib.check(ctx._lib.luzgpu_render_batch(ctx.handle, n, handles, arr, ptrs, ctypes.byref(st)))
        return images, RenderStats.of(st)

    @staticmethod
    def autotune(scenes, seeds, samples=None, max_bounces=None, resolutions=None) -> Tuning:
        """Autotunes the wavefront launch shape for these scenes and settings (one scene: the
        single-scene path; several: the merged batch). Later renders with tuning="auto" and
        the same scenes and settings use the result."""
        n = len(scenes)
        ctx = scenes[0].ctx
        arr = (_lib.Settings * n)()
        for i in range(n):
            arr[i] = _settings(seeds[i], samples[i] if samples else None, max_bounces[i] if max_bounces else None, 0, None,
                               "wavefront", resolution=resolutions[i] if resolutions else None)
        handles = (ctypes.c_void_p * n)(*[sc._h.value for sc in scenes])
        t = _lib.Tuning()
        _lib.check(ctx._lib.luzgpu_autotune(ctx.handle, n, handles, arr, ctypes.byref(t)))
        return Tuning(t.traversal_grid, t.extend_refill, t.shadow_refill, t.fetch_chunk, t.lds_stack, t.trials,
                      t.calibration_paths, t.default_ms, t.tuned_ms, t.tune_ms)

    NFOR_FIELDS = (("color_a", 3), ("color_b", 3), ("color_variance", 1), ("features_a", 11), ("features_b", 11),
                   ("feature_variance", 11))

    @staticmethod
    def _nfor_settings(nfor):
        if nfor is None:
            return None
        ns = _lib.NforSettings()
        ns.feature_prefilter_radius = int(nfor.get("feature_prefilter_radius", 1))
        ns.regression_radius = int(nfor.get("regression_radius", 2))
        ns.patch_radius = int(nfor.get("patch_radius", 1))
        ns.min_variance = float(nfor.get("min_variance", 1e-5))
        return ns

    @staticmethod
    def render_denoised(scene: Scene, seed: int, samples=None, max_bounces=None, resolution=None, nfor=None,
                        buffers=False, adaptive=None, integrator: str = "wavefront", postprocess=False):
        """Render with denoising on (luzgpu_render_denoised): returns (raw image, denoised image,
        RenderStats, NFOR buffers or None). ``nfor``: {"feature_prefilter_radius",
        "regression_radius", "patch_radius", "min_variance"} (defaults as NFORSettings);
        ``adaptive`` as for ``render``."""
        s = _settings(seed, samples, max_bounces, 0, None, integrator, resolution=resolution, adaptive=adaptive,
                      postprocess=postprocess)
        w, h = resolution if resolution is not None else (scene.width, scene.height)
        img = np.zeros((h, w, 3), dtype=np.float32)
        den = np.zeros((h, w, 3), dtype=np.float32)
        st = _lib.Stats()
        out = None
        bp = None
        if buffers:
            out = {"image": np.zeros((h, w, 3))}
            out.update({k: np.zeros((h, w, c)) for k, c in Renderer.NFOR_FIELDS})
            bp = _lib.NforBuffers()
            for k, a in out.items():
                setattr(bp, k, a.ctypes.data)
        ns = Renderer._nfor_settings(nfor)
        _lib.check(scene.ctx._lib.luzgpu_render_denoised(scene.ctx.handle, scene._h, ctypes.byref(s),
                                                         ctypes.byref(ns) if ns is not None else None, img.ctypes.data,
                                                         den.ctypes.data, ctypes.byref(bp) if bp is not None else None,
                                                         ctypes.byref(st)))
        scene.ctx._last_resolution = (w, h)
        return img, den, RenderStats.of(st), out

    @staticmethod
    def nfor(ctx, buffers: dict, nfor=None):
        """Denoise::applyNFOR on the device for host FP64 buffers {color_a, color_b,
        color_variance, features_a, features_b, feature_variance} (H, W, c); returns
        (denoised (H, W, 3) FP64, device ms)."""
        h, w = buffers["color_a"].shape[:2]
        arrs = {k: np.ascontiguousarray(buffers[k], dtype=np.float64).reshape(h, w, c) for k, c in Renderer.NFOR_FIELDS}
        bp = _lib.NforBuffers()
        bp.width, bp.height = w, h
        for k, a in arrs.items():
            setattr(bp, k, a.ctypes.data)
        out = np.zeros((h, w, 3))
        ms = ctypes.c_double()
        ns = Renderer._nfor_settings(nfor)
        _lib.check(ctx._lib.luzgpu_nfor(ctx.handle, ctypes.byref(bp), ctypes.byref(ns) if ns is not None else None,
                                        out.ctypes.data, ctypes.byref(ms)))
        return out, ms.value

    @staticmethod
    def postprocess(ctx, image: np.ndarray, post) -> np.ndarray:
        """applyPostProcessing on the device for an (H, W, 3) scene-linear image (FP64);
        ``post``: a Scene (its file's settings) or {"exposure", "contrast", "bloom",
        "view_transform" ("standard" | "agx" | "aces" | "raw")}."""
        ps = post.post_settings() if isinstance(post, Scene) else _post_settings(post)
        img = np.ascontiguousarray(image, dtype=np.float64)
        h, w = img.shape[:2]
        out = np.zeros_like(img)
        _lib.check(ctx._lib.luzgpu_postprocess(ctx.handle, ctypes.byref(ps), w, h, img.ctypes.data, out.ctypes.data, 0))
        return out

    @staticmethod
    def sample_counts(ctx) -> np.ndarray:
        """Per-pixel sample counts (H, W) of the context's last render / render_denoised."""
        if ctx._last_resolution is None:
            raise RuntimeError("sample_counts: no render on this context yet")
        w, h = ctx._last_resolution
        out = np.zeros((h, w), dtype=np.uint32)
        _lib.check(ctx._lib.luzgpu_sample_counts(ctx.handle, out.ctypes.data, w * h))
        return out

    @staticmethod
    def primary_hits(scene: Scene, seed: int, resolution=None):
        s = _settings(seed, None, None, 0, None, "megakernel", resolution=resolution)
        w, h = resolution if resolution is not None else (scene.width, scene.height)
        ids = np.zeros((h, w, 2), dtype=np.int64)
        t = np.zeros((h, w), dtype=np.float32)
        _lib.check(scene.ctx._lib.luzgpu_primary_hits(scene.ctx.handle, scene._h, ctypes.byref(s), ids.ctypes.data,
                                                      t.ctypes.data))
        return ids, t
