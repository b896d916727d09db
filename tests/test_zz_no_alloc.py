"""luzgpu_render_noalloc (the render entry point that allocates nothing: its scratch is one
fixed device array per GPU) against luzgpu_render on the plain kernel, both called through
the C ABI: the images and the ray counts must be identical, on two calls in a row, the
second with a smaller image and fewer samples (the fixed array keeps the first call's
contents, so stale sums or counters would show), then the limits it refuses."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _settings(seed, samples, res, on_device, integrator):
    from luz_amd import _lib
    s = _lib.Settings()
    s.seed = seed
    s.samples = samples
    s.max_bounces = -1
    s.sample_begin = 0
    s.sample_count = -1
    s.integrator = integrator
    s.output_on_device = 1 if on_device else 0
    s.width, s.height = res
    s.adaptive = _lib.ADAPTIVE_OFF
    return s


def _call(fn, ctx, scene, s, out_ptr):
    from luz_amd import _lib
    st = _lib.Stats()
    _lib.check(fn(ctx.handle, scene._h, ctypes.byref(s), ctypes.c_void_p(out_ptr), ctypes.byref(st)))
    return st


@pytest.mark.parametrize("scene_name", ["primitives", "cornell"])
def test_noalloc_render_equals_plain_kernel_render(ctx, request, scene_name):
    import torch
    from luz_amd import SceneFile, _lib
    wl = request.getfixturevalue(scene_name)
    lib = _lib.load()
    scene = SceneFile.read(ctx, wl.scene_path)
    # two calls in a row, the second smaller; device outputs, then a host output
    for (res, spp, seed, on_device) in (((160, 120), 6, 3, True), ((57, 31), 3, 4, True), ((40, 30), 2, 7, False)):
        n = res[0] * res[1] * 3
        print(f"{scene_name}: {res[0]}x{res[1]} x {spp} spp, seed {seed}, output on device {on_device}", flush=True)
        if on_device:
            a = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
            b = torch.full((n,), -2.0, dtype=torch.float32, device="cuda")
            pa, pb = a.data_ptr(), b.data_ptr()
        else:
            a = np.full(n, -1.0, dtype=np.float32)
            b = np.full(n, -2.0, dtype=np.float32)
            pa, pb = a.ctypes.data, b.ctypes.data
        sa = _call(lib.luzgpu_render, ctx, scene, _settings(seed, spp, res, on_device, _lib.INTEGRATOR_MEGAKERNEL), pa)
        sb = _call(lib.luzgpu_render_noalloc, ctx, scene, _settings(seed, spp, res, on_device, _lib.INTEGRATOR_MEGAKERNEL), pb)
        if on_device:
            torch.cuda.synchronize()
            a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.isfinite(b).all()
        assert np.array_equal(a, b), float(np.abs(a - b).max())
        assert sb.rendered_samples == sa.rendered_samples == res[0] * res[1] * spp
        assert (sb.closest_hit_rays, sb.shadow_rays) == (sa.closest_hit_rays, sa.shadow_rays)
    scene.release()


def test_noalloc_refuses_what_needs_allocation(ctx, primitives):
    import torch
    from luz_amd import SceneFile, _lib
    lib = _lib.load()
    scene = SceneFile.read(ctx, primitives.scene_path)
    out = torch.empty(1025 * 1024 * 3, dtype=torch.float32, device="cuda")
    big = _settings(1, 1, (1025, 1024), True, _lib.INTEGRATOR_MEGAKERNEL)
    with pytest.raises(_lib.LuzGpuError) as e:
        _call(lib.luzgpu_render_noalloc, ctx, scene, big, out.data_ptr())
    assert e.value.code == _lib.E_INVALID and "1048576" in str(e.value)
    wave = _settings(1, 1, (16, 16), True, _lib.INTEGRATOR_WAVEFRONT)
    with pytest.raises(_lib.LuzGpuError) as e:
        _call(lib.luzgpu_render_noalloc, ctx, scene, wave, out.data_ptr())
    assert e.value.code == _lib.E_UNSUPPORTED
    adaptive = _settings(1, 8, (16, 16), True, _lib.INTEGRATOR_MEGAKERNEL)
    adaptive.adaptive = _lib.ADAPTIVE_ON
    adaptive.adaptive_min_samples = 2
    adaptive.adaptive_check_interval = 2
    with pytest.raises(_lib.LuzGpuError) as e:
        _call(lib.luzgpu_render_noalloc, ctx, scene, adaptive, out.data_ptr())
    assert e.value.code == _lib.E_UNSUPPORTED
    # the largest image it accepts, 1024 x 1024, and the context still works after the refusals
    ok = _settings(1, 1, (1024, 1024), True, _lib.INTEGRATOR_MEGAKERNEL)
    st = _call(lib.luzgpu_render_noalloc, ctx, scene, ok, out.data_ptr())
    torch.cuda.synchronize()
    assert st.rendered_samples == 1024 * 1024
    assert torch.isfinite(out[:1024 * 1024 * 3]).all().item()
    scene.release()
