"""Edge cases of the render entry points on the device (GPU), against the oracle (the FP64
restatement pinned to the reference's renders): single-pixel and single-row images, no
bounce (camera rays only: emitters and sky), one sample, odd sizes that leave partial waves,
both integrators; and the argument errors the entry points must raise instead of running."""

import numpy as np
import pytest

import oracle_lib
from conftest import rel_mse

pytestmark = pytest.mark.gpu

CASES = [((1, 1), 4, 8), ((7, 1), 3, 8), ((1, 9), 2, 8), ((33, 17), 1, 8), ((64, 48), 8, 0), ((31, 29), 5, 1)]


@pytest.mark.parametrize("integrator", ["megakernel", "wavefront"])
@pytest.mark.parametrize("res,spp,depth", CASES)
def test_small_shapes_match_oracle(ctx, primitives, res, spp, depth, integrator):
    from luz_amd import Renderer, SceneFile
    scene = SceneFile.read(ctx, primitives.scene_path)
    img, st = Renderer.render(scene, seed=5, samples=spp, max_bounces=depth, resolution=res, integrator=integrator)
    ref, _ = oracle_lib.render(primitives.scene_path, 5, spp, res[0], res[1], max_bounces=depth)
    assert img.shape == (res[1], res[0], 3)
    assert st.rendered_samples == res[0] * res[1] * spp
    assert np.isfinite(img).all()
    e = rel_mse(img, ref) if np.abs(ref).sum() > 0 else float(np.abs(img).max())
    print(f"{res} x {spp} spp, depth {depth} [{integrator}]: rel-MSE {e:.3e}")
    assert e <= 1e-4


def test_depth_zero_casts_one_closest_ray_per_path(ctx, primitives):
    """maxlightbounces=0: the bounce loop runs once (ColorHelper.cpp:1165), so one
    closest-hit query per path; its light connections still cast shadow rays."""
    from luz_amd import Renderer, SceneFile
    scene = SceneFile.read(ctx, primitives.scene_path)
    for integrator in ("megakernel", "wavefront"):
        _, st = Renderer.render(scene, seed=1, samples=2, max_bounces=0, resolution=(40, 30), integrator=integrator)
        assert st.closest_hit_rays == 40 * 30 * 2
        assert 0 < st.shadow_rays <= 2 * 40 * 30 * 2


def test_argument_errors(ctx, primitives):
    from luz_amd import Renderer, SceneFile, _lib
    scene = SceneFile.read(ctx, primitives.scene_path)
    with pytest.raises(_lib.LuzGpuError):
        Renderer.render(scene, seed=1, samples=8, sample_begin=6, sample_count=4)  # beyond the total
    with pytest.raises(_lib.LuzGpuError):
        Renderer.render(scene, seed=1, samples=8, sample_begin=0, sample_count=0)  # empty range
    with pytest.raises(_lib.LuzGpuError):
        Renderer.render_denoised(scene, seed=1, samples=4, resolution=(16, 16), nfor={"regression_radius": 0})
    with pytest.raises(_lib.LuzGpuError):
        Renderer.postprocess(ctx, np.zeros((4, 4, 3)), {"contrast": -1.0})
    # the context stays usable after every refused call
    img, _ = Renderer.render(scene, seed=1, samples=1, resolution=(8, 8))
    assert np.isfinite(img).all()


def test_auto_integrator_picks_by_size(ctx, primitives, cornell):
    """LUZGPU_INTEGRATOR_AUTO runs the plain kernel up to 4M paths: the images equal the
    plain kernel's, single and batched."""
    from luz_amd import Renderer, SceneFile
    a = SceneFile.read(ctx, primitives.scene_path)
    b = SceneFile.read(ctx, cornell.scene_path)
    auto, _ = Renderer.render(a, seed=2, samples=4, resolution=(64, 48), integrator="auto")
    plain, _ = Renderer.render(a, seed=2, samples=4, resolution=(64, 48), integrator="megakernel")
    assert np.array_equal(auto, plain)
    imgs, _ = Renderer.render_batch([a, b], seeds=[2, 3], samples=[4, 2], resolutions=[(64, 48), (40, 40)],
                                    integrator="auto")
    ref, _ = Renderer.render_batch([a, b], seeds=[2, 3], samples=[4, 2], resolutions=[(64, 48), (40, 40)],
                                   integrator="megakernel")
    assert all(np.array_equal(x, y) for x, y in zip(imgs, ref))


def test_wrapping_sample_range_and_size_are_refused(ctx, primitives):
    """A sample_begin near 2^32 must not wrap round the 32-bit range check; an image whose
    pixel count overflows 32 bits must be refused, not sized modulo 2^32."""
    from luz_amd import Renderer, SceneFile, _lib
    scene = SceneFile.read(ctx, primitives.scene_path)
    with pytest.raises(_lib.LuzGpuError):
        Renderer.render(scene, seed=1, samples=8, sample_begin=0xFFFFFFF0, sample_count=32)
    with pytest.raises(_lib.LuzGpuError):
        Renderer.render(scene, seed=1, samples=1, resolution=(65536, 65537))
    img, _ = Renderer.render(scene, seed=1, samples=1, resolution=(8, 8))
    assert np.isfinite(img).all()


def test_bad_output_buffers_raise_before_the_call(ctx, primitives):
    """Caller tensors that are not float32, contiguous, H*W*3 on the context's device raise
    ValueError in Python; none reaches a kernel."""
    import torch
    from luz_amd import Renderer, SceneFile
    scene = SceneFile.read(ctx, primitives.scene_path)
    n = 8 * 8 * 3
    bad = [torch.zeros(n, dtype=torch.float32),                               # host memory
           torch.zeros(n, dtype=torch.float64, device="cuda"),                # wrong dtype
           torch.zeros(n * 2, dtype=torch.float32, device="cuda")[::2],       # not contiguous
           torch.zeros(n - 3, dtype=torch.float32, device="cuda")]            # too small
    for t in bad:
        with pytest.raises(ValueError):
            Renderer.render(scene, seed=1, samples=1, resolution=(8, 8), out=t)
        with pytest.raises(ValueError):
            Renderer.render_batch([scene], seeds=[1], samples=[1], resolutions=[(8, 8)], outs=[t])
    with pytest.raises(ValueError):
        scene.adjoint(np.zeros(n - 3, dtype=np.float32), seed=1, samples=1, resolution=(8, 8), adaptive=False)
    with pytest.raises(ValueError):
        scene.adjoint(torch.zeros(n - 3, device="cuda"), seed=1, samples=1, resolution=(8, 8), adaptive=False)
    good = torch.zeros(n, dtype=torch.float32, device="cuda")
    Renderer.render(scene, seed=1, samples=1, resolution=(8, 8), out=good)
    torch.cuda.synchronize()
    assert torch.isfinite(good).all()


def test_adjoint_refuses_adaptive_renders(ctx, workload_dir):
    """With adaptive sampling in effect (the scene default is on, 128 minimum samples) the
    forward image averages each pixel over its own sample count; the adjoint replays every
    sample, so it must refuse rather than return gradients of another image. With
    adaptive=False both agree, and the autograd wrapper passes that itself."""
    import os
    import torch
    from luz_amd import Renderer, SceneFile, _lib, diff, workloads
    d = os.path.join(workload_dir, "diff_adaptive")
    os.makedirs(d, exist_ok=True)
    kinds, albedo, roughness, emission = workloads.diff_parameters()
    text = workloads.diff_scene_text(kinds, albedo, roughness, emission, 16, 16, 200, 3)
    path = os.path.join(d, "adaptive_on.luz")
    with open(path, "w") as fh:
        fh.write(text.replace("adaptive=0", "adaptive=1"))
    scene = SceneFile.read(ctx, path)
    g = np.ones((16, 16, 3), dtype=np.float32)
    with pytest.raises(_lib.LuzGpuError):
        scene.adjoint(g, seed=1, samples=200)
    da, dr, de = scene.adjoint(g, seed=1, samples=200, adaptive=False)
    assert np.isfinite(da).all() and np.isfinite(dr).all() and np.isfinite(de).all()
    n = scene.info["materials"]
    alb = torch.full((n, 3), 0.5, dtype=torch.float64, requires_grad=True)
    rough = torch.full((n,), 0.3, dtype=torch.float64, requires_grad=True)
    emi = torch.tensor(emission, dtype=torch.float64, requires_grad=True)
    img = diff.render(scene, alb, rough, emi, seed=1, samples=200)
    img.sum().backward()
    assert torch.isfinite(alb.grad).all()
    ref, _ = Renderer.render(scene, seed=1, samples=200, adaptive=False)
    assert np.array_equal(img.detach().cpu().numpy(), ref)


def test_scene_outlives_its_context():
    """luzgpu_destroy frees the device arrays of scenes still loaded; the handles stay valid
    for host-side queries and release (a path that once crashed the host process)."""
    from luz_amd import Context, SceneFile, workloads
    import tempfile
    d = tempfile.mkdtemp(prefix="luz_outlive_")
    w = workloads.write_cornell(d)
    c = Context(0)
    scene = SceneFile.read(c, w.scene_path)
    topo = scene.tlas_topology()
    c.close()
    assert scene.tlas_topology() == topo
    assert scene.post_settings().view_transform == 3
    scene.release()
    scene.release()


def test_render_of_a_scene_whose_context_is_gone_is_refused(primitives):
    """After luzgpu_destroy the scene's info and topology still answer, a render of it is
    refused (LUZGPU_E_INVALID) rather than touching freed device memory, and a new context
    on the device works."""
    import ctypes
    from luz_amd import Context, SceneFile, _lib
    ctx = Context(0)
    scene = SceneFile.read(ctx, primitives.scene_path)
    topology = scene.tlas_topology()
    ctx.close()
    info = _lib.SceneInfo()
    _lib.check(scene._lib.luzgpu_scene_get_info(scene._h, ctypes.byref(info)))
    assert info.hittables == scene.info["hittables"]
    assert scene.tlas_topology() == topology
    out = np.zeros(16 * 16 * 3, dtype=np.float32)
    s = _lib.Settings()
    s.seed, s.samples, s.max_bounces, s.sample_count, s.width, s.height = 1, 1, -1, -1, 16, 16
    assert scene._lib.luzgpu_render(None, scene._h, ctypes.byref(s), out.ctypes.data, None) == _lib.E_INVALID
    scene.release()
    # a new context on the same device works after all this
    with Context(0) as ctx2:
        sc2 = SceneFile.read(ctx2, primitives.scene_path)
        from luz_amd import Renderer
        img, _ = Renderer.render(sc2, seed=1, samples=1, resolution=(16, 16))
        assert np.isfinite(img).all()
        sc2.release()
