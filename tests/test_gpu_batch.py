"""Multi-scene launches (luzgpu_render_batch) on the device (GPU).

A batch merges the scenes' device arrays and renders slices of every scene per pass; each
scene must come out as its own single-scene render does (same random numbers, same
per-pixel sample order) and so within the golden bars."""

import numpy as np
import pytest

from conftest import rel_mse
from test_gpu_parity import REL_MSE_BAR, _golden

pytestmark = pytest.mark.gpu


def test_batch_matches_single_scene_renders_and_goldens(ctx, cornell, primitives, volume, materials, meshlight):
    from luz_amd import Renderer, SceneFile
    wls = [cornell, primitives, volume, materials, meshlight]
    goldens = ["cornell_300_50spp_seed1.npz"] + [f"{w.name}_{w.width}x{w.height}_{w.spp}spp_seed{w.seed}.npz"
                                                 for w in wls[1:]]
    scenes = [SceneFile.read(ctx, w.scene_path) for w in wls]
    images, stats = Renderer.render_batch(scenes, seeds=[w.seed for w in wls])
    assert stats.rendered_samples == sum(w.width * w.height * w.spp for w in wls)
    for w, g, sc, img in zip(wls, goldens, scenes, images):
        single, _ = Renderer.render(sc, seed=w.seed)
        e_single = rel_mse(img, single)
        e_golden = rel_mse(img, _golden(g)["image"])
        print(f"batch {w.name}: vs single-scene render {e_single:.2e}, vs reference {e_golden:.3e}")
        assert e_single < 1e-10
        assert e_golden <= REL_MSE_BAR
    # the merged arrays are cached: a second call with other seeds reuses them
    again, _ = Renderer.render_batch(scenes, seeds=[w.seed + 1 for w in wls], samples=[4] * len(wls))
    for w, sc, img in zip(wls, scenes, again):
        single, _ = Renderer.render(sc, seed=w.seed + 1, samples=4)
        assert rel_mse(img, single) < 1e-10


def test_batch_of_many_uneven_scenes(ctx, cornell, primitives):
    """More scenes than fit one pass's path budget at once, mixed resolutions and spp."""
    from luz_amd import Renderer, SceneFile
    a = SceneFile.read(ctx, cornell.scene_path)
    b = SceneFile.read(ctx, primitives.scene_path)
    scenes = [a, b] * 6
    seeds = list(range(1, 13))
    spp = [40, 8, 24, 16, 64, 4, 12, 32, 20, 8, 48, 16]
    images, stats = Renderer.render_batch(scenes, seeds=seeds, samples=spp)
    for sc, sd, n, img in zip(scenes, seeds, spp, images):
        single, _ = Renderer.render(sc, seed=sd, samples=n)
        assert rel_mse(img, single) < 1e-10


def test_plain_kernel_batch_matches_single_scene_renders(ctx, cornell, primitives, volume, materials, meshlight):
    """The batched plain kernel (every scene's paths in one launch, one scene per wave)
    gives each scene its own single-scene plain-kernel render, with uneven sizes and spp."""
    from luz_amd import Renderer, SceneFile
    wls = [cornell, primitives, volume, materials, meshlight]
    scenes = [SceneFile.read(ctx, w.scene_path) for w in wls]
    res = [(37, 29), (64, 48), (16, 16), (50, 30), (33, 65)]
    spp = [5, 3, 8, 1, 2]
    seeds = [w.seed + 3 for w in wls]
    images, stats = Renderer.render_batch(scenes, seeds=seeds, samples=spp, resolutions=res, integrator="megakernel")
    assert stats.rendered_samples == sum(r[0] * r[1] * n for r, n in zip(res, spp))
    for w, sc, sd, r, n, img in zip(wls, scenes, seeds, res, spp, images):
        single, _ = Renderer.render(sc, seed=sd, samples=n, resolution=r, integrator="megakernel")
        e = rel_mse(img, single)
        print(f"plain-kernel batch {w.name}: vs single-scene render {e:.2e}")
        assert e < 1e-10
