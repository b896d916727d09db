"""The volume-and-denoise workload and the uneven batch workload on the device (GPU), at
reduced sizes the reference renders in seconds (tests/golden/make_golden.py,
terrain_fixtures / batch_fixtures).

* Terrain: the 1,002,528-triangle BLAS (the deep traversal stack and its spill path), the
  Henyey-Greenstein haze box and the directional sun; primary hits against the reference's
  brute force, images against the reference's render at 160 x 90, 8 spp, depth 16; NFOR with a
  radius-9 regression window on the reference's own denoise buffers of that scene, and the
  device's denoising render against those buffers.
* Batch: four of the 32 scenes (the largest at 474k triangles and three below it) in one
  luzgpu_render_batch call on both integrators, each image against the reference's render
  of that scene at its own seed."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_mse
from test_gpu_parity import REL_MSE_BAR, _primary_check

pytestmark = pytest.mark.gpu


def _golden(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.fixture(scope="module")
def terrain_scene(ctx, terrain):
    from luz_amd import SceneFile
    print(f"terrain: {terrain.triangles} triangles", flush=True)
    scene = SceneFile.read(ctx, terrain.scene_path)
    yield scene
    scene.release()


def test_terrain_primary_hits(terrain_scene):
    assert _primary_check(terrain_scene, "terrain_primary.npz", resolution=(160, 90)) <= 2


@pytest.mark.parametrize("integrator", ["megakernel", "wavefront"])
def test_terrain_image_parity(terrain_scene, integrator):
    from luz_amd import Renderer
    ref = _golden("terrain_160x90_8spp_seed1.npz")["image"]
    img, st = Renderer.render(terrain_scene, seed=1, samples=8, resolution=(160, 90), integrator=integrator)
    e = rel_mse(img, ref)
    print(f"terrain 160x90x8 [{integrator}]: rel-MSE {e:.3e}, {st.closest_hit_rays} closest / {st.shadow_rays} shadow rays")
    assert np.isfinite(img).all()
    assert e <= REL_MSE_BAR


def _nfor_settings(g):
    r = [int(v) for v in g["radii"]]
    return {"feature_prefilter_radius": r[0], "regression_radius": r[1], "patch_radius": r[2], "min_variance": 1e-5}


def test_terrain_nfor_radius9_on_reference_buffers(ctx):
    from luz_amd import Renderer
    g = _golden("nfor_terrain_96x64_8spp_r9.npz")
    out, ms = Renderer.nfor(ctx, {k: g[k] for k, _ in Renderer.NFOR_FIELDS}, _nfor_settings(g))
    e = rel_mse(out, g["denoised"])
    print(f"terrain NFOR (radius 9) vs reference: rel-MSE {e:.3e}, {ms:.2f} ms")
    assert np.isfinite(out).all()
    assert e < 1e-12


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_terrain_denoised_render_matches_reference(terrain_scene, integrator):
    from luz_amd import Renderer
    g = _golden("nfor_terrain_96x64_8spp_r9.npz")
    h, w = g["image"].shape[:2]
    img, den, st, buf = Renderer.render_denoised(terrain_scene, seed=int(g["seed"]), samples=int(g["spp"]), resolution=(w, h),
                                                 nfor=_nfor_settings(g), buffers=True, integrator=integrator)
    errors = {k: rel_mse(buf[k], g[k]) for k in ("image", "color_a", "color_b", "features_a", "features_b")}
    errors["denoised"] = rel_mse(den, g["denoised"])
    print(f"terrain denoised [{integrator}]: " + ", ".join(f"{k} {v:.2e}" for k, v in errors.items()))
    for k in ("image", "color_a", "color_b", "features_a", "features_b"):
        assert errors[k] <= REL_MSE_BAR, k
    assert errors["denoised"] <= 1e-3


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_batch_scenes_match_reference(ctx, batch, integrator):
    from luz_amd import Renderer, SceneFile
    g = _golden("batch_96x64_8spp.npz")
    picks = [int(i) for i in g["scenes"]]
    wls = [batch[i] for i in picks]
    assert [w.triangles for w in wls] == [int(t) for t in g["triangles"]]
    scenes = [SceneFile.read(ctx, w.scene_path) for w in wls]
    n = len(scenes)
    imgs, st = Renderer.render_batch(scenes, [int(s) for s in g["seeds"]], samples=[8] * n, resolutions=[(96, 64)] * n,
                                     integrator=integrator)
    for i, (img, ref) in enumerate(zip(imgs, g["images"])):
        e = rel_mse(img, ref)
        print(f"batch scene {picks[i]} ({wls[i].triangles} triangles) [{integrator}]: rel-MSE {e:.3e}")
        assert np.isfinite(img).all()
        assert e <= REL_MSE_BAR
    for sc in scenes:
        sc.release()
