"""Adaptive per-pixel stopping on the device (GPU).

Fixtures (tests/golden/make_golden.py): the reference's per-pixel loop with adaptive
sampling on at seed 1 (each pixel's sample count and the image), without and with
denoising. The device renders in rounds over the still-active pixels and checks the
stopping rule once per round, at exactly the sample counts where the reference's
per-sample check can stop a pixel.

The device's per-sample colours are fp32 where the reference's are FP64, so a pixel whose
confidence interval lands within rounding of the threshold may stop one check earlier or
later: the bar is that at least 99% of pixels take exactly the reference's sample count
and that the pixels which do match its image within rel-MSE 1e-4. Black pixels, constant
pixels and the schedule (min + k * interval, or the total) are exact."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_mse

pytestmark = pytest.mark.gpu

CASES = ["adaptive_primitives_64x48_64spp", "adaptive_cornell_64x48_96spp", "adaptive_cornell_dark_40x30_288spp"]


def _fixture(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def _rule(g):
    return {"min_samples": int(g["min_samples"]), "check_interval": int(g["check_interval"]),
            "threshold": float(g["threshold"])}


def _compare(name, img, counts, g, st=None):
    ref_c = g["counts"].astype(np.uint32)
    same = counts == ref_c
    frac = same.mean()
    e = rel_mse(img[same], g["image"][same])
    print(f"{name}: {frac * 100:.2f}% of pixels take the reference's sample count "
          f"(mean {counts.mean():.2f} vs {ref_c.mean():.2f}), rel-MSE on those {e:.3e}")
    assert frac >= 0.99
    assert e < 1e-4
    spp, mn, iv = int(g["spp"]), int(g["min_samples"]), int(g["check_interval"])
    c = counts.astype(np.int64)
    assert ((c == spp) | ((c >= mn) & ((c - mn) % iv == 0))).all()
    if st is not None:
        assert st.rendered_samples == int(counts.sum())


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
@pytest.mark.parametrize("name", CASES)
def test_adaptive_render_matches_reference(ctx, request, name, integrator):
    from luz_amd import Renderer, SceneFile
    g = _fixture(name)
    wl = request.getfixturevalue(str(g["scene"]))
    scene = SceneFile.read(ctx, wl.scene_path)
    h, w = g["counts"].shape
    img, st = Renderer.render(scene, seed=int(g["seed"]), samples=int(g["spp"]), resolution=(w, h),
                              integrator=integrator, adaptive=_rule(g))
    counts = Renderer.sample_counts(ctx)
    _compare(f"{name} [{integrator}]", img, counts, g, st)
    if name.endswith("dark_40x30_288spp"):
        black = (g["image"] == 0.0).all(axis=2)
        assert (counts[black] == 256).all()


def test_constant_background_converges_at_minimum(ctx, tmp_path):
    """Tests.cpp:4959-4990 on the device: every pixel stops at 2 samples, colour exact."""
    from luz_amd import Renderer, SceneFile
    path = tmp_path / "bg.luz"
    path.write_text("[settings]\nresolution=2,2\nsamples=4\nmaxlightbounces=1\nadaptive=1\nadaptiveminsamples=2\n"
                    "adaptivecheckinterval=1\nadaptivethreshold=0.5\nbackground=(0.1,0.2,0.3)\nview_transform=raw\n"
                    "bloom=0\ndenoise=0\nsky=none\n\n[scene]\ncamera main {\nposition=(0,0,1)\ndirection=(0,0,-1)\n"
                    "up=(0,1,0)\nfocal_length_mm=35\nsensor_width_mm=25\nsensor_height_mm=25\npinhole=1\n"
                    "focus_distance=10\n}\n")
    scene = SceneFile.read(ctx, str(path))
    img, st = Renderer.render(scene, seed=42)
    counts = Renderer.sample_counts(ctx)
    assert (counts == 2).all() and st.rendered_samples == 8
    assert np.abs(img - np.array([0.1, 0.2, 0.3], dtype=np.float32)).max() < 1e-7


def test_adaptive_off_and_sharding(ctx, cornell):
    """adaptive=False renders every sample; adaptive on refuses a partial sample range."""
    from luz_amd import Renderer, SceneFile, _lib
    scene = SceneFile.read(ctx, cornell.scene_path)
    full, st = Renderer.render(scene, seed=1, samples=32, resolution=(32, 24), adaptive=False)
    assert (Renderer.sample_counts(ctx) == 32).all() and st.rendered_samples == 32 * 32 * 24
    with pytest.raises(_lib.LuzGpuError):
        Renderer.render(scene, seed=1, samples=32, sample_begin=0, sample_count=16, resolution=(32, 24),
                        adaptive={"min_samples": 8})
    # the rule cannot stop a pixel when min samples >= samples: identical to a plain render
    same, st2 = Renderer.render(scene, seed=1, samples=32, resolution=(32, 24), adaptive={"min_samples": 32})
    assert np.array_equal(same, full) and st2.rendered_samples == st.rendered_samples


def test_adaptive_in_batch_matches_single(ctx, cornell, primitives):
    """A batch mixing an adaptive scene with a plain one: each image and the adaptive
    counts equal the single-scene renders."""
    from luz_amd import Renderer, SceneFile
    g = _fixture("adaptive_primitives_64x48_64spp")
    a = SceneFile.read(ctx, primitives.scene_path)
    b = SceneFile.read(ctx, cornell.scene_path)
    rule = _rule(g)
    single_a, sa = Renderer.render(a, seed=1, samples=64, resolution=(64, 48), adaptive=rule)
    counts_a = Renderer.sample_counts(ctx)
    single_b, sb = Renderer.render(b, seed=3, samples=16, resolution=(40, 40), adaptive=False)
    imgs, st = Renderer.render_batch([a, b], seeds=[1, 3], samples=[64, 16], resolutions=[(64, 48), (40, 40)],
                                     adaptive=[rule, False])
    assert rel_mse(imgs[0], single_a) < 1e-10
    assert rel_mse(imgs[1], single_b) < 1e-10
    assert st.rendered_samples == int(counts_a.sum()) + 16 * 40 * 40


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
def test_adaptive_denoised_render_matches_reference(ctx, cornell, integrator):
    from luz_amd import Renderer, SceneFile
    g = np.load(os.path.join(GOLDEN, "nfor_adaptive_cornell_48x48_48spp.npz"))
    scene = SceneFile.read(ctx, cornell.scene_path)
    h, w = g["image"].shape[:2]
    r = [int(v) for v in g["radii"]]
    nfor = {"feature_prefilter_radius": r[0], "regression_radius": r[1], "patch_radius": r[2], "min_variance": 1e-5}
    img, den, st, buf = Renderer.render_denoised(scene, seed=1, samples=int(g["spp"]), resolution=(w, h), nfor=nfor,
                                                 buffers=True, adaptive=_rule(g), integrator=integrator)
    counts = Renderer.sample_counts(ctx)
    _compare(f"adaptive denoise [{integrator}]", img, counts, g, st)
    same = counts == g["counts"]
    for k in ("color_a", "color_b", "features_a", "features_b"):
        e = rel_mse(buf[k][same], g[k][same])
        print(f"  {k} rel-MSE {e:.3e}")
        assert e < 1e-4
    e = rel_mse(den, g["denoised"])
    print(f"  denoised rel-MSE {e:.3e}")
    assert e < 1e-3
