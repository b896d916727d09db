"""Display post-processing on the device (GPU).

Against the reference's own post-processing of the same FP64 input (fixture
post_cornell_64x64.npz, six settings: defaults, ACES, the contrast branch, bloom on a raw
view, AgX with contrast, exposure alone), and the render path that applies it to the
device's FP64 mean image (luzgpu_settings.postprocess)."""

import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _fixture():
    return np.load(os.path.join(GOLDEN, "post_cornell_64x64.npz"))


@pytest.mark.parametrize("case", range(6))
def test_postprocess_matches_reference(ctx, case):
    from luz_amd import Renderer
    g = _fixture()
    post = {"exposure": float(g["exposure"][case]), "bloom": bool(g["bloom"][case]), "view_transform": str(g["view"][case]),
            "contrast": float(g["contrast"][case])}
    out = Renderer.postprocess(ctx, g["image"], post)
    ref = g["outputs"][case]
    d = np.abs(out - ref)
    print(f"post {post}: max |diff| {d.max():.3e}, pixels off by > 1e-12: {(d.max(axis=2) > 1e-12).sum()}")
    assert d.max() <= 1e-10


def test_render_with_postprocess(ctx, cornell, tmp_path):
    """render(postprocess=True) = the scene's post-processing of the raw render (the raw
    image here is fp32, the device's own input FP64: the difference is rounding)."""
    from luz_amd import Renderer, SceneFile, workloads
    text = workloads.cornell_scene_text(300, 300, 50, 8).replace("view_transform=raw\n", "view_transform=agx\n")
    text = text.replace("bloom=0\n", "bloom=1\n")
    path = tmp_path / "cornell_post.luz"
    path.write_text(text)
    scene = SceneFile.read(ctx, str(path))
    raw, _ = Renderer.render(scene, seed=1, samples=8, resolution=(96, 96))
    shown, st = Renderer.render(scene, seed=1, samples=8, resolution=(96, 96), postprocess=True)
    expect = Renderer.postprocess(ctx, raw.astype(np.float64), scene)
    d = np.abs(shown - expect)
    print(f"render + post vs post of the fp32 raw image: max |diff| {d.max():.3e}, "
          f"> 1e-4 at {(d.max(axis=2) > 1e-4).sum()} pixels")
    assert shown.min() >= 0.0 and shown.max() <= 1.0
    assert (d.max(axis=2) > 1e-4).mean() < 1e-3
    den_raw, den, _, _ = Renderer.render_denoised(scene, seed=1, samples=8, resolution=(96, 96), postprocess=True)
    assert den.min() >= 0.0 and den.max() <= 1.0 and den_raw.max() <= 1.0
