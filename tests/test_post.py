"""Display post-processing: fixtures and scene-file keys (CPU).

Fixture (tests/golden/make_golden.py, post_fixture): the reference's own post-processing
(oracle/ref_harness.cpp, post command: Image::applyExposure, extractBloom + Gaussian::blur,
the view transforms, applyContrast, gammaCorrect, suppressIsolatedFireflies, in
applyPostProcessing's order, Threads.cpp:55-99) of a 64 x 64 crop of the Cornell reference
render with fireflies, a saturated patch and non-finite input planted, under six settings.
There is no restatement in the oracle: the reference itself is the checker, and the device
is compared with it in tests/test_gpu_post.py. Here: what the mathematics fixes."""

import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _fixture():
    return np.load(os.path.join(GOLDEN, "post_cornell_64x64.npz"))


def test_exposure_alone_is_a_scale():
    """Raw view, no bloom: out = max(x, 0) * 2^exposure, non-finite -> 0 (Image.cpp:467-484)."""
    g = _fixture()
    img = g["image"]
    clean = np.where(np.isfinite(img) & (img > 0), img, 0.0)
    for k in range(len(g["view"])):
        if str(g["view"][k]) == "raw" and int(g["bloom"][k]) == 0:
            assert np.array_equal(g["outputs"][k], clean * 2.0 ** float(g["exposure"][k]))


def test_display_outputs_are_encoded():
    g = _fixture()
    for k in range(len(g["view"])):
        o = g["outputs"][k]
        assert np.isfinite(o).all()
        if str(g["view"][k]) != "raw":
            assert o.min() >= 0.0 and o.max() <= 1.0


def test_bloom_spreads_light_on_a_raw_view():
    """Bloom adds a blurred bright-pass: on the raw view it only ever adds light."""
    g = _fixture()
    img = g["image"]
    clean = np.where(np.isfinite(img) & (img > 0), img, 0.0)
    for k in range(len(g["view"])):
        if str(g["view"][k]) == "raw" and int(g["bloom"][k]) == 1:
            base = clean * 2.0 ** float(g["exposure"][k])
            added = g["outputs"][k] - base
            assert added.min() >= 0.0 and added.max() > 0.1


def _scene(tmp_path, lines):
    from luz_amd import workloads
    text = workloads.cornell_scene_text(8, 8, 4, 2)
    text = text.replace("view_transform=raw\n", "").replace("bloom=0\n", "")
    text = text.replace("[settings]\n", "[settings]\n" + "".join(f"{ln}\n" for ln in lines))
    path = tmp_path / "post.luz"
    path.write_text(text)
    return str(path)


def test_scene_file_post_settings(tmp_path):
    from luz_amd import SceneFile, _lib
    sc = SceneFile.read(None, _scene(tmp_path, []))
    ps = sc.post_settings()  # Scene.cpp:99-139 defaults
    assert (ps.view_transform, ps.bloom, ps.exposure, ps.contrast) == (_lib.VIEW_AGX, 1, 0.0, 1.0)
    sc.release()
    sc = SceneFile.read(None, _scene(tmp_path, ["view_transform=ACES", "bloom=0", "exposure=-1.25", "contrast=1.4"]))
    ps = sc.post_settings()
    assert (ps.view_transform, ps.bloom, ps.exposure, ps.contrast) == (_lib.VIEW_ACES, 0, -1.25, 1.4)
    sc.release()
    # photographic exposure (Scene.cpp:270-286): log2(t * ISO / 100 / N^2)
    sc = SceneFile.read(None, _scene(tmp_path, ["photographic_exposure=4,0.0166666666667,200"]))
    assert sc.post_settings().exposure == pytest.approx(math.log2(0.0166666666667 * 2.0 / 16.0), rel=1e-15)
    sc.release()


@pytest.mark.parametrize("line", ["view_transform=filmic", "bloom=2", "contrast=-1", "exposure=inf", "exposure=5000",
                                  "photographic_exposure=4,0,100", "photographic_exposure=4,1"])
def test_scene_file_rejects_invalid_post_settings(tmp_path, line):
    from luz_amd import SceneFile
    with pytest.raises(Exception):
        SceneFile.read(None, _scene(tmp_path, [line]))
