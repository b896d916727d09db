"""NFOR denoiser and the denoising render mode on the device (GPU).

Fixtures (tests/golden/make_golden.py, nfor_fixtures): the reference's own denoise
pipeline at seed 1 (oracle/ref_harness.cpp, denoise command): its NFOR input buffers
(even / odd half-buffer means on sampler streams 1 / 2, colour variance, 11 primary
features and their variance), the raw mean image and Denoise::applyNFOR's output, for the
Cornell box (default radii 1 / 2 / 1, and regression radius 12 with patch radius 2, a window
too wide for the LDS-staged regression) and the primitives scene (regression radius 9).

* Device NFOR on the reference's buffers reproduces its output to FP64 rounding (the
  summation order differs).
* The device's denoising render produces the reference's buffers within the image bar and
  its denoised image close to the reference's (NFOR amplifies fp32 path differences a
  little through its data-dependent weights)."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_mse

pytestmark = pytest.mark.gpu

CASES = ["nfor_cornell_48x48_8spp", "nfor_primitives_64x48_8spp_r9", "nfor_cornell_48x48_8spp_r12p2"]


def _fixture(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def _settings(g):
    r = [int(v) for v in g["radii"]]
    return {"feature_prefilter_radius": r[0], "regression_radius": r[1], "patch_radius": r[2], "min_variance": 1e-5}


@pytest.mark.parametrize("name", CASES)
def test_nfor_on_reference_buffers(ctx, name):
    from luz_amd import Renderer
    g = _fixture(name)
    out, ms = Renderer.nfor(ctx, {k: g[k] for k, _ in Renderer.NFOR_FIELDS}, _settings(g))
    ref = g["denoised"]
    e = rel_mse(out, ref)
    print(f"{name}: device NFOR vs reference rel-MSE {e:.3e}, max |diff| {np.abs(out - ref).max():.3e}, {ms:.2f} ms")
    assert np.isfinite(out).all()
    assert e < 1e-12


@pytest.mark.parametrize("integrator", ["wavefront", "megakernel"])
@pytest.mark.parametrize("name", CASES)
def test_denoised_render_matches_reference(ctx, request, name, integrator):
    from luz_amd import Renderer, SceneFile
    g = _fixture(name)
    wl = request.getfixturevalue(str(g["scene"]))
    scene = SceneFile.read(ctx, wl.scene_path)
    h, w = g["image"].shape[:2]
    img, den, st, buf = Renderer.render_denoised(scene, seed=int(g["seed"]), samples=int(g["spp"]), resolution=(w, h),
                                                 nfor=_settings(g), buffers=True, integrator=integrator)
    errors = {k: rel_mse(buf[k], g[k]) for k in ("image", "color_a", "color_b", "features_a", "features_b")}
    errors["denoised"] = rel_mse(den, g["denoised"])
    print(f"{name} [{integrator}]: " + ", ".join(f"{k} {v:.2e}" for k, v in errors.items()) + f"; render {st.render_ms:.1f} ms, "
          f"NFOR {st.denoise_ms:.2f} ms")
    assert np.array_equal(img, buf["image"].astype(np.float32))
    for k in ("image", "color_a", "color_b", "features_a", "features_b"):
        assert errors[k] <= 1e-4, k
    assert errors["denoised"] <= 1e-3
    # the variances agree in aggregate (they are noisy per pixel at 8 spp)
    for k in ("color_variance", "feature_variance"):
        a, b = buf[k].sum(), g[k].sum()
        assert abs(a - b) <= 1e-2 * abs(b) + 1e-12, (k, a, b)
