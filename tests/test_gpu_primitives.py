"""Analytic primitives, participating media and delta lights on the device (GPU).

Scenes: ``primitives`` (plane, spheres, cubes, sphere / rectangle / cube emitters, constant
background), ``volume`` (Henyey-Greenstein box fog, isotropic sphere smoke, directional sun,
sphere light), ``materials`` (glossy, diffuse-glossy and principled variants) and
``meshlight`` (emissive mesh instances under uniform and non-uniform scales, sampled by
triangle area and re-intersected for their MIS pdf) and ``optics`` (the reference's default
thin-lens camera, rough dielectrics, conductor-Fresnel metals, dispersive glass presets and
interior absorption), ``subsurface`` (Burley exits relocated by a probe ray inside the
shading kernels, the skin profile, the thin two-sided lobe) and ``invisible`` (an invisible
sphere light that lights the room through light sampling alone, an invisible ball before the
camera) and ``ies`` (IES-profiled sphere lights: a downlight emitting its profile's flux,
a tilted and turned wash light) and ``sunlit`` (the atmosphere scene lit by a `solar=`
directional light that is also the atmosphere's sun). Same bars as test_gpu_parity: primary-hit IDs against the reference's
FP64 brute force, images within relative MSE 1e-4 of the reference render at equal spp and
seed (tests/golden/make_golden.py)."""

import numpy as np
import pytest

from conftest import rel_mse
from test_gpu_parity import REL_MSE_BAR, _golden, _primary_check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["primitives", "volume", "materials", "meshlight", "skylinear", "atmosphere", "spectral", "atmosphere_env", "optics",
                                  "subsurface", "invisible", "ies", "sunlit"])
def test_primary_hits(ctx, request, name):
    from luz_amd import SceneFile
    wl = request.getfixturevalue(name)
    scene = SceneFile.read(ctx, wl.scene_path)
    # the atmosphere scene sits 6.36 units from the origin, where an FP32 origin carries
    # 4.8e-7 of rounding; over the ground plane's grazing rays (|d.y| down to ~0.005) that is
    # up to 1e-4 of t, so its distances are held to 2e-3 (its hit ids stay exact)
    # (the optics camera sees the floor plane out to the horizon: the same grazing rays)
    t_rtol = 2e-3 if name in ("atmosphere", "optics", "sunlit") else 2e-5
    assert _primary_check(scene, f"{name}_primary.npz", t_rtol=t_rtol) <= 2


@pytest.mark.parametrize("integrator", ["megakernel", "wavefront"])
@pytest.mark.parametrize("name", ["primitives", "volume", "materials", "meshlight", "skylinear", "atmosphere", "spectral", "atmosphere_env", "optics",
                                  "subsurface", "invisible", "ies", "sunlit"])
def test_image_parity(ctx, request, name, integrator):
    from luz_amd import Renderer, SceneFile
    wl = request.getfixturevalue(name)
    scene = SceneFile.read(ctx, wl.scene_path)
    img, stats = Renderer.render(scene, seed=wl.seed, integrator=integrator)
    ref = _golden(f"{name}_{wl.width}x{wl.height}_{wl.spp}spp_seed{wl.seed}.npz")["image"]
    e = rel_mse(img, ref)
    print(f"{name} {integrator}: rel-MSE {e:.3e}, {stats.render_ms:.1f} ms, "
          f"{stats.closest_hit_rays} closest / {stats.shadow_rays} shadow rays")
    assert np.isfinite(img).all()
    assert e <= REL_MSE_BAR


def test_volume_integrators_agree(ctx, volume):
    """The megakernel and the wavefront draw the same volume free-flight samples."""
    from luz_amd import Renderer, SceneFile
    scene = SceneFile.read(ctx, volume.scene_path)
    a, _ = Renderer.render(scene, seed=5, samples=8, integrator="megakernel")
    b, _ = Renderer.render(scene, seed=5, samples=8, integrator="wavefront")
    assert rel_mse(b, a) < 1e-5
