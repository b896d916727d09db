"""No statistically significant bias against the reference renderer (GPU).

For every small scene family, 64 device renders at seeds 1..64 against 64 reference renders
at the independent seeds 1001..1064 (tests/golden/make_golden.py, zstat_fixtures): the
Welch z of the per-channel image mean must satisfy |z| < 3, and of each of the 16 tile
means per channel |z| < 4 (48 comparisons per scene, so a per-test 3-sigma bar would fire
by chance).

The optics scene is not among them: its rough dielectrics make the reference's own estimator
heavy-tailed (single samples reach 1e8 where the transmission half vector the reference
forms, normalize(view + (eta_i / eta_t) * l), is nearly perpendicular to the normal), so a
64-seed mean is dominated by whichever seeds draw such a sample. It is compared at equal
seeds instead (test_gpu_primitives), where both sides draw the same samples."""

import os

import numpy as np
import pytest

from bias import tile_means, welch_z
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cornell", "spheres", "primitives", "materials", "volume"])
def test_no_per_channel_bias_over_64_seeds(ctx, request, name):
    from luz_amd import Renderer, SceneFile
    wl = request.getfixturevalue(name)
    g = np.load(os.path.join(GOLDEN, f"zstats_{name}.npz"))
    ref = g["tile_means"]
    spp, w, h = int(g["spp"]), int(g["width"]), int(g["height"])
    scene = SceneFile.read(ctx, wl.scene_path)
    dev = np.stack([tile_means(Renderer.render(scene, seed=s, samples=spp, resolution=(w, h))[0])
                    for s in range(1, len(ref) + 1)])
    z_image = welch_z(dev.mean(1), ref.mean(1))
    z_tiles = welch_z(dev, ref)
    print(f"{name}: image z {np.round(z_image, 2)}, max |tile z| {np.abs(z_tiles).max():.2f}, "
          f"device mean {dev.mean((0, 1))}, reference mean {ref.mean((0, 1))}")
    assert np.all(np.abs(z_image) < 3.0)
    assert np.abs(z_tiles).max() < 4.0
