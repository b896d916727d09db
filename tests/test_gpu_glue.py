"""The reference-side binding end to end on the device (GPU): a Scene built in code (no
file), and Scenes read by the reference's own loader, rendered through
luzgpu_glue::manageThreadsOnDevice -- flattened, handed over with luzgpu_scene_create,
rendered, the image left in the Scene -- against the same Scene rendered by the reference's
own Renderer::render on the host with the same seed (oracle/glue_driver.cpp, prebuilt by
`make -C oracle glue` where the reference sources are)."""

import os
import re
import subprocess

import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "oracle", "_ref", "glue_driver")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER), reason="glue driver not built (make -C oracle glue)")]


def test_scene_built_in_code_renders_on_device_like_the_reference():
    print("glue_driver render-inmemory 0 160 120 16 7", flush=True)
    out = subprocess.run([DRIVER, "render-inmemory", "0", "160", "120", "16", "7"], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=300)
    print(out.stdout, out.stderr[-2000:])
    assert out.returncode == 0
    tokens = out.stdout.split()
    rel = float(tokens[tokens.index("rel_mse") + 1])
    i = tokens.index("samples")
    assert tokens[i + 1] == tokens[i + 2] == str(160 * 120 * 16)
    assert rel <= 1e-4


@pytest.mark.parametrize("which", ["subsurface", "invisible", "ies", "sunlit", "cornell", "suzanne", "volumetric_godrays"])
def test_loaded_scene_renders_on_device_like_the_reference(request, which):
    """Principled subsurface (Burley and thin), invisible spheres, IES-profiled sphere lights
    (the profiles handed over as luzgpu_ies_desc records), and the reference's example
    scenes (spectral reflectances, area and point lights, sphere blocks, meshes, an invisible
    sphere light) read by the reference's loader, flattened and rendered on the device."""
    from test_examples import EXAMPLES
    path = request.getfixturevalue(which).scene_path if which in ("subsurface", "invisible", "ies", "sunlit") else \
        next(e[1] for e in EXAMPLES if e[0] == which)
    print("glue_driver render-file 0", path, 1, flush=True)
    out = subprocess.run([DRIVER, "render-file", "0", path, "1"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, timeout=300)
    print(out.stdout[-1500:], out.stderr[-2000:])
    assert out.returncode == 0
    # the reference's loader prints its own progress on the same stream
    samples = re.search(r"device_ms \S+ cpu_ms \S+ samples (\d+) (\d+)", out.stdout)
    rel = float(re.search(r"rel_mse (\S+)", out.stdout).group(1))
    assert samples and samples.group(1) == samples.group(2)
    assert rel <= 1e-4
