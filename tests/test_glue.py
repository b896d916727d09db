"""The reference-side binding (luz_amd/glue: a populated Scene flattened into
luzgpu_scene_desc records and handed to luzgpu_scene_create, no file involved), compiled
against the reference's own headers and linked with the reference (oracle/Makefile, glue
target; driver oracle/glue_driver.cpp).

For scenes read by the reference's own loader and for a scene built in code as the
reference's tests build theirs (Tests.cpp:221-242), the library's top-level BVH built from
the flattened records equals the reference's own BVHNode topology of that Scene, and equals
what luzgpu_scene_load builds from the same file; every mesh BVH (links, triangle order,
boxes) is identical between the two routes. Runs only where the reference sources are
(the driver links them)."""

import os
import re
import subprocess

import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "oracle", "_ref", "glue_driver")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="the glue driver needs the reference sources (make -C oracle glue)")


def _run(*args):
    out = subprocess.run([DRIVER, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    lines = {}
    for raw in out.stdout.splitlines():
        # the reference's loader prints progress (with colour codes) on the same stream
        m = re.search(r"(?:^|\x1b\[0m)((?:file|device|reference|blas|info) .*)$", raw)
        if not m:
            continue
        key, _, rest = m.group(1).partition(" ")
        if key == "blas" or key == "info":
            route, _, rest = rest.partition(" ")
            lines.setdefault((key, route), []).append(rest)
        else:
            lines[key] = rest
    return lines


@pytest.mark.parametrize("name", ["cornell", "primitives", "volume", "materials", "meshlight", "optics", "invisible", "ies", "sunlit", "spheres"])
def test_flattened_scene_matches_file_route_and_reference(request, name):
    wl = request.getfixturevalue(name)
    r = _run("topology", wl.scene_path)
    assert r["device"] == r["reference"], (r["device"][:200], r["reference"][:200])
    assert r["device"] == r["file"]
    assert r[("info", "device")] == r[("info", "file")]
    assert r.get(("blas", "device")) == r.get(("blas", "file"))
    print(f"{name}: {r[('info', 'device')]}, {len(r.get(('blas', 'device'), []))} mesh BVHs equal")


@pytest.mark.parametrize("example", __import__("test_examples").EXAMPLES, ids=lambda e: e[0])
def test_flattened_example_scene_matches_file_route_and_reference(example):
    """The reference's example scenes (spectra, area and point lights, sphere blocks, inline
    materials, [meshes]) read by the reference's loader and flattened, against our reader."""
    r = _run("topology", example[1])
    assert r["device"] == r["reference"]
    assert r["device"] == r["file"]
    # the file route also keeps named materials that no object uses (suzanne's floor_matte);
    # the flattened route holds only those the Scene's hittables refer to
    strip = lambda info: [re.sub(r" materials \d+", "", x) for x in info]
    assert strip(r[("info", "device")]) == strip(r[("info", "file")])
    assert r.get(("blas", "device")) == r.get(("blas", "file"))


def test_scene_built_in_code_matches_reference_topology():
    r = _run("inmemory", 64, 48)
    assert r["device"] == r["reference"]
    assert r[("info", "device")] == ["hittables 15 triangles 18 lights 1 materials 6"]
