"""The C-ABI library loads and exports every symbol include/luzgpu.h declares (CPU)."""

import os
import re

from conftest import ROOT


def test_library_exports_header_symbols():
    from luz_amd import _lib
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "luzgpu.h")).read()
    declared = set(re.findall(r"\b(luzgpu_[a-z_]+)\s*\(", header))
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        return
    from luz_amd import Context, _lib
    import pytest
    with pytest.raises(_lib.LuzGpuError):
        Context(0)


C_CALLER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "luzgpu.h"

int main(int argc, char** argv)
{
	luzgpu_scene* scene = NULL;
	luzgpu_scene_info info;
	luzgpu_post_settings post;
	size_t need = 0;
	if (argc < 2) return 2;
	/* host-only load (no device): parse, build the acceleration structures */
	if (luzgpu_scene_load(NULL, argv[1], &scene) != LUZGPU_OK)
	{
		fprintf(stderr, "load: %s\n", luzgpu_last_error());
		return 1;
	}
	if (luzgpu_scene_get_info(scene, &info) != LUZGPU_OK) return 1;
	if (luzgpu_scene_post_settings(scene, &post) != LUZGPU_OK) return 1;
	if (luzgpu_scene_tlas_topology(scene, NULL, 0, &need) != LUZGPU_OK || need == 0) return 1;
	printf("%u %u %u %u %u %.3f\n", info.hittables, info.triangles, info.materials, post.view_transform, post.bloom,
		post.contrast);
	/* an error path: a missing file returns a status and a message, never aborts */
	luzgpu_scene* missing = NULL;
	if (luzgpu_scene_load(NULL, "/nonexistent/scene.luz", &missing) == LUZGPU_OK) return 1;
	if (luzgpu_last_error()[0] == '\0') return 1;
	luzgpu_scene_release(scene);
	return 0;
}
"""


def test_c_caller_compiles_links_and_loads(tmp_path, cornell):
    """A plain C11 caller of include/luzgpu.h: the header is valid C, the library links with
    -lluzgpu, and a host-only scene load answers queries (no GPU involved)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "a C compiler is needed for this check"
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.join(ROOT, "luz_amd")
    subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-lluzgpu", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), cornell.scene_path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    fields = out.stdout.split()
    assert int(fields[1]) == 34          # the Cornell box's triangles
    assert int(fields[3]) == 3 and int(fields[4]) == 0   # view_transform=raw, bloom=0
