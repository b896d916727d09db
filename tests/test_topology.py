"""Acceleration-structure topology against the reference (CPU: host-side builders only).

Golden data: tests/golden/*_tlas.txt and spheres_blas_sha256.txt were written by the
reference's own BVHNode / Mesh builders (tests/golden/make_golden.py)."""

import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN
from luz_amd import SceneFile, _lib


def test_cornell_tlas_matches_reference(cornell):
    scene = SceneFile.read(None, cornell.scene_path)
    assert scene.tlas_topology() == open(os.path.join(GOLDEN, "cornell_tlas.txt")).read().strip()
    assert scene.info["hittables"] == 35 and scene.info["lights"] == 1


def test_spheres_tlas_and_all_blas_match_reference(spheres):
    scene = SceneFile.read(None, spheres.scene_path)
    assert scene.tlas_topology() == open(os.path.join(GOLDEN, "spheres_tlas.txt")).read().strip()
    digests = open(os.path.join(GOLDEN, "spheres_blas_sha256.txt")).read().split()
    assert scene.info["meshes"] == len(digests) == 512
    assert scene.info["triangles"] == 512 * 3840
    for mesh in range(scene.info["meshes"]):
        boxes, links, order = scene.blas(mesh)
        blob = np.concatenate([boxes.view(np.uint8), links.view(np.uint8)], axis=1).tobytes() + order.tobytes()
        assert hashlib.sha256(blob).hexdigest() == digests[mesh], f"mesh {mesh}"


def test_blas_invariants(spheres):
    """Each triangle exactly once; child boxes inside the parent; leaves hold <= 8."""
    scene = SceneFile.read(None, spheres.scene_path)
    boxes, links, order = scene.blas(0)
    assert sorted(order.tolist()) == list(range(len(order)))
    covered = np.zeros(len(order), dtype=int)
    for k, (left, right, start, count, leaf) in enumerate(links):
        if leaf:
            assert 1 <= count <= 8
            covered[start:start + count] += 1
        else:
            for c in (left, right):
                assert np.all(boxes[c, :3] >= boxes[k, :3]) and np.all(boxes[c, 3:] <= boxes[k, 3:])
    assert np.all(covered == 1)


def test_unsupported_feature_is_reported(tmp_path):
    import pytest
    from luz_amd import _lib
    p = tmp_path / "bad.luz"
    # material textures stay refused
    p.write_text("[settings]\nsky=none\n\n[materials]\nmaterial m {\ntype=lambertian\ntexture=a.png\n}\n\n"
                 "[scene]\ncamera c {\nposition=(0,0,0)\n}\n")
    with pytest.raises(_lib.LuzGpuError) as e:
        SceneFile.read(None, str(p))
    assert e.value.code == _lib.E_UNSUPPORTED


def test_missing_file_is_io_error():
    import pytest
    from luz_amd import _lib
    with pytest.raises(_lib.LuzGpuError) as e:
        SceneFile.read(None, "/nonexistent/scene.luz")
    assert e.value.code == _lib.E_IO


def test_primitive_and_volume_scenes_match_reference(primitives, volume):
    """Spheres, cubes and volumes enter the TLAS, planes and directional lights stay outside."""
    p = SceneFile.read(None, primitives.scene_path)
    assert p.tlas_topology() == open(os.path.join(GOLDEN, "primitives_tlas.txt")).read().strip()
    # emitters: the sphere, the rectangle and the cube (the cube never yields a light sample)
    assert p.info["hittables"] == 8 and p.info["lights"] == 3
    v = SceneFile.read(None, volume.scene_path)
    assert v.tlas_topology() == open(os.path.join(GOLDEN, "volume_tlas.txt")).read().strip()
    # emitters: the sphere lamp and the directional sun
    assert v.info["hittables"] == 8 and v.info["lights"] == 2


def test_mesh_light_scene_matches_reference(meshlight):
    """Emissive mesh instances are lights (MeshInstance::lightSelectionWeight) beside the
    rectangle; the TLAS over four instances and a rectangle (the plane stays outside)."""
    m = SceneFile.read(None, meshlight.scene_path)
    assert m.tlas_topology() == open(os.path.join(GOLDEN, "meshlight_tlas.txt")).read().strip()
    assert m.info["hittables"] == 6 and m.info["lights"] == 3 and m.info["instances"] == 4


def test_invisible_spheres_stay_in_the_tlas_and_the_light_list(invisible):
    """Invisible spheres keep their place in the reference's BVH (the builder sees every
    bounded hittable) and an invisible sphere light stays a light; only ray hits skip them."""
    s = SceneFile.read(None, invisible.scene_path)
    assert s.tlas_topology() == open(os.path.join(GOLDEN, "invisible_tlas.txt")).read().strip()
    # key light, fill light, hidden ball, the wall, the chrome ball, the cube and the haze
    # are bounded; the plane is not. Lights: both sphere lights.
    assert s.info["hittables"] == 8 and s.info["lights"] == 2


def test_ies_lights_load_and_bad_profiles_are_refused(tmp_path, ies):
    """Both IES lights are lights; profiles the reference refuses are refused with its
    messages (IESProfile.cpp:133-232), as are a zero axis and an unreadable file."""
    from luz_amd import workloads
    s = SceneFile.read(None, ies.scene_path)
    assert s.info["lights"] == 2 and s.info["hittables"] == 6
    base = open(ies.scene_path).read()
    good = workloads.ies_profiles()[0][1]
    bad = {
        "tilt": (good.replace("TILT=NONE", "TILT=INCLUDE"), "unsupported tilt"),
        "notilt": (good.replace("TILT=NONE\n", ""), "missing TILT=NONE"),
        "short": ("TILT=NONE\n1 -1 1 19\n", "incomplete lamp metadata"),
        "count": (good.replace("1 -1 1 19 5", "1 -1 1 0 5"), "vertical angle count"),
        "order": (good.replace("0 10 20 30", "0 20 10 30"), "strictly increasing"),
        "negative": (good.replace("\n" + good.split("\n")[8].split(" ")[0] + " ", "\n-5 ", 1), "non-negative"),
    }
    for name, (text, message) in bad.items():
        (tmp_path / "downlight.ies").write_text(text)
        (tmp_path / "wash.ies").write_text(workloads.ies_profiles()[1][1])
        p = tmp_path / f"{name}.luz"
        p.write_text(base)
        with pytest.raises(Exception, match=message):
            SceneFile.read(None, str(p))
    (tmp_path / "downlight.ies").write_text(good)
    p = tmp_path / "axis.luz"
    p.write_text(base.replace("ies_direction=(0.3,-1,0.2)", "ies_direction=(0,0,0)"))
    with pytest.raises(Exception, match="direction must be non-zero"):
        SceneFile.read(None, str(p))
    p = tmp_path / "missing.luz"
    p.write_text(base.replace("ies=wash.ies", "ies=nowhere.ies"))
    with pytest.raises(Exception, match="could not be opened"):
        SceneFile.read(None, str(p))


def test_volume_and_phase_validation(tmp_path):
    """HenyeyGreenstein.cpp:10-18 and Objects.cpp:491-497, 746-800 reject these inputs."""
    import pytest
    from luz_amd import _lib
    head = ("[settings]\nresolution=8,8\nsamples=1\n\n[materials]\nmaterial m {{\ntype=hg\ncolor=(1,1,1)\nanisotropy={g}\n}}\n\n"
            "[scene]\ncamera c {{\nposition=(0,0,0)\n}}\nobjects{{\n{obj}\n}}\n")
    bad_g = tmp_path / "g.luz"
    bad_g.write_text(head.format(g="1.5", obj="sphere=(0,0,5),1,material=m"))
    with pytest.raises(_lib.LuzGpuError) as e:
        SceneFile.read(None, str(bad_g))
    assert e.value.code == _lib.E_INVALID
    bad_v = tmp_path / "v.luz"
    bad_v.write_text(head.format(g="0.5", obj="volume v {\nshape=box\ndensity=-1\n}"))
    with pytest.raises(_lib.LuzGpuError) as e:
        SceneFile.read(None, str(bad_v))
    assert "density" in str(e.value)
    # a measured preset loads; overriding its density does not (Objects.cpp:768-775)
    preset = tmp_path / "p.luz"
    preset.write_text(head.format(g="0.5", obj="volume v {\nshape=box\npreset=fog\n}"))
    SceneFile.read(None, str(preset))
    preset.write_text(head.format(g="0.5", obj="volume v {\nshape=box\npreset=fog\ndensity=1\n}"))
    with pytest.raises(_lib.LuzGpuError) as e:
        SceneFile.read(None, str(preset))
    assert "density_scale" in str(e.value)


def test_material_validation(tmp_path):
    """Principled / glossy parameters outside [0,1] are rejected (Principled.cpp:35-88), and
    so are malformed subsurface properties (Materials.cpp:349-367, 805-818)."""
    import pytest
    from luz_amd import _lib
    text = ("[settings]\nresolution=8,8\nsamples=1\n\n[materials]\nmaterial m {{\n{props}\n}}\n\n"
            "[scene]\ncamera c {{\nposition=(0,0,0)\n}}\nobjects{{\nsphere=(0,0,5),1,material=m\n}}\n")
    bad = ["type=principled\nmetallic=1.5", "type=glossy\nroughness=-0.1", "type=principled\nsubsurface=1.5",
           "type=lambertian\nsubsurface=0.5", "type=lambertian\nsubsurface_scale=0.1",
           "type=principled\nsubsurface=0.5\nsubsurface_method=volumetric", "type=principled\nsubsurface_radius=(0,0,0)",
           "type=principled\nsubsurface_scale=0", "type=principled\nsubsurface_color=(-1,1,1)",
           "type=principled\nsubsurface_profile=marble"]
    for props in bad:
        f = tmp_path / "m.luz"
        f.write_text(text.format(props=props))
        with pytest.raises(_lib.LuzGpuError) as e:
            SceneFile.read(None, str(f))
        # invalid values (std::invalid_argument) and malformed files (std::runtime_error)
        assert e.value.code in (_lib.E_INVALID, _lib.E_IO), props
    for props in ("type=principled\nsubsurface=0.5", "type=principled\nsss_profile=skin",
                  "type=principled\nsubsurface=0.3\nsubsurface_method=thin_translucent"):
        f = tmp_path / "m.luz"
        f.write_text(text.format(props=props))
        SceneFile.read(None, str(f))
    for props, code in ():
        f = tmp_path / "m.luz"
        f.write_text(text.format(props=props))
        with pytest.raises(_lib.LuzGpuError) as e:
            SceneFile.read(None, str(f))
        assert e.value.code == code, props


ATMOSPHERE_BAD = [
    # Tests.cpp:1461-1463
    "sky=atmosphere\natmosphere=0,1,2,3,4,0,1,0",
    "sky=atmosphere\natmosphere=0,2,1,3,4,1,1,0",
    "sky=atmosphere\natmosphere_sun_scale=-1",
    # Settings.cpp:618-625: the key needs the atmosphere sky, and all eight values
    "sky=linear\natmosphere=0.25,12840000.0,12910000.0,7994.0,1200.0,12,6,0.1",
    "sky=atmosphere\natmosphere=0.25,12840000.0,12910000.0",
    "sky=atmosphere\natmosphere=0.25,12840000.0,12910000.0,0,1200.0,12,6,0.1",
    "sky=atmosphere\natmosphere=0.25,12840000.0,12910000.0,7994.0,1200.0,12,6,-1",
]


@pytest.mark.parametrize("settings", ATMOSPHERE_BAD)
def test_scene_file_rejects_invalid_atmosphere(tmp_path, settings):
    import pytest as _pytest
    p = tmp_path / "bad.luz"
    p.write_text(f"[settings]\n{settings}\n\n[scene]\ncamera c {{\nposition=(0,0,0)\n}}\n")
    with _pytest.raises(Exception):
        SceneFile.read(None, str(p))


def test_scene_file_accepts_atmosphere(tmp_path, atmosphere):
    """Tests.cpp:1466-1496: the default sky and a custom atmosphere= line load; the sun scale
    keys are accepted in all four spellings."""
    for settings in ("", "sky=atmosphere\natmosphere=0.25,12840000.0,12910000.0,7994.0,1200.0,12,6,0.1",
                     "atmospheresunscale=0.5", "atmosphere_sun_multiplier=2", "atmospheresunmultiplier=0"):
        p = tmp_path / "ok.luz"
        p.write_text(f"[settings]\n{settings}\n\n[scene]\ncamera c {{\nposition=(0,0,0)\n}}\n")
        SceneFile.read(None, str(p)).release()
    SceneFile.read(None, atmosphere.scene_path).release()


def test_atmosphere_march_length_is_capped(tmp_path):
    """The device marches SAMPLES x LIGHT_SAMPLES FP64 steps per miss ray in one launch; the
    loader refuses values above 4096 each as unsupported (the reference would only slow
    down) and accepts the cap itself."""
    line = "sky=atmosphere\natmosphere=0.25,12840000.0,12910000.0,7994.0,1200.0,{},{},0.1"
    for s, l in ((4097, 6), (12, 5000)):
        p = tmp_path / "long.luz"
        p.write_text(f"[settings]\n{line.format(s, l)}\n\n[scene]\ncamera c {{\nposition=(0,0,0)\n}}\n")
        with pytest.raises(_lib.LuzGpuError) as e:
            SceneFile.read(None, str(p))
        assert e.value.code == _lib.E_UNSUPPORTED
    p = tmp_path / "cap.luz"
    p.write_text(f"[settings]\n{line.format(4096, 4096)}\n\n[scene]\ncamera c {{\nposition=(0,0,0)\n}}\n")
    SceneFile.read(None, str(p)).release()


class _FakeTensor:
    """Stands in for a device tensor on a machine without a GPU."""

    def __init__(self, is_cuda=True, dtype="torch.float32", contiguous=True, numel=12, index=0):
        import types
        self.is_cuda, self.dtype, self._c, self._n = is_cuda, dtype, contiguous, numel
        self.device = types.SimpleNamespace(index=index)

    def data_ptr(self):
        return 4096

    def is_contiguous(self):
        return self._c

    def numel(self):
        return self._n


def test_device_buffer_checks_caller_tensors():
    """Renderer.render(out=), render_batch(outs=) and Scene.adjoint(dl_dimage) pass caller
    tensors to kernels as raw pointers; each property that makes such a pointer bad raises
    ValueError first."""
    import types
    from luz_amd.renderer import _device_buffer
    ctx = types.SimpleNamespace(device=0)
    assert _device_buffer(_FakeTensor(), ctx, 12, "out") == 4096
    for bad in (_FakeTensor(is_cuda=False), _FakeTensor(dtype="torch.float64"), _FakeTensor(contiguous=False),
                _FakeTensor(numel=9), _FakeTensor(index=1), np.zeros(12, dtype=np.float32)):
        with pytest.raises(ValueError):
            _device_buffer(bad, ctx, 12, "out")
