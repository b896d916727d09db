"""The reference's example scenes through the scene-file reader, and the reader's spectra,
meshes, light blocks, sphere blocks, inline materials and volume presets (CPU).

The reduced example scenes under tests/golden/scenes are the reference's own files
(examples/scenes) with their resolution and sample count cut, adaptive sampling and
denoising off (tests/golden/make_golden.py); the reference rendered each in FP64 at seed 1.
The oracle reads them with the product's reader, so a reader mistake (a spectrum integrated
wrongly, a light unit off by pi or 683, a preset's coefficients misapplied) shows as an image
that differs from the reference's. The identities below pin the reader's conversions
without the reference: equal scenes written two ways render the same image."""

import glob
import os
import re

import numpy as np
import pytest

import oracle_lib
from conftest import GOLDEN, rel_mse

REFERENCE_EXAMPLES = "/root/reference/examples/scenes"


def example_scenes():
    """(name, path, width, height, spp) of every reduced example scene."""
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "scenes", "*.luz"))):
        m = re.match(r"(.+)_(\d+)x(\d+)_(\d+)spp\.luz$", os.path.basename(path))
        out.append((m.group(1), path, int(m.group(2)), int(m.group(3)), int(m.group(4))))
    return out


EXAMPLES = example_scenes()


def test_example_fixtures_present():
    names = {e[0] for e in EXAMPLES}
    assert {"cornell", "suzanne", "rtiow", "volumetric_spotlight_fog", "volumetric_gray_sphere_white_bg",
            "volumetric_godrays"} <= names


@pytest.mark.parametrize("example", EXAMPLES, ids=[e[0] for e in EXAMPLES])
def test_example_scene_matches_reference(example):
    """cornell: [spectra] reflectance curves, conductor and glass presets, an area light in
    luminance units at meters_per_unit 0.01, sphere blocks; suzanne: [meshes], an area light
    with a blackbody colour, a point light in lumens, an inline material; rtiow: sphere
    blocks, a thin lens; volumetric_spotlight_fog: an area light in a Henyey-Greenstein box;
    volumetric_gray_sphere_white_bg: an isotropic sphere volume on a white background;
    volumetric_godrays: an invisible sphere light, a blackbody sun, a mist preset scaled by
    density_scale, a metal preset, inline materials."""
    name, path, w, h, spp = example
    ref = np.load(os.path.join(GOLDEN, f"example_{name}_{w}x{h}_{spp}spp_seed1.npz"))["image"]
    img, _ = oracle_lib.render(path, 1, spp, w, h)
    e = rel_mse(img, ref)
    print(f"oracle vs reference ({name}): rel-MSE {e:.3e}")
    assert e < 1e-12
    g = np.load(os.path.join(GOLDEN, f"example_{name}_primary.npz"))
    ids, _ = oracle_lib.primary(path, 1, w, h)
    assert np.array_equal(ids[..., 0], g["top"]) and np.array_equal(ids[..., 1], g["tri"])


@pytest.mark.skipif(not os.path.isdir(REFERENCE_EXAMPLES), reason="the reference's example scenes are not here")
def test_reference_example_scenes_load():
    """Every example scene either loads, or names the one feature the device path does not
    take (textures, photon-map caustics)."""
    from luz_amd import SceneFile
    for name in ("cornell", "suzanne", "rtiow", "volumetric_spotlight_fog", "volumetric_gray_sphere_white_bg",
                 "volumetric_godrays"):
        SceneFile.read(None, os.path.join(REFERENCE_EXAMPLES, f"{name}.luz"))
    refused = {"earth": "textures", "earth_atmosphere": "textures", "underwater_caustics": "caustics"}
    for name, what in refused.items():
        with pytest.raises(Exception, match=re.escape(what)):
            SceneFile.read(None, os.path.join(REFERENCE_EXAMPLES, f"{name}.luz"))


# A small scene the identities below rewrite: one Lambertian ball over a floor under a
# ceiling panel, rendered by the oracle at 24 x 18, 8 spp.
BASE = """[settings]
resolution=24,18
samples=8
maxlightbounces=4
view_transform=raw
sky=none
background=(0.05,0.05,0.05)
distanceblueness=0
{settings}
{spectra}
[materials]
material ball {{
type=lambertian
color={ball}
}}
material floor {{
type=lambertian
color=(0.6,0.6,0.6)
}}

[scene]
camera main {{
position=(0,1,-4)
direction=(0,-0.1,1)
focal_length_mm=35
sensor_width_mm=36
sensor_height_mm=27
pinhole=1
focus_distance=4
}}
{light}
objects{{
plane=0,(0,1,0),material=floor
{ball_object}
}}
"""

PANEL = """area_light panel {{
position=(0,3,0)
normal=(0,-1,0)
size=(2,1.5)
color=(1,0.9,0.8)
{unit}
}}"""


def _render(tmp_path, tag, **parts):
    fields = dict(settings="", spectra="", ball="(0.5,0.5,0.5)", light=PANEL.format(unit="radiance=5"),
                  ball_object="sphere=(0,1,0),1,material=ball")
    fields.update(parts)
    p = tmp_path / f"{tag}.luz"
    p.write_text(BASE.format(**fields))
    return oracle_lib.render(str(p), 3, 8, 24, 18)[0]


def test_flat_reflectance_is_its_value(tmp_path):
    """A flat spectrum reflects its value in every channel: the illuminant-weighted sum divided
    by the illuminant's own white is the value, whatever the illuminant and the CIE fits."""
    rgb = _render(tmp_path, "rgb", ball="(0.35,0.35,0.35)")
    spectra = "[spectra]\nreflectance flat {\n400 0.35\n550, 0.35\n700;0.35 # flat\n}\n"
    flat = _render(tmp_path, "flat", spectra=spectra, ball="reflectance(flat)")
    assert rel_mse(flat, rgb) < 1e-24
    assert not np.array_equal(flat, _render(tmp_path, "other", ball="(0.3,0.35,0.35)"))


def test_reflectance_file_equals_named_spectrum(tmp_path):
    """reflectance(PATH) reads the same samples from a file beside the scene (order, commas,
    comments and quotes do not matter); the spectrum's own shape gives a coloured result."""
    samples = [(400, 0.1), (480, 0.2), (560, 0.7), (640, 0.4), (700, 0.3)]
    body = "\n".join(f"{w} {v}" for w, v in samples)
    named = _render(tmp_path, "named", spectra=f"[spectra]\nreflectance s {{\n{body}\n}}\n", ball="reflectance(s)")
    (tmp_path / "curve.csv").write_text("# nm, value\n" + "\n".join(f"{w},{v}" for w, v in reversed(samples)) + "\n")
    from_file = _render(tmp_path, "file", ball="reflectance('curve.csv')")
    assert np.array_equal(named, from_file)
    # the curve is green-heavy: the ball's pixels are greener than the grey-ball render
    grey = _render(tmp_path, "grey", ball="(0.3,0.3,0.3)")
    diff = named - grey
    assert diff[..., 1].sum() > diff[..., 2].sum()


def test_spectrum_errors(tmp_path):
    from luz_amd import SceneFile
    bad = {
        "one sample": "reflectance s {\n500 0.5\n}\n",
        "repeated wavelength": "reflectance s {\n500 0.5\n500 0.6\n}\n",
        "out of range": "reflectance s {\n350 0.5\n500 0.6\n}\n",
        "value above one": "reflectance s {\n400 0.5\n500 1.2\n}\n",
        "three numbers": "reflectance s {\n400 0.5 1\n500 0.6\n}\n",
        "not closed": "reflectance s {\n400 0.5\n500 0.6\n",
        "duplicate": "reflectance s {\n400 0.5\n500 0.6\n}\nreflectance s {\n400 0.5\n500 0.6\n}\n",
    }
    for what, block in bad.items():
        p = tmp_path / "bad.luz"
        p.write_text(BASE.format(settings="", spectra="[spectra]\n" + block, ball="reflectance(s)",
                                 light=PANEL.format(unit="radiance=5"), ball_object="sphere=(0,1,0),1,material=ball"))
        with pytest.raises(Exception):
            SceneFile.read(None, str(p))
        print("refused:", what)
    p = tmp_path / "missing.luz"
    p.write_text(BASE.format(settings="", spectra="", ball="reflectance(nothing)", light=PANEL.format(unit="radiance=5"),
                             ball_object="sphere=(0,1,0),1,material=ball"))
    with pytest.raises(Exception, match="could not be opened"):
        SceneFile.read(None, str(p))


def test_area_light_units_agree(tmp_path):
    """luminance L is radiance L/683; power P over a w x h panel is radiance P/(pi w h);
    lumens F is luminance F/(pi w h); meters_per_unit scales the area in square metres."""
    area = 2.0 * 1.5
    base = _render(tmp_path, "radiance", light=PANEL.format(unit="radiance=5"))
    cases = {
        "luminance": f"luminance={5 * 683}",
        "power": f"power={float(5 * np.pi * area)!r}",
        "lumens": f"lumens={float(5 * 683 * np.pi * area)!r}",
    }
    for tag, unit in cases.items():
        img = _render(tmp_path, tag, light=PANEL.format(unit=unit))
        assert rel_mse(img, base) < 1e-24, tag
    # 0.5 m per unit: the same panel is a quarter of the square metres, so a quarter the power
    scaled = _render(tmp_path, "mpu", settings="meters_per_unit=0.5",
                     light=PANEL.format(unit=f"power={float(5 * np.pi * area * 0.25)!r}"))
    assert rel_mse(scaled, base) < 1e-24
    assert not np.array_equal(base, _render(tmp_path, "dimmer", light=PANEL.format(unit="radiance=4")))


def test_sphere_light_units_agree(tmp_path):
    """A point light of radius r: power P is radiance P/(4 pi^2 r^2); candela I is
    luminance 4I/(4 pi r^2); w_sr J is radiance 4J/(4 pi r^2)."""
    r = 0.25
    surface = 4.0 * np.pi * r * r
    light = "point_light bulb {{\nposition=(1.5,2.5,-0.5)\nradius=0.25\ncolor=(1,0.9,0.8)\n{unit}\n}}"
    base = _render(tmp_path, "radiance", light=light.format(unit="radiance=20"))
    for tag, unit in {"power": f"power={float(20 * np.pi * surface)!r}",
                      "candela": f"candela={float(20 * 683 * surface / 4)!r}",
                      "w_sr": f"w_sr={float(20 * surface / 4)!r}"}.items():
        img = _render(tmp_path, tag, light=light.format(unit=unit).replace("point_light", "sphere_light"))
        assert rel_mse(img, base) < 1e-24, tag


def test_light_block_errors(tmp_path):
    from luz_amd import SceneFile
    bad = {
        "two quantities": PANEL.format(unit="radiance=5\nluminance=10"),
        "no quantity": PANEL.format(unit=""),
        "zero size": PANEL.format(unit="radiance=5").replace("size=(2,1.5)", "size=(0,1.5)"),
        "zero normal": PANEL.format(unit="radiance=5").replace("normal=(0,-1,0)", "normal=(0,0,0)"),
        "black colour": PANEL.format(unit="radiance=5").replace("color=(1,0.9,0.8)", "color=(0,0,0)"),
        "negative": PANEL.format(unit="radiance=-1"),
        "intensity on a panel": PANEL.format(unit="candela=5"),
        "unknown key": PANEL.format(unit="radiance=5\nshape=disc"),
        "sphere light without a quantity": "sphere_light s {\nposition=(0,2,0)\n}",
        "sphere light radius": "point_light s {\nradius=0\nradiance=1\n}",
        "sphere light ies file missing": "point_light s {\nradiance=1\nies=lamp.ies\n}",
        "not closed": "area_light a {\nradiance=1\n",
    }
    for what, light in bad.items():
        p = tmp_path / "bad.luz"
        p.write_text(BASE.format(settings="", spectra="", ball="(0.5,0.5,0.5)", light=light,
                                 ball_object="sphere=(0,1,0),1,material=ball"))
        with pytest.raises(Exception):
            SceneFile.read(None, str(p))
        print("refused:", what)


def test_sphere_block_inline_material_and_mesh_names(tmp_path):
    """`sphere NAME {}` is the sphere= line; `material[ lambertian=(...) ]` is a named
    Lambertian of that colour; `mesh=NAME` from [meshes] is file=PATH."""
    base = _render(tmp_path, "line")
    block = _render(tmp_path, "block", ball_object="sphere ball {\ncenter=(0,1,0)\nradius=1\nmaterial=ball\n}")
    assert np.array_equal(block, base)
    inline = _render(tmp_path, "inline", ball_object="sphere=(0,1,0),1,material[\n# grey\nlambertian=(0.5,0.5,0.5)\n]")
    assert np.array_equal(inline, base)

    from luz_amd.workloads import _QUAD_OBJ
    (tmp_path / "quad.obj").write_text(_QUAD_OBJ)
    obj = "object q {\n%s\nposition=(0,0.5,1.5)\nscale=(0.8,1,0.8)\nmaterial=ball\n}"
    by_file = _render(tmp_path, "by_file", ball_object=obj % "file=quad.obj")
    text = BASE.format(settings="", spectra="", ball="(0.5,0.5,0.5)", light=PANEL.format(unit="radiance=5"),
                       ball_object=obj % "mesh=quad")
    text = text.replace("[scene]", "[meshes]\nmesh quad {\nfile=quad.obj\n}\n\n[scene]")
    p = tmp_path / "by_name.luz"
    p.write_text(text)
    assert np.array_equal(oracle_lib.render(str(p), 3, 8, 24, 18)[0], by_file)

    from luz_amd import SceneFile
    for what, obj_text in {"both": obj % "file=quad.obj\nmesh=quad", "unknown": obj % "mesh=nothing"}.items():
        p.write_text(BASE.format(settings="", spectra="", ball="(0.5,0.5,0.5)", light=PANEL.format(unit="radiance=5"),
                                 ball_object=obj_text).replace("[scene]", "[meshes]\nmesh quad {\nfile=quad.obj\n}\n\n[scene]"))
        with pytest.raises(Exception):
            SceneFile.read(None, str(p))
    for inline_bad in ("material[\n]", "material[\nlambertian=(1,1,1)\nlambertian=(1,1,1)\n]", "material[\nfoo=1\n]",
                       "material[\nlambertian=(1,1,1)\n"):
        p.write_text(BASE.format(settings="", spectra="", ball="(0.5,0.5,0.5)", light=PANEL.format(unit="radiance=5"),
                                 ball_object="sphere=(0,1,0),1," + inline_bad))
        with pytest.raises(Exception):
            SceneFile.read(None, str(p))


def test_volume_presets(tmp_path):
    """preset=NAME with density_scale=s is the explicit coefficients s*sigma_s, s*sigma_a of
    that preset; density is the luminance of the extinction and the albedo sigma_s/density,
    clamped to 1 (the mist preset: sigma_s (0.025,0.027,0.030), sigma_a 0.001, g 0.72)."""
    fog = "volume v {{\nshape=sphere\nposition=(0,1,0)\nradius=1.5\n{keys}\n}}"
    preset = _render(tmp_path, "preset", ball_object=fog.format(keys="preset=mist\ndensity_scale=20"))
    explicit = _render(tmp_path, "explicit", ball_object=fog.format(
        keys="sigma_s=(0.5,0.54,0.6)\nsigma_a=(0.02,0.02,0.02)\nanisotropy=0.72"))
    assert rel_mse(explicit, preset) < 1e-20
    lum = np.array([0.2722287168, 0.6740817658, 0.0536895174])
    density = float(lum @ np.array([0.52, 0.56, 0.62]))
    albedo = [min(1.0, float(v)) for v in np.array([0.5, 0.54, 0.6]) / density]  # blue clamps to 1
    plain = _render(tmp_path, "plain", ball_object=fog.format(
        keys=f"density={density!r}\ncolor=({albedo[0]!r},{albedo[1]!r},{albedo[2]!r})\nanisotropy=0.72"))
    assert rel_mse(plain, preset) < 1e-20
    from luz_amd import SceneFile
    for keys in ("preset=mist\ndensity=1", "preset=mist\nsigma_s=(1,1,1)", "preset=nothing", "preset=mist\ndensity_scale=0",
                 "sigma_s=(1,1,1)\ncolor=(1,1,1)", "sigma_s=(0,0,0)"):
        p = tmp_path / "bad.luz"
        p.write_text(BASE.format(settings="", spectra="", ball="(0.5,0.5,0.5)", light=PANEL.format(unit="radiance=5"),
                                 ball_object=fog.format(keys=keys)))
        with pytest.raises(Exception):
            SceneFile.read(None, str(p))
