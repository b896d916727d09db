"""Deterministic synthetic workloads written as Luz scene files (``.luz`` + ``.obj`` + ``.hdr``).

The files use the reference's own scene-file format (docs/scene-files.md), so the same
bytes feed the reference renderer and this package's loader.  Every generator is a pure
function of its arguments: numbers are written with ``repr``-exact formatting and all
randomness comes from ``random.Random(seed)`` (Mersenne Twister), so two machines produce
byte-identical files.

Workloads
---------
* ``cornell``      Cornell box: 5 Lambertian walls (10 triangles), 1 rectangle area light of
                   radiance (17, 12, 4), 2 Lambertian boxes of 12 triangles; 300x300, 50 spp,
                   depth 8, fp32 path state.
* ``spheres``      512 seeded displaced spheres of 3840 triangles each (1,966,080 triangles),
                   radius U[0.5, 3], centres uniform in a 200^3 box; 40 % Lambertian, 30 %
                   GGX metal (alpha U[0.05, 0.5]), 30 % smooth dielectric IOR 1.5; procedural
                   HDR sky (gradient + sun disk of 5e4 cd/m^2); 1920x1080, 64 spp, depth 12.
* ``primitives``   analytic primitives: a ground plane, spheres (glass, Lambertian, an
                   emissive sphere light), cubes (rough metal, an emissive cube), a rectangle
                   light, under a constant background; 160x120, 32 spp, depth 8.
* ``materials``    one sphere per glossy / diffuse-glossy / principled variant (plastic,
                   metallic, clear-coated sheen, rough transmission, smooth glass) on a plane,
                   lit by a rectangle light and a constant background; 160x120, 32 spp, depth 8.
* ``volume``       participating media: a Henyey-Greenstein box of fog (g 0.6) and an
                   isotropic sphere of smoke over a ground plane, lit by a directional sun and
                   a sphere light; 160x120, 32 spp, depth 16.
* ``batch``        32 scenes of seeded displaced-sphere meshes with triangle counts
                   log-uniform in [1e4, 5e5], resolution drawn from {256^2, 512^2, 768x432,
                   1024^2}, spp uniform in {16..128}, depth 10, alternating sky / background.
* ``diffcornell``  the differentiable Cornell-like box: 16 materials (12 Lambertian walls and
                   box faces with albedo U[0.1, 0.9]^3, 4 GGX metal faces with roughness
                   U[0.05, 0.8]) and one rectangle emitter; 512x512, 16 spp, depth 6.
* ``terrain``      a 1,002,528-triangle displaced grid terrain inside a box of
                   Henyey-Greenstein fog (sigma_t 0.05, albedo 0.9, g 0.6) under a directional
                   light; 1280x720, 32 spp, depth 16.
"""

from __future__ import annotations

import math
import os
import random
import struct
from dataclasses import dataclass, field

import numpy as np

LUMINANCE_WEIGHTS = (0.2722287168, 0.6740817658, 0.0536895174)  # ColorManagement.cpp:313-320
LUMENS_PER_RADIANT_WATT = 683.0


def _f(x: float) -> str:
    """Shortest round-trip decimal for a double (strtod on the reader side is exact)."""
    r = repr(float(x))
    return r[:-2] if r.endswith(".0") else r


def _v(v) -> str:
    return "(" + ",".join(_f(c) for c in v) + ")"


@dataclass
class Workload:
    name: str
    directory: str
    scene_path: str
    width: int
    height: int
    spp: int
    depth: int
    seed: int
    triangles: int
    extra: dict = field(default_factory=dict)


# ---------------------------------------------------------------------------------------
# Cornell box
# ---------------------------------------------------------------------------------------

_CORNELL_WALLS = [
    # (name, material, quad corners) -- each quad becomes two triangles (v0,v1,v2),(v0,v2,v3)
    ("floor", "white", [(0, 0, 0), (555, 0, 0), (555, 0, 555), (0, 0, 555)]),
    ("ceiling", "white", [(0, 555, 0), (0, 555, 555), (555, 555, 555), (555, 555, 0)]),
    ("back", "white", [(0, 0, 555), (555, 0, 555), (555, 555, 555), (0, 555, 555)]),
    ("left", "red", [(555, 0, 0), (555, 555, 0), (555, 555, 555), (555, 0, 555)]),
    ("right", "green", [(0, 0, 0), (0, 0, 555), (0, 555, 555), (0, 555, 0)]),
]

_SHORT_BOX_TOP = [(130, 165, 65), (82, 165, 225), (240, 165, 272), (290, 165, 114)]
_TALL_BOX_TOP = [(423, 330, 247), (265, 330, 296), (314, 330, 456), (472, 330, 406)]


def _box_quads(top):
    bottom = [(x, 0.0, z) for (x, _, z) in top]
    quads = [list(top), list(reversed(bottom))]
    for i in range(4):
        j = (i + 1) % 4
        quads.append([bottom[i], bottom[j], top[j], top[i]])
    return quads


def cornell_scene_text(width=300, height=300, spp=50, depth=8) -> str:
    light_color = (17.0, 12.0, 4.0)
    lum = sum(c * w for c, w in zip(light_color, LUMINANCE_WEIGHTS))
    lines = [
        "[settings]",
        f"resolution={width},{height}",
        f"samples={spp}",
        f"maxlightbounces={depth}",
        "adaptive=0",
        "denoise=0",
        "view_transform=raw",
        "bloom=0",
        "outputfilename=render.tiff",
        "sky=none",
        "background=(0,0,0)",
        "",
        "[materials]",
        "material white {", "type=lambertian", "color=(0.73,0.73,0.73)", "}",
        "material red {", "type=lambertian", "color=(0.65,0.05,0.05)", "}",
        "material green {", "type=lambertian", "color=(0.12,0.45,0.15)", "}",
        "material light {", "type=emissive", f"color={_v(light_color)}", f"radiance={_f(lum)}", "}",
        "",
        "[scene]",
        "camera main {",
        "position=(278,273,-800)",
        "direction=(0,0,1)",
        "up=(0,1,0)",
        "focal_length_mm=35",
        "sensor_width_mm=25",
        "sensor_height_mm=25",
        "pinhole=1",
        "focus_distance=10",
        "}",
        "objects{",
    ]
    for _, mat, q in _CORNELL_WALLS:
        lines.append(f"triangle={_v(q[0])},{_v(q[1])},{_v(q[2])},material={mat}")
        lines.append(f"triangle={_v(q[0])},{_v(q[2])},{_v(q[3])},material={mat}")
    lines.append("rectangle=(278,554,279.5),(0,-1,0),130,105,material=light")
    for top in (_SHORT_BOX_TOP, _TALL_BOX_TOP):
        for q in _box_quads(top):
            lines.append(f"triangle={_v(q[0])},{_v(q[1])},{_v(q[2])},material=white")
            lines.append(f"triangle={_v(q[0])},{_v(q[2])},{_v(q[3])},material=white")
    lines.append("}")
    return "\n".join(lines) + "\n"


def write_cornell(directory: str, width=300, height=300, spp=50, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "cornell.luz")
    _write_if_changed(path, cornell_scene_text(width, height, spp, depth).encode())
    return Workload("cornell", directory, path, width, height, spp, depth, seed, 34)


# ---------------------------------------------------------------------------------------
# 2M-triangle displaced spheres under a procedural HDR sky
# ---------------------------------------------------------------------------------------

SPHERE_SLICES = 48
SPHERE_STACKS = 41  # 2 * 48 * (41 - 1) = 3840 triangles per sphere


def _sphere_obj(rng: random.Random, radius: float) -> str:
    # Smooth radial displacement: a sum of 4 random planar waves over the unit direction.
    waves = []
    for _ in range(4):
        a = rng.uniform(0.0, 2.0 * math.pi)
        z = rng.uniform(-1.0, 1.0)
        r = math.sqrt(1.0 - z * z)
        waves.append((r * math.cos(a), z, r * math.sin(a), rng.uniform(2.0, 6.0), rng.uniform(0.0, 2 * math.pi)))
    amp = 0.12
    verts = []

    def vertex(x, y, z):
        d = sum(math.sin(f * (x * wx + y * wy + z * wz) + ph) for (wx, wy, wz, f, ph) in waves) * 0.25
        s = radius * (1.0 + amp * d)
        verts.append((x * s, y * s, z * s))

    vertex(0.0, 1.0, 0.0)
    for k in range(1, SPHERE_STACKS):
        theta = math.pi * k / SPHERE_STACKS
        st, ct = math.sin(theta), math.cos(theta)
        for j in range(SPHERE_SLICES):
            phi = 2.0 * math.pi * j / SPHERE_SLICES
            vertex(st * math.cos(phi), ct, st * math.sin(phi))
    vertex(0.0, -1.0, 0.0)
    south = len(verts)  # 1-based index of the last vertex
    out = ["v %.9g %.9g %.9g" % v for v in verts]

    def ring(k, j):
        return 2 + (k - 1) * SPHERE_SLICES + (j % SPHERE_SLICES)

    for j in range(SPHERE_SLICES):
        out.append(f"f 1 {ring(1, j + 1)} {ring(1, j)}")
    for k in range(1, SPHERE_STACKS - 1):
        for j in range(SPHERE_SLICES):
            a, b = ring(k, j), ring(k, j + 1)
            c, d = ring(k + 1, j), ring(k + 1, j + 1)
            out.append(f"f {a} {b} {d}")
            out.append(f"f {a} {d} {c}")
    for j in range(SPHERE_SLICES):
        out.append(f"f {south} {ring(SPHERE_STACKS - 1, j)} {ring(SPHERE_STACKS - 1, j + 1)}")
    return "\n".join(out) + "\n"


def _rgbe_bytes(rgb: np.ndarray) -> bytes:
    """Flat (non-RLE) Radiance RGBE pixels for an (H, W, 3) float array."""
    m = rgb.max(axis=2)
    mant, exp = np.frexp(m)
    scale = np.where(m > 1e-32, mant * 256.0 / np.where(m > 1e-32, m, 1.0), 0.0)
    out = np.zeros(rgb.shape[:2] + (4,), dtype=np.uint8)
    out[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(m > 1e-32, exp + 128, 0).astype(np.uint8)
    return out.tobytes()


def sky_hdr(width=1024, height=512, sun_luminance=5.0e4) -> bytes:
    """Gradient sky + sun disk, linear sRGB radiance, equirectangular (Luz convention)."""
    v = 1.0 - (np.arange(height) + 0.5) / height  # EnvironmentMap: row y <-> v = 1 - (y+0.5)/H
    u = (np.arange(width) + 0.5) / width
    uu, vv = np.meshgrid(u, v)
    elev = (vv - 0.5) * math.pi  # asin(y)
    phi = (uu - 0.5) * 2.0 * math.pi
    dy = np.sin(elev)
    rad = np.cos(elev)
    dx, dz = rad * np.cos(phi), rad * np.sin(phi)
    t = np.clip(dy, 0.0, 1.0)[..., None]
    zenith = np.array([0.25, 0.45, 1.0]) * 0.8
    horizon = np.array([1.0, 0.95, 0.9]) * 1.2
    ground = np.array([0.35, 0.3, 0.25]) * 0.25
    sky = horizon * (1.0 - t) + zenith * t
    img = np.where((dy >= 0.0)[..., None], sky, ground)
    sun_dir = np.array([math.cos(math.radians(35)) * math.cos(1.0), math.sin(math.radians(35)),
                        math.cos(math.radians(35)) * math.sin(1.0)])
    cosang = dx * sun_dir[0] + dy * sun_dir[1] + dz * sun_dir[2]
    sun_radius = math.radians(1.5)
    sun = (cosang >= math.cos(sun_radius))[..., None]
    sun_rgb = np.array([1.0, 0.95, 0.85])
    sun_rgb = sun_rgb * ((sun_luminance / LUMENS_PER_RADIANT_WATT) / float(np.dot(sun_rgb, [0.2126, 0.7152, 0.0722])))
    img = np.where(sun, sun_rgb, img)
    header = (b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {height} +X {width}\n".encode())
    return header + _rgbe_bytes(img.astype(np.float64))


def spheres_scene(directory: str, width=1920, height=1080, spp=64, depth=12, count=512,
                  seed=1002, render_seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    rng = random.Random(seed)
    mats, objs = [], []
    for i in range(count):
        radius = rng.uniform(0.5, 3.0)
        centre = (rng.uniform(-100.0, 100.0), rng.uniform(-100.0, 100.0), rng.uniform(-100.0, 100.0))
        pick = rng.random()
        if pick < 0.4:
            albedo = tuple(rng.uniform(0.1, 0.9) for _ in range(3))
            mats += [f"material m{i} {{", "type=lambertian", f"color={_v(albedo)}", "}"]
        elif pick < 0.7:
            alpha = rng.uniform(0.05, 0.5)
            tint = tuple(rng.uniform(0.5, 0.95) for _ in range(3))
            mats += [f"material m{i} {{", "type=metal", f"color={_v(tint)}", f"roughness={_f(math.sqrt(alpha))}", "}"]
        else:
            mats += [f"material m{i} {{", "type=dielectric", "color=(1,1,1)", "ior=1.5", "roughness=0", "}"]
        obj_name = f"sphere_{i:03d}.obj"
        _write_if_changed(os.path.join(directory, obj_name), _sphere_obj(rng, radius).encode())
        objs += [f"object s{i} {{", f"file={obj_name}", f"position={_v(centre)}", f"material=m{i}", "}"]
    _write_if_changed(os.path.join(directory, "sky.hdr"), sky_hdr())
    lines = [
        "[settings]",
        f"resolution={width},{height}",
        f"samples={spp}",
        f"maxlightbounces={depth}",
        "adaptive=0",
        "denoise=0",
        "view_transform=raw",
        "bloom=0",
        "outputfilename=render.tiff",
        "environment=sky.hdr",
        "",
        "[materials]",
        *mats,
        "",
        "[scene]",
        "camera main {",
        "position=(0,0,-135)",
        "direction=(0,0,1)",
        "up=(0,1,0)",
        "focal_length_mm=24",
        "sensor_width_mm=36",
        "sensor_height_mm=20.25",
        "pinhole=1",
        "focus_distance=10",
        "}",
        *objs,
    ]
    path = os.path.join(directory, "spheres.luz")
    _write_if_changed(path, ("\n".join(lines) + "\n").encode())
    return Workload("spheres", directory, path, width, height, spp, depth, render_seed,
                    count * 2 * SPHERE_SLICES * (SPHERE_STACKS - 1), {"instances": count})


_HEADER = ["adaptive=0", "denoise=0", "view_transform=raw", "bloom=0", "outputfilename=render.tiff"]


def primitives_scene_text(width=160, height=120, spp=32, depth=8) -> str:
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.05,0.06,0.08)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.6,0.6,0.6)", "}",
        "material red {", "type=lambertian", "color=(0.7,0.15,0.1)", "}",
        "material glass {", "type=dielectric", "color=(1,1,1)", "ior=1.5", "roughness=0", "}",
        "material brushed {", "type=metal", "color=(0.85,0.8,0.7)", "roughness=0.35", "}",
        "material mirror {", "type=metal", "color=(0.9,0.9,0.9)", "roughness=0", "}",
        "material bulb {", "type=emissive", "color=(1,0.85,0.6)", "radiance=12", "}",
        "material panel {", "type=emissive", "color=(0.6,0.7,1)", "radiance=4", "}",
        "material block {", "type=emissive", "color=(1,0.3,0.2)", "radiance=2", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,3,-12)", "direction=(0,-0.15,1)", "up=(0,1,0)", "focal_length_mm=35",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=10", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "sphere=(-2.5,1.5,2),1.5,material=glass",
        "sphere=(2.5,1,1),1,material=red",
        "sphere=(0.5,0.6,-1.5),0.6,material=mirror",
        "cube=(0,1,4),(0,0,-1),2,2,2,material=brushed",
        "sphere=(0,6,2),0.75,material=bulb",
        "rectangle=(-4,5,6),(0.5,-1,-0.3),2,2,material=panel",
        "cube=(4,3,6),(0,0,-1),1,1,1,material=block",
        "}",
    ]
    return "\n".join(lines) + "\n"


def skylinear_scene_text(width=128, height=96, spp=16, depth=6) -> str:
    """The primitives scene under the linear gradient sky, with colours given as sRGB-encoded,
    display and linear-sRGB values (converted to ACEScg by the loader)."""
    text = primitives_scene_text(width, height, spp, depth)
    text = text.replace("sky=none\nbackground=(0.05,0.06,0.08)\n", "sky=linear\ndistanceblueness=0\n")
    text = text.replace("color=(0.6,0.6,0.6)", "color=srgb(0.8,0.78,0.75)")
    text = text.replace("color=(0.7,0.15,0.1)", "color=linear_srgb(0.7,0.15,0.1)")
    text = text.replace("color=(0.85,0.8,0.7)", "color=display_srgb(0.9,0.85,0.7)")
    text = text.replace("color=(0.6,0.7,1)", "color=lin_srgb(0.6,0.7,1)")
    return text


def write_skylinear(directory: str, width=128, height=96, spp=16, depth=6, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "skylinear.luz")
    _write_if_changed(path, skylinear_scene_text(width, height, spp, depth).encode())
    return Workload("skylinear", directory, path, width, height, spp, depth, seed, 0)


def atmosphere_scene_text(width=96, height=64, spp=8, depth=6) -> str:
    """The atmosphere sky over a small landscape 2 km above a planet of the default size:
    scene units of 1000 km (meters_per_unit=1e6), a custom atmosphere= line (12 view and 6
    sun steps, sun 45 degrees up) and a scaled sun; half the frame is sky, the rest shows
    the aerial perspective over 50-300 km."""
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "meters_per_unit=1000000", "sky=atmosphere", "atmosphere_sun_scale=0.02",
        "atmosphere=0.25,6360000,6420000,7994,1200,12,6,0.5", "",
        "[materials]",
        "material ground {", "type=lambertian", "color=(0.35,0.3,0.22)", "}",
        "material clay {", "type=lambertian", "color=(0.7,0.35,0.2)", "}",
        "material steel {", "type=metal", "color=(0.8,0.8,0.85)", "roughness=0.2", "}",
        "material glow {", "type=emissive", "color=(1,0.8,0.5)", "radiance=6", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,6.362,0)", "direction=(0,-0.04,1)", "up=(0,1,0)", "focal_length_mm=35",
        "sensor_width_mm=36", "sensor_height_mm=24", "pinhole=1", "focus_distance=1", "}",
        "objects{",
        "plane=6.3605,(0,1,0),material=ground",
        "sphere=(-0.04,6.372,0.15),0.012,material=clay",
        "sphere=(0.05,6.375,0.3),0.015,material=steel",
        "sphere=(0.0,6.364,0.08),0.004,material=glow",
        "}",
    ]
    return "\n".join(lines) + "\n"


def sunlit_scene_text(width=96, height=64, spp=8, depth=6) -> str:
    """The atmosphere scene lit by a `solar=` directional light, which also becomes the
    atmosphere's sun: the sky's scattering follows the light's direction (low, from the
    side, not the atmosphere= line's 45 degrees) and its solar radiance."""
    text = atmosphere_scene_text(width, height, spp, depth)
    light = "directional_light sun {\ndirection=(-0.5,-0.35,0.6)\nsolar=0.02\n}\n"
    return text.replace("objects{\n", light + "objects{\n", 1)


def write_sunlit(directory: str, width=96, height=64, spp=8, depth=6, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "sunlit.luz")
    _write_if_changed(path, sunlit_scene_text(width, height, spp, depth).encode())
    return Workload("sunlit", directory, path, width, height, spp, depth, seed, 0)


def write_atmosphere(directory: str, width=96, height=64, spp=8, depth=6, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "atmosphere.luz")
    _write_if_changed(path, atmosphere_scene_text(width, height, spp, depth).encode())
    return Workload("atmosphere", directory, path, width, height, spp, depth, seed, 0)


def spectral_scene_text(width=96, height=72, spp=8, depth=6) -> str:
    """The primitives scene with colours given as spectral functions: blackbody temperatures
    (with and without the K suffix, under all three names), a single wavelength and the solar
    colour."""
    text = primitives_scene_text(width, height, spp, depth)
    text = text.replace("color=(0.6,0.6,0.6)", "color=temperature(6500K)")
    text = text.replace("color=(0.7,0.15,0.1)", "color=wavelength(610nm)")
    text = text.replace("color=(0.85,0.8,0.7)", "color=solar")
    text = text.replace("color=(1,0.85,0.6)", "color=blackbody(2700)")
    text = text.replace("color=(0.6,0.7,1)", "color=color_temperature(9000 k)")
    return text


def write_spectral(directory: str, width=96, height=72, spp=8, depth=6, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "spectral.luz")
    _write_if_changed(path, spectral_scene_text(width, height, spp, depth).encode())
    return Workload("spectral", directory, path, width, height, spp, depth, seed, 0)


def write_atmosphere_env(directory: str, width=96, height=64, spp=8, depth=6, seed=1) -> Workload:
    """A small spheres scene under the atmosphere sky with its environment map both seen
    through the atmosphere and sampled as a light: a 30 m planet at the origin inside a
    220 m atmosphere with scale heights of 8 m / 2 m."""
    wl = spheres_scene(directory, width=width, height=height, spp=spp, depth=depth, count=24, render_seed=seed)
    text = open(wl.scene_path).read()
    text = text.replace("environment=sky.hdr", "sky=atmosphere\natmosphere_sun_scale=0.5\n"
                        "atmosphere=0.3,30,220,8,2,8,4,0.5\nenvironment=sky.hdr", 1)
    path = os.path.join(directory, "atmosphere_env.luz")
    _write_if_changed(path, text.encode())
    return Workload("atmosphere_env", directory, path, width, height, spp, depth, seed, wl.triangles)


_QUAD_OBJ = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nf 1 2 3\nf 1 3 4\n"


def meshlight_scene_text(width=160, height=120, spp=32, depth=8) -> str:
    """Emissive mesh instances: a displaced sphere under a non-uniform scale (transformed
    area prefix sums), a quad under a uniform scale (the mesh's own prefix sums), a rectangle
    light beside them, a lit mesh and a floor plane."""
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.02,0.02,0.03)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.55,0.55,0.5)", "}",
        "material clay {", "type=lambertian", "color=(0.7,0.35,0.2)", "}",
        "material steel {", "type=metal", "color=(0.8,0.8,0.85)", "roughness=0.3", "}",
        "material bulb {", "type=emissive", "color=(1,0.8,0.55)", "radiance=9", "}",
        "material sheet {", "type=emissive", "color=(0.55,0.7,1)", "radiance=3", "}",
        "material strip {", "type=emissive", "color=(1,1,1)", "radiance=2", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,3,-12)", "direction=(0,-0.15,1)", "up=(0,1,0)", "focal_length_mm=35",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=10", "}",
        "object bulb {", "file=bulb.obj", "position=(-1.5,4.5,3)", "rotation=(20,35,0)", "scale=(0.9,0.6,1.2)",
        "material=bulb", "}",
        "object sheet {", "file=quad.obj", "position=(3,3,5)", "rotation=(70,-20,0)", "scale=(1.5,1.5,1.5)",
        "material=sheet", "}",
        "object blob {", "file=blob.obj", "position=(1,1.2,2)", "rotation=(0,40,0)", "scale=(1,1,1)", "material=clay", "}",
        "object ball {", "file=bulb.obj", "position=(-2.5,1,0.5)", "scale=(1,1,1)", "material=steel", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "rectangle=(0,6,7),(0,-1,-0.4),4,0.5,material=strip",
        "}",
    ]
    return "\n".join(lines) + "\n"


def write_meshlight(directory: str, width=160, height=120, spp=32, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    _write_if_changed(os.path.join(directory, "bulb.obj"), _sphere_obj(random.Random(1006), 1.0).encode())
    _write_if_changed(os.path.join(directory, "blob.obj"), _sphere_obj(random.Random(1007), 1.1).encode())
    _write_if_changed(os.path.join(directory, "quad.obj"), _QUAD_OBJ.encode())
    path = os.path.join(directory, "meshlight.luz")
    _write_if_changed(path, meshlight_scene_text(width, height, spp, depth).encode())
    return Workload("meshlight", directory, path, width, height, spp, depth, seed,
                    3 * 2 * SPHERE_SLICES * (SPHERE_STACKS - 1) + 2)


def write_primitives(directory: str, width=160, height=120, spp=32, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "primitives.luz")
    _write_if_changed(path, primitives_scene_text(width, height, spp, depth).encode())
    return Workload("primitives", directory, path, width, height, spp, depth, seed, 0)


def materials_scene_text(width=160, height=120, spp=32, depth=8) -> str:
    mats = [
        ("glossy", ["type=glossy", "color=(0.9,0.7,0.3)", "roughness=0.3"]),
        ("dglossy", ["type=diffuse_glossy", "color=(0.2,0.5,0.8)", "glossy_color=(1,1,1)", "glossy_weight=0.35",
                     "roughness=0.2"]),
        ("plastic", ["type=principled", "color=(0.8,0.1,0.1)", "roughness=0.4"]),
        ("brass", ["type=principled", "color=(0.9,0.75,0.4)", "metallic=1", "roughness=0.25"]),
        ("velvet", ["type=principled", "color=(0.3,0.6,0.3)", "roughness=0.6", "sheen=0.8", "clearcoat=0.7",
                    "clearcoat_roughness=0.1"]),
        ("frosted", ["type=principled", "color=(0.95,0.95,1)", "transmission=0.9", "roughness=0.3", "ior=1.45"]),
        ("crystal", ["type=principled", "color=(1,1,1)", "transmission=1", "roughness=0", "ior=1.5"]),
    ]
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.08,0.09,0.1)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.55,0.55,0.55)", "}",
        "material panel {", "type=emissive", "color=(1,0.95,0.9)", "radiance=8", "}",
    ]
    for name, props in mats:
        lines += [f"material {name} {{", *props, "}"]
    lines += [
        "",
        "[scene]",
        "camera main {", "position=(0,3.5,-11)", "direction=(0,-0.25,1)", "up=(0,1,0)", "focal_length_mm=30",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=10", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "rectangle=(0,7,2),(0,-1,0),6,3,material=panel",
    ]
    for i, (name, _) in enumerate(mats):
        x = -4.5 + 1.5 * i
        z = 1.0 + (i % 2) * 1.5
        lines.append(f"sphere=({_f(x)},0.7,{_f(z)}),0.7,material={name}")
    lines.append("}")
    return "\n".join(lines) + "\n"


def write_materials(directory: str, width=160, height=120, spp=32, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "materials.luz")
    _write_if_changed(path, materials_scene_text(width, height, spp, depth).encode())
    return Workload("materials", directory, path, width, height, spp, depth, seed, 0)


def optics_scene_text(width=160, height=120, spp=128, depth=10) -> str:
    """The reference's default thin-lens camera (f/2.8, focused at 12 units) over the
    materials no other fixture reaches: rough dielectrics, conductor-Fresnel metals (measured
    presets and explicit eta / k, smooth and rough), glass presets with Sellmeier / Abbe
    dispersion, interior absorption (coefficient and transmittance forms) in a dielectric and
    a transmissive principled material, and direct material lines."""
    mats = [
        ("frost", ["type=dielectric", "color=(1,1,1)", "ior=1.5", "roughness=0.3"]),
        ("gold", ["type=metal", "preset=gold", "roughness=0.2"]),
        ("copper", ["type=metal", "eta=(0.27,0.68,1.34)", "k=(3.57,2.62,2.29)", "roughness=0"]),
        ("chrome", ["type=metal", "conductor=Chrome", "fuzz=0.45"]),
        ("bk7", ["type=dielectric", "color=(1,1,1)", "glass=BK7", "ior_wavelength=550nm", "roughness=0.08",
                 "transmittance=(0.6,0.85,0.95)", "attenuation_distance=0.5"]),
        ("amber", ["type=dielectric", "color=(1,1,1)", "ior=1.55", "abbe=40", "ior_wavelength=500",
                   "absorption=(0.2,0.9,2.5)"]),
        ("tinted", ["type=principled", "color=(0.95,0.95,0.95)", "transmission=1", "roughness=0.15",
                    "preset=fused_silica", "absorption=(1.5,0.4,0.2)"]),
        ("brushed", ["metal=(0.8,0.75,0.7),0.3"]),
    ]
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.06,0.07,0.09)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.6,0.58,0.55)", "}",
        "material panel {", "type=emissive", "color=(1,0.95,0.85)", "radiance=9", "}",
    ]
    for name, props in mats:
        lines += [f"material {name} {{", *props, "}"]
    lines += [
        "",
        "[scene]",
        "camera main {", "position=(0,3.2,-12)", "direction=(0,-0.22,1)", "up=(0,1,0)", "focal_length_mm=35",
        "sensor_width_mm=36", "sensor_height_mm=27", "f_stop=2.8", "focus_distance=12", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "rectangle=(0,7,3),(0,-1,0),7,3,material=panel",
    ]
    for i, (name, _) in enumerate(mats):
        x = -5.25 + 1.5 * i
        z = 0.5 + (i % 3) * 2.0
        if name == "bk7":
            lines.append(f"cube=({_f(x)},0.75,{_f(z)}),(0,0,-1),1.4,1.5,1.4,material={name}")
        else:
            lines.append(f"sphere=({_f(x)},0.7,{_f(z)}),0.7,material={name}")
    lines.append("}")
    return "\n".join(lines) + "\n"


def write_optics(directory: str, width=160, height=120, spp=128, depth=10, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "optics.luz")
    _write_if_changed(path, optics_scene_text(width, height, spp, depth).encode())
    return Workload("optics", directory, path, width, height, spp, depth, seed, 0)


def subsurface_scene_text(width=160, height=120, spp=64, depth=8) -> str:
    """Principled subsurface: Burley diffusion on analytic spheres (the skin profile with a
    wider scale, and explicit radius / colour), on a displaced mesh sphere and on a cube,
    whose exits are relocated by the probe ray; and a thin translucent panel lit from behind
    (the two-sided lobe), under an overhead panel and a sphere light behind the leaf."""
    mats = [
        ("skin", ["type=principled", "color=(0.9,0.7,0.6)", "roughness=0.5", "subsurface_profile=skin",
                  "subsurface=1", "subsurface_scale=0.06"]),
        ("jade", ["type=principled", "color=(0.5,0.8,0.6)", "roughness=0.2", "subsurface=0.7",
                  "subsurface_radius=(0.4,1.0,0.6)", "subsurface_scale=0.08", "subsurface_color=(0.8,1,0.9)",
                  "sheen=0.3"]),
        ("wax", ["type=principled", "color=(0.95,0.9,0.7)", "roughness=0.35", "subsurface=0.9",
                 "subsurface_method=burley", "subsurface_radius=(1,0.8,0.5)", "subsurface_scale=0.1",
                 "clearcoat=0.5"]),
        ("marble", ["type=principled", "color=(0.9,0.9,0.88)", "roughness=0.25", "subsurface=0.6",
                    "sss_scale=0.05"]),
        ("leaf", ["type=principled", "color=(0.3,0.7,0.2)", "roughness=0.6", "subsurface=0.8",
                  "subsurface_method=thin", "subsurface_color=(0.9,1,0.6)"]),
    ]
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.03,0.03,0.04)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.55,0.55,0.52)", "}",
        "material panel {", "type=emissive", "color=(1,0.95,0.85)", "radiance=7", "}",
        "material back {", "type=emissive", "color=(1,0.9,0.7)", "radiance=40", "}",
    ]
    for name, props in mats:
        lines += [f"material {name} {{", *props, "}"]
    lines += [
        "",
        "[scene]",
        "camera main {", "position=(0,3,-10)", "direction=(0,-0.2,1)", "up=(0,1,0)", "focal_length_mm=35",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=10", "}",
        "object wax {", "file=wax.obj", "position=(0.2,1,1.5)", "rotation=(0,30,0)", "scale=(1,1,1)", "material=wax", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "rectangle=(0,7,2),(0,-1,0),6,3,material=panel",
        "sphere=(-3,1,0.5),1,material=skin",
        "sphere=(3,0.9,0.5),0.9,material=jade",
        "cube=(-1.5,0.6,-1.5),(0,0,-1),1.2,1.2,1.2,material=marble",
        "rectangle=(2,1.6,-2),(0.3,0,-1),1.6,2.4,material=leaf",
        "sphere=(2.4,1.6,-0.6),0.25,material=back",
        "}",
    ]
    return "\n".join(lines) + "\n"


def write_subsurface(directory: str, width=160, height=120, spp=64, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    _write_if_changed(os.path.join(directory, "wax.obj"), _sphere_obj(random.Random(1008), 0.9).encode())
    path = os.path.join(directory, "subsurface.luz")
    _write_if_changed(path, subsurface_scene_text(width, height, spp, depth).encode())
    return Workload("subsurface", directory, path, width, height, spp, depth, seed, 2 * SPHERE_SLICES * (SPHERE_STACKS - 1))


def invisible_scene_text(width=160, height=120, spp=32, depth=8, key_visible=0) -> str:
    """Invisible spheres (visible=0): the key light is an invisible sphere light, lighting the
    room, a rough chrome ball and a thin haze through light sampling alone (no ray ever hits
    it, so neither the camera nor the chrome shows it); an invisible Lambertian ball stands
    in front of the camera and must not appear or cast a shadow; a visible sphere light fills."""
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.02,0.02,0.03)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.6,0.58,0.55)", "}",
        "material wall {", "type=lambertian", "color=(0.4,0.45,0.5)", "}",
        "material chrome {", "type=metal", "color=(0.9,0.9,0.92)", "roughness=0.12", "}",
        "material red {", "type=lambertian", "color=(0.8,0.15,0.1)", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,1.6,-6)", "direction=(0,-0.12,1)", "up=(0,1,0)", "focal_length_mm=30",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=6", "}",
        "sphere_light key {", "position=(-1.4,2.2,-1.2)", "radius=0.35", "color=(1,0.9,0.75)", "radiance=25",
        f"visible={key_visible}", "}",
        "sphere_light fill {", "position=(2.2,1.2,1.8)", "radius=0.25", "color=(0.6,0.7,1)", "radiance=6", "visible=1", "}",
        "sphere hidden_ball {", "position=(0.3,0.9,-3.2)", "radius=0.7", "material=red", "visible=0", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "rectangle=(0,2.5,3.5),(0,0,-1),10,5,material=wall",
        "sphere=(0.6,0.8,0.8),0.8,material=chrome",
        "cube=(-1.6,0.4,1.2),(0,0,-1),0.8,0.8,0.8,material=wall",
        "volume haze {", "shape=box", "position=(0,2,0)", "size=(12,4,12)", "density=0.03",
        "color=(0.9,0.9,0.9)", "anisotropy=0.3", "}",
        "}",
    ]
    return "\n".join(lines) + "\n"


def write_invisible(directory: str, width=160, height=120, spp=32, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "invisible.luz")
    _write_if_changed(path, invisible_scene_text(width, height, spp, depth).encode())
    return Workload("invisible", directory, path, width, height, spp, depth, seed, 0)


def ies_text(vertical, horizontal, candela, name: str) -> str:
    """An LM-63 photometric file: header, TILT=NONE, the 13 lamp numbers (one lamp, absolute
    photometry, multiplier 1), the angles and the candela values for each horizontal angle."""
    lines = ["IESNA:LM-63-2002", f"[TEST] {name}", "[MANUFAC] synthetic profile for the parity tests", "TILT=NONE",
             f"1 -1 1 {len(vertical)} {len(horizontal)} 1 2 0.1 0.1 0.1", "1 1 40",
             " ".join(f"{v:g}" for v in vertical), " ".join(f"{h:g}" for h in horizontal)]
    for row in candela:
        lines.append(" ".join(f"{c:.6g}" for c in row))
    return "\n".join(lines) + "\n"


def ies_profiles():
    """(file name, text) of two profiles: a downlight over horizontal angles 0..360 (the last
    angle held), and an asymmetric wash over 0..240 (interpolated back to 0 past 240)."""
    vertical = list(range(0, 181, 10))
    down = [[900.0 * max(0.0, math.cos(math.radians(v))) ** 1.5 + 40.0 + 150.0 * math.sin(math.radians(v)) ** 2
             * (1.0 + 0.5 * math.cos(math.radians(h))) for v in vertical] for h in (0, 90, 180, 270, 360)]
    wash = [[60.0 + 500.0 * math.sin(math.radians(v)) ** 2 * (1.0 + 0.8 * math.cos(math.radians(h - 60)))
             for v in vertical] for h in (0, 60, 120, 180, 240)]
    return [("downlight.ies", ies_text(vertical, (0, 90, 180, 270, 360), down, "downlight")),
            ("wash.ies", ies_text(vertical, (0, 60, 120, 180, 240), wash, "wash"))]


def ies_scene_text(width=160, height=120, spp=32, depth=8) -> str:
    """IES-profiled sphere lights: a downlight that emits its profile's own flux, and a tilted,
    turned wash light with an explicit flux; light samples carry the profiles, rays that hit
    the lights see them unscaled (as the reference does)."""
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.01,0.01,0.015)", "",
        "[materials]",
        "material floor {", "type=lambertian", "color=(0.6,0.58,0.55)", "}",
        "material wall {", "type=lambertian", "color=(0.5,0.5,0.55)", "}",
        "material steel {", "type=metal", "color=(0.85,0.85,0.9)", "roughness=0.2", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,1.7,-6)", "direction=(0,-0.15,1)", "up=(0,1,0)", "focal_length_mm=30",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=6", "}",
        "sphere_light spot {", "position=(-1.2,2.6,0.6)", "radius=0.15", "color=(1,0.95,0.85)", "ies=downlight.ies", "}",
        "sphere_light wash {", "position=(1.6,2.2,-0.4)", "radius=0.12", "color=(0.8,0.9,1)", "lumens=1500",
        "ies=wash.ies", "ies_direction=(0.3,-1,0.2)", "ies_rotation=30", "}",
        "objects{",
        "plane=0,(0,1,0),material=floor",
        "rectangle=(0,2.5,3.5),(0,0,-1),10,5,material=wall",
        "sphere=(0.4,0.7,1.0),0.7,material=steel",
        "cube=(-1.8,0.5,1.4),(0,0,-1),1,1,1,material=wall",
        "}",
    ]
    return "\n".join(lines) + "\n"


def write_ies(directory: str, width=160, height=120, spp=32, depth=8, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    for name, text in ies_profiles():
        _write_if_changed(os.path.join(directory, name), text.encode())
    path = os.path.join(directory, "ies.luz")
    _write_if_changed(path, ies_scene_text(width, height, spp, depth).encode())
    return Workload("ies", directory, path, width, height, spp, depth, seed, 0)


def volume_scene_text(width=160, height=120, spp=32, depth=16) -> str:
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.02,0.02,0.03)", "",
        "[materials]",
        "material ground {", "type=lambertian", "color=(0.5,0.45,0.4)", "}",
        "material steel {", "type=metal", "color=(0.8,0.8,0.85)", "roughness=0.2", "}",
        "material glass {", "type=dielectric", "color=(1,1,1)", "ior=1.5", "roughness=0", "}",
        "material lamp {", "type=emissive", "color=(1,0.8,0.5)", "radiance=20", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,4,-20)", "direction=(0,-0.1,1)", "up=(0,1,0)", "focal_length_mm=35",
        "sensor_width_mm=36", "sensor_height_mm=27", "pinhole=1", "focus_distance=10", "}",
        "objects{",
        "plane=0,(0,1,0),material=ground",
        "sphere=(0,2,0),2,material=steel",
        "sphere=(-5,1.5,-3),1.5,material=glass",
        "cube=(-4,1,4),(0,0,-1),2,2,2,material=ground",
        "sphere=(4,5,-2),0.5,material=lamp",
        "volume fog {", "shape=box", "position=(0,5,0)", "size=(30,10,30)", "density=0.05",
        "color=(0.9,0.9,0.9)", "anisotropy=0.6", "}",
        "volume smoke {", "shape=sphere", "position=(5,3,3)", "radius=2", "density=0.5", "color=(0.8,0.7,0.6)", "}",
        "directional_light sun {", "direction=(-1,-2,-1)", "color=(1,0.95,0.9)", "irradiance=3", "}",
        "}",
    ]
    return "\n".join(lines) + "\n"


def write_volume(directory: str, width=160, height=120, spp=32, depth=16, seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "volume.luz")
    _write_if_changed(path, volume_scene_text(width, height, spp, depth).encode())
    return Workload("volume", directory, path, width, height, spp, depth, seed, 0)


TERRAIN_QUADS = 708  # 708 x 708 quads x 2 = 1,002,528 triangles


def _terrain_obj(rng: random.Random, n: int, extent: float) -> bytes:
    """Height field over an n x n grid: a sum of seeded sinusoidal ridges."""
    waves = [(rng.uniform(0.0, 2 * math.pi), rng.uniform(0.02, 0.12), rng.uniform(0.0, 2 * math.pi),
              rng.uniform(0.5, 3.0)) for _ in range(8)]
    xs = np.linspace(-extent, extent, n + 1)
    gx, gz = np.meshgrid(xs, xs)
    hgt = np.zeros_like(gx)
    for ang, freq, phase, amp in waves:
        hgt += amp * np.sin(freq * (gx * math.cos(ang) + gz * math.sin(ang)) + phase)
    verts = np.stack([gx.ravel(), hgt.ravel(), gz.ravel()], axis=1)
    vtxt = "\n".join("v %.6f %.6f %.6f" % tuple(v) for v in verts)
    i = np.arange(n)
    a = (i[None, :] + (n + 1) * i[:, None]).ravel() + 1  # 1-based corner (row r, column c)
    b, c, d = a + 1, a + (n + 1), a + (n + 2)
    faces = np.empty((2 * n * n, 3), dtype=np.int64)
    faces[0::2] = np.stack([a, c, b], axis=1)
    faces[1::2] = np.stack([b, c, d], axis=1)
    ftxt = "\n".join("f %d %d %d" % tuple(f) for f in faces)
    return (vtxt + "\n" + ftxt + "\n").encode()


def terrain_scene(directory: str, width=1280, height=720, spp=32, depth=16, quads=TERRAIN_QUADS, seed=1005,
                  render_seed=1) -> Workload:
    os.makedirs(directory, exist_ok=True)
    rng = random.Random(seed)
    obj = os.path.join(directory, f"terrain_{quads}.obj")
    if not os.path.exists(obj):
        _write_if_changed(obj, _terrain_obj(rng, quads, 100.0))
    lines = [
        "[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
        "sky=none", "background=(0.03,0.04,0.06)", "",
        "[materials]",
        "material rock {", "type=lambertian", "color=(0.45,0.4,0.35)", "}",
        "",
        "[scene]",
        "camera main {", "position=(0,25,-120)", "direction=(0,-0.25,1)", "up=(0,1,0)", "focal_length_mm=28",
        "sensor_width_mm=36", "sensor_height_mm=20.25", "pinhole=1", "focus_distance=10", "}",
        "objects{",
        f"object land {{", f"file={os.path.basename(obj)}", "material=rock", "}",
        "volume haze {", "shape=box", "position=(0,15,0)", "size=(210,50,210)", "density=0.05",
        "color=(0.9,0.9,0.9)", "anisotropy=0.6", "}",
        "directional_light sun {", "direction=(-0.4,-1,0.6)", "color=(1,0.95,0.88)", "irradiance=4", "}",
        "}",
    ]
    path = os.path.join(directory, f"terrain_{quads}.luz")
    _write_if_changed(path, ("\n".join(lines) + "\n").encode())
    return Workload("terrain", directory, path, width, height, spp, depth, render_seed, 2 * quads * quads,
                    {"instances": 1})


BATCH_RESOLUTIONS = [(256, 256), (512, 512), (768, 432), (1024, 1024)]


def _uv_sphere_obj(rng: random.Random, radius: float, slices: int, stacks: int) -> str:
    """Displaced UV sphere with 2 * slices * (stacks - 1) triangles."""
    waves = []
    for _ in range(3):
        a = rng.uniform(0.0, 2.0 * math.pi)
        z = rng.uniform(-1.0, 1.0)
        r = math.sqrt(1.0 - z * z)
        waves.append((r * math.cos(a), z, r * math.sin(a), rng.uniform(2.0, 6.0), rng.uniform(0.0, 2 * math.pi)))
    verts = []

    def vertex(x, y, z):
        d = sum(math.sin(f * (x * wx + y * wy + z * wz) + ph) for (wx, wy, wz, f, ph) in waves) / 3.0
        sc = radius * (1.0 + 0.1 * d)
        verts.append((x * sc, y * sc, z * sc))

    vertex(0.0, 1.0, 0.0)
    for k in range(1, stacks):
        th = math.pi * k / stacks
        for j in range(slices):
            ph = 2.0 * math.pi * j / slices
            vertex(math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph))
    vertex(0.0, -1.0, 0.0)
    south = len(verts)
    out = ["v %.7g %.7g %.7g" % v for v in verts]

    def ring(k, j):
        return 2 + (k - 1) * slices + (j % slices)

    for j in range(slices):
        out.append(f"f 1 {ring(1, j + 1)} {ring(1, j)}")
    for k in range(1, stacks - 1):
        for j in range(slices):
            a, b2 = ring(k, j), ring(k, j + 1)
            c, d = ring(k + 1, j), ring(k + 1, j + 1)
            out.append(f"f {a} {b2} {d}")
            out.append(f"f {a} {d} {c}")
    for j in range(slices):
        out.append(f"f {south} {ring(stacks - 1, j)} {ring(stacks - 1, j + 1)}")
    return "\n".join(out) + "\n"


def batch_scenes(directory: str, count=32, seed=1003, depth=10, spp_range=(16, 128)) -> list:
    """The uneven multi-scene batch: one Workload per scene."""
    os.makedirs(directory, exist_ok=True)
    rng = random.Random(seed)
    sky = os.path.join(directory, "sky.hdr")
    _write_if_changed(sky, sky_hdr(512, 256))
    out = []
    for si in range(count):
        target = math.exp(rng.uniform(math.log(1e4), math.log(5e5)))
        width, height = rng.choice(BATCH_RESOLUTIONS)
        spp = rng.randint(spp_range[0], spp_range[1])
        # meshes of 2 * 32 * 31 = 1984 triangles, instanced to reach the target count
        slices, stacks = 32, 32
        per = 2 * slices * (stacks - 1)
        instances = max(1, int(round(target / per)))
        mesh_count = min(4, instances)
        meshes = []
        for m in range(mesh_count):
            name = f"s{si:02d}_mesh{m}.obj"
            _write_if_changed(os.path.join(directory, name), _uv_sphere_obj(rng, 1.0, slices, stacks).encode())
            meshes.append(name)
        mats, objs = [], []
        extent = 6.0 * instances ** (1.0 / 3.0)
        for i in range(instances):
            pick = rng.random()
            if pick < 0.4:
                mats += [f"material m{i} {{", "type=lambertian", f"color={_v([rng.uniform(0.1, 0.9) for _ in range(3)])}", "}"]
            elif pick < 0.7:
                mats += [f"material m{i} {{", "type=metal", f"color={_v([rng.uniform(0.5, 0.95) for _ in range(3)])}",
                         f"roughness={_f(math.sqrt(rng.uniform(0.05, 0.5)))}", "}"]
            else:
                mats += [f"material m{i} {{", "type=dielectric", "color=(1,1,1)", "ior=1.5", "roughness=0", "}"]
            pos = [rng.uniform(-extent, extent) for _ in range(3)]
            objs += [f"object o{i} {{", f"file={meshes[i % mesh_count]}", f"position={_v(pos)}",
                     f"scale={_v([rng.uniform(0.5, 2.0)] * 3)}", f"material=m{i}", "}"]
        lines = ["[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER]
        if si % 2 == 0:
            lines.append("environment=sky.hdr")
        else:
            lines += ["sky=none", "background=(0.6,0.65,0.8)"]
        lines += ["", "[materials]", *mats, "", "[scene]",
                  "camera main {", f"position=(0,0,{_f(-3.2 * extent)})", "direction=(0,0,1)", "up=(0,1,0)",
                  "focal_length_mm=30", "sensor_width_mm=36", "sensor_height_mm=24", "pinhole=1", "focus_distance=10", "}",
                  *objs]
        path = os.path.join(directory, f"scene_{si:02d}.luz")
        _write_if_changed(path, ("\n".join(lines) + "\n").encode())
        out.append(Workload(f"batch{si:02d}", directory, path, width, height, spp, depth, si + 1, instances * per,
                            {"instances": instances}))
    return out


# the 16 materials of the differentiable box: 5 walls, the 6 faces of the short box, the
# top and bottom of the tall box (one material) and its 4 sides
def diff_parameters(seed=1004, metals=True):
    """Per-material parameters of the differentiable box: (kinds, albedo, roughness, emission).
    kinds[i] in {"lambertian", "metal", "emissive"}; material 16 is the emitter."""
    rng = random.Random(seed)
    kinds, albedo, roughness = [], [], []
    for i in range(16):
        metal = metals and i in (6, 9, 12, 15)
        kinds.append("metal" if metal else "lambertian")
        albedo.append([rng.uniform(0.1, 0.9) for _ in range(3)])
        roughness.append(rng.uniform(0.05, 0.8))
    kinds.append("emissive")
    albedo.append([0.0, 0.0, 0.0])
    roughness.append(0.0)
    emission = [[0.0, 0.0, 0.0]] * 16 + [[17.0, 12.0, 4.0]]
    return kinds, albedo, roughness, emission


DIFF_ROUGH_METALS = ((3, 0.25), (2, 0.3), (5, 0.2), (12, 0.4))


def diff_rough_parameters():
    """The differentiable box with four rough GGX metals on large, visible surfaces (the
    left wall, back wall, short box top and a tall box side), the rest Lambertian:
    (kinds, albedo, roughness, emission) as diff_parameters returns them."""
    kinds, albedo, roughness, emission = diff_parameters(metals=False)
    kinds, roughness = list(kinds), list(roughness)
    for m, r in DIFF_ROUGH_METALS:
        kinds[m] = "metal"
        roughness[m] = r
    return kinds, albedo, roughness, emission


def diff_scene_text(kinds, albedo, roughness, emission, width=512, height=512, spp=16, depth=6) -> str:
    lines = ["[settings]", f"resolution={width},{height}", f"samples={spp}", f"maxlightbounces={depth}", *_HEADER,
             "sky=none", "background=(0,0,0)", "", "[materials]"]
    for i, k in enumerate(kinds):
        if k == "lambertian":
            lines += [f"material d{i} {{", "type=lambertian", f"color={_v(albedo[i])}", "}"]
        elif k == "metal":
            lines += [f"material d{i} {{", "type=metal", f"color={_v(albedo[i])}", f"roughness={_f(roughness[i])}", "}"]
        else:
            lum = sum(c * w for c, w in zip(emission[i], LUMINANCE_WEIGHTS))
            lines += [f"material d{i} {{", "type=emissive", f"color={_v(emission[i])}", f"radiance={_f(lum)}", "}"]
    lines += ["", "[scene]", "camera main {", "position=(278,273,-800)", "direction=(0,0,1)", "up=(0,1,0)",
              "focal_length_mm=35", "sensor_width_mm=25", "sensor_height_mm=25", "pinhole=1", "focus_distance=10", "}",
              "objects{"]
    for wi, (_, _, q) in enumerate(_CORNELL_WALLS):
        lines.append(f"triangle={_v(q[0])},{_v(q[1])},{_v(q[2])},material=d{wi}")
        lines.append(f"triangle={_v(q[0])},{_v(q[2])},{_v(q[3])},material=d{wi}")
    lines.append(f"rectangle=(278,554,279.5),(0,-1,0),130,105,material=d{len(kinds) - 1}")
    faces = _box_quads(_SHORT_BOX_TOP)
    for fi, q in enumerate(faces):
        lines.append(f"triangle={_v(q[0])},{_v(q[1])},{_v(q[2])},material=d{5 + fi}")
        lines.append(f"triangle={_v(q[0])},{_v(q[2])},{_v(q[3])},material=d{5 + fi}")
    faces = _box_quads(_TALL_BOX_TOP)
    for fi, q in enumerate(faces):
        mat = 11 + max(0, fi - 1)  # top and bottom share material 11, the sides take 12..15
        lines.append(f"triangle={_v(q[0])},{_v(q[1])},{_v(q[2])},material=d{mat}")
        lines.append(f"triangle={_v(q[0])},{_v(q[2])},{_v(q[3])},material=d{mat}")
    lines.append("}")
    return "\n".join(lines) + "\n"


def write_diff_cornell(directory: str, params=None, width=512, height=512, spp=16, depth=6, seed=1, metals=True,
                       name="diffcornell.luz") -> Workload:
    os.makedirs(directory, exist_ok=True)
    kinds, albedo, roughness, emission = params if params is not None else diff_parameters(metals=metals)
    path = os.path.join(directory, name)
    _write_if_changed(path, diff_scene_text(kinds, albedo, roughness, emission, width, height, spp, depth).encode())
    return Workload("diffcornell", directory, path, width, height, spp, depth, seed, 34,
                    {"kinds": kinds, "albedo": albedo, "roughness": roughness, "emission": emission})


def _write_if_changed(path: str, data: bytes) -> None:
    try:
        with open(path, "rb") as fh:
            if fh.read() == data:
                return
    except OSError:
        pass
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(data)
    os.replace(tmp, path)
