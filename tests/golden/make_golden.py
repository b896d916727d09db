"""Writes the committed golden fixtures from the reference renderer.

Runs only in the container that holds the reference sources: it needs
``oracle/_ref/luzref`` (built from /root/reference by ``make -C oracle ref``), which wraps
the reference's own Sampler, BVH builders, intersection code and renderer
(oracle/ref_harness.cpp). Output: small ``.npz`` / ``.txt`` files next to this script.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from luz_amd import workloads  # noqa: E402

LUZREF = os.path.join(ROOT, "oracle", "_ref", "luzref")


def run(*args):
    proc = subprocess.run([LUZREF, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    if proc.returncode != 0:
        raise RuntimeError(f"luzref {args[0]} failed: {proc.stderr.decode()[-2000:]}")
    return proc.stdout.decode(errors="replace")


def read_image(path):
    data = open(path, "rb").read()
    w, h = struct.unpack("<II", data[:8])
    ms = struct.unpack("<d", data[8:16])[0]
    return np.frombuffer(data[16:], dtype="<f8").reshape(h, w, 3), ms


def sampler_fixture(tmp):
    out = os.path.join(tmp, "sampler.bin")
    run("sampler", out)
    rec = np.dtype([("k", "<u4", 7), ("s1", "<f8"), ("s2", "<f4", 2)])
    arr = np.fromfile(out, dtype=rec)
    # deterministic subset: every 5th record plus the reference test's known case
    keep = np.zeros(len(arr), dtype=bool)
    keep[::5] = True
    k = arr["k"]
    known = (k[:, 0] == 1234) & (k[:, 1] == 13) & (k[:, 2] == 17) & (k[:, 3] == 5) & (k[:, 4] == 0) & (k[:, 5] == 0)
    keep |= known
    arr = arr[keep]
    np.savez_compressed(os.path.join(HERE, "sampler.npz"), keys=arr["k"], s1=arr["s1"], s2=arr["s2"])
    print("sampler records", len(arr))


def read_primary(path):
    data = open(path, "rb").read()
    w, h = struct.unpack("<II", data[:8])
    rec = np.dtype([("top", "<i8"), ("tri", "<i8"), ("t", "<f8"), ("margin", "<f8"), ("top2", "<i8"), ("tri2", "<i8"),
                    ("t2", "<f8"), ("tref", "<f8"), ("refhit", "u1")])
    return np.frombuffer(data[8:], dtype=rec).reshape(h, w)


def primary_fixture(name, scene, seed, tmp, w=None, h=None):
    out = os.path.join(tmp, f"{name}_primary.bin")
    args = ["primary", scene, seed, out] + ([w, h] if w else [])
    run(*args)
    p = read_primary(out)
    # the brute-force closest hit must agree with the reference's own traversal
    assert np.all((p["top"] >= 0) == (p["refhit"] == 1))
    hit = p["top"] >= 0
    assert np.allclose(p["t"][hit], p["tref"][hit], rtol=1e-12, atol=0), "brute force disagrees with reference traversal"
    # fp32 resolution: barycentric margin or a second candidate within 1e-5 relative
    gap = np.where(np.isfinite(p["t2"]), (p["t2"] - p["t"]) / np.maximum(p["t"], 1e-300), np.inf)
    near_tie = hit & ((p["margin"] < 1e-5) | (gap < 1e-5))
    np.savez_compressed(os.path.join(HERE, f"{name}_primary.npz"), top=p["top"].astype(np.int32),
                        tri=p["tri"].astype(np.int32), t=p["t"].astype(np.float32), near_tie=near_tie)
    print(name, "primary near ties", int(near_tie.sum()), "of", near_tie.size)


def small_scene_fixtures(name, wl, tmp):
    """Render, TLAS topology and primary hits of a small CPU-runnable workload."""
    out = os.path.join(tmp, f"{name}.bin")
    run("render", wl.scene_path, wl.seed, os.cpu_count() or 8, out)
    img, ms = read_image(out)
    np.savez_compressed(os.path.join(HERE, f"{name}_{wl.width}x{wl.height}_{wl.spp}spp_seed{wl.seed}.npz"),
                        image=img.astype(np.float32), render_ms=ms, threads=os.cpu_count() or 8)
    tl = os.path.join(tmp, f"{name}_tlas.txt")
    run("tlas", wl.scene_path, tl)
    with open(os.path.join(HERE, f"{name}_tlas.txt"), "w") as fh:
        fh.write(open(tl).read())
    primary_fixture(name, wl.scene_path, wl.seed, tmp)


ZSTAT_SEEDS = (1001, 64)  # reference seeds of the bias fixtures; device renders use 1..64


def zstat_configs(tmp):
    """(name, workload, spp, width, height) of the per-seed bias fixtures: every small scene
    family at low spp, so 64 reference renders take seconds."""
    return [
        ("cornell", workloads.write_cornell(os.path.join(tmp, "z_cornell")), 4, 300, 300),
        ("spheres", workloads.spheres_scene(os.path.join(tmp, "z_spheres")), 2, 192, 108),
        ("primitives", workloads.write_primitives(os.path.join(tmp, "z_primitives")), 4, 160, 120),
        ("materials", workloads.write_materials(os.path.join(tmp, "z_materials")), 4, 160, 120),
        ("volume", workloads.write_volume(os.path.join(tmp, "z_volume")), 4, 160, 120),
    ]


def zstat_fixtures(tmp):
    """Per-seed tile means of reference renders at 64 independent seeds (4 x 4 tiles, RGB)."""
    first, count = ZSTAT_SEEDS
    for name, wl, spp, w, h in zstat_configs(tmp):
        out = os.path.join(tmp, f"z_{name}.bin")
        run("seeds", wl.scene_path, first, count, os.cpu_count() or 8, out, spp, w, h)
        head = np.fromfile(out, dtype="<u4", count=2)
        means = np.fromfile(out, dtype="<f8", offset=8).reshape(int(head[0]), int(head[1]) ** 2, 3)
        np.savez_compressed(os.path.join(HERE, f"zstats_{name}.npz"), tile_means=means, first_seed=first, spp=spp,
                            width=w, height=h)
        print("zstats", name, means.shape, means.mean((0, 1)))


NFOR_CASES = [
    # (name, workload writer, spp, width, height, (prefilter, regression, patch) radii)
    ("nfor_cornell_48x48_8spp", "cornell", 8, 48, 48, (1, 2, 1)),
    ("nfor_primitives_64x48_8spp_r9", "primitives", 8, 64, 48, (1, 9, 1)),
    # a window too wide for the device's LDS-staged regression (its unstaged form)
    ("nfor_cornell_48x48_8spp_r12p2", "cornell", 8, 48, 48, (1, 12, 2)),
    # adaptive sampling on: (min samples, check interval, threshold)
    ("nfor_adaptive_cornell_48x48_48spp", "cornell", 48, 48, 48, (1, 2, 1), (8, 4, 0.2)),
]


def nfor_fixtures(tmp):
    """The reference's denoise pipeline: its NFOR input buffers (half-buffer means and
    variances, primary features) and Denoise::applyNFOR's output for them, plus the raw
    mean image, at seed 1 (oracle/ref_harness.cpp, denoise command); with adaptive
    sampling on (min samples, check interval, threshold given) also each pixel's sample
    count."""
    writers = {"cornell": workloads.write_cornell, "primitives": workloads.write_primitives}
    for name, scene, spp, w, h, radii, *rule in NFOR_CASES:
        wl = writers[scene](os.path.join(tmp, f"n_{scene}"))
        out = os.path.join(tmp, f"{name}.bin")
        run("denoise", wl.scene_path, 1, os.cpu_count() or 8, out, spp, w, h, *radii,
            *([rule[0][0], rule[0][1], repr(rule[0][2])] if rule else []))
        head = np.fromfile(out, dtype="<u4", count=2)
        assert head[0] == w and head[1] == h
        n = w * h
        nf = n * (3 + 3 + 3 + 1 + 11 + 11 + 11 + 3)
        data = np.fromfile(out, dtype="<f8", count=nf, offset=8)
        fields, off = {}, 0
        for k, c in (("image", 3), ("color_a", 3), ("color_b", 3), ("color_variance", 1), ("features_a", 11),
                     ("features_b", 11), ("feature_variance", 11), ("denoised", 3)):
            fields[k] = data[off:off + n * c].reshape(h, w, c)
            off += n * c
        extra = {}
        if rule:
            counts = np.fromfile(out, dtype="<u4", offset=8 + 8 * nf).reshape(h, w)
            extra = dict(counts=counts.astype(np.uint16), min_samples=rule[0][0], check_interval=rule[0][1],
                         threshold=rule[0][2])
        else:
            assert os.path.getsize(out) == 8 + 8 * nf
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), seed=1, spp=spp, scene=scene,
                            radii=np.array(radii), **fields, **extra)
        print(name, {k: float(np.abs(v).mean()) for k, v in fields.items()})


ADAPTIVE_CASES = [
    # (name, workload writer, spp, width, height, min samples, check interval, threshold)
    ("adaptive_primitives_64x48_64spp", "primitives", 64, 64, 48, 8, 4, 0.1),
    ("adaptive_cornell_64x48_96spp", "cornell", 96, 64, 48, 16, 8, 0.2),
    # dark pixels (outside the box) stop only from darkMinimumSamples = 256 on
    ("adaptive_cornell_dark_40x30_288spp", "cornell", 288, 40, 30, 16, 16, 0.2),
]


def adaptive_fixtures(tmp):
    """The reference's per-pixel loop with adaptive sampling on (denoising off) at seed 1:
    each pixel's sample count (_threadRender's return value) and the image
    (oracle/ref_harness.cpp, adaptive command)."""
    writers = {"cornell": workloads.write_cornell, "primitives": workloads.write_primitives}
    for name, scene, spp, w, h, mn, interval, threshold in ADAPTIVE_CASES:
        wl = writers[scene](os.path.join(tmp, f"a_{scene}"))
        out = os.path.join(tmp, f"{name}.bin")
        run("adaptive", wl.scene_path, 1, os.cpu_count() or 8, out, spp, w, h, mn, interval, repr(threshold))
        head = np.fromfile(out, dtype="<u4", count=2)
        assert head[0] == w and head[1] == h
        counts = np.fromfile(out, dtype="<u4", count=w * h, offset=8).reshape(h, w)
        image = np.fromfile(out, dtype="<f8", offset=8 + 4 * w * h).reshape(h, w, 3)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), seed=1, spp=spp, scene=scene, min_samples=mn,
                            check_interval=interval, threshold=threshold, counts=counts.astype(np.uint16), image=image)
        print(name, "counts", np.unique(counts, return_counts=True))


POST_CASES = [
    # (exposure, bloom, view transform, contrast)
    (0.0, 1, "agx", 1.0),          # the reference's defaults
    (1.0, 0, "aces", 1.0),
    (-0.5, 1, "standard", 1.3),    # the contrast branch (applyViewTransform + applyContrast + gammaCorrect)
    (0.25, 1, "raw", 1.0),         # bloom on a raw view
    (2.0, 0, "agx", 0.8),
    (1.5, 0, "raw", 1.0),          # exposure alone: max(x, 0) * 2^1.5
]


def post_fixture(tmp):
    """The reference's display post-processing (oracle/ref_harness.cpp, post command) of a
    64 x 64 crop of the Cornell reference render around the light, with isolated fireflies
    and a saturated patch planted, under six settings."""
    src = np.load(os.path.join(HERE, "cornell_300_50spp_seed1.npz"))["image"].astype(np.float64)
    img = np.ascontiguousarray(src[0:64, 118:182])
    h, w = img.shape[:2]
    img[40, 10] = (30.0, 30.0, 30.0)      # isolated firefly (bloom bright-pass and display)
    img[20, 50] = (5.0, 4.0, 3.0)         # bright outlier
    img[50, 30] = (20.0, 20.0, 20.0)      # saturated display pixel ...
    img[50, 31] = (20.0, 20.0, 20.0)      # ... with one saturated neighbour
    img[10, 10] = (-1.0, np.nan, np.inf)  # non-finite and negative input
    inp = os.path.join(tmp, "post_in.bin")
    with open(inp, "wb") as fh:
        fh.write(struct.pack("<II", w, h))
        fh.write(img.astype("<f8").tobytes())
    outs = []
    for exposure, bloom, view, contrast in POST_CASES:
        out = os.path.join(tmp, "post_out.bin")
        run("post", inp, out, repr(exposure), bloom, view, repr(contrast))
        data = np.fromfile(out, dtype="<f8", offset=8).reshape(h, w, 3)
        outs.append(data)
    np.savez_compressed(os.path.join(HERE, "post_cornell_64x64.npz"), image=img, outputs=np.stack(outs),
                        exposure=np.array([c[0] for c in POST_CASES]), bloom=np.array([c[1] for c in POST_CASES]),
                        view=np.array([c[2] for c in POST_CASES]), contrast=np.array([c[3] for c in POST_CASES]))
    print("post", [float(np.abs(o).mean()) for o in outs])


def fd_fixture(tmp, name="diff_fd_128_8spp_depth2", width=128, height=128, spp=8, depth=2, metals=False, h_albedo=0.01,
               h_emission=0.1, seed=1):
    """FP64 central finite differences of L = sum(w * image) on the reference renderer, with
    the same seed on both sides, for every Lambertian albedo channel, metal roughness and
    the emitter radiance of the differentiable box."""
    kinds, albedo, roughness, emission = workloads.diff_parameters(metals=metals)
    w = np.random.default_rng(7).uniform(-1.0, 1.0, (height, width, 3)).astype(np.float32)

    def loss(a, r, e, tag):
        d = os.path.join(tmp, "fd")
        wl = workloads.write_diff_cornell(d, (kinds, a, r, e), width, height, spp, depth, name=f"{tag}.luz")
        out = os.path.join(d, f"{tag}.bin")
        run("render", wl.scene_path, seed, os.cpu_count() or 8, out)
        img, _ = read_image(out)
        return float((img * w.astype(np.float64)).sum()), img

    base, img0 = loss(albedo, roughness, emission, "base")
    n = len(kinds)
    d_albedo = np.zeros((n, 3))
    d_rough = np.zeros(n)
    d_emission = np.zeros((n, 3))
    for m, k in enumerate(kinds):
        if k == "lambertian":
            for c in range(3):
                ap = [list(x) for x in albedo]
                am = [list(x) for x in albedo]
                ap[m][c] += h_albedo
                am[m][c] -= h_albedo
                d_albedo[m, c] = (loss(ap, roughness, emission, "p")[0] - loss(am, roughness, emission, "m")[0]) / (2 * h_albedo)
        elif k == "metal":
            rp, rm = list(roughness), list(roughness)
            rp[m] += 0.01
            rm[m] -= 0.01
            d_rough[m] = (loss(albedo, rp, emission, "p")[0] - loss(albedo, rm, emission, "m")[0]) / 0.02
        else:
            for c in range(3):
                ep = [list(x) for x in emission]
                em = [list(x) for x in emission]
                ep[m][c] += h_emission
                em[m][c] -= h_emission
                d_emission[m, c] = (loss(albedo, roughness, ep, "p")[0] - loss(albedo, roughness, em, "m")[0]) / (2 * h_emission)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), weights=w, d_albedo=d_albedo, d_roughness=d_rough,
                        d_emission=d_emission, loss=base, image=img0.astype(np.float32), seed=seed, spp=spp, depth=depth,
                        width=width, height=height, metals=metals)
    print(name, "loss", base, "|d_albedo|", np.abs(d_albedo).sum(), "d_emission", d_emission[-1])


FD_ROUGH = dict(width=64, height=64, spp=8, depth=6, h=0.02, seeds=64)


def fd_rough_fixture(tmp):
    """Roughness gradients of the reference renderer at depth 6 (Russian roulette, mirror
    fallbacks, MIS): FP64 central differences (h = 0.02) of L_j = sum(w_j * image) for each
    metal j and each of 64 seeds, both sides of each difference at the same seed. A
    fixed-seed estimator jumps where a sampled direction crosses an edge or a discrete
    decision flips, so a single seed's difference is noisy; the per-seed values are kept so a
    test can form the mean over seeds and its standard error. Metal j's weight image w_j is
    the normalised change of a 256-spp render when that metal's roughness is scaled by 0.7,
    so each loss looks at that metal's highlight."""
    c = FD_ROUGH
    kinds, albedo, roughness, emission = workloads.diff_rough_parameters()
    metals = [m for m, _ in workloads.DIFF_ROUGH_METALS]
    d = os.path.join(tmp, "fdr")

    def render(r, seed, spp, tag):
        wl = workloads.write_diff_cornell(d, (kinds, albedo, r, emission), c["width"], c["height"], spp, c["depth"],
                                          name=f"{tag}.luz")
        out = os.path.join(d, f"{tag}.bin")
        run("render", wl.scene_path, seed, os.cpu_count() or 8, out)
        return read_image(out)[0]

    base = render(roughness, 999, 256, "base")
    w = []
    for m in metals:
        r = list(roughness)
        r[m] *= 0.7
        delta = render(r, 999, 256, "sharp") - base
        w.append(delta / np.abs(delta).max())
    w = np.stack(w)
    fd = np.zeros((c["seeds"], len(metals)))
    for si in range(c["seeds"]):
        for j, m in enumerate(metals):
            rp, rm = list(roughness), list(roughness)
            rp[m] += c["h"]
            rm[m] -= c["h"]
            fd[si, j] = ((render(rp, si + 1, c["spp"], "p") - render(rm, si + 1, c["spp"], "m")) * w[j]).sum() / (2 * c["h"])
    np.savez_compressed(os.path.join(HERE, "diff_fd_rough_64_8spp_depth6.npz"), weights=w.astype(np.float32), fd=fd,
                        seeds=np.arange(1, c["seeds"] + 1), metals=np.array(metals), h=c["h"], spp=c["spp"],
                        depth=c["depth"], width=c["width"], height=c["height"])
    sem = fd.std(0, ddof=1) / np.sqrt(c["seeds"])
    print("fd_rough mean", fd.mean(0), "sem", sem)


TERRAIN_GOLDEN = dict(width=160, height=90, spp=8)
TERRAIN_NFOR = dict(width=96, height=64, spp=8, radii=(1, 9, 1))
BATCH_GOLDEN = dict(width=96, height=64, spp=8)


def batch_golden_indices(wls):
    """The batch scenes the reduced fixture renders: the one with the most triangles and
    three more spread over the triangle range (the smallest, and two between)."""
    order = sorted(range(len(wls)), key=lambda i: wls[i].triangles)
    return sorted({order[-1], order[0], order[len(order) // 3], order[2 * len(order) // 3]})


def terrain_fixtures(tmp):
    """The volume workload's own scene (1,002,528-triangle terrain in Henyey-Greenstein haze
    under a directional sun) at a reduced size the reference renders in seconds: its render
    and primary hits, and its denoise pipeline with a radius-9 regression window (NFOR
    inputs and the reference's applyNFOR output)."""
    wl = workloads.terrain_scene(os.path.join(tmp, "terrain"))
    c = TERRAIN_GOLDEN
    out = os.path.join(tmp, "terrain.bin")
    run("render", wl.scene_path, 1, os.cpu_count() or 8, out, c["spp"], c["width"], c["height"])
    img, ms = read_image(out)
    np.savez_compressed(os.path.join(HERE, f"terrain_{c['width']}x{c['height']}_{c['spp']}spp_seed1.npz"),
                        image=img.astype(np.float32), render_ms=ms)
    primary_fixture("terrain", wl.scene_path, 1, tmp, c["width"], c["height"])
    n = TERRAIN_NFOR
    name = f"nfor_terrain_{n['width']}x{n['height']}_{n['spp']}spp_r9"
    out = os.path.join(tmp, f"{name}.bin")
    run("denoise", wl.scene_path, 1, os.cpu_count() or 8, out, n["spp"], n["width"], n["height"], *n["radii"])
    w, h = n["width"], n["height"]
    px = w * h
    data = np.fromfile(out, dtype="<f8", count=px * 48, offset=8)
    fields, off = {}, 0
    for k, ch in (("image", 3), ("color_a", 3), ("color_b", 3), ("color_variance", 1), ("features_a", 11),
                  ("features_b", 11), ("feature_variance", 11), ("denoised", 3)):
        fields[k] = data[off:off + px * ch].reshape(h, w, ch)
        off += px * ch
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), seed=1, spp=n["spp"], scene="terrain",
                        radii=np.array(n["radii"]), **fields)
    print("terrain fixtures", float(img.mean()), {k: float(np.abs(v).mean()) for k, v in fields.items()})


def batch_fixtures(tmp):
    """Four of the uneven batch's scenes (its largest, 500k triangles, and three spread below
    it), each rendered by the reference at a reduced size with the seed the batch gives it."""
    wls = workloads.batch_scenes(os.path.join(tmp, "batch"))
    c = BATCH_GOLDEN
    images = []
    picks = batch_golden_indices(wls)
    for i in picks:
        out = os.path.join(tmp, f"batch_{i}.bin")
        run("render", wls[i].scene_path, wls[i].seed, os.cpu_count() or 8, out, c["spp"], c["width"], c["height"])
        images.append(read_image(out)[0].astype(np.float32))
    np.savez_compressed(os.path.join(HERE, f"batch_{c['width']}x{c['height']}_{c['spp']}spp.npz"), scenes=np.array(picks),
                        seeds=np.array([wls[i].seed for i in picks]), triangles=np.array([wls[i].triangles for i in picks]),
                        images=np.stack(images))
    print("batch fixtures", picks, [wls[i].triangles for i in picks])


REF_EXAMPLES = "/root/reference/examples/scenes"
EXAMPLE_DIR = os.path.join(HERE, "scenes")
# the reference's example scenes rendered for the fixtures: name -> (width, height, spp)
EXAMPLE_GOLDEN = {
    "cornell": (128, 128, 64),
    "suzanne": (150, 100, 32),
    "rtiow": (160, 90, 32),
    "volumetric_spotlight_fog": (160, 90, 64),
    "volumetric_gray_sphere_white_bg": (96, 96, 64),
    "volumetric_godrays": (160, 90, 32),
}


def example_scene_path(name):
    width, height, spp = EXAMPLE_GOLDEN[name]
    return os.path.join(EXAMPLE_DIR, f"{name}_{width}x{height}_{spp}spp.luz")


def reduced_example(name):
    """Writes the reference's example scene ``name`` (examples/scenes) as a fixture input: the
    file as it stands, with its resolution and sample count cut to the fixture's, adaptive
    sampling and denoising off (the reference harness renders with denoising off; on, both
    it and the device draw each sample from one of two half-buffer streams) and mesh files
    pointed at copies under scenes/assets."""
    width, height, spp = EXAMPLE_GOLDEN[name]
    src = os.path.join(REF_EXAMPLES, f"{name}.luz")
    os.makedirs(os.path.join(EXAMPLE_DIR, "assets"), exist_ok=True)
    lines = []
    for line in open(src).read().splitlines():
        key, _, value = line.partition("=")
        key = key.strip().lower()
        if key == "[settings]":
            lines += [line, "adaptive=0", "denoise=0"]
            continue
        if key in ("adaptive", "denoise"):
            continue
        if key == "resolution":
            line = f"resolution={width},{height}"
        elif key == "samples":
            line = f"samples={spp}"
        elif key in ("file", "path") and value.strip().lower().endswith(".obj"):
            asset = os.path.normpath(os.path.join(REF_EXAMPLES, value.strip()))
            dst = os.path.join(EXAMPLE_DIR, "assets", os.path.basename(asset))
            with open(asset, "rb") as fin, open(dst, "wb") as fout:
                fout.write(fin.read())
            line = f"{key}=assets/{os.path.basename(asset)}"
        lines.append(line)
    path = example_scene_path(name)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def example_fixtures(tmp, names=None):
    """The reference's example scenes at reduced sizes: the reduced scene files (inputs),
    the reference's FP64 render of each at seed 1, and its primary hits."""
    for name in names or EXAMPLE_GOLDEN:
        width, height, spp = EXAMPLE_GOLDEN[name]
        path = reduced_example(name)
        out = os.path.join(tmp, f"example_{name}.bin")
        run("render", path, 1, os.cpu_count() or 8, out)
        img, ms = read_image(out)
        assert img.shape == (height, width, 3)
        np.savez_compressed(os.path.join(HERE, f"example_{name}_{width}x{height}_{spp}spp_seed1.npz"),
                            image=img.astype(np.float32), render_ms=ms)
        primary_fixture(f"example_{name}", path, 1, tmp)
        print("example", name, float(img.mean()), f"{ms:.0f} ms")


def main():
    if not os.path.exists(LUZREF):
        sys.exit("build the reference harness first: make -C oracle ref")
    tmp = tempfile.mkdtemp(prefix="luz_golden_")
    only = set(sys.argv[1:])
    if only:
        # e.g. `make_golden.py primitives volume`: regenerate just those fixtures
        if "primitives" in only:
            small_scene_fixtures("primitives", workloads.write_primitives(os.path.join(tmp, "primitives")), tmp)
        if "volume" in only:
            small_scene_fixtures("volume", workloads.write_volume(os.path.join(tmp, "volume")), tmp)
        if "materials" in only:
            small_scene_fixtures("materials", workloads.write_materials(os.path.join(tmp, "materials")), tmp)
        if "skylinear" in only:
            small_scene_fixtures("skylinear", workloads.write_skylinear(os.path.join(tmp, "skylinear")), tmp)
        if "atmosphere" in only:
            small_scene_fixtures("atmosphere", workloads.write_atmosphere(os.path.join(tmp, "atmosphere")), tmp)
        if "spectral" in only:
            small_scene_fixtures("spectral", workloads.write_spectral(os.path.join(tmp, "spectral")), tmp)
        if "atmosphere_env" in only:
            small_scene_fixtures("atmosphere_env", workloads.write_atmosphere_env(os.path.join(tmp, "atmosphere_env")), tmp)
        if "optics" in only:
            small_scene_fixtures("optics", workloads.write_optics(os.path.join(tmp, "optics")), tmp)
        if "subsurface" in only:
            small_scene_fixtures("subsurface", workloads.write_subsurface(os.path.join(tmp, "subsurface")), tmp)
        if "sunlit" in only:
            small_scene_fixtures("sunlit", workloads.write_sunlit(os.path.join(tmp, "sunlit")), tmp)
        if "ies" in only:
            small_scene_fixtures("ies", workloads.write_ies(os.path.join(tmp, "ies")), tmp)
        if "invisible" in only:
            small_scene_fixtures("invisible", workloads.write_invisible(os.path.join(tmp, "invisible")), tmp)
        if "meshlight" in only:
            small_scene_fixtures("meshlight", workloads.write_meshlight(os.path.join(tmp, "meshlight")), tmp)
        if "fd" in only:
            fd_fixture(tmp)
        if "fd_rough" in only:
            fd_rough_fixture(tmp)
        if "terrain" in only:
            terrain_fixtures(tmp)
        if "batch" in only:
            batch_fixtures(tmp)
        if "zstats" in only:
            zstat_fixtures(tmp)
        if "nfor" in only:
            nfor_fixtures(tmp)
        if "adaptive" in only:
            adaptive_fixtures(tmp)
        if "post" in only:
            post_fixture(tmp)
        if "examples" in only:
            example_fixtures(tmp)
        for name in EXAMPLE_GOLDEN:
            if name in only:
                example_fixtures(tmp, [name])
        return
    sampler_fixture(tmp)
    small_scene_fixtures("primitives", workloads.write_primitives(os.path.join(tmp, "primitives")), tmp)
    small_scene_fixtures("volume", workloads.write_volume(os.path.join(tmp, "volume")), tmp)
    small_scene_fixtures("materials", workloads.write_materials(os.path.join(tmp, "materials")), tmp)
    small_scene_fixtures("meshlight", workloads.write_meshlight(os.path.join(tmp, "meshlight")), tmp)
    small_scene_fixtures("skylinear", workloads.write_skylinear(os.path.join(tmp, "skylinear")), tmp)
    small_scene_fixtures("atmosphere", workloads.write_atmosphere(os.path.join(tmp, "atmosphere")), tmp)
    small_scene_fixtures("spectral", workloads.write_spectral(os.path.join(tmp, "spectral")), tmp)
    small_scene_fixtures("atmosphere_env", workloads.write_atmosphere_env(os.path.join(tmp, "atmosphere_env")), tmp)
    small_scene_fixtures("optics", workloads.write_optics(os.path.join(tmp, "optics")), tmp)
    small_scene_fixtures("subsurface", workloads.write_subsurface(os.path.join(tmp, "subsurface")), tmp)
    small_scene_fixtures("invisible", workloads.write_invisible(os.path.join(tmp, "invisible")), tmp)
    small_scene_fixtures("ies", workloads.write_ies(os.path.join(tmp, "ies")), tmp)
    small_scene_fixtures("sunlit", workloads.write_sunlit(os.path.join(tmp, "sunlit")), tmp)
    fd_fixture(tmp)
    fd_rough_fixture(tmp)
    terrain_fixtures(tmp)
    batch_fixtures(tmp)
    zstat_fixtures(tmp)
    nfor_fixtures(tmp)
    adaptive_fixtures(tmp)
    example_fixtures(tmp)

    # Cornell box, 300x300, 50 spp, depth 8, seed 1 (the plain CPU-runnable config)
    c = workloads.write_cornell(os.path.join(tmp, "cornell"))
    out = os.path.join(tmp, "cornell.bin")
    run("render", c.scene_path, 1, os.cpu_count() or 8, out)
    img, ms = read_image(out)
    np.savez_compressed(os.path.join(HERE, "cornell_300_50spp_seed1.npz"), image=img.astype(np.float32),
                        render_ms=ms, threads=os.cpu_count() or 8)
    with open(os.path.join(HERE, "cornell_tlas.txt"), "w") as fh:
        tl = os.path.join(tmp, "c_tlas.txt")
        run("tlas", c.scene_path, tl)
        fh.write(open(tl).read())
    primary_fixture("cornell", c.scene_path, 1, tmp)
    post_fixture(tmp)

    # 512 displaced spheres: topology hashes, a reduced-resolution render, primary hits
    s = workloads.spheres_scene(os.path.join(tmp, "spheres"))
    tl = os.path.join(tmp, "s_tlas.txt")
    run("tlas", s.scene_path, tl)
    with open(os.path.join(HERE, "spheres_tlas.txt"), "w") as fh:
        fh.write(open(tl).read())
    digests = []
    for i in range(s.extra["instances"]):
        b = os.path.join(tmp, f"blas_{i}.bin")
        run("blas", os.path.join(s.directory, f"sphere_{i:03d}.obj"), b)
        raw = open(b, "rb").read()
        n_nodes, n_tris = struct.unpack("<QQ", raw[:16])
        # topology part only: node boxes + links + triangle order (vertices excluded)
        topo = raw[16:16 + n_nodes * 88 + n_tris * 8]
        digests.append(hashlib.sha256(topo).hexdigest())
    with open(os.path.join(HERE, "spheres_blas_sha256.txt"), "w") as fh:
        fh.write("\n".join(digests) + "\n")
    out = os.path.join(tmp, "spheres_small.bin")
    run("render", s.scene_path, 1, os.cpu_count() or 8, out, 16, 192, 108)
    img, ms = read_image(out)
    np.savez_compressed(os.path.join(HERE, "spheres_192x108_16spp_seed1.npz"), image=img.astype(np.float32),
                        render_ms=ms, threads=os.cpu_count() or 8)
    primary_fixture("spheres", s.scene_path, 1, tmp, 240, 135)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
