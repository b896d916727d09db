"""CPU-baseline calibration (test infrastructure; runs only where the reference sources are).

The CPU baseline bench.py reports is the reference renderer built faithfully
(oracle/_ref/luzref: -O3, no -ffast-math, no -march=native, no LTO), because that build is
bit-reproducible and portable to any x86-64 host. The reference's own Makefile builds with
-O3 -ffast-math -march=native -flto (its defaults, NATIVE=1 LTO=1). This script times both
builds of the same harness on this host with identical settings -- the Cornell box at
300x300, 50 spp, depth 8 (the plain CPU-runnable workload) and the displaced-sphere frame at
480x270, 16 spp, depth 12 (a bounded slice of the headline workload) -- two warm-ups, median
of five, each timed by the reference's own render clock, and writes the speed ratio
(release / faithful) to oracle/calibration.json, with the relative MSE between the two
builds' images. bench.py multiplies the faithful build's throughput on the GPU host by the
workload's ratio when it differs from 1 by more than 5 % and the release build renders the
same image (relative MSE <= 1e-4); a release build that renders a different image is not
doing the same work, and its speed is not a baseline.

    make -C oracle ref ref-release && python oracle/calibrate.py
"""

from __future__ import annotations

import json
import os
import platform
import statistics
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from luz_amd import workloads  # noqa: E402

BUILDS = {"faithful": os.path.join(HERE, "_ref", "luzref"), "release": os.path.join(HERE, "_ref", "release", "luzref")}


def render(exe, scene, seed, threads, spp, width, height):
    """(render ms by the reference's clock, FP64 image)."""
    import numpy as np
    out = os.path.join(tempfile.mkdtemp(prefix="luz_cal_"), "img.bin")
    subprocess.run([exe, "render", scene, str(seed), str(threads), out, str(spp), str(width), str(height)], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    with open(out, "rb") as fh:
        data = fh.read()
    w, h = struct.unpack("<II", data[:8])
    return struct.unpack("<d", data[8:16])[0], np.frombuffer(data[16:], dtype="<f8").reshape(h, w, 3)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    for name, exe in BUILDS.items():
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: make -C oracle ref ref-release")
    threads = len(os.sched_getaffinity(0))
    tmp = tempfile.mkdtemp(prefix="luz_cal_wl_")
    cases = [
        ("cornell", workloads.write_cornell(os.path.join(tmp, "cornell")), 50, 300, 300),
        ("spheres", workloads.spheres_scene(os.path.join(tmp, "spheres")), 16, 480, 270),
    ]
    result = {"host_cpu": cpu_model(), "threads": threads, "cases": {}}
    for case, wl, spp, w, h in cases:
        times, images = {}, {}
        for build, exe in BUILDS.items():
            for _ in range(2):
                _, images[build] = render(exe, wl.scene_path, 1, threads, spp, w, h)
            runs = [render(exe, wl.scene_path, 1, threads, spp, w, h)[0] for _ in range(5)]
            times[build] = statistics.median(runs)
            print(f"{case} {build}: median {times[build]:.1f} ms of {[round(r, 1) for r in runs]}", flush=True)
        a, b = images["faithful"], images["release"]
        same = float(((b - a) ** 2).sum() / max((a ** 2).sum(), 1e-300))
        result["cases"][case] = {"spp": spp, "width": w, "height": h, "faithful_ms": round(times["faithful"], 3),
                                 "release_ms": round(times["release"], 3),
                                 "release_speedup": round(times["faithful"] / times["release"], 4),
                                 "release_image_rel_mse": same, "release_image_mean": float(b.mean()),
                                 "faithful_image_mean": float(a.mean())}
    with open(os.path.join(HERE, "calibration.json"), "w") as fh:
        json.dump(result, fh, indent=1)
        fh.write("\n")
    print(json.dumps(result, indent=1))


if __name__ == "__main__":
    main()
