"""Throughput benchmark of the radiance path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload spheres|cornell|batch|terrain|diff]
                    [--integrator megakernel|wavefront] [--no-cpu-baseline] [--no-extras]

A step is one full render of the workload at its configured resolution and spp
(Renderer::render worker loop, Threads.cpp:375-548), inputs resident in HBM and the image
left in HBM. Default workload: 512 displaced spheres of 3840 triangles (1,966,080
triangles), 40/30/30 Lambertian / GGX metal / dielectric, procedural HDR sky, 1920x1080,
64 spp, depth 12 (synthetic, luz_amd/workloads.py). Metric: Msamples/s =
W*H*spp*steps / seconds (docker/benchmark/benchmark.sh:150-151), with Mrays/s beside it.

The CPU baseline is the reference renderer itself (oracle/_ref, built from the reference
sources by oracle/Makefile) on every CPU this process may use, for a bounded sample of the
same workload (16 spp at full resolution), timed by its own render clock; the host's CPU
model and core counts are reported with it, and oracle/calibration.json says how the
reference's release flags compare. The same 16 samples rendered on the device are checked
against that reference image (`quality`: relative MSE, bar 1e-4). The default line also
carries the differentiable box's forward + adjoint iteration (`diff`) and the uneven
32-scene batch with its small-scene regime (`batch`), measured after the timed steps.
"""

from __future__ import annotations

import argparse
import json
import os
import struct
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU model and its physical / logical core counts (lscpu), and the threads
    this process may use (its CPU affinity)."""
    info = {"model": None, "sockets": None, "physical_cores": None, "logical_cpus": os.cpu_count(),
            "threads_available": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=30).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cores = int(kv.get("Core(s) per socket", "0") or 0)
        info["sockets"] = sockets or None
        info["physical_cores"] = sockets * cores or None
        info["logical_cpus"] = int(kv.get("CPU(s)", info["logical_cpus"]) or info["logical_cpus"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def calibration(workload: str):
    """Release-build speed ratio of the reference on its calibration host
    (oracle/calibration.json, written by oracle/calibrate.py) for this workload, or None."""
    path = os.path.join(ROOT, "oracle", "calibration.json")
    if not os.path.exists(path):
        return None
    try:
        cal = json.load(open(path))
        case = cal["cases"][workload]
        return {"release_speedup": case["release_speedup"], "host_cpu": cal.get("host_cpu"), "threads": cal.get("threads"),
                "faithful_ms": case["faithful_ms"], "release_ms": case["release_ms"],
                "release_image_rel_mse": case["release_image_rel_mse"],
                "case": f"{case['width']}x{case['height']}x{case['spp']}spp"}
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(scene_path: str, seed: int, spp: int, workload: str):
    """The reference's own Renderer::render (oracle/_ref/luzref, built from its sources) on
    every CPU this process may use, for `spp` of the workload's samples at full resolution,
    timed by its render clock. Returns (baseline dict, the reference's FP64 image)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "luzref")
    if not os.path.exists(exe):
        return None, None
    host = host_cpu()
    threads = host["threads_available"]
    out = os.path.join(tempfile.mkdtemp(prefix="luz_cpu_"), "img.bin")
    t0 = time.time()
    proc = subprocess.run([exe, "render", scene_path, str(seed), str(threads), out, str(spp)],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=900)
    if proc.returncode != 0:
        log("cpu baseline failed:", proc.stdout.decode(errors="replace")[-500:])
        return None, None
    import numpy as np
    data = open(out, "rb").read()
    w, h = struct.unpack("<II", data[:8])
    render_ms = struct.unpack("<d", data[8:16])[0]
    image = np.frombuffer(data[16:], dtype="<f8").reshape(h, w, 3)
    measured = w * h * spp / (render_ms / 1000.0) / 1e6
    log(f"cpu baseline: {w}x{h}x{spp} in {render_ms:.0f} ms render on {threads} threads "
        f"({time.time() - t0:.1f} s wall incl. load): {measured:.3f} Msamples/s; host {host}")
    base = {
        "value": round(measured, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"reference Renderer::render of the same scene, {w}x{h} at {spp} spp (of the workload's full spp), "
                  f"--threads {threads} (every CPU of this process), seed {seed}; faithful build (-O3, no fast-math / "
                  f"native / LTO)",
        "host": host,
    }
    cal = calibration(workload)
    if cal:
        base["calibration"] = cal
        if cal["release_image_rel_mse"] > 1e-4:
            # -ffast-math changes what the release build renders for this workload: its time
            # is not the time of the same work, so the faithful build's figure stands
            base["sample"] += (f"; the reference's release build (-ffast-math -march=native -flto) renders this workload "
                               f"differently (image rel-MSE {cal['release_image_rel_mse']:.3g} against the faithful build, "
                               f"oracle/calibration.json), so its speed is not used")
        elif abs(cal["release_speedup"] - 1.0) > 0.05:
            # the reference's release flags (-ffast-math -march=native -flto) were this much
            # faster on the calibration host: report the baseline scaled to them
            base["faithful_build_value"] = base["value"]
            base["value"] = round(measured * cal["release_speedup"], 4)
            base["sample"] += (f"; value scaled by the release build's measured speedup x{cal['release_speedup']:.3f} "
                               f"({cal['case']} on {cal['threads']} threads of {cal['host_cpu']}, oracle/calibration.json)")
    return base, image


def cpu_nfor_baseline(scene_path: str, threads: int, width: int, height: int, full_pixels: int):
    """The reference's own Denoise::applyNFOR (regression radius 9) on the host cores, on a
    bounded crop of the same scene (2 spp half-buffers from the reference's denoising
    render, untimed): its time per pixel, scaled to the full frame (the regression's cost is
    linear in the pixel count)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "luzref")
    if not os.path.exists(exe):
        return None
    out = os.path.join(tempfile.mkdtemp(prefix="luz_nfor_"), "nfor.bin")
    proc = subprocess.run([exe, "denoise", scene_path, "1", str(threads), out, "2", str(width), str(height), "1", "9", "1"],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=900)
    text = proc.stdout.decode(errors="replace")
    if proc.returncode != 0 or "nfor_ms=" not in text:
        log("cpu NFOR baseline failed:", text[-500:])
        return None
    ms = float(text.split("nfor_ms=")[1].split()[0])
    frame_ms = ms * full_pixels / (width * height)
    log(f"cpu NFOR baseline: {width}x{height} in {ms:.0f} ms on {threads} threads -> {frame_ms:.0f} ms per full frame")
    return {"nfor_ms_per_frame": round(frame_ms, 1), "cores": threads, "kind": "reference",
            "sample": f"reference Denoise::applyNFOR, regression radius 9, on a {width}x{height} render of the same scene "
                      f"(2 spp), --threads {threads}, scaled by pixel count to the full frame"}


def load_traffic(workload: str, kernel: str):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path))
    except Exception:
        return None


def bench_batch(args, wdir, steps=None, warmup=None):
    """32 uneven scenes in one multi-scene launch sequence against the same scenes rendered
    one at a time by the plain single-scene megakernel (and by the single-scene wavefront).
    Returns the result line."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    import torch
    from luz_amd import Context, Renderer, SceneFile, workloads
    wls = workloads.batch_scenes(os.path.join(wdir, "batch"))
    ctx = Context(0)
    scenes = [SceneFile.read(ctx, w.scene_path) for w in wls]
    seeds = [args.seed + i for i in range(len(wls))]
    outs = [torch.empty(w.width * w.height * 3, dtype=torch.float32, device="cuda") for w in wls]
    total = sum(w.width * w.height * w.spp for w in wls)
    log(f"batch: {len(wls)} scenes, {sum(w.triangles for w in wls)} triangles, {total / 1e6:.1f} M samples per step")
    tuning = args.explicit_tuning
    if args.tuning == "auto":
        tuned = Renderer.autotune(scenes, seeds)
        tuning = "auto"
        log(f"autotune: {tuned.as_dict()} in {tuned.tune_ms:.0f} ms over {tuned.trials} shapes "
            f"(calibration pass {tuned.default_ms:.2f} -> {tuned.tuned_ms:.2f} ms)")
    for i in range(warmup):
        _, st = Renderer.render_batch(scenes, seeds, outs=outs, tuning=tuning)
        log(f"warmup {i}: {st.render_ms:.1f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        _, st = Renderer.render_batch(scenes, seeds, outs=outs, tuning=tuning)
        log(f"step {i}: {st.render_ms:.1f} ms, {st.kernel_launches} passes")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = total * steps / elapsed / 1e6
    # the same scenes one by one (untimed for the headline; each timed by its own events)
    plain_ms = wave_ms = 0.0
    for w, sc, sd, o in zip(wls, scenes, seeds, outs):
        _, a = Renderer.render(sc, seed=sd, integrator="megakernel", out=o)
        _, b = Renderer.render(sc, seed=sd, integrator="wavefront", out=o)
        plain_ms += a.render_ms
        wave_ms += b.render_ms
    batch_ms = elapsed * 1000.0 / steps
    log(f"one at a time: megakernel {plain_ms:.0f} ms, wavefront {wave_ms:.0f} ms; batched {batch_ms:.0f} ms "
        f"({plain_ms / batch_ms:.2f}x / {wave_ms / batch_ms:.2f}x)")
    # The small-scene regime: the same 32 scenes at sizes where one scene cannot fill the GPU.
    # Device time (HIP events) of the plain kernel per scene, summed, against one batched
    # render on the multi-scene wavefront and on the batched plain kernel (every scene's paths
    # in one launch): launch gaps and host syncs are left out of all three.
    small = []
    for (sw, sh), sspp in (((128, 128), 4), ((64, 64), 1), ((32, 32), 1)):
        n = len(wls)
        kw = dict(samples=[sspp] * n, resolutions=[(sw, sh)] * n)
        batched = {}
        for integ in ("wavefront", "megakernel"):
            for _ in range(2):
                Renderer.render_batch(scenes, seeds, integrator=integ, **kw)
            batched[integ] = min(Renderer.render_batch(scenes, seeds, integrator=integ, **kw)[1].render_ms for _ in range(3))
        plain = 0.0
        for sc, sd in zip(scenes, seeds):
            Renderer.render(sc, seed=sd, samples=sspp, resolution=(sw, sh), integrator="megakernel")
            plain += min(Renderer.render(sc, seed=sd, samples=sspp, resolution=(sw, sh),
                                         integrator="megakernel")[1].render_ms for _ in range(3))
        log(f"small scenes ({n} x {sw}x{sh} x {sspp} spp): plain kernel one at a time {plain:.3f} ms of device time; "
            f"batched wavefront {batched['wavefront']:.3f} ms ({plain / batched['wavefront']:.2f}x), "
            f"batched plain kernel {batched['megakernel']:.3f} ms ({plain / batched['megakernel']:.2f}x)")
        small.append({"scenes": n, "resolution": f"{sw}x{sh}", "spp": sspp, "plain_megakernel_device_ms": round(plain, 3),
                      "batched_wavefront_device_ms": round(batched["wavefront"], 3),
                      "batched_plain_kernel_device_ms": round(batched["megakernel"], 3),
                      "batched_wavefront_vs_plain": round(plain / batched["wavefront"], 3),
                      "batched_plain_kernel_vs_plain": round(plain / batched["megakernel"], 3)})
    line = {
        "metric": "Msamples/s", "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(batch_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch: {len(wls)} scenes of displaced sphere meshes, triangles log-uniform [1e4, 5e5], "
                               f"resolutions from 256^2 to 1024^2, spp 16..128, depth 10",
                   "scenes": len(wls), "samples_per_step": total, "integrator": "wavefront, multi-scene, hipGraph passes"},
        "batched_vs_plain_megakernel": round(plain_ms / batch_ms, 3),
        "batched_vs_single_scene_wavefront": round(wave_ms / batch_ms, 3),
        "small_scenes": small,
        "roofline": None, "cpu_baseline": None,
    }
    for sc in scenes:
        sc.release()
    ctx.close()
    return line


def bench_diff(args, wdir, steps=None, warmup=None):
    """Differentiable box, 512x512, 16 materials, 16 spp, depth 6: one iteration = forward
    render + L2 loss against a target rendered at perturbed parameters + path-replay adjoint
    (through luz_amd.diff, a torch.autograd.Function). Returns the result line: the
    iteration time and its ratio to the forward alone."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    import numpy as np
    import torch
    from luz_amd import Context, SceneFile, diff, workloads
    kinds, albedo, roughness, emission = workloads.diff_parameters(metals=True)
    wl = workloads.write_diff_cornell(os.path.join(wdir, "diff"), (kinds, albedo, roughness, emission), 512, 512, 16, 6)
    ctx = Context(0)
    scene = SceneFile.read(ctx, wl.scene_path)
    rng = np.random.default_rng(args.seed)
    ta = np.clip(np.array(albedo) * rng.uniform(0.8, 1.2, (len(kinds), 3)), 0.02, 0.98)
    tr = np.clip(np.array(roughness) * rng.uniform(0.8, 1.2, len(kinds)), 0.02, 1.0)
    dev = torch.device("cuda")
    with torch.no_grad():
        target = diff.render(scene, torch.tensor(ta), torch.tensor(tr), torch.tensor(emission), seed=args.seed + 1000)
    a = torch.tensor(albedo, dtype=torch.float64, requires_grad=True)
    r = torch.tensor(roughness, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(emission, dtype=torch.float64, requires_grad=True)

    def forward():
        with torch.no_grad():
            return diff.render(scene, a, r, e, seed=args.seed)

    def iteration():
        a.grad = r.grad = e.grad = None
        img = diff.render(scene, a, r, e, seed=args.seed)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        return loss

    def timed(fn, label):
        for i in range(warmup):
            fn()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for i in range(steps):
            fn()
        torch.cuda.synchronize(dev)
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        log(f"{label}: {ms:.2f} ms")
        return ms

    fwd_ms = timed(forward, "forward")
    it_ms = timed(iteration, "forward + adjoint")
    loss = iteration()
    log(f"loss {loss.item():.6g}; |d albedo| {a.grad.abs().sum().item():.4g}, |d roughness| {r.grad.abs().sum().item():.4g}")
    samples = 512 * 512 * 16
    line = {
        "metric": "fwd+adjoint ms/iter", "value": round(it_ms, 3), "unit": "ms", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(it_ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "differentiable box: 512x512, 16 materials (12 Lambertian, 4 GGX metals) + 1 emitter, "
                               "16 spp, depth 6, L2 loss against a target at perturbed parameters",
                   "samples_per_iteration": samples},
        "forward_ms": round(fwd_ms, 3), "iteration_over_forward": round(it_ms / fwd_ms, 3),
        "forward_msamples_per_s": round(samples / fwd_ms / 1e3, 3),
        "roofline": None, "cpu_baseline": None,
    }
    scene.release()
    ctx.close()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spheres", choices=["spheres", "cornell", "batch", "terrain", "diff"])
    ap.add_argument("--integrator", default="wavefront", choices=["megakernel", "wavefront"])
    ap.add_argument("--tuning", default="auto",
                    help="wavefront launch shape: auto (untimed calibration first), default (built-in), or an "
                         "explicit shape grid=G,extend_refill=E,shadow_refill=S,fetch_chunk=C,lds_stack=L "
                         "(as the autotuner reports it, so a profile can run the timed shape alone)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=16)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the forward+adjoint and batch figures attached to the default line")
    ap.add_argument("--probe-copy", action="store_true", help="run one plain 1 GiB copy kernel (PMC scale probe)")
    args = ap.parse_args()
    args.explicit_tuning = None
    if args.tuning not in ("auto", "default"):
        keys = {"grid", "extend_refill", "shadow_refill", "fetch_chunk", "lds_stack"}
        try:
            shape = {k: int(v) for k, v in (item.split("=") for item in args.tuning.split(","))}
        except ValueError:
            ap.error(f"--tuning {args.tuning!r}: expected auto, default or key=value pairs")
        if not set(shape) <= keys:
            ap.error(f"--tuning keys must be among {sorted(keys)}")
        args.explicit_tuning = shape

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from luz_amd import Context, Renderer, SceneFile, workloads

    wdir = os.path.join(tempfile.gettempdir(), "luz_bench_workloads")
    t0 = time.time()
    if args.workload == "batch":
        print(json.dumps(bench_batch(args, wdir)), flush=True)
        return
    if args.workload == "diff":
        print(json.dumps(bench_diff(args, wdir)), flush=True)
        return
    if args.workload == "terrain":
        wl = workloads.terrain_scene(os.path.join(wdir, "terrain"))
    elif args.workload == "spheres":
        wl = workloads.spheres_scene(os.path.join(wdir, "spheres"))
    else:
        wl = workloads.write_cornell(os.path.join(wdir, "cornell"))
    log(f"workload {wl.name}: {wl.width}x{wl.height} {wl.spp} spp depth {wl.depth}, {wl.triangles} triangles "
        f"(generated in {time.time() - t0:.1f} s)")

    if args.probe_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB device-to-device copy done")
        return

    ctx = Context(0)
    t0 = time.time()
    scene = SceneFile.read(ctx, wl.scene_path)
    log(f"scene upload {time.time() - t0:.1f} s: {scene.info}")
    pixels = scene.width * scene.height
    spp = scene.info["samples"]
    out = torch.empty(pixels * 3, dtype=torch.float32, device="cuda")

    wavefront = args.integrator == "wavefront"
    # visit counts for the roofline (untimed, a 4-spp slice of the same workload)
    _, cst = Renderer.render(scene, seed=args.seed, sample_count=min(4, spp), integrator=args.integrator, out=out,
                             count_visits=True)
    log(f"visit counts: {cst.nodes_visited / max(cst.rendered_samples, 1):.1f} nodes, "
        f"{cst.prims_tested / max(cst.rendered_samples, 1):.1f} primitives, "
        f"{(cst.closest_hit_rays + cst.shadow_rays) / max(cst.rendered_samples, 1):.2f} rays per sample")

    tuning = args.explicit_tuning if wavefront else None
    tuned = None
    if wavefront and args.tuning == "auto":
        # the autotuner's calibration pass (untimed; its result is reused by every step)
        tuned = Renderer.autotune([scene], [args.seed])
        tuning = "auto"
        log(f"autotune: {tuned.as_dict()} in {tuned.tune_ms:.0f} ms over {tuned.trials} shapes "
            f"(calibration pass {tuned.default_ms:.2f} -> {tuned.tuned_ms:.2f} ms)")
    for i in range(args.warmup):
        _, st = Renderer.render(scene, seed=args.seed, integrator=args.integrator, out=out, stage_times=wavefront,
                                tuning=tuning)
        log(f"warmup {i}: {st.render_ms:.1f} ms")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = Renderer.render(scene, seed=args.seed, integrator=args.integrator, out=out, stage_times=wavefront,
                                tuning=tuning)
        stats.append(st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for i, st in enumerate(stats):
        extra = (f", extend {st.extend_ms:.1f} / shade {st.shade_ms:.1f} / shadow {st.shadow_ms:.1f} ms"
                 if wavefront else "")
        log(f"step {i}: {st.render_ms:.1f} ms render, {st.kernel_ms:.1f} ms integrator kernels{extra}")
    samples = pixels * spp * args.steps
    value = samples / elapsed / 1e6
    rays_total = sum(s.closest_hit_rays + s.shadow_rays for s in stats)
    if wavefront:
        # dominant kernel: wf_extend (closest hit). Algorithmic bytes per closest-hit ray:
        # ray in (2 x 16 B), hit out (16 B + 4 B), BVH nodes (64 B) and primitives (48 B)
        # visited along the reference-order traversal.
        kernel = "wf_extend"
        rays_c = max(cst.closest_hit_rays, 1)
        per_ray = 32.0 + 20.0 + cst.closest_nodes / rays_c * 64.0 + cst.closest_prims / rays_c * 48.0
        closest = sum(s.closest_hit_rays for s in stats)
        launches = sum(s.extend_launches for s in stats)
        kernel_s = sum(s.extend_ms for s in stats) / 1000.0
        bytes_per_launch = per_ray * closest / max(launches, 1)
    else:
        # dominant kernel: render_mega. Algorithmic bytes per sample: BVH nodes (64 B) and
        # primitives (48 B) of every ray of the path, plus the 12 B colour written.
        kernel = "render_mega"
        per_sample = (cst.nodes_visited * 64.0 + cst.prims_tested * 48.0) / max(cst.rendered_samples, 1) + 12.0
        launches = sum(s.kernel_launches for s in stats)
        kernel_s = sum(s.kernel_ms for s in stats) / 1000.0
        bytes_per_launch = per_sample * samples / max(launches, 1)
    achieved = bytes_per_launch / (kernel_s / max(launches, 1)) / 1e9
    traffic = load_traffic(wl.name, kernel)
    line = {
        "metric": "Msamples/s",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: {wl.triangles} triangles in {wl.extra.get('instances', 0)} instances, "
                        f"{wl.width}x{wl.height}, {wl.spp} spp, depth {wl.depth}, seed {args.seed}",
            "width": wl.width, "height": wl.height, "spp": wl.spp, "max_depth": wl.depth,
            "triangles": wl.triangles, "integrator": args.integrator,
            "launch_shape": tuned.as_dict() if tuned else (args.explicit_tuning or "default"),
        },
        "mrays_per_s": round(rays_total / elapsed / 1e6, 3),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "traffic_source": (f"profiles/traffic_{kernel}_{wl.name}.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                               "per dispatch of this kernel on this workload, scaled by a 1 GiB copy probe")
            if traffic else None,
            "kernel": kernel,
            "kernel_ms_per_launch": round(kernel_s * 1000.0 / max(launches, 1), 3),
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
        },
        "cpu_baseline": None,
    }
    if args.workload == "terrain":
        # the denoise stage of the volume workload: the reference's NFOR with a radius-9
        # regression window over the two half-buffers (untimed for the headline)
        nfor = {"feature_prefilter_radius": 1, "regression_radius": 9, "patch_radius": 1}
        Renderer.render_denoised(scene, seed=args.seed, nfor=nfor)  # first call: graph capture, buffers
        _, _, dst, _ = Renderer.render_denoised(scene, seed=args.seed, nfor=nfor)
        log(f"denoised render: {dst.render_ms:.1f} ms render (wavefront, half-buffers + features), "
            f"NFOR {dst.denoise_ms:.1f} ms")
        line["denoise"] = {"nfor_ms": round(dst.denoise_ms, 3), "render_ms": round(dst.render_ms, 3),
                           "regression_radius": 9, "patch_radius": 1, "feature_prefilter_radius": 1}
        if not args.no_cpu_baseline:
            line["denoise"]["cpu_baseline"] = cpu_nfor_baseline(wl.scene_path, len(os.sched_getaffinity(0)), 320, 180,
                                                                wl.width * wl.height)
    # the quality gate's device image: samples [0, q) of the same frame (untimed), set
    # against the reference's render of those samples from the CPU leg below
    qspp = min(args.cpu_spp, wl.spp)
    quality_img = None
    if not args.no_cpu_baseline:
        quality_img, _ = Renderer.render(scene, seed=args.seed, samples=qspp, integrator=args.integrator, tuning=tuning)
    scene.release()
    ctx.close()
    if args.workload == "spheres" and not args.no_extras:
        # the other two headline figures, outside the timed region: the differentiable box's
        # forward + adjoint iteration and the uneven batch (with its small-scene regime)
        d = bench_diff(args, wdir, steps=5, warmup=2)
        line["diff"] = {k: d[k] for k in ("metric", "value", "unit", "forward_ms", "iteration_over_forward",
                                          "forward_msamples_per_s", "steps", "warmup")}
        line["diff"]["config"] = d["config"]["workload"]
        b = bench_batch(args, wdir, steps=2, warmup=1)
        line["batch"] = {k: b[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "warmup",
                                           "batched_vs_plain_megakernel", "batched_vs_single_scene_wavefront",
                                           "small_scenes")}
        line["batch"]["config"] = b["config"]["workload"]
    if not args.no_cpu_baseline:
        base, ref_img = cpu_baseline(wl.scene_path, args.seed, qspp, wl.name)
        line["cpu_baseline"] = base
        if ref_img is not None and quality_img is not None:
            import numpy as np
            r = ref_img.astype(np.float64)
            e = float(((quality_img.astype(np.float64) - r) ** 2).sum() / max((r ** 2).sum(), 1e-300))
            line["quality"] = {"rel_mse": e, "bar": 1e-4, "pass": e <= 1e-4,
                               "against": f"the reference's FP64 render of samples [0, {qspp}) of the same frame "
                                          f"({wl.width}x{wl.height}, seed {args.seed}) from the CPU leg"}
            log(f"quality: rel-MSE {e:.3e} against the reference's {wl.width}x{wl.height}x{qspp} render")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
