"""Autotuner and launch-shape knobs (GPU).

The traversal launch shape (persistent grid, lane-refill thresholds, rays per queue claim,
LDS stack depth) only decides which lane traverses which ray and where its stack entries
live; every ray is traversed in the same order, so images must be bit-identical under every
shape, and an autotuned render must equal a default one."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    {"lds_stack": 8}, {"lds_stack": 24}, {"fetch_chunk": 64}, {"fetch_chunk": 1024},
    {"extend_refill": 1, "shadow_refill": 64}, {"extend_refill": 64, "shadow_refill": 1},
    {"grid": 256, "lds_stack": 8, "fetch_chunk": 128},
]


def test_launch_shapes_do_not_change_the_image(ctx, spheres):
    from luz_amd import Renderer, SceneFile
    scene = SceneFile.read(ctx, spheres.scene_path)
    ref, _ = Renderer.render(scene, seed=3, samples=2, resolution=(320, 180))
    for shape in SHAPES:
        img, _ = Renderer.render(scene, seed=3, samples=2, resolution=(320, 180), tuning=shape)
        assert np.array_equal(img, ref), shape


def test_autotune_single_scene(ctx, cornell):
    from luz_amd import Renderer, SceneFile
    scene = SceneFile.read(ctx, cornell.scene_path)
    t = Renderer.autotune([scene], [5], samples=[8])
    print(f"cornell autotune: {t}")
    assert t.trials >= 10 and t.tuned_ms <= t.default_ms
    assert t.lds_stack in (8, 16, 24) and 64 <= t.fetch_chunk <= 1024 and 1 <= t.grid <= 2048
    ref, _ = Renderer.render(scene, seed=5, samples=8)
    img, st = Renderer.render(scene, seed=5, samples=8, tuning="auto")
    assert st.tune_ms == 0.0  # the tuning above is reused, not searched again
    assert np.array_equal(img, ref)


def test_autotune_on_first_render_and_batch(ctx, cornell, primitives, volume):
    from luz_amd import Renderer, SceneFile
    scenes = [SceneFile.read(ctx, w.scene_path) for w in (cornell, primitives, volume)]
    seeds = [1, 2, 3]
    ref, _ = Renderer.render_batch(scenes, seeds, samples=[4, 4, 4])
    img, st = Renderer.render_batch(scenes, seeds, samples=[4, 4, 4], tuning="auto")
    print(f"batch: first auto render tuned in {st.tune_ms:.1f} ms")
    assert st.tune_ms > 0.0
    for a, b in zip(img, ref):
        assert np.array_equal(a, b)
    again, st2 = Renderer.render_batch(scenes, seeds, samples=[4, 4, 4], tuning="auto")
    assert st2.tune_ms == 0.0
    for a, b in zip(again, ref):
        assert np.array_equal(a, b)
