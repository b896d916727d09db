"""Path-replay adjoint against FP64 central finite differences of the reference renderer (GPU).

Fixture (tests/golden/make_golden.py, fd_fixture): the differentiable box with Lambertian
materials and one rectangle emitter at depth 2 (no Russian roulette, one light), where the
estimator at a fixed seed is a polynomial in each albedo and linear in the radiance, so the
reference's finite differences are the exact derivatives of the estimator the adjoint
differentiates. Loss L = sum(w * image) with a fixed random weight image w, so dL/dimage = w."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_mse

pytestmark = pytest.mark.gpu


def _fd():
    return np.load(os.path.join(GOLDEN, "diff_fd_128_8spp_depth2.npz"))


@pytest.fixture(scope="module")
def diff_scene(ctx, workload_dir):
    from luz_amd import SceneFile, workloads
    g = _fd()
    wl = workloads.write_diff_cornell(os.path.join(workload_dir, "diff"), None, int(g["width"]), int(g["height"]),
                                      int(g["spp"]), int(g["depth"]), metals=bool(g["metals"]))
    return SceneFile.read(ctx, wl.scene_path), wl


def test_adjoint_matches_reference_finite_differences(diff_scene):
    from luz_amd import Renderer
    scene, wl = diff_scene
    g = _fd()
    img, _ = Renderer.render(scene, seed=int(g["seed"]))
    e = rel_mse(img, g["image"])
    print(f"forward image vs reference: rel-MSE {e:.3e}")
    assert e < 1e-4
    da, dr, de = scene.adjoint(g["weights"], seed=int(g["seed"]))
    for name, got, ref in (("albedo", da, g["d_albedo"]), ("emission", de, g["d_emission"])):
        mask = np.abs(ref) > 1e-3 * np.abs(ref).max()
        rel = np.abs(got[mask] - ref[mask]) / np.abs(ref[mask])
        print(f"{name}: max relative error {rel.max():.2e} over {int(mask.sum())} gradients")
        assert rel.max() < 2e-3, (name, got, ref)
        assert np.allclose(got[~mask], ref[~mask], atol=2e-3 * np.abs(ref).max())


def test_autograd_function_and_set_materials(diff_scene):
    """torch.autograd.Function round trip: gradients of a torch loss equal the adjoint's, and
    set_materials changes the render the way the scene file would."""
    import torch
    from luz_amd import diff, workloads
    scene, wl = diff_scene
    g = _fd()
    kinds, albedo, roughness, emission = workloads.diff_parameters(metals=bool(g["metals"]))
    a = torch.tensor(albedo, dtype=torch.float64, requires_grad=True)
    r = torch.tensor(roughness, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(emission, dtype=torch.float64, requires_grad=True)
    w = torch.from_numpy(g["weights"]).cuda()
    img = diff.render(scene, a, r, e, seed=int(g["seed"]))
    loss = (img * w).sum()
    loss.backward()
    assert abs(loss.item() - float(g["loss"])) < 1e-4 * abs(float(g["loss"]))
    mask = np.abs(g["d_albedo"]) > 1e-3 * np.abs(g["d_albedo"]).max()
    assert np.allclose(a.grad.numpy()[mask], g["d_albedo"][mask], rtol=2e-3)
    # doubling the light doubles the image
    e2 = e.detach() * 2.0
    img2 = diff.render(scene, a.detach(), r.detach(), e2, seed=int(g["seed"]))
    assert torch.allclose(img2, 2.0 * img.detach(), rtol=1e-4, atol=1e-6)
    scene.set_materials(albedo, roughness, emission)


@pytest.mark.parametrize("depth", [2, 6])
def test_adjoint_matches_oracle_adjoint_with_metals(ctx, workload_dir, depth):
    """With rough metals (roughness gradients, mirror-fallback score terms) and, at depth 6,
    Russian roulette, the device adjoint replays the oracle's paths with the same random
    numbers: per-material gradients agree up to fp32 path divergence."""
    import oracle_lib
    from luz_amd import SceneFile, workloads
    g = _fd()
    wl = workloads.write_diff_cornell(os.path.join(workload_dir, f"diff_metal_{depth}"), None, 128, 128, 8, depth,
                                      metals=True)
    scene = SceneFile.read(ctx, wl.scene_path)
    w = g["weights"]
    da, dr, de = scene.adjoint(w, seed=5)
    ref = oracle_lib.adjoint(wl.scene_path, 5, 8, 128, 128, w, 17)
    for name, got, want in (("albedo", da, ref[:, :3]), ("roughness", dr, ref[:, 3]), ("emission", de, ref[:, 4:])):
        err = np.abs(got - want)
        scale = np.abs(want).max()
        print(f"depth {depth} {name}: max |device - oracle| {err.max():.3e} of max {scale:.3e}")
        assert scale > 0.0
        assert np.all(err <= 2e-2 * scale + 2e-2 * np.abs(want)), (name, got, want)


def test_roughness_gradients_match_reference_finite_differences(ctx, workload_dir):
    """Depth-6 roughness gradients of four rough metals (Russian roulette, mirror fallbacks,
    MIS) against the reference's FP64 central differences at 64 seeds
    (make_golden.py fd_rough_fixture): the paired mean difference over the seeds is within 4
    standard errors for every metal (test_oracle_adjoint holds the oracle to the same bar)."""
    from luz_amd import SceneFile, workloads
    from test_oracle_adjoint import _fd_rough, paired_z
    g = _fd_rough()
    params = workloads.diff_rough_parameters()
    w, h, spp, depth = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"])
    wl = workloads.write_diff_cornell(os.path.join(workload_dir, "diff_rough"), params, w, h, spp, depth)
    scene = SceneFile.read(ctx, wl.scene_path)
    fd = g["fd"]
    adj = np.zeros_like(fd)
    for si, seed in enumerate(g["seeds"]):
        for j, m in enumerate(g["metals"]):
            _, dr, _ = scene.adjoint(g["weights"][j], seed=int(seed))
            adj[si, j] = dr[m]
    z, _ = paired_z(adj, fd)
    print("reference FD", fd.mean(0), "\ndevice adjoint", adj.mean(0), "\npaired z", z)
    assert np.all(np.abs(z) < 4), z
    scene.release()
