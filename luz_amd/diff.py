"""Differentiable rendering through the C ABI: a torch.autograd.Function whose forward is
``luzgpu_render`` (wavefront) and whose backward is the path-replay adjoint
``luzgpu_adjoint`` (no counterpart in the reference renderer).

    img = render(scene, albedo, roughness, emission, seed=1, samples=16)
    loss = ((img - target) ** 2).mean(); loss.backward()

albedo (n, 3), roughness (n,) and emission (n, 3) are per-material tensors in scene-file
order (n = scene.info["materials"]); each is applied to the materials of its kind
(Lambertian, metal, emissive). The forward and the adjoint draw the same random numbers.
"""

from __future__ import annotations

import torch

from .renderer import Renderer


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, roughness, emission, scene, seed, samples):
        scene.set_materials(albedo, roughness, emission)
        out = torch.empty(scene.height * scene.width * 3, dtype=torch.float32, device="cuda")
        # adaptive sampling off: the adjoint differentiates the image of every sample
        Renderer.render(scene, seed=seed, samples=samples, out=out, adaptive=False)
        ctx.scene, ctx.seed, ctx.samples = scene, seed, samples
        ctx.dtypes = (albedo.dtype, roughness.dtype, emission.dtype)
        ctx.devices = (albedo.device, roughness.device, emission.device)
        return out.view(scene.height, scene.width, 3)

    @staticmethod
    def backward(ctx, grad):
        da, dr, de = ctx.scene.adjoint(grad.contiguous(), seed=ctx.seed, samples=ctx.samples, adaptive=False)
        # each gradient lives where its parameter does
        return (torch.as_tensor(da, dtype=ctx.dtypes[0], device=ctx.devices[0]),
                torch.as_tensor(dr, dtype=ctx.dtypes[1], device=ctx.devices[1]),
                torch.as_tensor(de, dtype=ctx.dtypes[2], device=ctx.devices[2]), None, None, None)


def render(scene, albedo, roughness, emission, seed: int = 1, samples=None):
    """Differentiable render: image (H, W, 3) on the device."""
    return RenderFunction.apply(albedo, roughness, emission, scene, seed, samples)
