"""MI355X-native radiance path for the Luz renderer (BVH traversal, ray-primitive
intersection and the MIS path integrator) behind the reference's scene-loader /
render-settings interface. See include/luzgpu.h for the C ABI."""

from .renderer import Context, Renderer, RenderStats, Scene, SceneFile, Tuning  # noqa: F401
