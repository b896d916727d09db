// Reference-side binding of the device library: the code a Luz maintainer adds to the
// reference's renderer (beside src/renderer/Threads.cpp) so that Renderer::render(Scene&)
// (include/luz/Renderer/Renderer.hpp:7) can run its worker loop on the GPU with the Scene it
// already holds, no file involved.
//
// It reads the Scene's hittables, materials, meshes and environment map, several of which
// keep their parameters private (Rectangle::_transform, MeshInstance::_transform,
// Mesh::_triangles, EnvironmentMap::_pixels ...). Each such class needs
//     friend struct luzgpu_glue::Access;
// in its declaration; nothing else in the reference changes.
#pragma once

#include "../../include/luzgpu.h"

#include <cstdint>
#include <memory>
#include <vector>

class	Scene;
class	Hittable;

namespace luzgpu_glue
{
	struct	Access;

	// The flattened records of one Scene and the arrays they point into.
	struct	FlatScene
	{
		luzgpu_scene_desc	desc{};
		std::vector<luzgpu_material_desc>	materials;
		std::vector<luzgpu_mesh_desc>	meshes;
		std::vector<std::vector<double>>	meshVertices;
		std::vector<std::vector<double>>	meshNormals;
		std::vector<luzgpu_hittable_desc>	hittables;
		std::vector<double>	environment;
		std::vector<luzgpu_ies_desc>	ies;	// point into the Scene's IESProfile objects

		FlatScene() = default;
		FlatScene(const FlatScene&) = delete;
		FlatScene& operator=(const FlatScene&) = delete;
	};

	// The scene-file loader wraps each object it reads in a lazily loading hittable
	// (DeferredHittable, src/io/SceneFile/Objects.cpp:223-288, private to the loader) that
	// forwards every query to the object it loads. A resolver returns that object (or the
	// hittable itself when it is not such a wrapper); inside the reference it is a one-line
	// accessor on the wrapper. Scenes built in code need none.
	using	Resolver = const Hittable* (*)(const std::shared_ptr<Hittable>&);

	// Flattens `scene` into luzgpu_scene_desc records (Scene::getHittables() order, each
	// distinct Material and Mesh once). Throws std::runtime_error naming what the device path
	// does not take (textures, photon-map caustics, legacy meshes).
	void	flatten(const Scene& scene, FlatScene& out, Resolver resolve = nullptr);

	// Renderer::internal::_manageThreads (Threads.cpp:375-548) on the GPU: the render seed,
	// the Scene's raw image (or, with denoising on, its half-buffer render and the NFOR
	// denoised image), the display post-processing of both, and the render statistics, all
	// left in the Scene as the worker loop leaves them. Throws std::runtime_error with the
	// library's message on failure.
	void	manageThreadsOnDevice(Scene& scene, int device = 0, Resolver resolve = nullptr);
}
