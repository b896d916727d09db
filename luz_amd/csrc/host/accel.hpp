// Host-side acceleration-structure build and device packing.
#pragma once

#include "scene.hpp"
#include "../common/layout.h"

#include <string>
#include <vector>

namespace luzgpu
{
	struct	AABBd
	{
		dvec3	lo{};
		dvec3	hi{};
	};

	// One node of a mesh BVH exactly as the reference's Mesh::PackedBVHNode (Mesh.hpp:37-45).
	struct	BlasNode
	{
		AABBd	box;
		std::size_t	left = 0, right = 0, start = 0, count = 0;
		bool	leaf = false;
	};

	struct	Blas
	{
		std::vector<BlasNode>	nodes;
		std::vector<std::size_t>	triangleIndices;
	};

	Blas	buildBlas(const Mesh& mesh);

	// The scene file's adaptive sampling settings (Settings.cpp:210-261, defaults
	// Defaults.hpp:13-16); whether a render stops pixels early is decided per call.
	struct	AdaptiveSettings
	{
		bool	enabled = true;
		std::uint32_t	minSamples = 128;
		std::uint32_t	checkInterval = 32;
		double	threshold = 0.03;
	};

	// The scene file's display post-processing settings (Settings.cpp:274-323), applied by
	// the caller's request after a render (applyPostProcessing, Threads.cpp:77-99).
	struct	PostSettings
	{
		int	viewTransform = 1;	// 0 standard, 1 AgX, 2 ACES, 3 raw
		bool	bloom = true;
		double	exposure = 0.0;
		double	contrast = 1.0;
	};

	struct	PackedScene
	{
		std::vector<BvhNode>	tlasNodes;	// binary device form (intermediate)
		std::vector<BvhNode>	blasNodes;
		std::vector<Bvh4Node>	tlasWide;	// what the kernels traverse
		std::vector<Bvh4Node>	blasWide;
		std::vector<GpuTri>	topTris;
		std::vector<GpuTri>	meshTris;
		std::vector<float>	meshNormals;	// 9 floats per mesh triangle slot (leaf order)
		std::vector<std::uint32_t>	tlasPrims;
		std::vector<std::uint32_t>	tlasLeaf;	// per tlasPrims item: hittable index, kind << 28 | index
		std::vector<GpuSphere>	spheres;
		std::vector<GpuRect>	cubeFaces;	// six per cube
		std::vector<GpuVolume>	volumes;
		std::vector<GpuPlane>	planes;
		std::vector<GpuHittable>	hittables;
		std::vector<GpuRect>	rects;
		std::vector<GpuInstance>	instances;
		std::vector<GpuMaterial>	materials;
		std::vector<GpuLight>	lights;
		std::vector<double>	lightCdf;
		std::vector<double>	lightTriCdf;			// emissive instances: triangle-area prefix sums
		std::vector<std::uint32_t>	lightTriSlot;	// ... and each entry's meshTris slot
		std::vector<float>	envPixels;	// float4 per texel
		std::vector<float>	envPdf;		// (weight / total) / solid angle per texel
		std::vector<double>	envCdf;
		std::vector<std::uint32_t>	envGuide;	// texel-aligned with envCdf (layout.h, envGuideBuckets)
		std::vector<double>	iesData;	// IES profiles of sphere lights (layout.h, GpuLight::ies)
		std::vector<std::uint32_t>	unbounded;
		GpuParams	params{};
		std::string	tlasTopology;	// nested "(...)" form over top-level hittable indices
		std::vector<Blas>	blas;		// reference-shaped mesh BVHs, for topology checks
		std::size_t	triangleCount = 0;
		int	tlasDeviceDepth = 0;	// device-layout (4-wide) node levels
		int	blasDeviceDepth = 0;
		int	tlasStackBound = 0;	// worst-case deferred traversal entries
		int	blasStackBound = 0;
		Camera	camera;
		double	metersPerUnit = 1.0;
		int	features = 0;	// kFeat* classes present (kernel specialisation)
		std::vector<double>	lightScale;	// light selection weight per unit luminance (area, or 1)
		std::vector<dvec3>	materialColor;	// FP64 material colours (weights are rebuilt from these)
		AdaptiveSettings	adaptive;
		PostSettings	post;
	};

	// Rebuilds light emission, selection weights, CDF and probabilities after material
	// colours changed (Scene::updateLights, Scene.cpp:615-638).
	void	rebuildLights(PackedScene& p);

	// RenderCamera for an image of width x height (RayHelper.cpp:50-82).
	GpuCamera	makeCamera(const Camera& cam, std::uint32_t width, std::uint32_t height, double metersPerUnit);

	// Builds the TLAS/BLAS exactly as the reference does (BVHNode.cpp:142-185,
	// Mesh.cpp:217-521), repacks them into the 64-byte device layout and derives
	// light and environment sampling tables. Throws on unsupported content.
	PackedScene	packScene(const Scene& scene);

	// Concatenates packed scenes into one set of device arrays for a multi-scene launch;
	// params receives each scene's parameters with its root and table bases.
	PackedScene	mergeScenes(const std::vector<const PackedScene*>& parts, std::vector<GpuParams>& params);
}
