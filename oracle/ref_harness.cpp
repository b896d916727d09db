// Test-infrastructure harness linked against the reference renderer (built from its
// own sources into oracle/_ref by oracle/Makefile). It never ships: it only writes
// golden vectors that tests/golden/make_golden.py turns into committed fixtures.
//
// Commands (all outputs little-endian binary, layouts documented per command):
//   luzref render  <scene.luz> <seed> <threads> <out.bin> [spp] [W H]
//   luzref sampler <out.bin>
//   luzref blas    <file.obj> <out.bin>
//   luzref tlas    <scene.luz> <out.txt>
//   luzref primary <scene.luz> <seed> <out.bin> [W H]
//   luzref env     <scene.luz> <out.bin>
//   luzref lights  <scene.luz>

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <future>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <sstream>
#include <string>
#include <thread>
#include <typeinfo>
#include <vector>

// The harness reads private members (packed BVH nodes, TLAS children, mesh
// triangles). Standard headers are included above, before this switch.
#define private public
#define protected public
#include "Scene/Scene.hpp"
#include "SceneFile/SceneFile.hpp"
#include "Renderer/Renderer.hpp"
#include "RendererInternal.hpp"
#include "Sampler.hpp"
#include "Random.hpp"
#include "Defaults.hpp"
#include "Utilities.hpp"
#include "OBJReader.hpp"
#include "Denoise/NFOR.hpp"
#include "Hittables/BVHNode.hpp"
#include "Hittables/Mesh.hpp"
#include "Hittables/MeshInstance.hpp"
#include "Hittables/Triangle.hpp"
#include "Hittables/Rectangle.hpp"
#include "Hittables/Sphere.hpp"
#include "Materials/Lambertian.hpp"
#include "EnvironmentMap.hpp"
#include "Image.hpp"
#include "Blur/Gaussian.hpp"
#undef private
#undef protected

namespace
{
	template <typename T>
	void	put(std::ofstream& out, const T& value)
	{
		out.write(reinterpret_cast<const char*>(&value), sizeof(T));
	}

	// Layout mirror of the scene loader's lazily-loaded mesh wrapper
	// (src/io/SceneFile/Objects.cpp:223-288): vptr, shared_future, once_flag,
	// shared_ptr<Hittable>. Used only after the wrapper has resolved, and every
	// unwrap is checked (dynamic_cast + bounding-box equality) before use.
	struct	DeferredMirror
	{
		void*	vptr;
		std::shared_future<std::shared_ptr<Hittable>>	future;
		std::once_flag	once;
		std::shared_ptr<Hittable>	hittable;
	};

	const Hittable*	unwrap(const std::shared_ptr<Hittable>& h)
	{
		const std::string name = typeid(*h).name();
		if (name.find("DeferredHittable") == std::string::npos)
		{
			return (h.get());
		}
		AABB box;
		h->createBoundingBox(box);	// forces the deferred load
		const DeferredMirror* mirror = reinterpret_cast<const DeferredMirror*>(h.get());
		const Hittable* inner = mirror->hittable.get();
		AABB innerBox;
		if (!inner || !inner->createBoundingBox(innerBox)
			|| !(innerBox.getMinimum() == box.getMinimum()) || !(innerBox.getMaximum() == box.getMaximum()))
		{
			std::cerr << "unwrap check failed" << std::endl;
			std::exit(3);
		}
		return (inner);
	}

	void	loadScene(Scene& scene, const std::string& path)
	{
		scene.setBenchmarkMode(true);
		SceneFile::read(scene, path);
		scene.setBenchmarkMode(true);
		scene.setViewTransform(ViewTransform::Raw);
		scene.setBloom(false);
		scene.setDenoise(false);
		scene.setAdaptiveSampling(false);
		scene.setExposure(0.0);
	}

	int	cmdRender(int argc, char** argv)
	{
		if (argc < 6)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		setRandomSeed(std::stoi(argv[3]));
		scene.setRenderingThreads(static_cast<std::size_t>(std::stoul(argv[4])));
		if (argc >= 7)
		{
			scene.setSampleCount(std::stoi(argv[6]));
		}
		if (argc >= 9)
		{
			scene.getImage()->setWidth(std::stoul(argv[7]));
			scene.getImage()->setHeight(std::stoul(argv[8]));
			scene.getImage()->initialize();
		}
		if (!Renderer::render(scene))
		{
			return (1);
		}
		const auto& image = scene.getImage();
		std::ofstream out(argv[5], std::ios::binary);
		const std::uint32_t w = image->getWidth();
		const std::uint32_t h = image->getHeight();
		put(out, w);
		put(out, h);
		put(out, scene.getRenderStats().renderMS);
		for (std::size_t i = 0; i < static_cast<std::size_t>(w) * h; i++)
		{
			const Color c = image->pixels()[i];
			put(out, c.getRed());
			put(out, c.getGreen());
			put(out, c.getBlue());
		}
		std::cout << "render_ms=" << scene.getRenderStats().renderMS << std::endl;
		return (0);
	}

	// seeds SCENE FIRST COUNT THREADS OUT SPP W H: renders seeds FIRST..FIRST+COUNT-1 of one
	// loaded scene; writes u32 count, u32 tiles (4), then per seed the per-channel means of
	// each of 4 x 4 image tiles (row-major, f64).
	int	cmdSeeds(int argc, char** argv)
	{
		if (argc < 10)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		const int first = std::stoi(argv[3]);
		const std::uint32_t count = static_cast<std::uint32_t>(std::stoul(argv[4]));
		scene.setRenderingThreads(static_cast<std::size_t>(std::stoul(argv[5])));
		scene.setSampleCount(std::stoi(argv[7]));
		std::ofstream out(argv[6], std::ios::binary);
		const std::uint32_t tiles = 4;
		put(out, count);
		put(out, tiles);
		for (std::uint32_t k = 0; k < count; k++)
		{
			scene.getImage()->setWidth(std::stoul(argv[8]));
			scene.getImage()->setHeight(std::stoul(argv[9]));
			scene.getImage()->initialize();
			setRandomSeed(first + static_cast<int>(k));
			if (!Renderer::render(scene))
			{
				return (1);
			}
			const auto& image = scene.getImage();
			const std::size_t w = image->getWidth(), h = image->getHeight();
			for (std::uint32_t ty = 0; ty < tiles; ty++)
			{
				for (std::uint32_t tx = 0; tx < tiles; tx++)
				{
					double sum[3] = {0.0, 0.0, 0.0};
					std::size_t n = 0;
					for (std::size_t y = ty * h / tiles; y < (ty + 1) * h / tiles; y++)
					{
						for (std::size_t x = tx * w / tiles; x < (tx + 1) * w / tiles; x++)
						{
							const Color c = image->pixels()[y * w + x];
							sum[0] += c.getRed();
							sum[1] += c.getGreen();
							sum[2] += c.getBlue();
							n++;
						}
					}
					for (double v : sum) put(out, v / static_cast<double>(n));
				}
			}
		}
		return (0);
	}

	// denoise SCENE SEED THREADS OUT SPP W H PREFILTER REGRESSION PATCH: the worker loop of
	// Threads.cpp:375-548 with denoising on (half-buffers on sample streams 1 / 2, primary
	// features), then Denoise::applyNFOR with the given radii. Writes u32 W, H, then f64
	// arrays: image (W*H*3, the raw mean), colorA, colorB (W*H*3), colorVariance (W*H),
	// featuresA, featuresB, featureVariance (W*H*11), denoised (W*H*3).
	int	cmdDenoise(int argc, char** argv)
	{
		if (argc < 12)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		setRandomSeed(std::stoi(argv[3]));
		const std::size_t threads = static_cast<std::size_t>(std::stoul(argv[4]));
		scene.setSampleCount(std::stoi(argv[6]));
		scene.getImage()->setWidth(std::stoul(argv[7]));
		scene.getImage()->setHeight(std::stoul(argv[8]));
		scene.getImage()->initialize();
		scene.setDenoise(true);
		const bool adaptive = argc >= 15;	// optional MIN INTERVAL THRESHOLD: adaptive sampling on
		if (adaptive)
		{
			scene.setAdaptiveSampling(true);
			scene.setAdaptiveMinSamples(std::stoi(argv[12]));
			scene.setAdaptiveCheckInterval(std::stoi(argv[13]));
			scene.setAdaptiveThreshold(std::stod(argv[14]));
		}
		scene.updateLights();
		scene.updateAccelerationStructure();
		const std::size_t w = scene.getImage()->getWidth(), h = scene.getImage()->getHeight();
		const Renderer::internal::RenderCamera camera = Renderer::internal::_prepareRenderCamera(scene);
		Sampler::setRenderSeed(static_cast<std::uint32_t>(randomSeedValue()));
		scene.initializeDenoiseBuffers(w, h);
		std::vector<std::uint32_t> counts(w * h, 0);
		std::atomic<std::size_t> next(0);
		std::vector<std::thread> pool;
		for (std::size_t t = 0; t < std::max<std::size_t>(1, threads); t++)
		{
			pool.emplace_back([&]() {
				for (std::size_t i = next++; i < w * h; i = next++)
				{
					counts[i] = Renderer::internal::_threadRender(scene, camera, i % w, i / w);
				}
			});
		}
		for (std::thread& th : pool) th.join();
		Denoise::NFORSettings settings;
		settings.featurePrefilterRadius = static_cast<unsigned int>(std::stoul(argv[9]));
		settings.regressionRadius = static_cast<unsigned int>(std::stoul(argv[10]));
		settings.patchRadius = static_cast<unsigned int>(std::stoul(argv[11]));
		settings.threadCount = std::max<std::size_t>(1, threads);
		const Denoise::NFORBuffers& b = *scene.getDenoiseBuffers();
		const auto nforStart = std::chrono::steady_clock::now();
		std::unique_ptr<Image> denoised = Denoise::applyNFOR(b, settings);
		std::cout << "nfor_ms=" << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - nforStart).count()
			<< std::endl;
		std::ofstream out(argv[5], std::ios::binary);
		put(out, static_cast<std::uint32_t>(w));
		put(out, static_cast<std::uint32_t>(h));
		auto putColor = [&](const Color& c) { put(out, c.getRed()); put(out, c.getGreen()); put(out, c.getBlue()); };
		for (std::size_t i = 0; i < w * h; i++) putColor(scene.getImage()->pixels()[i]);
		for (const Color& c : b.colorA) putColor(c);
		for (const Color& c : b.colorB) putColor(c);
		for (double v : b.colorVariance) put(out, v);
		for (const auto* fs : {&b.featuresA, &b.featuresB, &b.featureVariance})
		{
			for (const Denoise::FeatureVector& f : *fs)
			{
				for (std::size_t k = 0; k < Denoise::NFOR_FEATURE_COUNT; k++) put(out, f[k]);
			}
		}
		for (std::size_t i = 0; i < w * h; i++) putColor(denoised->pixels()[i]);
		if (adaptive)
		{
			for (std::uint32_t c : counts) put(out, c);
		}
		return (0);
	}

	// adaptive SCENE SEED THREADS OUT SPP W H MIN INTERVAL THRESHOLD: the per-pixel loop of
	// _threadRender with adaptive sampling on and denoising off; writes u32 W, H, then W*H
	// u32 sample counts (its return values) and the W*H*3 f64 image.
	int	cmdAdaptive(int argc, char** argv)
	{
		if (argc < 11)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		setRandomSeed(std::stoi(argv[3]));
		const std::size_t threads = static_cast<std::size_t>(std::stoul(argv[4]));
		scene.setSampleCount(std::stoi(argv[6]));
		scene.getImage()->setWidth(std::stoul(argv[7]));
		scene.getImage()->setHeight(std::stoul(argv[8]));
		scene.getImage()->initialize();
		scene.setAdaptiveSampling(true);
		scene.setAdaptiveMinSamples(std::stoi(argv[9]));
		scene.setAdaptiveCheckInterval(std::stoi(argv[10]));
		scene.setAdaptiveThreshold(argc >= 12 ? std::stod(argv[11]) : 0.03);
		scene.updateLights();
		scene.updateAccelerationStructure();
		const std::size_t w = scene.getImage()->getWidth(), h = scene.getImage()->getHeight();
		const Renderer::internal::RenderCamera camera = Renderer::internal::_prepareRenderCamera(scene);
		Sampler::setRenderSeed(static_cast<std::uint32_t>(randomSeedValue()));
		std::vector<std::uint32_t> counts(w * h, 0);
		std::atomic<std::size_t> next(0);
		std::vector<std::thread> pool;
		for (std::size_t t = 0; t < std::max<std::size_t>(1, threads); t++)
		{
			pool.emplace_back([&]() {
				for (std::size_t i = next++; i < w * h; i = next++)
				{
					counts[i] = Renderer::internal::_threadRender(scene, camera, i % w, i / w);
				}
			});
		}
		for (std::thread& th : pool) th.join();
		std::ofstream out(argv[5], std::ios::binary);
		put(out, static_cast<std::uint32_t>(w));
		put(out, static_cast<std::uint32_t>(h));
		for (std::uint32_t c : counts) put(out, c);
		for (std::size_t i = 0; i < w * h; i++)
		{
			const Color c = scene.getImage()->pixels()[i];
			put(out, c.getRed());
			put(out, c.getGreen());
			put(out, c.getBlue());
		}
		return (0);
	}

	// post IN OUT EXPOSURE BLOOM VIEW CONTRAST: the reference's display post-processing of an
	// FP64 image (u32 W, H, then W*H*3 f64, written back in the same layout). The steps are
	// applyPostProcessing's (Threads.cpp:55-99, an unexported function there), called through
	// the public Image / Gaussian API; VIEW: standard, agx, aces or raw.
	int	cmdPost(int argc, char** argv)
	{
		if (argc < 8)
		{
			return (2);
		}
		std::ifstream in(argv[2], std::ios::binary);
		std::uint32_t w = 0, h = 0;
		in.read(reinterpret_cast<char*>(&w), 4);
		in.read(reinterpret_cast<char*>(&h), 4);
		Image image(w, h);
		image.initialize();
		for (std::size_t i = 0; i < static_cast<std::size_t>(w) * h; i++)
		{
			double c[3];
			in.read(reinterpret_cast<char*>(c), sizeof(c));
			image.pixels()[i] = Color(c[0], c[1], c[2]);
		}
		const double exposure = std::stod(argv[4]);
		const bool bloom = std::stoi(argv[5]) != 0;
		const std::string view = argv[6];
		const double contrast = std::stod(argv[7]);
		const ViewTransform vt = view == "standard" ? ViewTransform::Standard : view == "aces" ? ViewTransform::ACES
			: view == "raw" ? ViewTransform::Raw : ViewTransform::AgX;
		image.applyExposure(exposure);
		if (bloom)
		{
			auto brightness = image.extractBloom(D_BLOOM_THRESHOLD, D_BLOOM_SOFT_KNEE);
			Image blurred(image.getWidth(), image.getHeight());
			blurred.initialize();
			Gaussian::blur(*brightness, blurred, 9, 2.0);
			for (std::size_t i = 0; i < static_cast<std::size_t>(w) * h; i++)
			{
				image.pixels()[i] += blurred.pixels()[i] * D_BLOOM_INTENSITY;
			}
		}
		if (vt != ViewTransform::Raw)
		{
			if (contrast == D_CONTRAST)
			{
				image.applyViewTransformAndEncodeSRGB(vt);
				image.suppressIsolatedFireflies();
			}
			else
			{
				image.applyViewTransform(vt);
				image.applyContrast(contrast);
				image.gammaCorrect();
				image.suppressIsolatedFireflies();
			}
		}
		std::ofstream out(argv[3], std::ios::binary);
		put(out, w);
		put(out, h);
		for (std::size_t i = 0; i < static_cast<std::size_t>(w) * h; i++)
		{
			const Color c = image.pixels()[i];
			put(out, c.getRed());
			put(out, c.getGreen());
			put(out, c.getBlue());
		}
		return (0);
	}

	// record: u32 seed, x, y, sample, stream, bounce, dim; f64 s1; f32 s2x, s2y
	int	cmdSampler(int argc, char** argv)
	{
		if (argc < 3)
		{
			return (2);
		}
		std::ofstream out(argv[2], std::ios::binary);
		const std::uint32_t seeds[] = {1u, 1234u, 2027u, 424242424u};
		const std::uint32_t pixels[][2] = {{0, 0}, {13, 17}, {299, 299}, {1919, 1079}};
		const std::uint32_t samples[] = {0, 1, 5, 63};
		const std::uint32_t bounces[] = {0, 1, 7, 16};
		std::uint32_t count = 0;
		for (std::uint32_t seed : seeds)
		for (const auto& px : pixels)
		for (std::uint32_t s : samples)
		for (std::uint32_t stream = 0; stream < 3; stream++)
		for (std::uint32_t b : bounces)
		for (std::uint32_t dim = 0; dim < 20; dim++)
		{
			Sampler::setRenderSeed(seed);
			Sampler::beginPixelSample(px[0], px[1], s, stream);
			Sampler::setBounce(b);
			const double v1 = Sampler::sample1D(dim);
			const Sampler::Sample2D v2 = Sampler::sample2D(dim);
			Sampler::endPixelSample();
			put(out, seed); put(out, px[0]); put(out, px[1]); put(out, s);
			put(out, stream); put(out, b); put(out, dim);
			put(out, v1); put(out, v2.x); put(out, v2.y);
			count++;
		}
		std::cout << "records=" << count << std::endl;
		return (0);
	}

	void	writeBLAS(std::ofstream& out, const Mesh& mesh)
	{
		const std::uint64_t nodeCount = mesh._packedBVHNodes.size();
		const std::uint64_t triCount = mesh._triangles.size();
		put(out, nodeCount);
		put(out, triCount);
		for (const auto& node : mesh._packedBVHNodes)
		{
			for (int a = 0; a < 3; a++) put(out, node.boundingBox.getMinimum()[a]);
			for (int a = 0; a < 3; a++) put(out, node.boundingBox.getMaximum()[a]);
			const std::int64_t links[5] = {
				static_cast<std::int64_t>(node.left), static_cast<std::int64_t>(node.right),
				static_cast<std::int64_t>(node.start), static_cast<std::int64_t>(node.count),
				node.isLeaf ? 1 : 0
			};
			for (std::int64_t v : links) put(out, v);
		}
		for (std::size_t idx : mesh._triangleIndices)
		{
			put(out, static_cast<std::int64_t>(idx));
		}
		for (const Triangle& t : mesh._triangles)
		{
			for (int a = 0; a < 3; a++) put(out, t._vertex0[a]);
			for (int a = 0; a < 3; a++) put(out, t._vertex1[a]);
			for (int a = 0; a < 3; a++) put(out, t._vertex2[a]);
		}
	}

	// out: u64 nodeCount, u64 triCount; per node 6 f64 box + 5 i64 (left,right,start,count,leaf);
	// triCount i64 triangle indices; triCount*9 f64 vertices (after degenerate removal)
	int	cmdBLAS(int argc, char** argv)
	{
		if (argc < 4)
		{
			return (2);
		}
		const ObjMeshData data = readObjMeshData(argv[2], ObjReadOptions{true, nullptr});
		const Mesh mesh = buildObjMesh(data, Vector3(), Vector3(), Vector3(1.0, 1.0, 1.0),
			std::make_shared<Lambertian>(Color(0.6, 0.6, 0.6)));
		std::ofstream out(argv[3], std::ios::binary);
		writeBLAS(out, mesh);
		std::cout << "nodes=" << mesh._packedBVHNodes.size() << " tris=" << mesh._triangles.size() << std::endl;
		return (0);
	}

	void	dumpTLAS(std::ostream& out, const Hittable* node, const std::map<const Hittable*, std::size_t>& index)
	{
		const BVHNode* bvh = dynamic_cast<const BVHNode*>(node);
		if (bvh == nullptr || index.count(node))
		{
			out << index.at(node);
			return;
		}
		out << "(";
		for (std::size_t i = 0; i < bvh->_childs.size(); i++)
		{
			if (i)
			{
				out << ",";
			}
			dumpTLAS(out, bvh->_childs[i].get(), index);
		}
		out << ")";
	}

	// out: one line, nested parentheses over top-level hittable indices (scene order)
	int	cmdTLAS(int argc, char** argv)
	{
		if (argc < 4)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		scene.updateLights();
		scene.updateAccelerationStructure();
		std::map<const Hittable*, std::size_t> index;
		for (std::size_t i = 0; i < scene.getHittables().size(); i++)
		{
			index[scene.getHittables()[i].get()] = i;
		}
		std::ofstream out(argv[3]);
		if (scene.getAccelerationStructure())
		{
			dumpTLAS(out, scene.getAccelerationStructure().get(), index);
		}
		out << "\n";
		return (0);
	}

	struct	Candidate
	{
		double	t = std::numeric_limits<double>::infinity();
		std::int64_t	top = -1;
		std::int64_t	tri = -1;
		double	margin = std::numeric_limits<double>::infinity();
	};

	// Möller–Trumbore exactly as Triangle::hit (Triangle.cpp:196-240), returning the
	// barycentric margin min(u, v, 1-u-v) and the determinant-scaled edge distance.
	bool	triangleT(const Triangle& tri, const Ray& ray, double tMin, double tMax, double& t, double& margin)
	{
		if (tri._isDegenerate)
		{
			return (false);
		}
		Vector3 p = Utilities::cross(ray.getDirection(), tri._edge2);
		double det = Utilities::dot(tri._edge1, p);
		if (std::fabs(det) < 1e-12)
		{
			return (false);
		}
		const double invDet = 1.0 / det;
		Vector3 tVec = ray.getOrigin() - tri._vertex0;
		double u = Utilities::dot(tVec, p) * invDet;
		Vector3 q = Utilities::cross(tVec, tri._edge1);
		double v = Utilities::dot(ray.getDirection(), q) * invDet;
		margin = std::min(u, std::min(v, 1.0 - u - v));
		if (u < 0.0 || u > 1.0 || v < 0.0 || u + v > 1.0)
		{
			return (false);
		}
		t = Utilities::dot(tri._edge2, q) * invDet;
		return (t >= tMin && t <= tMax);
	}

	// Brute-force FP64 closest and second-closest hits over every primitive, in the
	// reference's own intersection code. IDs: (top-level index, triangle index or -1).
	void	primaryCandidates(const Scene& scene, Ray ray, Candidate& best, Candidate& second)
	{
		auto offer = [&](const Candidate& c) {
			if (c.t < best.t)
			{
				second = best;
				best = c;
			}
			else if (c.t < second.t)
			{
				second = c;
			}
		};
		const auto& hittables = scene.getHittables();
		for (std::size_t i = 0; i < hittables.size(); i++)
		{
			const Hittable* h = unwrap(hittables[i]);
			if (const Triangle* tri = dynamic_cast<const Triangle*>(h))
			{
				Candidate c;
				double t, m;
				if (triangleT(*tri, ray, T_MIN, T_MAX, t, m))
				{
					c.t = t; c.top = i; c.tri = -1; c.margin = m;
					offer(c);
				}
				continue;
			}
			if (const Mesh* mesh = dynamic_cast<const Mesh*>(h))
			{
				for (std::size_t k = 0; k < mesh->_triangles.size(); k++)
				{
					double t, m;
					if (triangleT(mesh->_triangles[k], ray, T_MIN, T_MAX, t, m))
					{
						Candidate c;
						c.t = t; c.top = i; c.tri = k; c.margin = m;
						offer(c);
					}
				}
				continue;
			}
			if (const MeshInstance* inst = dynamic_cast<const MeshInstance*>(h))
			{
				Ray local;
				double scale = 1.0;
				if (!inst->_localRay(ray, local, scale))
				{
					continue;
				}
				const Mesh& mesh = *inst->_geometry;
				double rootNear;
				if (mesh._packedBVHNodes.empty())
				{
					continue;
				}
				AABB box = mesh._packedBVHNodes[0].boundingBox;
				HitRecord dummy;
				if (!box.hit(local, dummy, T_MAX))
				{
					continue;
				}
				(void)rootNear;
				for (std::size_t k = 0; k < mesh._triangles.size(); k++)
				{
					double t, m;
					if (triangleT(mesh._triangles[k], local, T_MIN * scale, T_MAX, t, m))
					{
						Candidate c;
						c.t = t / scale; c.top = i; c.tri = k; c.margin = m;
						if (c.t >= T_MIN)
						{
							offer(c);
						}
					}
				}
				continue;
			}
			HitRecord rec;
			if (h->hit(ray, rec, T_MIN, T_MAX))
			{
				Candidate c;
				c.t = rec.t0; c.top = i; c.tri = -1; c.margin = std::numeric_limits<double>::infinity();
				offer(c);
			}
		}
	}

	// record per pixel (row-major, sample 0, stream 0):
	//   i64 top, i64 tri, f64 t, f64 margin, i64 top2, i64 tri2, f64 t2, f64 tRef, u8 refHit
	int	cmdPrimary(int argc, char** argv)
	{
		if (argc < 5)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		if (argc >= 7)
		{
			scene.getImage()->setWidth(std::stoul(argv[5]));
			scene.getImage()->setHeight(std::stoul(argv[6]));
			scene.getImage()->initialize();
		}
		scene.updateLights();
		scene.updateAccelerationStructure();
		const std::uint32_t seed = static_cast<std::uint32_t>(std::stoi(argv[3]));
		const auto camera = Renderer::internal::_prepareRenderCamera(scene);
		const std::size_t w = scene.getImage()->getWidth();
		const std::size_t h = scene.getImage()->getHeight();
		struct Rec { Candidate a, b; double tRef; bool refHit; };
		std::vector<Rec> recs(w * h);
		const unsigned threads = std::max(1u, std::thread::hardware_concurrency());
		std::vector<std::thread> pool;
		for (unsigned ti = 0; ti < threads; ti++)
		{
			pool.emplace_back([&, ti]() {
				Sampler::setRenderSeed(seed);
				for (std::size_t idx = ti; idx < w * h; idx += threads)
				{
					const std::size_t x = idx % w, y = idx / w;
					Sampler::beginPixelSample(x, y, 0, 0);
					Ray ray = Renderer::internal::_generateRay(camera, x, y);
					Sampler::setBounce(0);
					Candidate a, b;
					primaryCandidates(scene, ray, a, b);
					HitRecord rec;
					Ray probe = ray;
					const bool refHit = Renderer::internal::_checkHits(scene, probe, rec);
					Sampler::endPixelSample();
					recs[idx] = Rec{a, b, refHit ? rec.t0 : -1.0, refHit};
				}
			});
		}
		for (auto& t : pool)
		{
			t.join();
		}
		std::ofstream out(argv[4], std::ios::binary);
		put(out, static_cast<std::uint32_t>(w));
		put(out, static_cast<std::uint32_t>(h));
		for (const Rec& r : recs)
		{
			put(out, r.a.top); put(out, r.a.tri); put(out, r.a.t); put(out, r.a.margin);
			put(out, r.b.top); put(out, r.b.tri); put(out, r.b.t); put(out, r.tRef);
			const std::uint8_t flag = r.refHit ? 1 : 0;
			put(out, flag);
		}
		return (0);
	}

	// out: u32 W, u32 H, W*H*3 f64 ACEScg texels as loaded, then W*H f64 cdf, f64 total
	int	cmdEnv(int argc, char** argv)
	{
		if (argc < 4)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		const auto& env = scene.getEnvironmentMap();
		if (!env)
		{
			return (1);
		}
		std::ofstream out(argv[3], std::ios::binary);
		put(out, static_cast<std::uint32_t>(env->_width));
		put(out, static_cast<std::uint32_t>(env->_height));
		for (const Color& c : env->_pixels)
		{
			put(out, c.getRed()); put(out, c.getGreen()); put(out, c.getBlue());
		}
		for (double c : env->_cdf)
		{
			put(out, c);
		}
		put(out, env->_totalWeight);
		std::cout << "strength=" << scene.getEnvironmentStrength() << std::endl;
		return (0);
	}

	int	cmdLights(int argc, char** argv)
	{
		if (argc < 3)
		{
			return (2);
		}
		Scene scene;
		loadScene(scene, argv[2]);
		scene.updateLights();
		std::printf("lights=%zu total=%.17g\n", scene.getLights().size(), scene.getLightSelectionTotalWeight());
		for (std::size_t i = 0; i < scene.getLights().size(); i++)
		{
			const Color e = scene.getLights()[i]->getMaterial()->emitted();
			std::printf("%zu cdf=%.17g emitted=%.17g,%.17g,%.17g\n", i,
				scene.getLightSelectionCumulativeWeights().empty() ? 0.0 : scene.getLightSelectionCumulativeWeights()[i],
				e.getRed(), e.getGreen(), e.getBlue());
		}
		return (0);
	}
}

int	main(int argc, char** argv)
{
	if (argc < 2)
	{
		std::cerr << "usage: luzref render|seeds|denoise|adaptive|post|sampler|blas|tlas|primary|env|lights ..." << std::endl;
		return (2);
	}
	try
	{
		const std::string cmd = argv[1];
		if (cmd == "render") return (cmdRender(argc, argv));
		if (cmd == "sampler") return (cmdSampler(argc, argv));
		if (cmd == "blas") return (cmdBLAS(argc, argv));
		if (cmd == "tlas") return (cmdTLAS(argc, argv));
		if (cmd == "primary") return (cmdPrimary(argc, argv));
		if (cmd == "env") return (cmdEnv(argc, argv));
		if (cmd == "lights") return (cmdLights(argc, argv));
		if (cmd == "seeds") return (cmdSeeds(argc, argv));
		if (cmd == "denoise") return (cmdDenoise(argc, argv));
		if (cmd == "adaptive") return (cmdAdaptive(argc, argv));
		if (cmd == "post") return (cmdPost(argc, argv));
	}
	catch (const std::exception& e)
	{
		std::cerr << e.what() << std::endl;
		return (1);
	}
	return (2);
}
