// Test infrastructure: drives the reference-side binding (luz_amd/glue) against the
// reference itself. Linked with the reference (oracle/_ref/libluzref.a), the binding
// (oracle/glue_access.cpp) and luz_amd/libluzgpu.so.
//
//   glue_driver topology <scene.luz>
//       Scene from the reference's own loader (SceneFile::read), flattened by the binding
//       and handed to luzgpu_scene_create on the host only. Prints the library's TLAS
//       topology ("device ..."), the reference's own TLAS of that Scene ("reference ..."),
//       the library's topology of the same file through luzgpu_scene_load ("file ..."), and
//       per mesh BVH digests of both routes ("blas <route> <mesh> <nodes> <digest>").
//   glue_driver inmemory <W> <H>
//       The same for a Scene built in code (no file), without the "file" route.
//   glue_driver render-inmemory <device> <W> <H> <spp> <seed>
//       The Scene built in code rendered twice with the same seed: through the binding's
//       manageThreadsOnDevice and through the reference's Renderer::render on the CPU;
//       prints "rel_mse <value>" of the device image against the reference's.
//   glue_driver render-file <device> <scene.luz> <seed>
//       The same for a Scene read by the reference's own loader (twice, one per route),
//       with the raw view transform and no bloom on both.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <future>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <typeinfo>
#include <vector>

#define private public
#define protected public
#include "Scene/Scene.hpp"
#include "SceneFile/SceneFile.hpp"
#include "Renderer/Renderer.hpp"
#include "Random.hpp"
#include "Camera.hpp"
#include "Transform.hpp"
#include "Hittables/BVHNode.hpp"
#include "Hittables/Cube.hpp"
#include "Hittables/Mesh.hpp"
#include "Hittables/MeshInstance.hpp"
#include "Hittables/Rectangle.hpp"
#include "Hittables/Sphere.hpp"
#include "Hittables/Triangle.hpp"
#include "Materials/Dielectric.hpp"
#include "Materials/Emissive.hpp"
#include "Materials/Lambertian.hpp"
#include "Materials/Metal.hpp"
#undef private
#undef protected

#include "../luz_amd/glue/luzgpu_glue.hpp"

namespace
{
	// Layout mirror of the loader's lazily-loaded object wrapper (Objects.cpp:223-288), used
	// only after the wrapper has loaded; every unwrap is checked by bounding-box equality.
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
		if (name.find("DeferredHittable") == std::string::npos) return h.get();
		AABB box;
		h->createBoundingBox(box);
		const Hittable* inner = reinterpret_cast<const DeferredMirror*>(h.get())->hittable.get();
		AABB innerBox;
		if (!inner || !inner->createBoundingBox(innerBox) || !(innerBox.getMinimum() == box.getMinimum())
			|| !(innerBox.getMaximum() == box.getMaximum()))
		{
			std::cerr << "unwrap check failed" << std::endl;
			std::exit(3);
		}
		return inner;
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
			if (i) out << ",";
			dumpTLAS(out, bvh->_childs[i].get(), index);
		}
		out << ")";
	}

	void	check(int rc, const char* what)
	{
		if (rc != LUZGPU_OK)
		{
			std::cerr << what << ": " << luzgpu_last_error() << std::endl;
			std::exit(4);
		}
	}

	std::string	deviceTopology(const luzgpu_scene* s)
	{
		std::size_t need = 0;
		luzgpu_scene_tlas_topology(s, nullptr, 0, &need);
		std::string buf(need, '\0');
		check(luzgpu_scene_tlas_topology(s, buf.data(), buf.size(), &need), "luzgpu_scene_tlas_topology");
		buf.resize(need ? need - 1 : 0);
		return buf;
	}

	// FNV-1a over each mesh BVH's links and triangle order (boxes are derived from them)
	void	printBlas(const char* route, const luzgpu_scene* s)
	{
		luzgpu_scene_info info{};
		check(luzgpu_scene_get_info(s, &info), "luzgpu_scene_get_info");
		for (std::uint32_t m = 0; m < info.meshes; m++)
		{
			std::uint64_t nodes = 0, tris = 0;
			check(luzgpu_scene_blas(s, m, nullptr, nullptr, nullptr, &nodes, &tris), "luzgpu_scene_blas");
			std::vector<double> boxes(6 * nodes);
			std::vector<int64_t> links(5 * nodes), order(tris);
			check(luzgpu_scene_blas(s, m, boxes.data(), links.data(), order.data(), &nodes, &tris), "luzgpu_scene_blas");
			std::uint64_t h = 1469598103934665603ull;
			auto mix = [&](std::uint64_t v) {
				for (int b = 0; b < 8; b++)
				{
					h ^= (v >> (8 * b)) & 0xffu;
					h *= 1099511628211ull;
				}
			};
			for (int64_t v : links) mix(static_cast<std::uint64_t>(v));
			for (int64_t v : order) mix(static_cast<std::uint64_t>(v));
			for (double v : boxes)
			{
				std::uint64_t bits;
				std::memcpy(&bits, &v, 8);
				mix(bits);
			}
			std::cout << "blas " << route << " " << m << " " << nodes << " " << std::hex << h << std::dec << "\n";
		}
		std::cout << "info " << route << " hittables " << info.hittables << " triangles " << info.triangles << " lights "
			<< info.lights << " materials " << info.materials << "\n";
	}

	void	printRoutes(Scene& scene, luzgpu_glue::Resolver resolve)
	{
		luzgpu_glue::FlatScene flat;
		luzgpu_glue::flatten(scene, flat, resolve);
		luzgpu_scene* s = nullptr;
		check(luzgpu_scene_create(nullptr, &flat.desc, &s), "luzgpu_scene_create");
		std::cout << "device " << deviceTopology(s) << "\n";
		printBlas("device", s);
		luzgpu_scene_release(s);
		scene.updateLights();
		scene.updateAccelerationStructure();
		std::map<const Hittable*, std::size_t> index;
		for (std::size_t i = 0; i < scene.getHittables().size(); i++) index[scene.getHittables()[i].get()] = i;
		std::ostringstream ref;
		if (scene.getAccelerationStructure()) dumpTLAS(ref, scene.getAccelerationStructure().get(), index);
		std::cout << "reference " << ref.str() << "\n";
	}

	// A scene built in code, as the reference's own tests build theirs (Tests.cpp:221-242):
	// a box of Lambertian triangles, a rectangle light, a glass sphere, a rough metal mesh
	// instance (rotated, scaled) and a cube.
	void	buildScene(Scene& scene, std::size_t width, std::size_t height, int spp)
	{
		scene.getImage()->setWidth(width);
		scene.getImage()->setHeight(height);
		scene.getImage()->initialize();
		scene.setSampleCount(spp);
		scene.setAdaptiveSampling(false);
		scene.setMaxLightBounces(6);
		scene.setViewTransform(ViewTransform::Raw);
		scene.setBloom(false);
		scene.setDenoise(false);
		scene.setRenderSky(SKY_NONE);
		scene.setBackgroundColor(Color(0.02, 0.03, 0.05));
		scene.setRenderingThreads(std::max(1u, std::thread::hardware_concurrency()));
		scene.setBenchmarkMode(true);
		Camera cam;
		cam.setPosition(Vector3(0.0, 2.0, 6.0));
		cam.setDirection(Vector3(0.0, -0.1, -1.0));
		cam.setFocalLengthMeters(0.03);
		cam.setSensorWidthMeters(0.036);
		cam.setSensorHeightMeters(0.027);
		cam.setPinhole(true);
		scene.addCamera(cam);
		auto white = std::make_shared<Lambertian>(Color(0.73, 0.73, 0.73));
		auto red = std::make_shared<Lambertian>(Color(0.65, 0.05, 0.05));
		auto green = std::make_shared<Lambertian>(Color(0.12, 0.45, 0.15));
		auto light = std::make_shared<Emissive>(Color(14.0, 12.0, 9.0));
		auto glass = std::make_shared<Dielectric>(Color(1.0, 1.0, 1.0), 1.5, 0.0);
		auto metal = std::make_shared<Metal>(Color(0.9, 0.8, 0.6), 0.25);
		auto quad = [&](Vector3 a, Vector3 b, Vector3 c, Vector3 d, std::shared_ptr<Material> m) {
			scene.addHittable(std::make_shared<Triangle>(a, b, c, m));
			scene.addHittable(std::make_shared<Triangle>(a, c, d, m));
		};
		quad(Vector3(-2, 0, 2), Vector3(2, 0, 2), Vector3(2, 0, -2), Vector3(-2, 0, -2), white);	// floor
		quad(Vector3(-2, 4, -2), Vector3(2, 4, -2), Vector3(2, 4, 2), Vector3(-2, 4, 2), white);	// ceiling
		quad(Vector3(-2, 0, -2), Vector3(2, 0, -2), Vector3(2, 4, -2), Vector3(-2, 4, -2), white);	// back
		quad(Vector3(-2, 0, 2), Vector3(-2, 0, -2), Vector3(-2, 4, -2), Vector3(-2, 4, 2), red);	// left
		quad(Vector3(2, 0, -2), Vector3(2, 0, 2), Vector3(2, 4, 2), Vector3(2, 4, -2), green);		// right
		scene.addHittable(std::make_shared<Rectangle>(Transform(Vector3(0, 3.99, -0.5), Vector3(0, -1, 0), Vector3(1, 1, 1)),
			1.2, 0.9, light));
		scene.addHittable(std::make_shared<Sphere>(Vector3(-0.9, 0.7, -0.6), 0.7, glass));
		std::vector<Triangle> tris;
		const Vector3 v[6] = {Vector3(1, 0, 0), Vector3(-1, 0, 0), Vector3(0, 1, 0), Vector3(0, -1, 0), Vector3(0, 0, 1), Vector3(0, 0, -1)};
		const int f[8][3] = {{0, 2, 4}, {2, 1, 4}, {1, 3, 4}, {3, 0, 4}, {2, 0, 5}, {1, 2, 5}, {3, 1, 5}, {0, 3, 5}};
		for (const auto& t : f) tris.emplace_back(v[t[0]], v[t[1]], v[t[2]], metal);
		auto mesh = std::make_shared<Mesh>(Vector3(), metal, tris);
		scene.addHittable(std::make_shared<MeshInstance>(mesh, Vector3(0.9, 0.8, -0.2), Vector3(20.0, 35.0, 10.0),
			Vector3(0.6, 0.8, 0.6), metal));
		scene.addHittable(std::make_shared<MeshInstance>(mesh, Vector3(0.2, 0.35, 0.9), Vector3(0.0, 45.0, 0.0),
			Vector3(0.35, 0.35, 0.35), red));
		scene.addHittable(std::make_shared<Cube>(Transform(Vector3(1.1, 0.4, -1.3), Vector3(0, 0, -1), Vector3(1, 1, 1)),
			0.8, 0.8, 0.8, white));
	}

	double	relMse(const Image& a, const Image& r)
	{
		double num = 0.0, den = 0.0;
		for (std::size_t i = 0; i < r.getWidth() * r.getHeight(); i++)
		{
			const Color x = a.pixels()[i], y = r.pixels()[i];
			const double d[3] = {x.getRed() - y.getRed(), x.getGreen() - y.getGreen(), x.getBlue() - y.getBlue()};
			num += d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
			den += y.getRed() * y.getRed() + y.getGreen() * y.getGreen() + y.getBlue() * y.getBlue();
		}
		return num / std::max(den, 1e-300);
	}
}

int	main(int argc, char** argv)
{
	if (argc < 2) return 2;
	const std::string cmd = argv[1];
	try
	{
		if (cmd == "topology" && argc >= 3)
		{
			Scene scene;
			scene.setBenchmarkMode(true);
			SceneFile::read(scene, argv[2]);
			luzgpu_scene* s = nullptr;
			check(luzgpu_scene_load(nullptr, argv[2], &s), "luzgpu_scene_load");
			std::cout << "file " << deviceTopology(s) << "\n";
			printBlas("file", s);
			luzgpu_scene_release(s);
			printRoutes(scene, unwrap);
			return 0;
		}
		if (cmd == "inmemory" && argc >= 4)
		{
			Scene scene;
			buildScene(scene, std::stoul(argv[2]), std::stoul(argv[3]), 1);
			printRoutes(scene, nullptr);
			return 0;
		}
		if (cmd == "render-file" && argc >= 5)
		{
			const int device = std::stoi(argv[2]);
			setRandomSeed(std::stoi(argv[4]));
			Scene onDevice, onCpu;
			for (Scene* sc : {&onDevice, &onCpu})
			{
				sc->setBenchmarkMode(true);
				SceneFile::read(*sc, argv[3]);
				sc->setRenderingThreads(std::max(1u, std::thread::hardware_concurrency()));
				// compare the radiance images: the display transform (tested on its own)
				// turns fp32 noise in dark pixels into large relative differences
				sc->setViewTransform(ViewTransform::Raw);
				sc->setBloom(false);
			}
			luzgpu_glue::manageThreadsOnDevice(onDevice, device, unwrap);
			if (!Renderer::render(onCpu)) return 5;
			std::cout << "device_ms " << onDevice.getRenderStats().renderMS << " cpu_ms " << onCpu.getRenderStats().renderMS
				<< " samples " << onDevice.getRenderStats().renderedSamples << " " << onCpu.getRenderStats().renderedSamples << "\n";
			std::cout << "rel_mse " << relMse(*onDevice.getImage(), *onCpu.getImage()) << "\n";
			return 0;
		}
		if (cmd == "render-inmemory" && argc >= 7)
		{
			const int device = std::stoi(argv[2]);
			const std::size_t w = std::stoul(argv[3]), h = std::stoul(argv[4]);
			const int spp = std::stoi(argv[5]);
			setRandomSeed(std::stoi(argv[6]));
			Scene onDevice, onCpu;
			buildScene(onDevice, w, h, spp);
			buildScene(onCpu, w, h, spp);
			luzgpu_glue::manageThreadsOnDevice(onDevice, device);
			if (!Renderer::render(onCpu)) return 5;
			std::cout << "device_ms " << onDevice.getRenderStats().renderMS << " cpu_ms " << onCpu.getRenderStats().renderMS
				<< " samples " << onDevice.getRenderStats().renderedSamples << " " << onCpu.getRenderStats().renderedSamples << "\n";
			std::cout << "rel_mse " << relMse(*onDevice.getImage(), *onCpu.getImage()) << "\n";
			return 0;
		}
	}
	catch (const std::exception& e)
	{
		std::cerr << "error: " << e.what() << std::endl;
		return 1;
	}
	return 2;
}
