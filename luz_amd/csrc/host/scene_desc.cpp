// The host scene from the C ABI's flattened records (luzgpu_scene_desc): the in-memory
// counterpart of readSceneFile for callers that hold a populated Scene
// (Renderer::render(Scene&), include/luz/Renderer/Renderer.hpp:7). The records carry the
// values the reference's objects hold after construction, so only their consistency is
// checked here: indices in range, kinds known, finite positive sizes where the reference's
// constructors would have refused others.
#include "scene.hpp"

#include "../../../include/luzgpu.h"

#include <cmath>
#include <stdexcept>
#include <string>

namespace luzgpu
{
	namespace
	{
		dvec3	d3(const double* v) { return {v[0], v[1], v[2]}; }

		void	requireFinite(const dvec3& v, const std::string& what)
		{
			if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2]))
			{
				throw std::invalid_argument(what + " must be finite");
			}
		}

		void	requirePositive(double v, const std::string& what)
		{
			if (!std::isfinite(v) || v <= 0.0)
			{
				throw std::invalid_argument(what + " must be finite and positive");
			}
		}

		Material	materialOf(const luzgpu_material_desc& d, std::uint32_t index)
		{
			if (d.kind < LUZGPU_MAT_LAMBERTIAN || d.kind > LUZGPU_MAT_PRINCIPLED)
			{
				throw std::invalid_argument("material " + std::to_string(index) + ": unknown kind " + std::to_string(d.kind));
			}
			Material m;
			m.kind = static_cast<MaterialKind>(d.kind);
			m.name = "material" + std::to_string(index);
			m.color = d3(d.color);
			m.roughness = d.roughness;
			m.ior = d.ior;
			m.conductor = d.kind == LUZGPU_MAT_METAL && d.conductor != 0;
			m.eta = d3(d.eta);
			m.k = d3(d.k);
			m.absorption = d3(d.absorption);
			m.anisotropy = d.anisotropy;
			m.glossyColor = d3(d.glossy_color);
			m.glossyWeight = d.glossy_weight;
			m.metallic = d.metallic;
			m.transmission = d.transmission;
			m.clearcoat = d.clearcoat;
			m.clearcoatRoughness = d.clearcoat_roughness;
			m.sheen = d.sheen;
			if (d.kind == LUZGPU_MAT_PRINCIPLED && d.subsurface != 0.0)
			{
				const std::string at = "material " + std::to_string(index) + ": subsurface ";
				if (!std::isfinite(d.subsurface) || d.subsurface < 0.0 || d.subsurface > 1.0)
				{
					throw std::invalid_argument(at + "weight must be finite and in [0,1]");
				}
				bool positive = false;
				for (int i = 0; i < 3; i++)
				{
					if (!std::isfinite(d.subsurface_radius[i]) || d.subsurface_radius[i] < 0.0
						|| !std::isfinite(d.subsurface_color[i]) || d.subsurface_color[i] < 0.0)
					{
						throw std::invalid_argument(at + "radius and color must have finite non-negative channels");
					}
					positive = positive || d.subsurface_radius[i] > 0.0;
				}
				if (!positive) throw std::invalid_argument(at + "radius must have at least one positive channel");
				if (!std::isfinite(d.subsurface_scale) || d.subsurface_scale <= 0.0)
				{
					throw std::invalid_argument(at + "scale must be finite and positive");
				}
				m.subsurface = d.subsurface;
				m.subsurfaceRadius = d3(d.subsurface_radius);
				m.subsurfaceScale = d.subsurface_scale;
				m.subsurfaceColor = d3(d.subsurface_color);
				m.subsurfaceThin = d.subsurface_thin != 0;
			}
			requireFinite(m.color, "material color");
			if (!std::isfinite(m.roughness) || m.roughness < 0.0 || m.roughness > 1.0)
			{
				throw std::invalid_argument("material " + std::to_string(index) + ": roughness must be finite and in [0,1]");
			}
			if ((d.kind == LUZGPU_MAT_DIELECTRIC || d.kind == LUZGPU_MAT_PRINCIPLED) && !(std::isfinite(m.ior) && m.ior > 0.0))
			{
				throw std::invalid_argument("material " + std::to_string(index) + ": refractive index must be finite and positive");
			}
			if (m.conductor)
			{
				for (int i = 0; i < 3; i++)
				{
					if (!std::isfinite(m.eta[i]) || m.eta[i] <= 0.0 || !std::isfinite(m.k[i]) || m.k[i] < 0.0)
					{
						throw std::invalid_argument("material " + std::to_string(index) + ": conductor eta must be positive, k non-negative");
					}
				}
			}
			return m;
		}
	}

	Scene	sceneFromDesc(const luzgpu_scene_desc& d)
	{
		Scene sc;
		Settings& s = sc.settings;
		if (d.width == 0 || d.height == 0)
		{
			throw std::invalid_argument("scene image size must be positive");
		}
		s.width = d.width;
		s.height = d.height;
		s.samples = d.samples;
		s.maxBounces = d.max_bounces;
		s.adaptive = d.adaptive != 0;
		s.adaptiveMinSamples = d.adaptive_min_samples;
		s.adaptiveCheckInterval = d.adaptive_check_interval;
		s.adaptiveThreshold = d.adaptive_threshold;
		s.denoise = d.denoise != 0;
		s.viewTransform = static_cast<int>(d.view_transform);
		s.bloom = d.bloom != 0;
		s.exposure = d.exposure;
		s.contrast = d.contrast;
		if (d.sky > LUZGPU_SKY_ATMOSPHERE)
		{
			throw std::invalid_argument("unknown sky kind " + std::to_string(d.sky));
		}
		s.sky = static_cast<SkyKind>(d.sky);
		s.skySet = true;
		s.sunAngle = d.sun_angle;
		s.earthRadius = d.earth_radius;
		s.atmosphereRadius = d.atmosphere_radius;
		s.atmosphereHR = d.rayleigh_height;
		s.atmosphereHM = d.mie_height;
		s.atmosphereSamples = d.atmosphere_samples;
		s.atmosphereLightSamples = d.atmosphere_light_samples;
		s.starsBrightness = d.stars_brightness;
		s.sunRadianceScale = d.sun_radiance_scale;
		s.background = d3(d.background);
		requirePositive(d.meters_per_unit, "meters per unit");
		s.metersPerUnit = d.meters_per_unit;
		if (s.sky == SKY_ATMOSPHERE && (s.atmosphereSamples <= 0 || s.atmosphereLightSamples <= 0 || s.atmosphereSamples > 4096
			|| s.atmosphereLightSamples > 4096))
		{
			throw std::invalid_argument("atmosphere samples and light samples must be in 1..4096");
		}

		Camera& c = sc.camera;
		c.position = d3(d.camera.position);
		c.direction = d3(d.camera.direction);
		c.up = d3(d.camera.up);
		c.focalLength = d.camera.focal_length;
		c.sensorWidth = d.camera.sensor_width;
		c.sensorHeight = d.camera.sensor_height;
		c.fNumber = d.camera.f_number;
		c.focusDistance = d.camera.focus_distance;
		c.pinhole = d.camera.pinhole != 0;
		requirePositive(c.focalLength, "camera focal length");
		requirePositive(c.sensorWidth, "camera sensor width");
		requirePositive(c.sensorHeight, "camera sensor height");
		requirePositive(c.fNumber, "camera f-number");
		requirePositive(c.focusDistance, "camera focus distance");
		sc.hasCamera = true;

		if (d.env_width > 0 || d.env_height > 0)
		{
			if (d.env_width == 0 || d.env_height == 0 || !d.env_texels)
			{
				throw std::invalid_argument("environment map needs a size and texels");
			}
			Environment& e = sc.environment;
			e.width = d.env_width;
			e.height = d.env_height;
			const std::size_t n = static_cast<std::size_t>(d.env_width) * d.env_height;
			e.pixels.resize(n);
			for (std::size_t i = 0; i < n; i++) e.pixels[i] = d3(d.env_texels + 3 * i);
			e.strength = d.env_strength;
			e.rotation = d.env_rotation;
			e.lighting = d.env_lighting != 0;
			sc.hasEnvironment = true;
		}
		if (s.sky == SKY_ENVIRONMENT && !sc.hasEnvironment)
		{
			throw std::invalid_argument("environment sky without an environment map");
		}

		if ((d.material_count && !d.materials) || (d.mesh_count && !d.meshes) || (d.hittable_count && !d.hittables))
		{
			throw std::invalid_argument("NULL material, mesh or hittable array with a non-zero count");
		}
		for (std::uint32_t i = 0; i < d.material_count; i++) sc.materials.push_back(materialOf(d.materials[i], i));
		if (d.ies_count && !d.ies_profiles) throw std::invalid_argument("NULL IES profile array with a non-zero count");
		for (std::uint32_t i = 0; i < d.ies_count; i++)
		{
			// IESProfile::load's checks (IESProfile.cpp:180-229) on the profile as handed over
			const luzgpu_ies_desc& id = d.ies_profiles[i];
			const std::string who = "IES profile " + std::to_string(i);
			if (id.vertical_count == 0 || id.horizontal_count == 0 || !id.vertical || !id.horizontal || !id.candela)
			{
				throw std::invalid_argument(who + ": empty angle or candela arrays");
			}
			IesProfile pr;
			pr.vertical.assign(id.vertical, id.vertical + id.vertical_count);
			pr.horizontal.assign(id.horizontal, id.horizontal + id.horizontal_count);
			pr.candela.assign(id.candela, id.candela + static_cast<std::size_t>(id.vertical_count) * id.horizontal_count);
			for (const std::vector<double>* a : {&pr.vertical, &pr.horizontal})
			{
				for (std::size_t k = 0; k < a->size(); k++)
				{
					if (!std::isfinite((*a)[k]) || (k > 0 && (*a)[k] <= (*a)[k - 1]))
					{
						throw std::invalid_argument(who + ": angles must be finite and strictly increasing");
					}
				}
			}
			for (const double c : pr.candela)
			{
				if (!std::isfinite(c) || c < 0.0) throw std::invalid_argument(who + ": candela values must be finite and non-negative");
			}
			if (!std::isfinite(id.total_lumens) || id.total_lumens <= 0.0) throw std::invalid_argument(who + ": zero luminous flux");
			pr.totalLumens = id.total_lumens;
			pr.meanCandela = id.total_lumens / (4.0 * 3.14159265358979323846);
			sc.iesProfiles.push_back(std::move(pr));
		}
		for (std::uint32_t i = 0; i < d.mesh_count; i++)
		{
			const luzgpu_mesh_desc& md = d.meshes[i];
			if (md.triangle_count && !md.vertices)
			{
				throw std::invalid_argument("mesh " + std::to_string(i) + ": NULL vertices");
			}
			Mesh mesh;
			mesh.source = "mesh" + std::to_string(i);
			mesh.triangles.resize(md.triangle_count);
			for (std::uint32_t t = 0; t < md.triangle_count; t++)
			{
				Triangle& tri = mesh.triangles[t];
				const double* v = md.vertices + 9 * static_cast<std::size_t>(t);
				tri.v0 = d3(v);
				tri.v1 = d3(v + 3);
				tri.v2 = d3(v + 6);
				if (md.normals)
				{
					const double* nn = md.normals + 9 * static_cast<std::size_t>(t);
					tri.n0 = d3(nn);
					tri.n1 = d3(nn + 3);
					tri.n2 = d3(nn + 6);
					tri.hasNormals = true;
				}
			}
			sc.meshes.push_back(std::move(mesh));
		}
		for (std::uint32_t i = 0; i < d.hittable_count; i++)
		{
			const luzgpu_hittable_desc& h = d.hittables[i];
			const std::string who = "hittable " + std::to_string(i);
			if (h.material >= d.material_count)
			{
				throw std::invalid_argument(who + ": material index out of range");
			}
			Hittable ht;
			ht.material = h.material;
			switch (h.kind)
			{
				case LUZGPU_HIT_TRIANGLE:
				{
					Triangle t;
					t.v0 = d3(h.a);
					t.v1 = d3(h.b);
					t.v2 = d3(h.c);
					if (h.flags & 1u)
					{
						t.n0 = d3(h.normals);
						t.n1 = d3(h.normals + 3);
						t.n2 = d3(h.normals + 6);
						t.hasNormals = true;
					}
					ht.kind = HIT_TRIANGLE;
					ht.index = static_cast<std::uint32_t>(sc.triangles.size());
					sc.triangles.push_back(t);
					break;
				}
				case LUZGPU_HIT_RECTANGLE:
				{
					Rectangle r;
					r.position = d3(h.a);
					r.orientation = d3(h.b);
					r.width = h.size[0];
					r.height = h.size[1];
					ht.kind = HIT_RECTANGLE;
					ht.index = static_cast<std::uint32_t>(sc.rectangles.size());
					sc.rectangles.push_back(r);
					break;
				}
				case LUZGPU_HIT_INSTANCE:
				{
					if (h.mesh >= d.mesh_count)
					{
						throw std::invalid_argument(who + ": mesh index out of range");
					}
					Instance in;
					in.mesh = h.mesh;
					in.position = d3(h.a);
					in.scale = d3(h.c);
					in.hasTrig = true;
					in.cosRotation = d3(h.cos_rotation);
					in.sinRotation = d3(h.sin_rotation);
					ht.kind = HIT_INSTANCE;
					ht.index = static_cast<std::uint32_t>(sc.instances.size());
					sc.instances.push_back(in);
					break;
				}
				case LUZGPU_HIT_SPHERE:
				{
					requirePositive(h.radius, who + " radius");
					Sphere sp;
					sp.position = d3(h.a);
					sp.radius = h.radius;
					sp.visible = (h.flags & 1u) == 0;
					if (h.flags & 2u)
					{
						if (h.mesh >= d.ies_count) throw std::invalid_argument(who + " IES profile index out of range");
						const double dl2 = h.b[0] * h.b[0] + h.b[1] * h.b[1] + h.b[2] * h.b[2];
						if (!std::isfinite(dl2) || dl2 <= 0.0) throw std::invalid_argument(who + " IES profile direction must be non-zero");
						if (!std::isfinite(h.size[0])) throw std::invalid_argument(who + " IES profile rotation must be finite");
						sp.ies = static_cast<std::int32_t>(h.mesh);
						sp.iesDirection = d3(h.b);
						sp.iesRotation = h.size[0];
					}
					ht.kind = HIT_SPHERE;
					ht.index = static_cast<std::uint32_t>(sc.spheres.size());
					sc.spheres.push_back(sp);
					break;
				}
				case LUZGPU_HIT_CUBE:
				{
					Cube cu;
					cu.position = d3(h.a);
					cu.width = h.size[0];
					cu.height = h.size[1];
					cu.depth = h.size[2];
					ht.kind = HIT_CUBE;
					ht.index = static_cast<std::uint32_t>(sc.cubes.size());
					sc.cubes.push_back(cu);
					break;
				}
				case LUZGPU_HIT_VOLUME:
				{
					requirePositive(h.density, who + " density");
					Volume v;
					v.sphere = (h.flags & 1u) != 0;
					v.position = d3(h.a);
					v.radius = h.radius;
					v.size = d3(h.size);
					v.density = h.density;
					ht.kind = HIT_VOLUME;
					ht.index = static_cast<std::uint32_t>(sc.volumes.size());
					sc.volumes.push_back(v);
					break;
				}
				case LUZGPU_HIT_PLANE:
				{
					Plane pl;
					pl.y = h.plane_y;
					pl.orientation = d3(h.b);
					ht.kind = HIT_PLANE;
					ht.index = static_cast<std::uint32_t>(sc.planes.size());
					sc.planes.push_back(pl);
					break;
				}
				case LUZGPU_HIT_DIRECTIONAL:
				{
					DirectionalLight dl;
					dl.direction = d3(h.a);
					if (h.flags & 1u)
					{
						dl.hasSunRadiance = true;
						dl.sunRadiance = d3(h.b);
						for (int k = 0; k < 3; k++)
						{
							if (!std::isfinite(dl.sunRadiance[k]) || dl.sunRadiance[k] < 0.0)
							{
								throw std::invalid_argument(who + ": atmosphere sun radiance must be finite and non-negative");
							}
						}
					}
					ht.kind = HIT_DIRECTIONAL;
					ht.index = static_cast<std::uint32_t>(sc.directionals.size());
					sc.directionals.push_back(dl);
					break;
				}
				default:
					throw std::invalid_argument(who + ": unknown kind " + std::to_string(h.kind));
			}
			sc.hittables.push_back(ht);
		}
		return sc;
	}
}
