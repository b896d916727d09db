// Reference-side binding (see luzgpu_glue.hpp): flattens a populated Scene into the C ABI's
// records and runs the worker loop of Renderer::render on the device. Compiled with the
// reference's headers (-I include/luz) and linked with the reference and libluzgpu.so.
#include "luzgpu_glue.hpp"

#include "Camera.hpp"
#include "ColorManagement.hpp"
#include "EnvironmentMap.hpp"
#include "IESProfile.hpp"
#include "Image.hpp"
#include "Random.hpp"
#include "Scene/Scene.hpp"
#include "Transform.hpp"
#include "Hittables/ConstantVolume.hpp"
#include "Hittables/Cube.hpp"
#include "Hittables/DirectionalLight.hpp"
#include "Hittables/Mesh.hpp"
#include "Hittables/MeshInstance.hpp"
#include "Hittables/Plane.hpp"
#include "Hittables/Rectangle.hpp"
#include "Hittables/Sphere.hpp"
#include "Hittables/Triangle.hpp"
#include "Materials/Dielectric.hpp"
#include "Materials/DiffuseGlossy.hpp"
#include "Materials/Glossy.hpp"
#include "Materials/HenyeyGreenstein.hpp"
#include "Materials/Metal.hpp"
#include "Materials/Principled.hpp"

#include <map>
#include <memory>
#include <stdexcept>
#include <string>

namespace luzgpu_glue
{
	// Read access to the members the reference keeps private (each class declares
	// `friend struct luzgpu_glue::Access;`).
	struct	Access
	{
		static bool	vertexNormals(const Triangle& t, double out[9])
		{
			if (!t._hasVertexNormals) return false;
			const Vector3* n[3] = {&t._normal0, &t._normal1, &t._normal2};
			for (int i = 0; i < 3; i++)
			{
				out[3 * i] = n[i]->getX();
				out[3 * i + 1] = n[i]->getY();
				out[3 * i + 2] = n[i]->getZ();
			}
			return true;
		}
		static const Transform&	transform(const Rectangle& r) { return r._transform; }
		static double	width(const Rectangle& r) { return r._width; }
		static double	height(const Rectangle& r) { return r._height; }
		static bool	explicitAxes(const Rectangle& r) { return r._hasAxes; }
		static const Transform&	transform(const Cube& c) { return c._transform; }
		static const IESProfile*	ies(const Sphere& s) { return s._iesProfile.get(); }
		static const Vector3&	iesDirection(const Sphere& s) { return s._iesDirection; }
		static double	iesRotation(const Sphere& s) { return s._iesRotationDegrees; }
		static const std::vector<double>&	verticalAngles(const IESProfile& p) { return p._verticalAngles; }
		static const std::vector<double>&	horizontalAngles(const IESProfile& p) { return p._horizontalAngles; }
		static const std::vector<double>&	candela(const IESProfile& p) { return p._candelaValues; }
		static void	size(const Cube& c, double out[3]) { out[0] = c._width; out[1] = c._height; out[2] = c._depth; }
		static double	y(const Plane& p) { return p._y; }
		static const Vector3&	orientation(const Plane& p) { return p._orientation; }
		static const std::shared_ptr<Hittable>&	boundary(const ConstantVolume& v) { return v._boundary; }
		static Material*	phase(const ConstantVolume& v) { return v._phaseFunction.get(); }
		static bool	packed(const Mesh& m) { return m._usesPackedTriangles; }
		static const std::vector<Triangle>&	triangles(const Mesh& m) { return m._triangles; }
		static void	transform(const MeshInstance& i, double position[3], double scale[3], double c[3], double s[3])
		{
			const auto& x = i._transform;
			position[0] = x.position.getX(); position[1] = x.position.getY(); position[2] = x.position.getZ();
			scale[0] = x.scale.getX(); scale[1] = x.scale.getY(); scale[2] = x.scale.getZ();
			c[0] = x.cosX; c[1] = x.cosY; c[2] = x.cosZ;
			s[0] = x.sinX; s[1] = x.sinY; s[2] = x.sinZ;
		}
		static bool	textured(const Material& m) { return m._texture != nullptr; }
		static const std::vector<Color>&	pixels(const EnvironmentMap& e) { return e._pixels; }
		static double	hR(const Atmosphere& a) { return a._hR; }
		static double	hM(const Atmosphere& a) { return a._hM; }
		static int	samples(const Atmosphere& a) { return a._samples; }
		static int	lightSamples(const Atmosphere& a) { return a._lightSamples; }
	};

	namespace
	{
		void	put(const Vector3& v, double out[3])
		{
			out[0] = v.getX();
			out[1] = v.getY();
			out[2] = v.getZ();
		}

		void	put(const Color& c, double out[3])
		{
			out[0] = c.getRed();
			out[1] = c.getGreen();
			out[2] = c.getBlue();
		}

		luzgpu_material_desc	materialOf(Material& m)
		{
			if (Access::textured(m))
			{
				throw std::runtime_error("Unsupported for the device path: material textures");
			}
			luzgpu_material_desc d{};
			put(m.getColor(), d.color);
			switch (m.getType())
			{
				case LAMBERTIAN: d.kind = LUZGPU_MAT_LAMBERTIAN; break;
				case EMISSIVE: d.kind = LUZGPU_MAT_EMISSIVE; break;
				case ISOTROPIC: d.kind = LUZGPU_MAT_ISOTROPIC; break;
				case METAL:
				{
					const Metal& x = static_cast<const Metal&>(m);
					d.kind = LUZGPU_MAT_METAL;
					d.roughness = x.getRoughness();
					d.conductor = x.usesConductorFresnel() ? 1u : 0u;
					put(x.getEta(), d.eta);
					put(x.getExtinctionCoefficient(), d.k);
					break;
				}
				case DIELECTRIC:
				{
					const Dielectric& x = static_cast<const Dielectric&>(m);
					d.kind = LUZGPU_MAT_DIELECTRIC;
					d.ior = x.getRefractiveIndex();
					d.roughness = x.getRoughness();
					put(x.getAbsorptionCoefficient(), d.absorption);
					break;
				}
				case HENYEY_GREENSTEIN:
					d.kind = LUZGPU_MAT_HENYEY_GREENSTEIN;
					d.anisotropy = static_cast<const HenyeyGreenstein&>(m).getAnisotropy();
					break;
				case GLOSSY:
					d.kind = LUZGPU_MAT_GLOSSY;
					d.roughness = static_cast<const Glossy&>(m).getRoughness();
					break;
				case DIFFUSE_GLOSSY:
				{
					const DiffuseGlossy& x = static_cast<const DiffuseGlossy&>(m);
					d.kind = LUZGPU_MAT_DIFFUSE_GLOSSY;
					put(x.getGlossyColor(), d.glossy_color);
					d.glossy_weight = x.getGlossyWeight();
					d.roughness = x.getGlossyRoughness();
					break;
				}
				case PRINCIPLED:
				{
					const Principled& x = static_cast<const Principled&>(m);
					d.kind = LUZGPU_MAT_PRINCIPLED;
					d.subsurface = x.getSubsurface();
					put(x.getSubsurfaceRadius(), d.subsurface_radius);
					d.subsurface_scale = x.getSubsurfaceScale();
					put(x.getSubsurfaceColor(), d.subsurface_color);
					d.subsurface_thin = x.getSubsurfaceMethod() == SUBSURFACE_THIN ? 1u : 0u;
					d.metallic = x.getMetallic();
					d.roughness = x.getRoughness();
					d.transmission = x.getTransmission();
					d.ior = x.getRefractiveIndex();
					d.clearcoat = x.getClearcoat();
					d.clearcoat_roughness = x.getClearcoatRoughness();
					d.sheen = x.getSheen();
					put(x.getAbsorptionCoefficient(), d.absorption);
					break;
				}
				default:
					throw std::runtime_error("Unsupported for the device path: a material of the base Material class");
			}
			return d;
		}

		struct	Flattener
		{
			FlatScene&	out;
			Resolver	resolve;
			std::map<const Material*, std::uint32_t>	materialIndex;
			std::map<const Mesh*, std::uint32_t>	meshIndex;
			std::map<const IESProfile*, std::uint32_t>	iesIndex;

			// each IESProfile once (lights may share one)
			std::uint32_t	ies(const IESProfile* p)
			{
				const auto it = iesIndex.find(p);
				if (it != iesIndex.end()) return it->second;
				const std::uint32_t i = static_cast<std::uint32_t>(out.ies.size());
				luzgpu_ies_desc d{};
				d.vertical_count = static_cast<std::uint32_t>(Access::verticalAngles(*p).size());
				d.horizontal_count = static_cast<std::uint32_t>(Access::horizontalAngles(*p).size());
				d.vertical = Access::verticalAngles(*p).data();
				d.horizontal = Access::horizontalAngles(*p).data();
				d.candela = Access::candela(*p).data();
				d.total_lumens = p->totalLumens();
				out.ies.push_back(d);
				iesIndex.emplace(p, i);
				return i;
			}

			std::uint32_t	material(Material* m)
			{
				if (!m) throw std::runtime_error("hittable without a material");
				const auto it = materialIndex.find(m);
				if (it != materialIndex.end()) return it->second;
				const std::uint32_t i = static_cast<std::uint32_t>(out.materials.size());
				out.materials.push_back(materialOf(*m));
				materialIndex[m] = i;
				return i;
			}

			std::uint32_t	mesh(const Mesh* m)
			{
				const auto it = meshIndex.find(m);
				if (it != meshIndex.end()) return it->second;
				if (!Access::packed(*m))
				{
					throw std::runtime_error("Unsupported for the device path: a Mesh of non-triangle hittables");
				}
				const std::vector<Triangle>& tris = Access::triangles(*m);
				std::vector<double> v(9 * tris.size()), n(9 * tris.size(), 0.0);
				bool anyNormals = false;
				for (std::size_t t = 0; t < tris.size(); t++)
				{
					put(tris[t].getVertex0(), &v[9 * t]);
					put(tris[t].getVertex1(), &v[9 * t + 3]);
					put(tris[t].getVertex2(), &v[9 * t + 6]);
					anyNormals |= Access::vertexNormals(tris[t], &n[9 * t]);
				}
				const std::uint32_t i = static_cast<std::uint32_t>(out.meshes.size());
				out.meshVertices.push_back(std::move(v));
				out.meshNormals.push_back(anyNormals ? std::move(n) : std::vector<double>());
				out.meshes.push_back(luzgpu_mesh_desc{static_cast<std::uint32_t>(tris.size()), 0u, nullptr, nullptr});
				meshIndex[m] = i;
				return i;
			}

			luzgpu_hittable_desc	hittable(const std::shared_ptr<Hittable>& hp)
			{
				const Hittable* h = resolve ? resolve(hp) : hp.get();
				luzgpu_hittable_desc d{};
				if (const auto* t = dynamic_cast<const Triangle*>(h))
				{
					d.kind = LUZGPU_HIT_TRIANGLE;
					put(t->getVertex0(), d.a);
					put(t->getVertex1(), d.b);
					put(t->getVertex2(), d.c);
					d.flags = Access::vertexNormals(*t, d.normals) ? 1u : 0u;
				}
				else if (const auto* r = dynamic_cast<const Rectangle*>(h))
				{
					if (Access::explicitAxes(*r))
					{
						throw std::runtime_error("Unsupported for the device path: a rectangle with explicit axes");
					}
					d.kind = LUZGPU_HIT_RECTANGLE;
					put(Access::transform(*r).getPosition(), d.a);
					put(Access::transform(*r).getOrientation(), d.b);
					d.size[0] = Access::width(*r);
					d.size[1] = Access::height(*r);
				}
				else if (const auto* mi = dynamic_cast<const MeshInstance*>(h))
				{
					d.kind = LUZGPU_HIT_INSTANCE;
					d.mesh = mesh(mi->getGeometry().get());
					Access::transform(*mi, d.a, d.c, d.cos_rotation, d.sin_rotation);
				}
				else if (const auto* sp = dynamic_cast<const Sphere*>(h))
				{
					d.kind = LUZGPU_HIT_SPHERE;
					put(sp->getPosition(), d.a);
					d.radius = sp->getRadius();
					d.flags = sp->isVisible() ? 0u : 1u;
					if (const IESProfile* p = Access::ies(*sp))
					{
						d.flags |= 2u;
						d.mesh = ies(p);
						put(Access::iesDirection(*sp), d.b);
						d.size[0] = Access::iesRotation(*sp);
					}
				}
				else if (const auto* cu = dynamic_cast<const Cube*>(h))
				{
					d.kind = LUZGPU_HIT_CUBE;
					put(Access::transform(*cu).getPosition(), d.a);
					Access::size(*cu, d.size);
				}
				else if (const auto* v = dynamic_cast<const ConstantVolume*>(h))
				{
					d.kind = LUZGPU_HIT_VOLUME;
					d.density = v->getDensity();
					const Hittable* b = Access::boundary(*v).get();
					if (const auto* bs = dynamic_cast<const Sphere*>(b))
					{
						d.flags = 1u;
						put(bs->getPosition(), d.a);
						d.radius = bs->getRadius();
					}
					else if (const auto* bc = dynamic_cast<const Cube*>(b))
					{
						put(Access::transform(*bc).getPosition(), d.a);
						Access::size(*bc, d.size);
					}
					else
					{
						throw std::runtime_error("Unsupported for the device path: a volume boundary other than a sphere or box");
					}
					d.material = material(Access::phase(*v));
					return d;
				}
				else if (const auto* p = dynamic_cast<const Plane*>(h))
				{
					d.kind = LUZGPU_HIT_PLANE;
					d.plane_y = Access::y(*p);
					put(Access::orientation(*p), d.b);
				}
				else if (const auto* dl = dynamic_cast<const DirectionalLight*>(h))
				{
					d.kind = LUZGPU_HIT_DIRECTIONAL;
					put(dl->getDirection(), d.a);
					if (dl->hasAtmosphereSunRadiance())
					{
						d.flags = 1u;
						const Color c = dl->getAtmosphereSunRadiance();
						d.b[0] = c.getRed();
						d.b[1] = c.getGreen();
						d.b[2] = c.getBlue();
					}
				}
				else
				{
					throw std::runtime_error("Unsupported for the device path: a top-level hittable of another kind "
						"(a bare Mesh, a BVHNode or a procedural sphere)");
				}
				d.material = material(h->getMaterial());
				return d;
			}
		};

		std::uint32_t	skyOf(SkyTypes sky)
		{
			switch (sky)
			{
				case SKY_NONE: return LUZGPU_SKY_NONE;
				case SKY_LINEAR: return LUZGPU_SKY_LINEAR;
				case SKY_ENVIRONMENT: return LUZGPU_SKY_ENVIRONMENT;
				case SKY_ATMOSPHERE: return LUZGPU_SKY_ATMOSPHERE;
			}
			throw std::runtime_error("unknown sky type");
		}

		void	check(int rc)
		{
			if (rc != LUZGPU_OK) throw std::runtime_error(luzgpu_last_error());
		}
	}

	void	flatten(const Scene& scene, FlatScene& out, Resolver resolve)
	{
		if (!scene.hasCamera())
		{
			throw std::runtime_error("No camera on the Scene.");	// Renderer.cpp:12-16 refuses it too
		}
		if (scene.getCausticsEnabled())
		{
			throw std::runtime_error("Unsupported for the device path: caustics (the photon map)");
		}
		Flattener f{out, resolve, {}, {}, {}};
		out.ies.clear();
		out.hittables.clear();
		for (const std::shared_ptr<Hittable>& h : scene.getHittables()) out.hittables.push_back(f.hittable(h));
		luzgpu_scene_desc& d = out.desc;
		d = luzgpu_scene_desc{};
		const Scene& s = scene;
		d.width = static_cast<std::uint32_t>(const_cast<Scene&>(s).getImage()->getWidth());
		d.height = static_cast<std::uint32_t>(const_cast<Scene&>(s).getImage()->getHeight());
		d.samples = static_cast<std::uint32_t>(s.getSampleCount());
		d.max_bounces = s.getMaxLightBounces();
		d.adaptive = s.getAdaptiveSampling() ? 1u : 0u;
		d.adaptive_min_samples = static_cast<std::uint32_t>(s.getAdaptiveMinSamples());
		d.adaptive_check_interval = static_cast<std::uint32_t>(s.getAdaptiveCheckInterval());
		d.adaptive_threshold = s.getAdaptiveThreshold();
		d.denoise = s.getDenoise() ? 1u : 0u;
		d.meters_per_unit = s.getMetersPerUnit();
		d.sky = skyOf(s.getRenderSky());
		d.view_transform = static_cast<std::uint32_t>(s.getViewTransform());
		d.bloom = s.getBloom() ? 1u : 0u;
		d.exposure = s.getExposure();
		d.contrast = s.getContrast();
		put(s.getBackgroundColor(), d.background);
		const Atmosphere& a = s.getAtmosphere();
		d.sun_angle = a.getSunAngle();
		d.earth_radius = a.getEarthRadius();
		d.atmosphere_radius = a.getAtmosphereRadius();
		d.rayleigh_height = Access::hR(a);
		d.mie_height = Access::hM(a);
		d.atmosphere_samples = Access::samples(a);
		d.atmosphere_light_samples = Access::lightSamples(a);
		d.stars_brightness = a.getStarsBrightness();
		d.sun_radiance_scale = a.getSunRadianceScale();
		if (s.hasEnvironmentMap())
		{
			const EnvironmentMap& e = *s.getEnvironmentMap();
			const std::vector<Color>& px = Access::pixels(e);
			out.environment.resize(3 * px.size());
			for (std::size_t i = 0; i < px.size(); i++) put(px[i], &out.environment[3 * i]);
			d.env_width = static_cast<std::uint32_t>(e.getWidth());
			d.env_height = static_cast<std::uint32_t>(e.getHeight());
			d.env_texels = out.environment.data();
			d.env_strength = s.getEnvironmentStrength();
			d.env_rotation = s.getEnvironmentRotation();
			d.env_lighting = s.getEnvironmentLighting() ? 1u : 0u;
		}
		const Camera c = s.getActiveCamera();
		put(c.getPosition(), d.camera.position);
		put(c.getDirection(), d.camera.direction);
		put(c.getUpDirection(), d.camera.up);
		d.camera.focal_length = c.getFocalLengthMeters();
		d.camera.sensor_width = c.getSensorWidthMeters();
		d.camera.sensor_height = c.getSensorHeightMeters();
		d.camera.f_number = c.getFNumber();
		d.camera.focus_distance = c.getFocusDistanceMeters();
		d.camera.pinhole = c.getPinhole() ? 1u : 0u;
		for (std::size_t i = 0; i < out.meshes.size(); i++)
		{
			out.meshes[i].vertices = out.meshVertices[i].data();
			out.meshes[i].normals = out.meshNormals[i].empty() ? nullptr : out.meshNormals[i].data();
		}
		d.material_count = static_cast<std::uint32_t>(out.materials.size());
		d.mesh_count = static_cast<std::uint32_t>(out.meshes.size());
		d.hittable_count = static_cast<std::uint32_t>(out.hittables.size());
		d.materials = out.materials.data();
		d.meshes = out.meshes.data();
		d.hittables = out.hittables.data();
		d.ies_count = static_cast<std::uint32_t>(out.ies.size());
		d.ies_profiles = out.ies.empty() ? nullptr : out.ies.data();
	}

	void	manageThreadsOnDevice(Scene& scene, int device, Resolver resolve)
	{
		// the render seed of _manageThreads (Threads.cpp:383-388)
		const std::uint32_t renderSeed = hasRandomSeed()
			? static_cast<std::uint32_t>(randomSeedValue())
			: randomEngine.integer();
		FlatScene flat;
		flatten(scene, flat, resolve);
		scene.resetRenderStats();
		scene.clearDenoisedImage();
		scene.clearDenoiseBuffers();
		luzgpu_ctx* ctx = nullptr;
		luzgpu_scene* dev = nullptr;
		struct Cleanup
		{
			luzgpu_ctx*& ctx;
			luzgpu_scene*& dev;
			~Cleanup()
			{
				luzgpu_scene_release(dev);
				luzgpu_destroy(ctx);
			}
		} cleanup{ctx, dev};
		check(luzgpu_create(&ctx, device));
		check(luzgpu_scene_create(ctx, &flat.desc, &dev));
		luzgpu_settings st{};
		st.seed = renderSeed;
		st.samples = -1;
		st.max_bounces = -1;
		st.sample_count = -1;
		st.integrator = LUZGPU_INTEGRATOR_AUTO;
		st.adaptive = LUZGPU_ADAPTIVE_SCENE;
		st.postprocess = 1;	// applyPostProcessing on both images, as _manageThreads ends
		const std::size_t w = flat.desc.width, h = flat.desc.height;
		std::vector<float> rgb(w * h * 3), den;
		luzgpu_stats stats{};
		if (flat.desc.denoise)
		{
			den.resize(rgb.size());
			check(luzgpu_render_denoised(ctx, dev, &st, nullptr, rgb.data(), den.data(), nullptr, &stats));
		}
		else
		{
			check(luzgpu_render(ctx, dev, &st, rgb.data(), &stats));
		}
		// a post-processed image holds display sRGB unless the view is raw (Image.cpp:540-600)
		const ImageColorEncoding encoding = scene.getViewTransform() == ViewTransform::Raw
			? ImageColorEncoding::SceneLinearACEScg : ImageColorEncoding::DisplayEncodedSRGB;
		auto fill = [&](Image& img, const std::vector<float>& px) {
			for (std::size_t y = 0; y < h; y++)
			{
				for (std::size_t x = 0; x < w; x++)
				{
					const float* p = &px[(y * w + x) * 3];
					img.setPixelUnchecked(x, y, Color(p[0], p[1], p[2]));
				}
			}
			img.setColorEncoding(encoding);
		};
		fill(*scene.getImage(), rgb);
		if (flat.desc.denoise)
		{
			auto image = std::make_unique<Image>(w, h);
			image->initialize();
			fill(*image, den);
			scene.setDenoisedImage(std::move(image));
		}
		SceneRenderStats rs;
		rs.renderMS = stats.render_ms;
		rs.renderedSamples = static_cast<std::size_t>(stats.rendered_samples);
		rs.averageSamplesPerPixel = w * h ? static_cast<double>(stats.rendered_samples) / static_cast<double>(w * h) : 0.0;
		rs.denoiseMS = stats.denoise_ms;
		scene.setRenderStats(rs);
	}
}
