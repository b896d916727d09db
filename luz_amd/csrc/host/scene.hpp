// Host-side scene model: the flattened form of the reference's polymorphic Scene
// (include/luz/Scene/Scene.hpp:26-182) that crosses into the device library.
// Everything here is FP64, as in the reference, so host-side derived data (BVH
// topology, light CDFs, environment CDF) can be built bit-identically to it.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

struct	luzgpu_scene_desc;	// include/luzgpu.h

namespace luzgpu
{
	using dvec3 = std::array<double, 3>;

	inline dvec3	operator+(const dvec3& a, const dvec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
	inline dvec3	operator-(const dvec3& a, const dvec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
	// Vector3 * double is computed as f * v (Vector3.hpp:82-89).
	inline dvec3	operator*(const dvec3& a, double f) { return {f * a[0], f * a[1], f * a[2]}; }
	inline double	dot(const dvec3& a, const dvec3& b) { return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]); }
	inline dvec3	cross(const dvec3& a, const dvec3& b)
	{
		return {(a[1] * b[2]) - (a[2] * b[1]), (a[2] * b[0]) - (a[0] * b[2]), (a[0] * b[1]) - (a[1] * b[0])};
	}

	constexpr double	kTMin = 0.001;	// Defaults.hpp T_MIN
	constexpr double	kPi = 3.14159265358979323846;

	enum MaterialKind : std::uint32_t
	{
		MAT_LAMBERTIAN = 1,
		MAT_METAL = 2,
		MAT_DIELECTRIC = 3,
		MAT_EMISSIVE = 4,
		MAT_ISOTROPIC = 5,	// Isotropic.cpp:15-24
		MAT_HG = 6,			// HenyeyGreenstein.cpp:54-63
		MAT_GLOSSY = 7,		// Glossy.cpp
		MAT_DIFFUSE_GLOSSY = 8,	// DiffuseGlossy.cpp
		MAT_PRINCIPLED = 9,	// Principled.cpp (no subsurface)
	};

	struct	Material
	{
		MaterialKind	kind = MAT_LAMBERTIAN;
		dvec3	color = {0.6, 0.6, 0.6};	// albedo / Schlick f0 / dielectric tint / radiance
		double	roughness = 0.0;
		double	ior = 1.5;
		bool	conductor = false;
		dvec3	eta = {1.0, 1.0, 1.0};
		dvec3	k = {0.0, 0.0, 0.0};
		dvec3	absorption = {0.0, 0.0, 0.0};	// dielectric interior absorption, 1/m
		double	anisotropy = 0.0;			// Henyey-Greenstein g
		dvec3	glossyColor = {1.0, 1.0, 1.0};
		double	glossyWeight = 0.5;
		double	metallic = 0.0;
		double	transmission = 0.0;
		double	clearcoat = 0.0;
		double	clearcoatRoughness = 0.03;
		double	sheen = 0.0;
		// principled subsurface (Principled.cpp:39-88, 450-454): the lobe weight, the Burley
		// diffusion radius per channel times its scale in metres, the tint, thin or Burley
		double	subsurface = 0.0;
		dvec3	subsurfaceRadius = {1.0, 1.0, 1.0};
		double	subsurfaceScale = 0.001;
		dvec3	subsurfaceColor = {1.0, 1.0, 1.0};
		bool	subsurfaceThin = false;
		std::string	name;
	};

	struct	Triangle
	{
		dvec3	v0, v1, v2;
		dvec3	n0{}, n1{}, n2{};
		bool	hasNormals = false;
	};

	struct	Mesh
	{
		std::vector<Triangle>	triangles;	// after the reference's zero-area removal
		std::string	source;
	};

	enum HittableKind : std::uint32_t
	{
		HIT_TRIANGLE = 0,
		HIT_RECTANGLE = 1,
		HIT_INSTANCE = 2,
		HIT_SPHERE = 3,
		HIT_CUBE = 4,
		HIT_VOLUME = 5,
		HIT_PLANE = 6,			// unbounded
		HIT_DIRECTIONAL = 7,	// unbounded, never hit
	};

	// Sphere::hit / hitAny (Sphere.cpp:227-233, 271-277): an invisible sphere is never hit
	// by any ray, but stays in the acceleration structure and, emissive, in the light list.
	// An IES-profiled sphere light (Sphere::setIESProfile, Sphere.cpp:188-202) scales the
	// radiance of its light samples by the profile's relative intensity towards the shaded
	// point, about its vertical axis `iesDirection` and turned by `iesRotation` degrees.
	struct	Sphere
	{
		dvec3	position{};
		double	radius = 1.0;
		bool	visible = true;
		std::int32_t	ies = -1;	// index into Scene::iesProfiles
		dvec3	iesDirection = {0.0, -1.0, 0.0};
		double	iesRotation = 0.0;
	};

	// IESProfile (IESProfile.cpp:133-240): vertical and horizontal angles in degrees (strictly
	// increasing), candela values (times the file's multiplier) for each horizontal angle over
	// all vertical angles, and the luminous flux integrated over the sphere.
	struct	IesProfile
	{
		std::vector<double>	vertical, horizontal, candela;
		double	totalLumens = 0.0;
		double	meanCandela = 0.0;	// totalLumens / 4 pi
	};

	// Cube.cpp:72-95: six axis-aligned rectangles around the position (the orientation
	// of the cube's transform is not used by its faces).
	struct	Cube
	{
		dvec3	position{};
		double	width = 1.0, height = 1.0, depth = 1.0;
	};

	// Plane.cpp: the plane through (0, y, 0) with normal `orientation` (not normalised).
	struct	Plane
	{
		double	y = 0.0;
		dvec3	orientation = {0.0, -1.0, 0.0};
	};

	// ConstantVolume (ConstantVolume.cpp) over a sphere or box boundary; density in scene
	// units (Objects.cpp:733-739), phase function = the hittable's material.
	struct	Volume
	{
		bool	sphere = false;
		dvec3	position{};
		double	radius = 1.0;
		dvec3	size = {1.0, 1.0, 1.0};
		double	density = 0.1;
	};

	// With the atmosphere sky the scene's first directional light is also the atmosphere's sun
	// (Scene::syncAtmosphereSunDirection, Scene.cpp:457-475): the sun shines along its
	// direction with its solar radiance (a `solar=` light, DirectionalLight::
	// setAtmosphereSunRadiance) or else its emission.
	struct	DirectionalLight
	{
		dvec3	direction = {0.0, -1.0, 0.0};	// normalised (DirectionalLight.cpp:9-17)
		bool	hasSunRadiance = false;
		dvec3	sunRadiance{};
	};

	struct	Rectangle
	{
		dvec3	position{};
		dvec3	orientation = {0.0, 0.0, 1.0};
		double	width = 1.0;
		double	height = 1.0;
	};

	struct	Instance
	{
		std::uint32_t	mesh = 0;
		dvec3	position{};
		dvec3	rotationDegrees{};
		dvec3	scale = {1.0, 1.0, 1.0};
		// an instance handed over from an in-memory Scene carries the cosines and sines its
		// MeshInstance::InstanceTransform holds (used instead of rotationDegrees)
		bool	hasTrig = false;
		dvec3	cosRotation = {1.0, 1.0, 1.0};
		dvec3	sinRotation{};
	};

	// One entry per reference top-level hittable, in Scene::getHittables() order.
	struct	Hittable
	{
		HittableKind	kind = HIT_TRIANGLE;
		std::uint32_t	index = 0;		// into triangles / rectangles / instances
		std::uint32_t	material = 0;
	};

	enum SkyKind : std::uint32_t
	{
		SKY_NONE = 0,
		SKY_ENVIRONMENT = 1,
		SKY_LINEAR = 2,		// white-to-blue gradient (Atmospherics.cpp:44-53)
		SKY_ATMOSPHERE = 3,	// the reference's default (Scene.cpp:118)
	};

	struct	Camera
	{
		dvec3	position{};
		dvec3	direction = {0.0, 0.0, -1.0};
		dvec3	up = {0.0, 1.0, 0.0};
		double	focalLength = 0.050;
		double	sensorWidth = 0.036;
		double	sensorHeight = 0.02025;
		double	fNumber = 8.0;
		bool	pinhole = false;
		double	focusDistance = 10.0;
	};

	struct	Environment
	{
		std::uint32_t	width = 0;
		std::uint32_t	height = 0;
		std::vector<dvec3>	pixels;		// scene-linear ACEScg, row 0 = top (v = 1)
		double	strength = 1.0;
		double	rotation = 0.0;
		bool	lighting = true;
	};

	struct	Settings
	{
		std::uint32_t	width = 1920;
		std::uint32_t	height = 1080;
		std::uint32_t	samples = 48;
		std::int32_t	maxBounces = 6;
		bool	adaptive = true;
		std::uint32_t	adaptiveMinSamples = 128;	// Defaults.hpp:13-16
		std::uint32_t	adaptiveCheckInterval = 32;
		double	adaptiveThreshold = 0.03;
		bool	denoise = true;
		// display post-processing (Scene.cpp:99-139 defaults): view transform 0 standard,
		// 1 AgX, 2 ACES, 3 raw
		int	viewTransform = 1;
		bool	bloom = true;
		double	exposure = 0.0;
		double	contrast = 1.0;
		SkyKind	sky = SKY_ATMOSPHERE;
		bool	skySet = false;
		// the atmosphere sky (Atmosphere.cpp:67-82 defaults, Defaults.hpp:22-25)
		double	sunAngle = -0.4;
		double	earthRadius = 6360e3;
		double	atmosphereRadius = 6420e3;
		double	atmosphereHR = 7994.0;
		double	atmosphereHM = 1200.0;
		int	atmosphereSamples = 16;
		int	atmosphereLightSamples = 8;
		double	starsBrightness = 0.5;
		double	sunRadianceScale = 1.0;
		dvec3	background{};
		double	metersPerUnit = 1.0;
	};

	struct	Scene
	{
		Settings	settings;
		Camera	camera;
		bool	hasCamera = false;
		std::vector<Material>	materials;
		std::vector<Triangle>	triangles;
		std::vector<Rectangle>	rectangles;
		std::vector<Mesh>	meshes;
		std::vector<Instance>	instances;
		std::vector<Sphere>	spheres;
		std::vector<Cube>	cubes;
		std::vector<Plane>	planes;
		std::vector<Volume>	volumes;
		std::vector<DirectionalLight>	directionals;
		std::vector<IesProfile>	iesProfiles;
		std::vector<Hittable>	hittables;
		Environment	environment;
		bool	hasEnvironment = false;
	};

	// Builds the host scene from the C ABI's flattened records (luzgpu_scene_desc,
	// include/luzgpu.h); throws std::invalid_argument on inconsistent records.
	Scene	sceneFromDesc(const ::luzgpu_scene_desc& desc);

	// Reads the subset of the reference's .luz format listed in scene_file.cpp.
	// Throws std::runtime_error with the reference's style of message on bad input.
	Scene	readSceneFile(const std::string& path);
	Mesh	readObjMesh(const std::string& path);
	Environment	readHDREnvironment(const std::string& path);
	IesProfile	readIesProfile(const std::string& path);
	// IESProfile::candelaAt (IESProfile.cpp:245-300)
	double	iesCandelaAt(const IesProfile& profile, double verticalDegrees, double horizontalDegrees);
	double	luminance(const dvec3& c);
	// ColorScience::blackbody / ::wavelength (ColorScience.cpp:182-214): ACEScg colours of
	// unit maximum channel; throw outside 1000-40000 K / 360-830 nm
	dvec3	blackbodyColor(double kelvin);
	dvec3	wavelengthColor(double nanometers);
}
