// Device-resident scene layout (HBM), shared by the host packer and the HIP kernels.
// Plain PODs only: every array is a flat, 16-byte aligned buffer.
//
//   BvhNode   64 B  two child boxes (fp32, rounded outward) + two child references.
//                   One cache-line fetch tests both children of a binary node; leaves
//                   are folded into their parent's reference (count > 0).
//   GpuTri    48 B  v0 | e1 | e2 in fp32 (e1, e2 formed in FP64, then rounded), with
//                   the triangle's index in its mesh (after degenerate removal) in v0.w.
//   Hittable  8 B   kind:4 | index:28, material id — the reference's top-level list.
#pragma once

#include <cstdint>

namespace luzgpu
{
	struct alignas(16) BvhNode
	{
		float	lo0[3];
		float	hi0[3];
		float	lo1[3];
		float	hi1[3];
		std::int32_t	child[2];	// count == 0: node index; count > 0: first primitive
		std::int32_t	count[2];	// 0 = internal child, > 0 = leaf primitive count, -1 = empty
	};
	static_assert(sizeof(BvhNode) == 64, "BvhNode must be 64 bytes");

	// 4-wide node in one 64-byte line. Child boxes are quantised to 8 bits per plane
	// against the node's own box: plane = fma(q, 2^exp, origin), with q chosen on the host so
	// that the decoded fp32 box contains the child's box (conservative by construction).
	struct alignas(64) Bvh4Node
	{
		float	origin[3];
		std::int8_t	exp[3];
		std::uint8_t	pad0;
		std::uint8_t	qlo[3][4];	// [axis][child]
		std::uint8_t	qhi[3][4];
		std::int32_t	child[4];	// count == 0: node index; count 1..16: first primitive
		std::uint8_t	count[4];	// 0 internal, 1..16 leaf, 255 empty
		std::uint32_t	pad1;
	};
	static_assert(sizeof(Bvh4Node) == 64, "Bvh4Node must be 64 bytes");

	struct alignas(16) GpuTri
	{
		float	v0[3];
		std::int32_t	prim;
		float	e1[3];
		float	pad1;
		float	e2[3];
		float	pad2;
	};
	static_assert(sizeof(GpuTri) == 48, "GpuTri must be 48 bytes");

	enum : std::uint32_t
	{
		GPU_HIT_TRIANGLE = 0,
		GPU_HIT_RECTANGLE = 1,
		GPU_HIT_INSTANCE = 2,
		GPU_HIT_SPHERE = 3,
		GPU_HIT_CUBE = 4,		// six GpuRect faces at cubeFaces[6 * index]
		GPU_HIT_VOLUME = 5,
		GPU_HIT_PLANE = 6,		// unbounded: tested after the BVH
		GPU_HIT_DIRECTIONAL = 7,	// unbounded light, never hit
		GPU_HIT_HIDDEN_SPHERE = 8,	// BVH leaf entries only: an invisible sphere, bounded but
									// never hit (its hittable and light entries say SPHERE)
	};
	constexpr std::uint32_t	kKindShift = 28;
	constexpr std::uint32_t	kIndexMask = (1u << kKindShift) - 1u;

	// Guide table of an environment map's CDF: for bucket g of G, the first texel whose
	// running weight reaches (g / G) x total (std::lower_bound), g = 0..G. A sample u falls
	// in bucket floor(u G), and its texel lies between that bucket's entry and the next one's,
	// so the search covers that range only and finds the texel the whole-table search finds.
	// G is the largest power of two at most a quarter of the texels; maps of fewer than 8
	// texels are searched whole. The table is texel-aligned with the CDF (its first G + 1
	// entries are used).
	constexpr std::uint32_t	kEnvGuideMinTexels = 8;
	constexpr std::uint32_t	envGuideBuckets(std::uint32_t texels)
	{
		return 1u << (31 - __builtin_clz(texels / 4u));
	}

	struct GpuHittable
	{
		std::uint32_t	kindIndex;	// kind << 28 | index
		std::uint32_t	material;
	};

	struct alignas(16) GpuSphere
	{
		float	center[3];
		float	radius;
	};

	// Plane.cpp: plane through (0, y, 0) with the (unnormalised) orientation as normal.
	struct alignas(16) GpuPlane
	{
		float	point[3];
		float	pad0;
		float	orientation[3];
		float	pad1;
	};

	// ConstantVolume.cpp: boundary (sphere: center/radius; box: lo/hi) and -1/density.
	struct alignas(16) GpuVolume
	{
		float	center[3];
		float	radius;
		float	lo[3];
		float	negInvDensity;
		float	hi[3];
		std::uint32_t	sphere;
	};

	struct alignas(16) GpuRect
	{
		float	position[3];
		float	halfWidth;
		float	normal[3];
		float	halfHeight;
		float	widthAxis[3];
		float	area;
		float	heightAxis[3];
		float	width;
		float	height;
		float	pad[3];
	};

	// Traversal reads the first 64 bytes (the 3x4 map and the mesh links) in one group of
	// 16-byte loads; shading reads the normal map.
	struct alignas(16) GpuInstance
	{
		float	worldToLocal[12];	// rows of a 3x4 affine map (inverse rotate, inverse scale)
		std::int32_t	blasRoot;	// global Bvh4Node index of the mesh's root
		std::int32_t	triBase;	// first GpuTri of the mesh
		std::int32_t	normalBase;	// first vertex-normal triple, -1 if the mesh has none
		std::int32_t	pad0;
		float	normalToWorld[9];	// rotate(n / scale), renormalised after use
		float	pad[3];
	};
	static_assert(sizeof(GpuInstance) == 112, "GpuInstance must be 112 bytes");

	enum : std::uint32_t
	{
		GPU_MAT_LAMBERTIAN = 1,
		GPU_MAT_METAL = 2,
		GPU_MAT_DIELECTRIC = 3,
		GPU_MAT_EMISSIVE = 4,
		GPU_MAT_ISOTROPIC = 5,
		GPU_MAT_HG = 6,
		GPU_MAT_GLOSSY = 7,
		GPU_MAT_DIFFUSE_GLOSSY = 8,
		GPU_MAT_PRINCIPLED = 9,
	};

	struct alignas(16) GpuMaterial
	{
		std::uint32_t	kind;
		float	color[3];		// albedo / f0 / tint / emitted radiance
		float	alpha;			// GGX alpha = max(1e-4, roughness^2)
		float	roughness;
		float	ior;
		std::uint32_t	flags;	// bit0 conductor Fresnel, bit1 medium absorption
		float	eta[3];
		float	anisotropy;		// Henyey-Greenstein g
		float	k[3];
		float	pad1;
		float	absorption[3];
		float	pad2;
		// glossy / diffuse-glossy / principled (Glossy.cpp, DiffuseGlossy.cpp, Principled.cpp)
		float	glossyColor[3];
		float	glossyWeight;	// diffuse-glossy mix
		float	metallic;
		float	transmission;	// principled transmission (1 - alpha folded in)
		float	clearcoat;
		float	clearcoatAlpha;
		float	sheen;
		float	diffuseScale;	// principled diffuseScale(), Principled.cpp:141-156
		float	subsurface;		// principled subsurface weight (the diffuse lobe keeps 1 - subsurface)
		float	subsurfaceRadius;	// Burley diffusion length: max channel of radius * scale, metres
		float	lobe[8];		// principled component weights / total (diffuse, subsurface, sheen,
								// dielectric specular, metal, transmission, clearcoat), Principled.cpp:158-202
		float	subsurfaceAlbedo[3];	// clamp01(base colour) * clamp01(subsurface colour)
		float	pad4;
	};
	static_assert(sizeof(GpuMaterial) == 176, "GpuMaterial must be 176 bytes");

	enum : std::uint32_t
	{
		GPU_MATF_CONDUCTOR = 1u,
		GPU_MATF_ABSORPTION = 2u,
		GPU_MATF_SMOOTH_TRANSMISSION = 4u,	// principled isSmoothFullTransmission (Principled.cpp:107-124)
		GPU_MATF_THIN_SUBSURFACE = 8u,		// principled usesThinSubsurface (Principled.cpp:550-556)
		GPU_MATF_BURLEY_SUBSURFACE = 16u,	// subsurface > 0 with the Burley method: exits are relocated
	};

	struct alignas(16) GpuLight
	{
		std::uint32_t	kind;		// GPU_HIT_*: triangle, rectangle, sphere, directional, instance; others never sample
		std::uint32_t	index;		// into top-level triangles / rectangles / spheres / instances
		std::uint32_t	hittable;
		std::uint32_t	material;
		float	emitted[3];
		float	area;
		float	selectionProbability;	// weight / total (or 1/n when unweighted)
		float	direction[3];		// directional: the sampled direction (minus the light's)
		// emissive mesh instance: its span of the triangle-area prefix sums (FP64) and of the
		// matching triangle slots, and its local -> world map (rotate(p * scale) + position)
		std::uint32_t	triBase;
		std::uint32_t	triCount;
		// IES-profiled sphere light: 1 + the offset of its profile in the IES table, else 0.
		// A profile is [mean candela, unit vertical axis xyz, rotation (degrees), vertical
		// count nV, horizontal count nH, nV vertical angles, nH horizontal angles, nH x nV
		// candela values], all FP64.
		std::uint32_t	ies;
		std::uint32_t	pad;
		float	localToWorld[12];
	};
	static_assert(sizeof(GpuLight) == 112, "GpuLight must be 112 bytes");

	// Everything the per-pixel camera needs, in FP64 exactly as the reference's
	// RenderCamera (src/renderer/RendererInternal.hpp:13-26).
	struct GpuCamera
	{
		double	position[3];
		double	u[3];
		double	v[3];
		double	horizontal[3];
		double	vertical[3];
		double	lowerLeftCorner[3];
		double	lensRadius;
		double	inverseWidthMinusOne;
		double	inverseHeightMinusOne;
	};

	enum : std::uint32_t
	{
		GPU_SKY_NONE = 0,
		GPU_SKY_ENVIRONMENT = 1,
		GPU_SKY_LINEAR = 2,
		GPU_SKY_ATMOSPHERE = 3,
	};

	// The atmosphere sky's parameters (Atmosphere.hpp:52-63) with the derived values its
	// march uses: unit sun direction, sun radiance x its scale, 1 / scale heights.
	struct	GpuAtmosphere
	{
		double	sunDirection[3];
		double	sunRadiance[3];
		double	earthRadius;		// metres
		double	atmosphereRadius;	// metres
		double	inverseHR;
		double	inverseHM;
		double	starsBrightness;
		double	metersPerUnit;
		std::uint32_t	samples;
		std::uint32_t	lightSamples;
	};

	// Scalars of one render: all device pointers live in GpuSceneView.
	struct GpuParams
	{
		GpuCamera	camera;
		std::uint32_t	width;
		std::uint32_t	height;
		std::uint32_t	renderSeed;
		std::int32_t	maxBounces;
		std::uint32_t	sampleBegin;
		std::uint32_t	sampleCount;	// samples in this launch
		std::uint32_t	totalSamples;	// divisor for the final mean
		std::uint32_t	sky;
		float	background[3];
		float	backgroundLuminance;
		std::uint32_t	lightCount;
		std::int32_t	tlasRoot;		// -1: no bounded hittables
		std::uint32_t	envWidth;
		std::uint32_t	envHeight;
		float	envStrength;
		float	envRotation;
		std::uint32_t	envLighting;
		std::uint32_t	hasEnvironment;
		double	envTotalWeight;
		float	metersPerUnit;
		std::uint32_t	unboundedCount;	// hittables without a box (planes), tested after the BVH
		std::uint32_t	hasVolumes;		// rays carry a volume free-flight sample
		// where this scene's tables start in the (possibly multi-scene) device arrays
		std::uint32_t	lightBase;		// lights / lightCdf
		std::uint32_t	lightWeighted;	// lightCdf holds weights (else uniform selection)
		std::uint32_t	envBase;		// envPixels / envPdf / envCdf texel
		std::uint32_t	unboundedBase;
		std::uint32_t	denoiseStreams;	// 1: sample k draws from stream (k % 2) + 1 (Threads.cpp:594-598)
		GpuAtmosphere	atmosphere;		// sky == GPU_SKY_ATMOSPHERE
	};

	// Scene content classes the kernels are specialised on (a scene without them runs
	// kernels that do not carry their code or registers).
	constexpr int	kFeatAnalytic = 1;	// spheres, cubes, volumes, planes, directional and mesh-instance lights
	constexpr int	kFeatRich = 2;		// glossy, diffuse-glossy, principled, isotropic, Henyey-Greenstein; the atmosphere sky
	constexpr int	kFeatAll = 3;
	constexpr int	kFeatSubsurface = 4;	// Burley subsurface exits (a closest-hit probe ray while shading)
	constexpr int	kFeatEvery = 7;

	// Per-scene traversal entry of a multi-scene launch: TLAS root and the unbounded
	// (plane) list as base << 16 | count.
	struct	GpuSceneSlot
	{
		std::int32_t	tlasRoot;
		std::uint32_t	unbounded;
	};
}
