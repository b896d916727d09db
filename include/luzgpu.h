/*
 * luzgpu — C ABI of the MI355X radiance path for the Luz renderer.
 *
 * Drop-in boundary. The reference has no FFI: its seam is the C++ call
 * `bool Renderer::render(Scene&)` (include/luz/Renderer/Renderer.hpp:7), whose worker
 * loop `Renderer::internal::_manageThreads` (src/renderer/Threads.cpp:375-548) runs
 * `_threadRender` (Threads.cpp:551-641) over 16-pixel blocks. These entry points
 * replace that worker loop: the scene (as loaded by SceneFile::read,
 * include/luz/SceneFile/SceneFile.hpp:8) is uploaded once, then each render call
 * produces the raw scene-linear image that `_manageThreads` leaves in
 * Scene::getImage() before post-processing.
 *
 * Conventions: every function returns an int status (0 = OK, negative = error) and
 * never throws; the message of the last failure on the calling thread is available
 * from luzgpu_last_error(). Output buffers are caller-owned. A context owns one HIP
 * device and stream; calls on one context must be serialised by the caller (the
 * reference's render is not re-entrant either: global render seed, Sampler.cpp:28).
 */
#ifndef LUZGPU_H
#define LUZGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUZGPU_OK 0
#define LUZGPU_E_INVALID (-1)      /* bad argument (reference: std::invalid_argument)   */
#define LUZGPU_E_UNSUPPORTED (-2)  /* scene feature outside the device path             */
#define LUZGPU_E_OOM (-3)          /* device or host allocation failed                  */
#define LUZGPU_E_DEVICE (-4)       /* HIP runtime error                                 */
#define LUZGPU_E_IO (-5)           /* file missing or malformed (reference: runtime_error) */

#define LUZGPU_INTEGRATOR_WAVEFRONT 0u
#define LUZGPU_INTEGRATOR_MEGAKERNEL 1u   /* the plain kernel: one thread per path                */
#define LUZGPU_INTEGRATOR_AUTO 2u         /* the plain kernel for calls of at most 4M paths        *
                                           * (pixels x samples, summed over a batch's scenes),     *
                                           * where per-bounce launches cost more than they save;   *
                                           * the wavefront above                                    */

typedef struct luzgpu_ctx luzgpu_ctx;
typedef struct luzgpu_scene luzgpu_scene;

/* Render settings. Fields left at their "scene" value (< 0 / 0) use what the scene
 * file set, as the reference's FlagsParser overrides only the flags it is given
 * (src/cli/FlagsParser.cpp:87-117). */
typedef struct luzgpu_settings
{
	uint32_t	seed;            /* --seed: Sampler render seed = (uint32_t)seed, Threads.cpp:383-388 */
	int32_t		samples;         /* spp, < 0: scene value (Scene::getSampleCount)                     */
	int32_t		max_bounces;     /* < 0: scene value (Scene::getMaxLightBounces)                      */
	uint32_t	sample_begin;    /* first sample index of this call (sample-range sharding)           */
	int32_t		sample_count;    /* samples in this call, < 0: samples - sample_begin                 */
	uint32_t	integrator;      /* LUZGPU_INTEGRATOR_*                                               */
	uint32_t	output_on_device;/* 1: rgb_out is a device pointer on the context's device            */
	int32_t		width;           /* <= 0: scene value (--resolution, FlagsParser.cpp)                 */
	int32_t		height;          /* <= 0: scene value                                                 */
	uint32_t	count_visits;    /* 1: also count BVH nodes / primitives visited (slower)             */
	uint32_t	stage_times;     /* 1: time each wavefront stage with HIP events                      */
	/* wavefront launch shape (0 = built-in default); no effect on the image */
	uint32_t	traversal_grid;  /* persistent blocks per traversal launch, 1..2048 (default 1280)    */
	int32_t		extend_refill;   /* idle lanes (1..64) at which a wave takes new closest-hit rays;    *
	                              * 64 (default): whole-wave claims, each lane traverses to the end   */
	int32_t		shadow_refill;   /* the same for shadow rays (default 64)                             */
	uint32_t	fetch_chunk;     /* most rays a wave claims per queue atomic, 64..1024 (default 256;
	                              * a claim is an eighth of a wave's even share of the queue,
	                              * rounded up to 64, at most this)                             */
	uint32_t	lds_stack;       /* LDS traversal-stack entries per lane: 8, 16 (default) or 24       */
	uint32_t	autotune;        /* 1: use the context's tuned launch shape for these scenes and      *
	                              * settings, tuning it first if there is none (luzgpu_autotune);     *
	                              * the fields above are then ignored                                 */
	/* adaptive per-pixel stopping (_threadRender, Threads.cpp:551-583; rule :194-262). It applies
	 * when enabled and adaptive_min_samples < samples (adaptiveSamplingCanStop, :265-271), and
	 * then needs the whole sample range in one call. */
	uint32_t	adaptive;        /* LUZGPU_ADAPTIVE_*: the scene's adaptive= setting, or forced       */
	uint32_t	adaptive_min_samples;    /* 0: scene value (adaptiveminsamples, default 128)      */
	uint32_t	adaptive_check_interval; /* 0: scene value (adaptivecheckinterval, default 32)    */
	uint32_t	reserved0;
	double		adaptive_threshold;      /* <= 0: scene value (adaptivethreshold, default 0.03)   */
	uint32_t	postprocess;     /* 1: rgb_out (and the denoised image) receive the scene's display  *
	                              * post-processing of the raw image, as Renderer::render leaves     *
	                              * Scene::getImage() (applyPostProcessing, Threads.cpp:77-99)       */
	uint32_t	reserved1;
} luzgpu_settings;

enum
{
	LUZGPU_ADAPTIVE_SCENE = 0,   /* as the scene file says (--adaptive / --no-adaptive unset) */
	LUZGPU_ADAPTIVE_OFF = 1,     /* --no-adaptive (FlagsParser)                               */
	LUZGPU_ADAPTIVE_ON = 2       /* --adaptive                                                */
};

/* Launch shape chosen by luzgpu_autotune, with the calibration that chose it. */
typedef struct luzgpu_tuning
{
	uint32_t	traversal_grid;
	int32_t		extend_refill;
	int32_t		shadow_refill;
	uint32_t	fetch_chunk;
	uint32_t	lds_stack;
	uint32_t	trials;              /* candidate shapes timed                                  */
	uint64_t	calibration_paths;   /* paths of the calibration render (one pass)              */
	double		default_ms;          /* calibration render with the built-in defaults           */
	double		tuned_ms;            /* ... with the chosen shape                               */
	double		tune_ms;             /* wall time of the whole search                           */
} luzgpu_tuning;

typedef struct luzgpu_stats
{
	uint64_t	rendered_samples;    /* SceneRenderStats::renderedSamples (Scene.hpp:16-24) */
	uint64_t	closest_hit_rays;
	uint64_t	shadow_rays;
	double		render_ms;           /* SceneRenderStats::renderMS (all device work)       */
	double		kernel_ms;           /* integrator kernels only (HIP events)               */
	uint64_t	nodes_visited;       /* with count_visits                                  */
	uint64_t	prims_tested;        /* with count_visits                                  */
	uint32_t	kernel_launches;
	uint32_t	extend_launches;     /* wavefront: closest-hit (extend) launches          */
	uint64_t	closest_nodes;       /* with count_visits: nodes / primitives of closest-hit rays */
	uint64_t	closest_prims;
	double		extend_ms;           /* wavefront, with stage_times                        */
	double		shade_ms;
	double		shadow_ms;
	double		tune_ms;             /* autotune: search time spent by this call (not in render_ms) */
	double		denoise_ms;          /* luzgpu_render_denoised: NFOR (SceneRenderStats::denoiseMS) */
} luzgpu_stats;

/* Display post-processing (applyPostProcessing, Threads.cpp:77-99; Scene defaults
 * Scene.cpp:109-139): exposure in stops, bloom on / off, the view transform and the display
 * contrast (1 = the default curve-free path). */
typedef struct luzgpu_post_settings
{
	double		exposure;
	double		contrast;
	uint32_t	bloom;
	uint32_t	view_transform;   /* LUZGPU_VIEW_* */
} luzgpu_post_settings;

enum
{
	LUZGPU_VIEW_STANDARD = 0,
	LUZGPU_VIEW_AGX = 1,         /* the reference's default */
	LUZGPU_VIEW_ACES = 2,
	LUZGPU_VIEW_RAW = 3
};

/* Denoise::NFORSettings (NFOR.hpp:39-48); NULL = the reference's defaults 1, 2, 1, 1e-5. */
typedef struct luzgpu_nfor_settings
{
	uint32_t	feature_prefilter_radius;
	uint32_t	regression_radius;     /* 1..32 */
	uint32_t	patch_radius;          /* 1..8 */
	uint32_t	pad;
	double		min_variance;
} luzgpu_nfor_settings;

/* Denoise::NFORBuffers (NFOR.hpp:23-37) as host FP64 arrays of width x height pixels:
 * colours 3 per pixel, colour variance 1, features (NFOR_FEATURE_COUNT = 11) per pixel;
 * `image` is the raw mean image of the same render (an output only). Any pointer may be
 * NULL where the buffers are an output. */
typedef struct luzgpu_nfor_buffers
{
	uint32_t	width, height;
	double*		image;
	double*		color_a;
	double*		color_b;
	double*		color_variance;
	double*		features_a;
	double*		features_b;
	double*		feature_variance;
} luzgpu_nfor_buffers;

typedef struct luzgpu_scene_info
{
	uint32_t	width, height, samples;
	int32_t		max_bounces;
	uint32_t	hittables, triangles, meshes, instances, lights;
	uint32_t	tlas_nodes, blas_nodes;
	uint32_t	env_width, env_height;
	uint32_t	tlas_depth, blas_depth;
	uint32_t	materials;
} luzgpu_scene_info;

/* Context: one HIP device + stream. (Threads.cpp:375-548 worker pool)
 * luzgpu_destroy frees the device arrays of scenes still loaded on the context; those
 * handles stay valid for host-side queries and luzgpu_scene_release. */
int		luzgpu_create(luzgpu_ctx** out, int device);
void	luzgpu_destroy(luzgpu_ctx* ctx);
const char*	luzgpu_last_error(void);

/* Scene upload from a .luz file: SceneFile::read (SceneFile.cpp:256-321) followed by the
 * pre-render sync of Renderer::render (Renderer.cpp:18-20: updateLights,
 * updateAccelerationStructure). With ctx == NULL the scene is parsed and its acceleration
 * structures built on the host only (for inspection; it cannot be rendered). */
int		luzgpu_scene_load(luzgpu_ctx* ctx, const char* path, luzgpu_scene** out);
void	luzgpu_scene_release(luzgpu_scene* scene);

/* --- an in-memory Scene, flattened (no file) ---
 * The seam Renderer::render(Scene&) (Renderer.hpp:7) -> _manageThreads(Scene&)
 * (Threads.cpp:375) holds a populated Scene, not a path: scenes built in code (the
 * reference's tests, --benchmark scenes, SceneHelpers.cpp:602) never see a file. These
 * records carry what that Scene holds, in its own units and FP64, so the host builds the
 * same acceleration structures and light tables as luzgpu_scene_load does from a file. */

enum	/* material kinds (the reference's Material subclasses) */
{
	LUZGPU_MAT_LAMBERTIAN = 1,        /* Lambertian(color)                                        */
	LUZGPU_MAT_METAL = 2,             /* Metal(color, roughness) or Metal(eta, k, roughness)      */
	LUZGPU_MAT_DIELECTRIC = 3,        /* Dielectric(color, ior, roughness) + absorption           */
	LUZGPU_MAT_EMISSIVE = 4,          /* Emissive: color = the emitted radiance                   */
	LUZGPU_MAT_ISOTROPIC = 5,         /* Isotropic(color)                                         */
	LUZGPU_MAT_HENYEY_GREENSTEIN = 6, /* HenyeyGreenstein(color, anisotropy)                      */
	LUZGPU_MAT_GLOSSY = 7,            /* Glossy(color, roughness)                                 */
	LUZGPU_MAT_DIFFUSE_GLOSSY = 8,    /* DiffuseGlossy(color, glossy color, weight, roughness)    */
	LUZGPU_MAT_PRINCIPLED = 9         /* Principled without subsurface                            */
};

enum	/* top-level hittable kinds (Scene::getHittables() entries) */
{
	LUZGPU_HIT_TRIANGLE = 0,
	LUZGPU_HIT_RECTANGLE = 1,
	LUZGPU_HIT_INSTANCE = 2,          /* MeshInstance of meshes[mesh]                             */
	LUZGPU_HIT_SPHERE = 3,
	LUZGPU_HIT_CUBE = 4,
	LUZGPU_HIT_VOLUME = 5,            /* ConstantVolume over a sphere or box boundary             */
	LUZGPU_HIT_PLANE = 6,
	LUZGPU_HIT_DIRECTIONAL = 7        /* DirectionalLight                                         */
};

enum	/* Scene::getRenderSky() (SkyTypes.hpp) */
{
	LUZGPU_SKY_NONE = 0,              /* the background colour                                    */
	LUZGPU_SKY_ENVIRONMENT = 1,       /* the environment map                                      */
	LUZGPU_SKY_LINEAR = 2,            /* the white-to-blue gradient                               */
	LUZGPU_SKY_ATMOSPHERE = 3         /* the physical atmosphere (the reference's default)        */
};

/* One material: the fields its kind reads (the others are ignored). */
typedef struct luzgpu_material_desc
{
	uint32_t	kind;                /* LUZGPU_MAT_*                                             */
	uint32_t	conductor;           /* metal: 1 = conductor Fresnel from eta / k                */
	double		color[3];            /* albedo / Schlick f0 / tint / emitted radiance / base     */
	double		roughness;           /* metal, dielectric, glossy, diffuse-glossy, principled    */
	double		ior;                 /* dielectric, principled                                   */
	double		eta[3], k[3];        /* conductor metal                                          */
	double		absorption[3];       /* dielectric / principled interior, 1/m                    */
	double		anisotropy;          /* Henyey-Greenstein g                                      */
	double		glossy_color[3];     /* diffuse-glossy                                           */
	double		glossy_weight;
	double		metallic, transmission, clearcoat, clearcoat_roughness, sheen;  /* principled   */
	/* principled subsurface (Principled.cpp:39-88): weight in [0,1] (0 = none), the Burley
	 * radius per channel times its scale in metres, the tint, and 1 for the thin method */
	double		subsurface;
	double		subsurface_radius[3];
	double		subsurface_scale;
	double		subsurface_color[3];
	uint32_t	subsurface_thin;
	uint32_t	pad;
} luzgpu_material_desc;

/* One mesh (Mesh::_triangles, after its degenerate-triangle removal): vertices[9*t..] =
 * v0, v1, v2 of triangle t; normals[9*t..] the vertex normals, or NULL. */
typedef struct luzgpu_mesh_desc
{
	uint32_t		triangle_count;
	uint32_t		pad;
	const double*	vertices;
	const double*	normals;
} luzgpu_mesh_desc;

/* One top-level hittable; per kind:
 *   TRIANGLE     a, b, c = vertices; normals = vertex normals when flags & 1
 *   RECTANGLE    a = position, b = orientation (normal), size[0..1] = width, height
 *   INSTANCE     mesh; a = position, c = scale; cos_rotation / sin_rotation of the X, Y, Z
 *                Euler angles (MeshInstance::InstanceTransform)
 *   SPHERE       a = centre, radius; flags & 1: invisible (Sphere::isVisible false: never hit,
 *                still a light when emissive); flags & 2: IES-profiled light (Sphere::
 *                setIESProfile): mesh = index into ies_profiles, b = vertical axis,
 *                size[0] = horizontal rotation (degrees)
 *   CUBE         a = position, size = width, height, depth
 *   VOLUME       flags & 1: sphere boundary (a = centre, radius), else box (a = position,
 *                size); density (per scene unit); material = the phase function
 *   PLANE        plane_y, b = orientation
 *   DIRECTIONAL  a = the light's direction of travel (normalised); flags & 1: b = its
 *                atmosphere sun radiance (DirectionalLight::getAtmosphereSunRadiance, a
 *                `solar=` light); with the atmosphere sky the first directional light is the
 *                atmosphere's sun (Scene::syncAtmosphereSunDirection) */
typedef struct luzgpu_hittable_desc
{
	uint32_t	kind;                /* LUZGPU_HIT_*                                             */
	uint32_t	material;            /* index into materials                                     */
	uint32_t	mesh;                /* INSTANCE: index into meshes                              */
	uint32_t	flags;
	double		a[3], b[3], c[3];
	double		normals[9];
	double		size[3];
	double		radius;
	double		plane_y;
	double		density;
	double		cos_rotation[3];
	double		sin_rotation[3];
} luzgpu_hittable_desc;

/* IESProfile (IESProfile.hpp): the angle grids in degrees (strictly increasing), the candela
 * values horizontal-major (vertical_count values per horizontal angle, the file's multiplier
 * applied), and the luminous flux IESProfile::totalLumens() integrated from them. */
typedef struct luzgpu_ies_desc
{
	uint32_t		vertical_count, horizontal_count;
	const double*	vertical;
	const double*	horizontal;
	const double*	candela;
	double			total_lumens;
} luzgpu_ies_desc;

/* Scene::getActiveCamera() (Camera.hpp), lengths in metres. */
typedef struct luzgpu_camera_desc
{
	double		position[3], direction[3], up[3];
	double		focal_length, sensor_width, sensor_height, f_number, focus_distance;
	uint32_t	pinhole;
	uint32_t	pad;
} luzgpu_camera_desc;

typedef struct luzgpu_scene_desc
{
	/* Scene getters (Scene.hpp:33-131) and the image size */
	uint32_t	width, height, samples;
	int32_t		max_bounces;
	uint32_t	adaptive, adaptive_min_samples, adaptive_check_interval, denoise;
	double		adaptive_threshold;
	double		meters_per_unit;
	uint32_t	sky;                 /* LUZGPU_SKY_*                                             */
	uint32_t	view_transform;      /* LUZGPU_VIEW_*                                            */
	uint32_t	bloom;
	uint32_t	pad0;
	double		exposure, contrast;
	double		background[3];
	/* Scene::getAtmosphere() (Atmosphere.hpp): sun angle (radians), radii and scale heights
	 * (metres), march samples, stars brightness, and the sun's radiance scale */
	double		sun_angle, earth_radius, atmosphere_radius, rayleigh_height, mie_height;
	double		stars_brightness, sun_radiance_scale;
	int32_t		atmosphere_samples, atmosphere_light_samples;
	/* Scene::getEnvironmentMap(): env_width x env_height texels, 3 doubles each, scene-linear,
	 * row 0 at the top (EnvironmentMap::_pixels); strength, rotation (degrees), lighting */
	uint32_t	env_width, env_height;
	const double*	env_texels;
	double		env_strength, env_rotation;
	uint32_t	env_lighting;
	uint32_t	pad1;
	luzgpu_camera_desc	camera;
	uint32_t	material_count, mesh_count, hittable_count;
	uint32_t	pad2;
	const luzgpu_material_desc*	materials;
	const luzgpu_mesh_desc*		meshes;
	const luzgpu_hittable_desc*	hittables;
	uint32_t	ies_count;
	uint32_t	pad3;
	const luzgpu_ies_desc*		ies_profiles;   /* profiles of IES-profiled sphere lights */
} luzgpu_scene_desc;

/* The same as luzgpu_scene_load for a Scene given as records: the scene's pre-render sync
 * (Renderer.cpp:18-20: updateLights, updateAccelerationStructure) and the upload. All
 * arrays are read during the call only. With ctx == NULL the scene is built on the host
 * only. Replaces the SceneFile::read step for callers that hold a Scene (Renderer.hpp:7). */
int		luzgpu_scene_create(luzgpu_ctx* ctx, const luzgpu_scene_desc* desc, luzgpu_scene** out);
int		luzgpu_scene_get_info(const luzgpu_scene* scene, luzgpu_scene_info* out);

/* Renders samples [sample_begin, sample_begin + sample_count) and writes
 * rgb_out[W*H*3] (fp32, row-major, y down) = mean over those samples of the cleaned
 * per-sample radiance — Scene::getImage() after _threadRender with denoising off and a raw
 * view transform (Threads.cpp:551-583, 77-88). With adaptive sampling in effect (the
 * settings' adaptive fields) each pixel stops where _threadRender's loop stops it and its
 * mean is over its own samples; stats->rendered_samples is then the sum of the per-pixel
 * counts (SceneRenderStats::renderedSamples, Threads.cpp:439,491). */
int		luzgpu_render(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings,
			float* rgb_out, luzgpu_stats* stats);

/* luzgpu_render on the plain kernel without allocating anything. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Its sums, counters and
 * per-sample colours live in one fixed device array per GPU (about 48 MiB, held until the
 * process ends), so an image has at most 1,048,576 pixels (LUZGPU_E_INVALID beyond), and
 * adaptive sampling and post-processing (which need per-pixel state) are refused with
 * LUZGPU_E_UNSUPPORTED, as is settings->integrator = LUZGPU_INTEGRATOR_WAVEFRONT. The image
 * equals luzgpu_render's with LUZGPU_INTEGRATOR_MEGAKERNEL. Every call on one GPU, from any
 * context, shares that array: calls that overlap on one GPU race. stats->render_ms is host
 * wall time (no HIP events are created). With output_on_device = 0 the image is copied to
 * host memory through the runtime's staging; pass a device pointer to avoid it. */
int		luzgpu_render_noalloc(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings,
			float* rgb_out, luzgpu_stats* stats);

/* Renders n scenes (1..127) in one multi-scene wavefront launch sequence: the scenes'
 * device arrays are merged (and cached while the same scenes are passed again), every pass
 * covers a slice of samples of every scene, and the pass's launches are replayed from a
 * captured hipGraph. Scenes with adaptive sampling in effect are rendered in rounds over
 * their still-active pixels, all scenes' rounds sharing the passes. Replaces n concurrent Renderer::render calls, which the reference
 * cannot run (its render seed is global, Sampler.cpp:28). settings[i] (or NULL: scene
 * defaults) and rgb_out[i] are per scene, as for luzgpu_render; launch tuning, visit
 * counting and stage timing are taken from settings[0]. The integrator (the same for every
 * scene) is the wavefront one, or the plain kernel: then every scene's paths run in one
 * launch, one scene per wave, which suits scenes too small to fill the GPU alone (no
 * adaptive sampling there). stats, if given, receives one aggregate record for the batch. */
int		luzgpu_render_batch(luzgpu_ctx* ctx, uint32_t n, luzgpu_scene* const* scenes, const luzgpu_settings* settings,
			float* const* rgb_out, luzgpu_stats* stats);

/* Autotuner (no counterpart in the reference, whose CPU workers have no launch shape):
 * times a calibration render of these scenes and settings (one pass of at most 16M paths)
 * under candidate traversal launch shapes (LDS stack depth, rays per queue claim,
 * lane-refill thresholds, persistent grid) and returns the fastest; the context keeps it
 * for later luzgpu_render / luzgpu_render_batch calls with the same scenes and settings
 * and autotune = 1. n = 1 tunes the single-scene path, n > 1 the merged batch. The images
 * do not depend on the launch shape. */
int		luzgpu_autotune(luzgpu_ctx* ctx, uint32_t n, luzgpu_scene* const* scenes, const luzgpu_settings* settings,
			luzgpu_tuning* out);

/* Renders with denoising on, as Renderer::render does for a scene with denoise enabled
 * (Threads.cpp:375-548): sample k draws from sampler stream (k % 2) + 1, the even and odd
 * samples fill two half-buffers with their colours and primary-hit features
 * (_threadRender, Threads.cpp:585-633), and Denoise::applyNFOR (NFOR.cpp:977-1102) runs on
 * the device. rgb_out receives Scene::getImage() (the raw mean), denoised_out the denoised
 * image (Scene::getDenoisedImage(), before post-processing); both W*H*3 fp32, host or
 * device pointers as settings->output_on_device says. buffers_out (optional, host FP64)
 * receives the NFOR inputs. The whole sample range is rendered in one call, by the
 * integrator settings->integrator names (features come from each path's primary hit). */
int		luzgpu_render_denoised(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings,
			const luzgpu_nfor_settings* nfor, float* rgb_out, float* denoised_out, luzgpu_nfor_buffers* buffers_out,
			luzgpu_stats* stats);

/* Denoise::applyNFOR on caller-supplied buffers (host FP64, `image` unused); denoised_out
 * receives W*H*3 FP64 on the host, elapsed_ms (optional) the device time. */
int		luzgpu_nfor(luzgpu_ctx* ctx, const luzgpu_nfor_buffers* in, const luzgpu_nfor_settings* nfor, double* denoised_out,
			double* elapsed_ms);

/* The scene file's post-processing settings (view_transform, bloom, exposure /
 * photographic_exposure, contrast keys; Settings.cpp:274-323). */
int		luzgpu_scene_post_settings(const luzgpu_scene* scene, luzgpu_post_settings* out);

/* applyPostProcessing (Threads.cpp:77-99) on a W x H x 3 FP64 scene-linear image:
 * Image::applyExposure, bloom (Image::extractBloom + Gaussian::blur(9, 2.0) added back at
 * 0.25, Threads.cpp:55-75), then unless the view is raw the view transform to display sRGB
 * (Image::applyViewTransformAndEncodeSRGB, or applyViewTransform + applyContrast +
 * gammaCorrect when contrast != 1) and Image::suppressIsolatedFireflies. image_in and
 * image_out are host pointers, or device pointers when on_device = 1; they may be equal. */
int		luzgpu_postprocess(luzgpu_ctx* ctx, const luzgpu_post_settings* post, uint32_t width, uint32_t height,
			const double* image_in, double* image_out, uint32_t on_device);

/* Path-replay adjoint (no counterpart in the reference): gradients of an image loss with
 * respect to per-material parameters, for the render luzgpu_render produces with the same
 * settings (samples, seed, range, resolution). dl_dimage[W*H*3] = dLoss/dImage (a device
 * pointer when settings->output_on_device, else host). Outputs (host, any may be NULL),
 * indexed by material in scene-file order: d_albedo[3*m+c] for Lambertian albedo,
 * d_roughness[m] for metal GGX roughness, d_emission[3*m+c] for emitter radiance; other
 * materials get 0. Sampling, MIS weights and Russian roulette are held fixed (detached). */
int		luzgpu_adjoint(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings, const float* dl_dimage,
			double* d_albedo, double* d_roughness, double* d_emission);

/* Sets per-material parameters (count = the scene's material count; any array may be NULL
 * to keep those values): Lambertian albedo[3*m..], metal roughness[m], emitter
 * emission[3*m..] (radiance); other material kinds ignore them. Light selection weights are
 * rebuilt (Scene::updateLights, Scene.cpp:615-638). */
int		luzgpu_scene_set_materials(luzgpu_scene* scene, uint32_t count, const double* albedo, const double* roughness,
			const double* emission);

/* --- introspection used by the parity tests (no counterpart in the reference API) --- */

/* Per-pixel sample counts (the values _threadRender returns, Threads.cpp:582,640) of the
 * context's last luzgpu_render or luzgpu_render_denoised: count = W*H of that render. */
int		luzgpu_sample_counts(luzgpu_ctx* ctx, uint32_t* counts_out, size_t count);

/* Primary-ray closest hits of sample 0, stream 0 (RayHelper.cpp:85-106 + HitHelper.cpp:5-32):
 * ids_out[W*H*2] = (top-level hittable index in Scene::getHittables() order, triangle index
 * within the mesh or -1), t_out[W*H] = hit distance (or -1 on a miss). */
int		luzgpu_primary_hits(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings,
			int64_t* ids_out, float* t_out);

/* Evaluates the device sampler (Sampler.cpp:161-237) for n keys; keys[7*i..] =
 * (seed, x, y, sample, stream, bounce, dimension). s1[i] = sample1D, s2[2*i..] = sample2D. */
int		luzgpu_sampler_eval(luzgpu_ctx* ctx, size_t n, const uint32_t* keys, double* s1, float* s2);

/* Top-level BVH topology in nested-parenthesis form over hittable indices
 * (BVHNode.cpp:142-185). needed receives the length including the terminator. */
int		luzgpu_scene_tlas_topology(const luzgpu_scene* scene, char* buf, size_t cap, size_t* needed);

/* Mesh BVH of mesh `mesh` (Mesh::_packedBVHNodes / _triangleIndices, Mesh.hpp:37-67):
 * boxes[6*k] (min xyz, max xyz, FP64), links[5*k] (left, right, start, count, leaf),
 * tri_index[t]. Pass NULL buffers to query node_count / tri_count. */
int		luzgpu_scene_blas(const luzgpu_scene* scene, uint32_t mesh, double* boxes, int64_t* links,
			int64_t* tri_index, uint64_t* node_count, uint64_t* tri_count);

#ifdef __cplusplus
}
#endif

#endif
