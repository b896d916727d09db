// Per-path integrator pieces shared by the megakernel and the wavefront stages:
// camera rays (FP64, RayHelper.cpp:85-106), direct lighting (ColorHelper.cpp:973-1033),
// environment lighting (ColorHelper.cpp:1035-1098) and the bounce loop
// (calculateLightRaysColor, ColorHelper.cpp:1136-1326).
#pragma once

#include "shade.hpp"
#include "atmosphere.hpp"

namespace luzgpu
{
	// Thin-lens disk sample (Sampler.cpp:267-292), fp32 as in the reference.
	__device__ __forceinline__ void	unitDisk(const PixelSampler& ps, float& dx, float& dy)
	{
		float sx, sy;
		ps.sample2D(DIM_LENS, sx, sy);
		sx = 2.0f * sx - 1.0f;
		sy = 2.0f * sy - 1.0f;
		if (sx == 0.0f && sy == 0.0f) { dx = 0.0f; dy = 0.0f; return; }
		float r, th;
		if (fabsf(sx) > fabsf(sy)) { r = sx; th = (kPiF / 4.0f) * (sy / sx); }
		else { r = sy; th = (kPiF / 2.0f) - ((kPiF / 4.0f) * (sx / sy)); }
		dx = r * cosf(th);
		dy = r * sinf(th);
	}

	// Camera ray in FP64 with the reference's operation order, then rounded to fp32.
	__device__ __forceinline__ Ray	cameraRay(const GpuParams& p, std::uint32_t x, std::uint32_t y, const PixelSampler& ps)
	{
#pragma clang fp contract(off)
		const GpuCamera& c = p.camera;
		float cx, cy;
		ps.sample2D(DIM_CAMERA, cx, cy);
		const double xU = (static_cast<double>(x) + static_cast<double>(cx)) * c.inverseWidthMinusOne;
		const double yV = (static_cast<double>(y) + static_cast<double>(cy)) * c.inverseHeightMinusOne;
		double off[3] = {0.0, 0.0, 0.0};
		if (c.lensRadius > 0.0)
		{
			float dx, dy;
			unitDisk(ps, dx, dy);
			const double rx = c.lensRadius * static_cast<double>(dx);
			const double ry = c.lensRadius * static_cast<double>(dy);
			for (int k = 0; k < 3; k++) off[k] = (rx * c.u[k]) + (ry * c.v[k]);
		}
		double d[3], o[3];
		for (int k = 0; k < 3; k++)
		{
			const double hx = xU * c.horizontal[k];
			const double vy = yV * c.vertical[k];
			d[k] = (((c.lowerLeftCorner[k] - hx) - vy) - c.position[k]) - off[k];
			o[k] = c.position[k] + off[k];
		}
		const double l = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
		f3 dir = mk(0, 0, 0);
		if (isfinite(l) && l > 0.0)
		{
			const double inv = 1.0 / l;
			dir = mk(static_cast<float>(inv * d[0]), static_cast<float>(inv * d[1]), static_cast<float>(inv * d[2]));
		}
		return makeRay(mk(static_cast<float>(o[0]), static_cast<float>(o[1]), static_cast<float>(o[2])), dir);
	}

	__device__ __forceinline__ float	directProbability(int bounce)
	{
		return bounce < 2 ? 1.0f : 0.5f;
	}

	__device__ __forceinline__ bool	samplesVisibleInfinite(const GpuParams& p, const Scatter& sc)
	{
		return ((p.sky == GPU_SKY_ENVIRONMENT || p.sky == GPU_SKY_ATMOSPHERE) && hasEnvLight(p)) || (hasConstantBackgroundLight(p) && sc.pdfType != PDF_BSDF);
	}

	// Unoccluded contribution of one light sample and the shadow ray that gates it.
	struct	Connection
	{
		f3	contribution;
		f3	direction;
		float	tMax;
		bool	active;
		std::uint32_t	lightMaterial;	// emissive material of the sampled light
	};

	template <int Feat = kFeatAll>
	__device__ __forceinline__ Connection	connectLight(const SceneView& s, const GpuParams& p, const Surface& sf, const GpuMaterial& m,
		const Scatter& sc, int bounce, const PixelSampler& ps)
	{
		Connection c;
		c.active = false;
		const float prob = directProbability(bounce);
		if (prob < 1.0f && ps.sample1D(DIM_LIGHT_STRATEGY) >= static_cast<double>(prob)) return c;
		const LightSample ls = sampleLight<Feat>(s, p, sf.position, ps);
		if (!ls.valid || !leavesOpaque(sf, m, ls.direction)) return c;
		const float spdf = scatterPdf<Feat>(m, sf, sc, ls.direction);
		const f3 f = scatterBsdfCos<Feat>(m, sf, sc, ls.direction);
		if (spdf <= 0.0f || !isfinite(spdf) || maxc(f) <= 0.0f) return c;
		const float w = powerHeuristic(ls.pdf, spdf);
		c.contribution = f * ls.emitted * (w / (ls.pdf * prob));
		c.direction = ls.direction;
		c.lightMaterial = ls.material;
		c.tMax = ls.tMax - kTMinF;
		c.active = c.tMax > kTMinF;	// isShadowOccluded: tMax <= T_MIN is never occluded
		if (!c.active)
		{
			c.active = true;
			c.tMax = -1.0f;	// no shadow test needed
		}
		return c;
	}

	template <int Feat = kFeatAll>
	__device__ __forceinline__ Connection	connectInfinite(const SceneView& s, const GpuParams& p, const Surface& sf, const GpuMaterial& m,
		const Scatter& sc, int bounce, const PixelSampler& ps)
	{
		Connection c;
		c.active = false;
		const bool constBg = hasConstantBackgroundLight(p);
		if (constBg && sc.pdfType == PDF_BSDF) return c;
		const float prob = directProbability(bounce);
		if (prob < 1.0f && ps.sample1D(DIM_ENVIRONMENT_STRATEGY) >= static_cast<double>(prob)) return c;
		const LightSample ls = sampleInfinite(s, p, ps);
		if (!ls.valid || !leavesOpaque(sf, m, ls.direction)) return c;
		const float spdf = scatterPdf<Feat>(m, sf, sc, ls.direction);
		const f3 f = scatterBsdfCos<Feat>(m, sf, sc, ls.direction);
		if (spdf <= 0.0f || !isfinite(spdf) || maxc(f) <= 0.0f) return c;
		const float w = powerHeuristic(ls.pdf, spdf);
		f3 contrib = f * ls.emitted * (w / (ls.pdf * prob));
		if (constBg) contrib = clampConstantBackground(p, contrib);
		c.contribution = contrib;
		c.direction = ls.direction;
		c.tMax = INFINITY;
		c.active = true;
		return c;
	}

	template <int Feat>
	__device__ __forceinline__ f3	missRadiance(const SceneView& s, const GpuParams& p, const Ray& ray, const PixelSampler& ps,
		float envWeight)
	{
		const f3 dir = ray.d;
		f3 sky;
		if constexpr ((Feat & kFeatRich) != 0)
		{
			if (p.sky == GPU_SKY_ATMOSPHERE)
			{
				// a visible environment map replaces the stars (ColorHelper.cpp:590-605)
				if (p.hasEnvironment && p.envStrength > 0.0f)
				{
					return atmosphereOver(p.atmosphere, ray.o, dir, envLookup(s, p, dir) * p.envStrength * envWeight);
				}
				return atmosphereSky(p.atmosphere, ray.o, dir, ps);
			}
		}
		if (p.sky == GPU_SKY_LINEAR)
		{
			// _calculateSkyInterpolation (Atmospherics.cpp:44-53), skyline 0.5 (Scene.cpp:117); no
			// MIS weight: the linear sky is never sampled as a light (ColorHelper.cpp:1196-1198)
			const float t = 0.5f * (dir.y + 1.0f);
			return mk(1.0f - t + 0.5f * t, 1.0f - t + 0.7f * t, 1.0f - t + t);
		}
		if (p.sky == GPU_SKY_ENVIRONMENT && p.hasEnvironment && p.envStrength > 0.0f)
		{
			sky = envLookup(s, p, dir) * p.envStrength;
		}
		else
		{
			sky = mk(p.background[0], p.background[1], p.background[2]);
		}
		return sky * envWeight;
	}

	__device__ __forceinline__ f3	cosineDirection(const Surface& sf, const PixelSampler& ps)
	{
		// Sampler::cosineHemisphere (Sampler.cpp:239-246) in the ONB of the shading normal
		float sx, sy;
		ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
		const float phi = (2.0f * kPiF) * sx;
		const float r = sqrtf(sy);
		const float z = sqrtf(fmaxf(0.0f, 1.0f - sy));
		float sp, cp;
		sincosf(phi, &sp, &cp);
		return Onb(sf.normal).local(r * cp, r * sp, z);
	}

	// Sampler::sphereDirection (Sampler.cpp:249-257)
	__device__ __forceinline__ f3	sphereDirection(const PixelSampler& ps)
	{
		float sx, sy;
		ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
		const float z = 1.0f - (2.0f * sy);
		const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
		const float phi = (2.0f * kPiF) * sx;
		return mk(r * cosf(phi), r * sinf(phi), z);
	}

	// henyeyGreensteinDirection (ColorHelper.cpp:115-148)
	__device__ __forceinline__ f3	hgDirection(const Scatter& sc, const PixelSampler& ps)
	{
		const float pl2 = len2(sc.cosineW);
		if (pl2 <= 0.0f || !isfinite(pl2)) return sphereDirection(ps);
		const f3 pd = sc.cosineW * (1.0f / sqrtf(pl2));
		const float g = fmaxf(-0.99f, fminf(0.99f, sc.g));
		float sx, sy;
		ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
		const float phi = 2.0f * kPiF * sx;
		float c;
		if (fabsf(g) < 1e-3f)
		{
			c = 1.0f - (2.0f * sy);
		}
		else
		{
			const float rm = (1.0f - (g * g)) / (1.0f - g + (2.0f * g * sy));
			c = (1.0f + (g * g) - (rm * rm)) / (2.0f * g);
			c = fmaxf(-1.0f, fminf(1.0f, c));
		}
		const float st = sqrtf(fmaxf(0.0f, 1.0f - (c * c)));
		return Onb(pd).local(st * cosf(phi), st * sinf(phi), c);
	}

	// scatterPDFGenerate (ColorHelper.cpp:150-166)
	template <int Feat = kFeatAll>
	__device__ __forceinline__ f3	scatterDirection(const Surface& sf, const Scatter& sc, const PixelSampler& ps)
	{
		if (sc.pdfType == PDF_COSINE) return cosineDirection(sf, ps);
		if ((Feat & kFeatRich) && sc.pdfType == PDF_SPHERE) return sphereDirection(ps);
		if ((Feat & kFeatRich) && sc.pdfType == PDF_HG) return hgDirection(sc, ps);
		return sc.direction;
	}

	// log(max(u, 1e-12)) of the bounce's DIM_VOLUME_DISTANCE sample: every ConstantVolume
	// test of this bounce's rays (closest hit and shadow) draws the same value
	// (ConstantVolume.cpp:96).
	__device__ __forceinline__ float	volumeLogU(const GpuParams& p, const PixelSampler& ps)
	{
		return p.hasVolumes ? static_cast<float>(log(fmax(ps.sample1D(DIM_VOLUME_DISTANCE), 1e-12))) : 0.0f;
	}

	__device__ __forceinline__ f3	cleanColor(f3 c)
	{
		return mk(isfinite(c.x) ? c.x : 0.0f, isfinite(c.y) ? c.y : 0.0f, isfinite(c.z) ? c.z : 0.0f);
	}

	// Whole path in one thread: the plain (megakernel) integrator.
	// Hooks of the path-replay adjoint; the forward render uses this no-op set.
	struct	NoAdjoint
	{
		static constexpr bool	enabled = false;
		__device__ __forceinline__ void	emission(std::uint32_t, f3, f3) {}
		__device__ __forceinline__ void	arrival(f3, float) {}
		__device__ __forceinline__ void	specular() {}
		__device__ __forceinline__ void	vertex(const GpuMaterial&, std::uint32_t, const Surface&, const Scatter&, f3) {}
		__device__ __forceinline__ void	connection(const GpuMaterial&, std::uint32_t, const Surface&, const Scatter&, f3, f3, std::int32_t) {}
		__device__ __forceinline__ void	continuation(const GpuMaterial&, std::uint32_t, const Surface&, const Scatter&, f3, f3) {}
	};

	// Denoising features of a primary hit or miss (ColorHelper.cpp:725-799), 11 values:
	// pixel coordinates, hit flag, t / (1 + t), normal, log-normalised albedo, material type.
	__device__ __forceinline__ float	normalizedColorChannel(float v)
	{
		if (!isfinite(v) || v <= 0.0f) return 0.0f;
		return fminf(1.0f, log1pf(v) / log1pf(1.0e12f));
	}

	// MaterialType of the reference (MaterialTypes.hpp): LAMBERTIAN 1 ... PRINCIPLED 7, GLOSSY 8,
	// DIFFUSE_GLOSSY 9
	__device__ __forceinline__ float	referenceMaterialType(std::uint32_t kind)
	{
		const std::uint32_t t = kind == GPU_MAT_GLOSSY ? 8u : kind == GPU_MAT_DIFFUSE_GLOSSY ? 9u : kind == GPU_MAT_PRINCIPLED ? 7u : kind;
		return static_cast<float>(t) / 7.0f;
	}

	__device__ __forceinline__ void	primaryFeatures(float* f, const GpuParams& p, std::uint32_t x, std::uint32_t y, bool hit, float t,
		f3 normal, f3 albedo, std::uint32_t kind)
	{
		f[0] = p.width > 1u ? static_cast<float>(static_cast<double>(x) / (static_cast<double>(p.width) - 1.0)) : 0.0f;
		f[1] = p.height > 1u ? static_cast<float>(static_cast<double>(y) / (static_cast<double>(p.height) - 1.0)) : 0.0f;
		if (!hit)
		{
			f[2] = 0.0f;
			f[3] = 1.0f;
			f[4] = f[5] = f[6] = 0.5f;
			f[10] = 0.0f;
		}
		else
		{
			const float l2 = len2(normal);
			if (l2 > 0.0f) normal = normal * (1.0f / sqrtf(l2));
			f[2] = 1.0f;
			f[3] = t > 0.0f ? t / (1.0f + t) : 0.0f;
			f[4] = normal.x * 0.5f + 0.5f;
			f[5] = normal.y * 0.5f + 0.5f;
			f[6] = normal.z * 0.5f + 0.5f;
			f[10] = referenceMaterialType(kind);
		}
		f[7] = normalizedColorChannel(albedo.x);
		f[8] = normalizedColorChannel(albedo.y);
		f[9] = normalizedColorChannel(albedo.z);
	}

	// relocateSubsurfaceExit (ColorHelper.cpp:376-504): after a Burley subsurface scatter, a
	// point a sampled radius away on the tangent disk is probed along -normal from above
	// (a closest-hit ray); where the probe meets the same material within reach, the vertex
	// moves there and its cosine direction is drawn again about the exit's normal
	// (refreshSubsurfaceScatterAtExit, :394-426, the same DIM_BSDF_DIRECTION sample). `h`,
	// `sf` and `sc` become the exit's on success and are left as they were otherwise.
	template <int Feat, typename Stack>
	__device__ __forceinline__ void	relocateSubsurface(const SceneView& s, const GpuParams& p, const GpuMaterial& m, HitInfo& h,
		Surface& sf, Scatter& sc, const PixelSampler& ps, Stack& stack, float logU, std::uint32_t& closestRays)
	{
		// subsurfaceSampleRadiusMeters (:376-392): one of two exponentials, 1 : 3 the
		// diffusion length, picked 1 : 3
		const double diffusion = m.subsurfaceRadius;
		if (!(diffusion > 0.0) || !isfinite(diffusion)) return;
		const double pick = ps.sample1D(DIM_SUBSURFACE_PROFILE_COMPONENT);
		float sx, sy;
		ps.sample2D(DIM_SUBSURFACE_SPATIAL, sx, sy);
		const double u = fmax(1e-12, fmin(1.0 - 1e-12, static_cast<double>(sy)));
		const double radiusM = -(pick < 0.25 ? diffusion : 3.0 * diffusion) * log(1.0 - u);
		const double mpu = p.metersPerUnit;
		if (!(radiusM > 0.0) || !isfinite(radiusM) || !(mpu > 0.0) || !isfinite(mpu)) return;
		const float radius = static_cast<float>(radiusM / mpu);
		if (radius <= kTMinF) return;
		f3 n = sf.normal;
		if (len2(n) <= 1e-12f) n = sf.geometricNormal;
		if (len2(n) <= 1e-12f) return;
		n = normalize(n);
		float sa, ca;
		sincosf((2.0f * kPiF) * sx, &sa, &ca);
		const f3 candidate = sf.position + Onb(n).local(ca * radius, sa * radius, 0.0f);
		const float reach = fmaxf(radius + static_cast<float>(4.0 * diffusion / mpu), 16.0f * kTMinF);
		const Ray probe = makeRay(candidate + n * reach, -n);
		HitInfo eh;
		closestRays++;
		if (!traverse<false, false, Feat>(s, p, probe, kTMinF, INFINITY, eh, stack, nullptr, logU, kNoOrigin)) return;
		if (eh.t <= kTMinF || eh.t > 2.0f * reach + kTMinF) return;
		const Surface ex = surfaceAt<Feat>(s, probe, eh);
		if (ex.material != sf.material) return;
		const f3 dir = cosineHemisphereDir(ex.normal, ps);
		const f3 view = safeNormalize(-sc.incident, ex.normal);
		const float pdf = principledPdf(m, ex, view, dir);
		const f3 f = principledEval(m, ex, view, dir);
		if (pdf <= 0.0f || !isfinite(pdf) || maxc(f) <= 0.0f) return;
		h = eh;
		sf = ex;
		sc.direction = dir;
		sc.attenuation = f * (1.0f / pdf);
	}

	// The sampler stream of a sample: 0, or with denoising (k % 2) + 1, the two half-buffers
	// (Threads.cpp:594-598).
	__device__ __forceinline__ std::uint32_t	sampleStream(const GpuParams& p, std::uint32_t sample)
	{
		return p.denoiseStreams ? (sample & 1u) + 1u : 0u;
	}

	// One path. With `features` (denoising), the primary hit's or miss's features are written
	// there.
	template <bool Count, int Feat, typename Stack, typename Adj = NoAdjoint>
	__device__ f3	tracePath(const SceneView& s, const GpuParams& p, std::uint32_t x, std::uint32_t y, std::uint32_t sample,
		Stack& stack, std::uint32_t& closestRays, std::uint32_t& shadowRays, VisitCounter& visits, Adj* adj = nullptr,
		float* features = nullptr)
	{
		PixelSampler ps;
		ps.begin(p.renderSeed, x, y, sample, sampleStream(p, sample));
		Ray ray = cameraRay(p, x, y, ps);
		f3 acc = mk(0, 0, 0), thr = mk(1, 1, 1), prevOrigin = mk(0, 0, 0);
		RayOrigin from = kNoOrigin;
		bool prevSpecular = true, prevInfinite = false;
		float prevPdf = 0.0f;
		const bool sampledInfinite = hasEnvLight(p) || hasConstantBackgroundLight(p);
		for (int b = 0; b <= p.maxBounces; b++)
		{
			ps.bounce = static_cast<std::uint32_t>(b);
			const float logU = volumeLogU(p, ps);
			HitInfo h;
			closestRays++;
			if (!traverse<false, Count, Feat>(s, p, ray, kTMinF, INFINITY, h, stack, &visits, logU, from))
			{
				float w = 1.0f;
				if (b > 0 && !prevSpecular && prevInfinite)
				{
					w = powerHeuristic(prevPdf, infiniteLightPdf(s, p, ray.d));
				}
				if (b == 0 && features)
				{
					primaryFeatures(features, p, x, y, false, 0.0f, mk(0, 0, 0), missRadiance<Feat>(s, p, ray, ps, 1.0f), 0u);
				}
				f3 c = thr * missRadiance<Feat>(s, p, ray, ps, w);
				if (b > 0 && hasConstantBackgroundLight(p)) c = clampConstantBackground(p, c);
				c = clampRay(c);
				if constexpr (Adj::enabled) adj->arrival(c, w);
				return acc + c;
			}
			Surface sf = surfaceAt<Feat>(s, ray, h);
			from = originOf(s, h);
			const GpuMaterial& m = s.materials[sf.material];
			if (b == 0 && features)
			{
				primaryFeatures(features, p, x, y, true, h.t, sf.normal, ld3(m.color), m.kind);
			}
			if constexpr ((Feat & kFeatRich) != 0)
			{
				if (b == 0 && p.sky == GPU_SKY_ATMOSPHERE)
				{
					atmospherePrimary(p.atmosphere, ray.o, ray.d, h.t, acc, thr);
					if (terminated(thr)) return acc;
				}
			}
			const f3 emitted = m.kind == GPU_MAT_EMISSIVE ? ld3(m.color) : mk(0, 0, 0);
			Scatter sc;
			if (!scatter<Feat>(m, sf, ray.d, ps, sc))
			{
				const float lp = (!prevSpecular && p.lightCount > 0) ? lightsPdf<Feat>(s, p, prevOrigin, ray.d) : 0.0f;
				const float w = prevSpecular ? 1.0f : powerHeuristic(prevPdf, lp);
				const f3 inc = clampRay((thr * emitted) * w);
				if constexpr (Adj::enabled)
				{
					adj->emission(sf.material, inc, emitted);
					adj->arrival(inc, w);
				}
				return acc + inc;
			}
			if constexpr ((Feat & kFeatSubsurface) != 0)
			{
				if (sc.subsurfaceExit)
				{
					relocateSubsurface<Feat>(s, p, m, h, sf, sc, ps, stack, logU, closestRays);
					from = originOf(s, h);
				}
			}
			f3 att = sc.attenuation;
			if (!sf.frontFace && sc.absorbing)
			{
				const float dm = h.t * p.metersPerUnit;
				if (isfinite(dm) && dm > 0.0f)
				{
					att = att * mk(expf(-m.absorption[0] * dm), expf(-m.absorption[1] * dm), expf(-m.absorption[2] * dm));
				}
			}
			if (sc.specular)
			{
				if (!leavesOpaque(sf, m, sc.direction)) return acc;
				if constexpr (Adj::enabled) adj->specular();
				thr = clampRay(thr * att);
				if (terminated(thr) || !russianRoulette(thr, b, ps)) return acc;
				ray = makeRay(sf.position, sc.direction);
				prevSpecular = true;
				prevPdf = 0.0f;
				continue;
			}
			acc = acc + thr * emitted;
			if constexpr (Adj::enabled) adj->vertex(m, sf.material, sf, sc, acc);
			if (terminated(thr * att)) return acc;
			if (p.lightCount > 0)
			{
				const Connection c = connectLight<Feat>(s, p, sf, m, sc, b, ps);
				if (c.active)
				{
					bool occluded = false;
					if (c.tMax > 0.0f)
					{
						HitInfo sh;
						shadowRays++;
						occluded = traverse<true, Count, Feat>(s, p, makeRay(sf.position, c.direction), kTMinF, c.tMax, sh, stack, &visits,
							logU, from);
					}
					if (!occluded)
					{
						const f3 inc = clampRay(thr * c.contribution);
						acc = acc + inc;
						if constexpr (Adj::enabled) adj->connection(m, sf.material, sf, sc, c.direction, inc, static_cast<std::int32_t>(c.lightMaterial));
					}
				}
			}
			if (sampledInfinite)
			{
				const Connection c = connectInfinite<Feat>(s, p, sf, m, sc, b, ps);
				if (c.active)
				{
					HitInfo sh;
					shadowRays++;
					if (!traverse<true, Count, Feat>(s, p, makeRay(sf.position, c.direction), kTMinF, INFINITY, sh, stack, &visits, logU,
						from))
					{
						const f3 inc = clampRay(thr * c.contribution);
						acc = acc + inc;
						if constexpr (Adj::enabled) adj->connection(m, sf.material, sf, sc, c.direction, inc, -1);
					}
				}
			}
			const f3 dir = scatterDirection<Feat>(sf, sc, ps);
			if (!leavesOpaque(sf, m, dir)) return acc;
			const float pdf = scatterPdf<Feat>(m, sf, sc, dir);
			if (pdf <= 0.0f || !isfinite(pdf)) return acc;
			if constexpr (Adj::enabled) adj->continuation(m, sf.material, sf, sc, dir, acc);
			thr = clampRay(thr * att);
			if (terminated(thr) || !russianRoulette(thr, b, ps)) return acc;
			ray = makeRay(sf.position, dir);
			prevSpecular = false;
			prevOrigin = sf.position;
			prevPdf = pdf;
			prevInfinite = samplesVisibleInfinite(p, sc);
		}
		return acc;
	}
}
