// The atmosphere sky (Atmosphere.cpp:311-476, ColorHelper.cpp:568-659): single-scattering
// Rayleigh + Mie in-scattering and view transmittance along a ray segment, marched in FP64
// with `samples` view steps and `lightSamples` sun steps per view step, plus the procedural
// stars drawn where no environment is visible.
//
// Rays carry FP32 origins and unit directions; the march widens them to FP64 and works in
// metres (scene units x metersPerUnit) around a planet centred at the world origin. Only
// the kernels specialised for rich scenes (kFeatRich) carry this code: an atmosphere scene
// is classed as rich, so the FP64 loops add no registers to the kernels of other scenes.
#pragma once

#include "sampler.hpp"
#include "shade.hpp"

namespace luzgpu
{
	constexpr double	kAtmosphereTMin = 0.001;	// T_MIN (Defaults.hpp:20)
	constexpr double	kAtmosphereTMax = 1.7976931348623157e308;	// T_MAX = DBL_MAX (Defaults.hpp:19)

	struct	AtmosphereSegment
	{
		double	in[3];
		double	tr[3];
	};

	__device__ __forceinline__ double	atmDot(const double* a, const double* b)
	{
#pragma clang fp contract(off)
		return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
	}

	// planetaryHit (Atmosphere.cpp:311-343): both roots of a sphere of `radius` at the origin
	__device__ __forceinline__ bool	planetaryHit(double radius, const double* o, const double* d, double& t0, double& t1)
	{
#pragma clang fp contract(off)
		if (!isfinite(radius) || radius <= 0.0) return false;
		const double a = atmDot(d, d);
		if (!isfinite(a) || a <= 0.0) return false;
		const double halfB = atmDot(o, d);
		const double c = atmDot(o, o) - radius * radius;
		const double disc = halfB * halfB - a * c;
		if (!isfinite(disc) || disc < 0.0) return false;
		const double sq = sqrt(disc < 0.0 ? 0.0 : disc);
		t0 = (-halfB - sq) / a;
		t1 = (-halfB + sq) / a;
		if (t0 > t1)
		{
			const double t = t0;
			t0 = t1;
			t1 = t;
		}
		return true;
	}

	// densityAtHeight (Atmosphere.cpp:23-30)
	__device__ __forceinline__ double	atmDensity(double height, double inverseScaleHeight)
	{
#pragma clang fp contract(off)
		if (height < 0.0 || !isfinite(height)) return 0.0;
		return exp(-height * inverseScaleHeight);
	}

	// Atmosphere::sceneUnitsToMeters (Atmosphere.cpp:156-172)
	__device__ __forceinline__ double	atmToMeters(const GpuAtmosphere& A, double units)
	{
#pragma clang fp contract(off)
		if (!isfinite(units)) return units;
		const double maxUnits = kAtmosphereTMax / A.metersPerUnit;
		if (units >= maxUnits) return kAtmosphereTMax;
		if (units <= -maxUnits) return -kAtmosphereTMax;
		return units * A.metersPerUnit;
	}

	// Atmosphere::sampleSegment (Atmosphere.cpp:346-476) over scene-unit distance [T_MIN, tMax]
	__device__ inline AtmosphereSegment	atmosphereSegment(const GpuAtmosphere& A, f3 origin, f3 dir, double tMax)
	{
#pragma clang fp contract(off)
		AtmosphereSegment out{{0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}};
		if (!isfinite(tMax) || tMax <= kAtmosphereTMin) return out;
		const double tMaxM0 = atmToMeters(A, tMax), tMinM = atmToMeters(A, kAtmosphereTMin);
		if (!isfinite(tMaxM0) || tMaxM0 <= tMinM) return out;
		const double o[3] = {static_cast<double>(origin.x) * A.metersPerUnit, static_cast<double>(origin.y) * A.metersPerUnit,
			static_cast<double>(origin.z) * A.metersPerUnit};
		const double d[3] = {dir.x, dir.y, dir.z};
		double a0, a1;
		if (!planetaryHit(A.atmosphereRadius, o, d, a0, a1) || a1 <= tMinM) return out;
		double tHi = tMaxM0, e0, e1;
		if (planetaryHit(A.earthRadius, o, d, e0, e1) && e1 > tMinM) tHi = fmin(tHi, e0 < 0.0 ? 0.0 : e0);
		const double tLo = tMinM < a0 ? a0 : tMinM;
		tHi = tHi < a1 ? tHi : a1;
		if (!isfinite(tLo) || !isfinite(tHi) || tHi <= tLo) return out;
		const double seg = (tHi - tLo) / static_cast<double>(A.samples);
		if (!isfinite(seg) || seg <= 0.0) return out;

		constexpr double g = 0.76;
		constexpr double rayleighScale = 3.0 / (16.0 * 3.14159265358979323846);
		constexpr double mieScale = 3.0 / (8.0 * 3.14159265358979323846);
		const double betaR[3] = {3.8e-6, 13.5e-6, 33.1e-6};
		const double betaM = 21e-6, mieAbsorption = 1.1;
		double mu = atmDot(d, A.sunDirection);
		mu = mu < -1.0 ? -1.0 : (mu > 1.0 ? 1.0 : mu);
		const double mu2 = mu * mu;
		const double phaseR = rayleighScale * (1.0 + mu2);
		double base = 1.0 + g * g - 2.0 * g * mu;
		base = base < 1.0e-6 ? 1.0e-6 : base;
		const double phaseM = mieScale * ((1.0 - g * g) * (1.0 + mu2) / ((2.0 + g * g) * (base * sqrt(base))));
		const double betaMA = betaM * mieAbsorption;

		double sumR[3] = {0.0, 0.0, 0.0}, sumM[3] = {0.0, 0.0, 0.0};
		double depthR = 0.0, depthM = 0.0;
		for (std::uint32_t i = 0; i < A.samples; i++)
		{
			const double t = tLo + (static_cast<double>(i) + 0.5) * seg;
			const double pos[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
			const double h = sqrt(atmDot(pos, pos)) - A.earthRadius;
			const double dR = atmDensity(h, A.inverseHR), dM = atmDensity(h, A.inverseHM);
			if (dR == 0.0 && dM == 0.0) continue;
			const double sR = dR * seg, sM = dM * seg;
			depthR += sR;
			depthM += sM;
			double l0, l1;
			if (!planetaryHit(A.atmosphereRadius, pos, A.sunDirection, l0, l1) || l1 <= 0.0) continue;
			const double segL = l1 / static_cast<double>(A.lightSamples);
			if (!isfinite(segL) || segL <= 0.0) continue;
			double lightR = 0.0, lightM = 0.0;
			bool reaches = true;
			for (std::uint32_t j = 0; j < A.lightSamples; j++)
			{
				const double lt = (static_cast<double>(j) + 0.5) * segL;
				const double q[3] = {pos[0] + lt * A.sunDirection[0], pos[1] + lt * A.sunDirection[1], pos[2] + lt * A.sunDirection[2]};
				const double hl = sqrt(atmDot(q, q)) - A.earthRadius;
				if (hl < 0.0)
				{
					reaches = false;
					break;
				}
				lightR += atmDensity(hl, A.inverseHR) * segL;
				lightM += atmDensity(hl, A.inverseHM) * segL;
			}
			if (reaches)
			{
				const double accR = depthR + lightR, accM = depthM + lightM;
				for (int k = 0; k < 3; k++)
				{
					const double att = exp(-(betaR[k] * accR + betaMA * accM));
					sumR[k] += att * sR;
					sumM[k] += att * sM;
				}
			}
		}
		for (int k = 0; k < 3; k++)
		{
			out.tr[k] = exp(-(betaR[k] * depthR + betaMA * depthM));
			out.in[k] = (sumR[k] * betaR[k] * phaseR + sumM[k] * betaM * phaseM) * A.sunRadiance[k];
		}
		return out;
	}

	// sampleAtmosphereSky without a visible environment (ColorHelper.cpp:568-605): the
	// segment to infinity plus its transmittance times the procedural stars
	__device__ inline f3	atmosphereSky(const GpuAtmosphere& A, f3 origin, f3 dir, const PixelSampler& ps)
	{
#pragma clang fp contract(off)
		const AtmosphereSegment a = atmosphereSegment(A, origin, dir, kAtmosphereTMax);
		double star = 0.0;
		if (ps.sample1D(DIM_ATMOSPHERE) >= 0.9996)
		{
			const double s = ps.sample1D(DIM_ATMOSPHERE + 1);
			star = (A.starsBrightness - 0.2 + 0.4 * s) - (a.in[0] + a.in[1] + a.in[2]) / 3.0;
			star = star < 0.0 ? 0.0 : (star > 1.0 ? 1.0 : star);
		}
		return mk(static_cast<float>(a.in[0] + a.tr[0] * star), static_cast<float>(a.in[1] + a.tr[1] * star),
			static_cast<float>(a.in[2] + a.tr[2] * star));
	}

	// sampleAtmosphereSky with a visible environment (ColorHelper.cpp:590-605): the
	// (MIS-weighted) environment radiance seen through the atmosphere
	__device__ inline f3	atmosphereOver(const GpuAtmosphere& A, f3 origin, f3 dir, f3 background)
	{
#pragma clang fp contract(off)
		const AtmosphereSegment a = atmosphereSegment(A, origin, dir, kAtmosphereTMax);
		return mk(static_cast<float>(a.in[0] + a.tr[0] * static_cast<double>(background.x)),
			static_cast<float>(a.in[1] + a.tr[1] * static_cast<double>(background.y)),
			static_cast<float>(a.in[2] + a.tr[2] * static_cast<double>(background.z)));
	}

	// compositePrimaryAtmosphereSegment with primaryAtmosphereTMax (ColorHelper.cpp:624-659):
	// the aerial perspective between the camera and the first surface (extended to the
	// ground when the planet surface lies beyond it).
	__device__ inline void	atmospherePrimary(const GpuAtmosphere& A, f3 origin, f3 dir, float surfaceT, f3& acc, f3& thr)
	{
#pragma clang fp contract(off)
		double tMax = surfaceT;
		if (isfinite(tMax) && tMax > kAtmosphereTMin)
		{
			const double o[3] = {origin.x, origin.y, origin.z}, d[3] = {dir.x, dir.y, dir.z};
			double t0, t1;
			if (planetaryHit(A.earthRadius / A.metersPerUnit, o, d, t0, t1) && t1 > kAtmosphereTMin)
			{
				const double ground = t0 < 0.0 ? 0.0 : t0;
				if (isfinite(ground) && ground > kAtmosphereTMin) tMax = tMax < ground ? ground : tMax;
			}
		}
		const AtmosphereSegment a = atmosphereSegment(A, origin, dir, tMax);
		acc = acc + clampRay(thr * mk(static_cast<float>(a.in[0]), static_cast<float>(a.in[1]), static_cast<float>(a.in[2])));
		thr = clampRay(thr * mk(static_cast<float>(a.tr[0]), static_cast<float>(a.tr[1]), static_cast<float>(a.tr[2])));
	}
}
