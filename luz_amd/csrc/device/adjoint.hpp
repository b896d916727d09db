// Path-replay adjoint: gradients of an image-space loss with respect to per-material
// Lambertian albedo, GGX roughness (metal) and emitter radiance, without storing paths.
//
// Each (pixel, sample) path is traced again with the forward render's random numbers
// (the sampler is stateless: PixelSampler keys on seed, pixel, sample, bounce and
// dimension), so it makes the same discrete choices. Sampling is detached: a derivative
// differentiates the path's contributions f / p with the sampled directions, the pdfs in
// MIS weights and the Russian-roulette survival held fixed, which is an unbiased estimator
// of the gradient of the expected image (E[f / p] = the integral for any strategy).
//
//   Lambertian albedo rho at a vertex multiplies that vertex's light and environment
//     connections and everything after it: dL/drho = (L - acc_i) / rho, acc_i the path's
//     radiance before those connections (recorded per vertex, resolved when L is known);
//   GGX roughness r (alpha = r^2) at a rough metal vertex: the connections there scale by
//     d ln f(l) / dr and the continuation by d ln f(dir) / dr (ggxDlogAlpha * 2r);
//     when the metal sampler falls back to the mirror direction (its GGX sample reflected
//     below the horizon), the continuation is a point mass of probability P(alpha) carrying
//     f / p at the mirror (= F G1(v)^2): its derivative is the score d ln P / dr plus
//     2 d ln G1(v) / dr, and the MIS weight w of an emitter the mirror ray reaches moves with
//     p(mirror): d ln w / dr = 2 (1 - w) d ln D(n.h) / dr;
//   emitter radiance Le: every emission or connection term it produced is linear in Le.
//
// The loss enters through dLoss/dImage: a sample of pixel p carries weight dLdI[p] / S.
// Gradients accumulate per block in LDS (7 floats per material: albedo rgb, roughness,
// emission rgb) and then into FP64 device sums.
#pragma once

#include "integrator.hpp"

namespace luzgpu
{
	constexpr int	kGradPerMaterial = 7;
	constexpr int	kAdjointRecords = 32;	// vertices per path (maxBounces + 1)

	// d ln G1(c) / d alpha of Smith GGX (BSDF.cpp:220-236)
	__device__ __forceinline__ float	ggxDlogG1(float c, float a)
	{
		c = clamp01(c);
		a = fmaxf(1e-4f, a);
		const float sq = sqrtf(a * a + (1.0f - a * a) * c * c);
		const float d = -(a * (1.0f - c * c) / sq) / (c + sq);
		return isfinite(d) ? d : 0.0f;
	}

	// Probability P that the metal sampler (Metal.cpp:126-170) falls back to the mirror for a
	// view at cos mu, and dP / dalpha. With view polar angle tv and half-vector polar angle th
	// the fallback happens iff cos(phi) <= K = -cot(tv) cot(2 th): never below
	// th = pi/4 - tv/2, always above pi/4 + tv/2. The sampler's t = cos^2(th) has CDF
	// C(t) = a^2 t / (1 + (a^2 - 1) t); the band between is a 16-point Gauss-Legendre rule in s
	// with th = mid + half sin(pi s / 2), which absorbs the square-root edges of the
	// azimuth fraction 1 - acos(K) / pi.
	__device__ inline float	ggxMirrorFallback(float mu, float a, float& dPda)
	{
		constexpr float kNode[8] = {0.095012510f, 0.281603551f, 0.458016778f, 0.617876244f, 0.755404408f, 0.865631202f,
			0.944575023f, 0.989400935f};
		constexpr float kWeight[8] = {0.189450610f, 0.182603415f, 0.169156519f, 0.149595989f, 0.124628971f, 0.095158512f,
			0.062253524f, 0.027152459f};
		constexpr float kPi = 3.14159265358979f;
		a = fmaxf(1e-4f, a);
		mu = fminf(1.0f, fmaxf(0.0f, mu));
		const float a2 = a * a, tv = acosf(mu), cotv = mu / sqrtf(fmaxf(1e-30f, 1.0f - mu * mu));
		const float th2 = 0.25f * kPi + 0.5f * tv, c2 = cosf(th2), t2 = c2 * c2, d2 = 1.0f + (a2 - 1.0f) * t2;
		float P = a2 * t2 / d2, dP = 2.0f * a * t2 * (1.0f - t2) / (d2 * d2);
		const float mid = 0.25f * kPi, half = 0.5f * tv;
		for (int i = 0; i < 16; i++)
		{
			const float x = i < 8 ? -kNode[7 - i] : kNode[i - 8];
			const float w = i < 8 ? kWeight[7 - i] : kWeight[i - 8];
			const float th = mid + half * sinf(0.5f * kPi * x);
			const float jac = half * 0.5f * kPi * cosf(0.5f * kPi * x);
			const float ct = cosf(th), t = ct * ct, d = 1.0f + (a2 - 1.0f) * t, s2 = sinf(2.0f * th);
			const float K = -cotv * cosf(2.0f * th) / s2;
			const float frac = 1.0f - acosf(fminf(1.0f, fmaxf(-1.0f, K))) * (1.0f / kPi);
			P += w * jac * frac * (a2 / (d * d)) * s2;
			dP += w * jac * frac * (2.0f * a * (1.0f - t - a2 * t) / (d * d * d)) * s2;
		}
		dPda = dP;
		return P;
	}

	struct	AdjointRecorder
	{
		static constexpr bool	enabled = true;
		f3	weight;				// dLoss / dL of this sample
		float*	grads;			// LDS, kGradPerMaterial per material
		std::uint32_t	count;
		std::uint32_t	param[kAdjointRecords];	// material << 1 | (0 albedo, 1 roughness)
		f3	value[kAdjointRecords];			// path radiance when recorded
		float	dlog[kAdjointRecords];			// roughness: d ln(continuation) / dr
		std::int32_t	pendingMaterial;	// mirror fallback: material whose alpha moves the next MIS weight
		float	pendingScale;				// d ln w / dr = (1 - w) * pendingScale

		__device__ __forceinline__ void	add(std::uint32_t material, int slot, float v)
		{
			if (v != 0.0f && isfinite(v)) atomicAdd(grads + material * kGradPerMaterial + slot, v);
		}

		__device__ __forceinline__ static bool	roughMetal(const GpuMaterial& m)
		{
			return m.kind == GPU_MAT_METAL && m.roughness > 1e-4f;
		}

		__device__ __forceinline__ static float	dAlphaDr(const GpuMaterial& m)
		{
			const float r = clamp01(m.roughness);
			return r * r > 1e-4f ? 2.0f * r : 0.0f;
		}

		// an emitter hit: the term is linear in the emitter's radiance
		__device__ __forceinline__ void	emission(std::uint32_t material, f3 inc, f3 le)
		{
			if (le.x > 0.0f) add(material, 4, weight.x * inc.x / le.x);
			if (le.y > 0.0f) add(material, 5, weight.y * inc.y / le.y);
			if (le.z > 0.0f) add(material, 6, weight.z * inc.z / le.z);
		}

		__device__ __forceinline__ float	wsum(f3 c) const { return weight.x * c.x + weight.y * c.y + weight.z * c.z; }

		// an emitter or the sky reached by a BSDF-sampled ray with MIS weight w
		__device__ __forceinline__ void	arrival(f3 inc, float w)
		{
			if (pendingMaterial >= 0) add(static_cast<std::uint32_t>(pendingMaterial), 3, wsum(inc) * (1.0f - w) * pendingScale);
			pendingMaterial = -1;
		}

		__device__ __forceinline__ void	specular() { pendingMaterial = -1; }

		// a scattering vertex, before its connections: acc = path radiance so far
		__device__ __forceinline__ void	vertex(const GpuMaterial& m, std::uint32_t material, const Surface&, const Scatter&, f3 acc)
		{
			pendingMaterial = -1;
			if (m.kind == GPU_MAT_LAMBERTIAN && count < kAdjointRecords)
			{
				param[count] = material << 1;
				value[count] = acc;
				dlog[count] = 0.0f;
				count++;
			}
		}

		// a visible light (lightMaterial >= 0) or environment connection of radiance inc
		__device__ __forceinline__ void	connection(const GpuMaterial& m, std::uint32_t material, const Surface& sf, const Scatter& sc,
			f3 dir, f3 inc, std::int32_t lightMaterial)
		{
			if (lightMaterial >= 0)
			{
				const GpuMaterial& le = materialOf(lightMaterial);
				const f3 e = ld3(le.color);
				const std::uint32_t lm = static_cast<std::uint32_t>(lightMaterial);
				if (e.x > 0.0f) add(lm, 4, weight.x * inc.x / e.x);
				if (e.y > 0.0f) add(lm, 5, weight.y * inc.y / e.y);
				if (e.z > 0.0f) add(lm, 6, weight.z * inc.z / e.z);
			}
			if (roughMetal(m))
			{
				const f3 view = safeNormalize(-sc.incident, sf.normal);
				const float d = ggxDlogAlpha(sf.normal, view, dir, m.alpha) * dAlphaDr(m);
				add(material, 3, (weight.x * inc.x + weight.y * inc.y + weight.z * inc.z) * d);
			}
		}

		// the sampled continuation of a non-specular vertex: acc = radiance so far
		__device__ __forceinline__ void	continuation(const GpuMaterial& m, std::uint32_t material, const Surface& sf, const Scatter& sc,
			f3 dir, f3 acc)
		{
			if (roughMetal(m) && count < kAdjointRecords)
			{
				const f3 view = safeNormalize(-sc.incident, sf.normal);
				param[count] = (material << 1) | 1u;
				value[count] = acc;
				if (!sc.mirrorFallback)
				{
					dlog[count] = ggxDlogAlpha(sf.normal, view, dir, m.alpha) * dAlphaDr(m);
				}
				else
				{
					const float mu = clamp01(dot(sf.normal, view));
					float dP;
					const float P = ggxMirrorFallback(mu, m.alpha, dP);
					dlog[count] = ((P > 0.0f ? dP / P : 0.0f) + 2.0f * ggxDlogG1(mu, m.alpha)) * dAlphaDr(m);
					const float a = fmaxf(1e-4f, m.alpha);
					const float noH = clamp01(dot(sf.normal, safeNormalize(view + dir, sf.normal)));
					const float dlogD = 2.0f / a - 4.0f * a * noH * noH / (noH * noH * (a * a - 1.0f) + 1.0f);
					pendingMaterial = static_cast<std::int32_t>(material);
					pendingScale = 2.0f * dlogD * dAlphaDr(m);
				}
				count++;
			}
		}

		// L: the path's radiance (what the forward render added for this sample)
		__device__ __forceinline__ void	finish(f3 L)
		{
			for (std::uint32_t i = 0; i < count; i++)
			{
				const std::uint32_t material = param[i] >> 1;
				const f3 rest = L - value[i];
				if (param[i] & 1u)
				{
					add(material, 3, (weight.x * rest.x + weight.y * rest.y + weight.z * rest.z) * dlog[i]);
				}
				else
				{
					const f3 rho = ld3(materialOf(static_cast<std::int32_t>(material)).color);
					if (rho.x > 1e-6f) add(material, 0, weight.x * rest.x / rho.x);
					if (rho.y > 1e-6f) add(material, 1, weight.y * rest.y / rho.y);
					if (rho.z > 1e-6f) add(material, 2, weight.z * rest.z / rho.z);
				}
			}
			pendingMaterial = -1;
		}

		const GpuMaterial*	materials;
		__device__ __forceinline__ const GpuMaterial&	materialOf(std::int32_t i) const { return materials[i]; }
	};
}
