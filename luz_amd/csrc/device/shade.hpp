// Surface interaction, BSDFs, light and environment sampling: fp32 device restatement of
// the reference's per-vertex work. Random numbers come from the bit-exact PixelSampler,
// in the same dimensions and bounce as the reference, so each path makes the same
// discrete choices as the reference's FP64 path up to fp32 rounding.
//   HitRecord / setFaceNormal        Hittable.hpp:12-36
//   triangle and instance normals     Triangle.cpp:241-264, MeshInstance.cpp:274-301
//   Lambertian                        Lambertian.cpp:15-36
//   Metal (Schlick / conductor, GGX)  Metal.cpp:126-221, BSDF.cpp:49-315
//   Dielectric (smooth / rough GGX)   Dielectric.cpp:127-250, BSDF.cpp:93-125,317-458
//   Isotropic / Henyey-Greenstein     Isotropic.cpp:15-24, HenyeyGreenstein.cpp:54-63,
//                                     ColorHelper.cpp:39-93,115-148
//   sphere / cube / plane / volume    Sphere.cpp:226-466, Cube.cpp:97-118, Plane.cpp:51-70,
//   records and light sampling        ConstantVolume.cpp:103-115, DirectionalLight.cpp:74-103
//   light sampling + MIS              ColorHelper.cpp:176-257,307-325,914-1033
//   environment map                   EnvironmentMap.cpp:468-616, ColorHelper.cpp:1035-1098
#pragma once

#include "sampler.hpp"
#include "traverse.hpp"

namespace luzgpu
{
	constexpr float	kPiF = 3.14159265358979323846f;

	struct	Surface
	{
		f3	position;
		f3	normal;		// shading normal, facing the incoming ray
		f3	geometricNormal;
		bool	frontFace;
		std::uint32_t	material;
	};

	__device__ __forceinline__ f3	faceNormalOf(const GpuTri& tri)
	{
		const f3 n = cross(ld3(tri.e1), ld3(tri.e2));
		return n * (1.0f / sqrtf(dot(n, n)));
	}

	// Rebuilds the reference's HitRecord from a traversal result.
	template <int Feat = kFeatAll>
	__device__ __forceinline__ Surface	surfaceAt(const SceneView& s, const Ray& ray, const HitInfo& h)
	{
		Surface sf;
		const GpuHittable gh = s.hittables[h.hittable];
		const std::uint32_t kind = gh.kindIndex >> kKindShift;
		const std::uint32_t idx = gh.kindIndex & kIndexMask;
		sf.material = gh.material;
		sf.position = ray.o + ray.d * h.t;
		if ((Feat & kFeatAnalytic) && kind == GPU_HIT_SPHERE)
		{
			const GpuSphere sp = s.spheres[idx];
			const f3 outward = (sf.position - ld3(sp.center)) * (1.0f / sp.radius);
			sf.frontFace = dot(ray.d, outward) < 0.0f;
			sf.normal = sf.frontFace ? outward : -outward;
			sf.geometricNormal = sf.normal;
			return sf;
		}
		if ((Feat & kFeatAnalytic) && (kind == GPU_HIT_CUBE || kind == GPU_HIT_PLANE))
		{
			const f3 n = kind == GPU_HIT_CUBE ? ld3(s.cubeFaces[6 * idx + h.tri].normal) : ld3(s.planes[idx].orientation);
			sf.frontFace = dot(ray.d, n) < 0.0f;
			sf.normal = sf.frontFace ? n : -n;
			sf.geometricNormal = sf.normal;
			return sf;
		}
		if ((Feat & kFeatAnalytic) && kind == GPU_HIT_VOLUME)
		{
			// ConstantVolume::hit: arbitrary normal; the phase functions ignore it
			sf.normal = mk(1.0f, 0.0f, 0.0f);
			sf.geometricNormal = sf.normal;
			sf.frontFace = true;
			return sf;
		}
		if (kind == GPU_HIT_TRIANGLE)
		{
			const f3 fn = faceNormalOf(s.topTris[idx]);
			sf.frontFace = dot(ray.d, fn) < 0.0f;
			sf.normal = sf.frontFace ? fn : -fn;
			sf.geometricNormal = sf.normal;
		}
		else if (kind == GPU_HIT_RECTANGLE)
		{
			const f3 n = ld3(s.rects[idx].normal);
			sf.frontFace = dot(ray.d, n) < 0.0f;
			sf.normal = sf.frontFace ? n : -n;
			sf.geometricNormal = sf.normal;
		}
		else
		{
			const GpuInstance& in = s.instances[idx];
			const GpuTri& tri = s.meshTris[h.tri];
			const Ray local = toLocal(in, ray);
			const f3 fn = faceNormalOf(tri);
			f3 ln = fn;
			if (in.normalBase >= 0)
			{
				const float* nn = s.meshNormals + static_cast<size_t>(h.tri - in.triBase + in.normalBase) * 9;
				const f3 n0 = ld3(nn), n1 = ld3(nn + 3), n2 = ld3(nn + 6);
				if (len2(n0) > 0.0f && len2(n1) > 0.0f && len2(n2) > 0.0f)
				{
					const f3 in3 = n0 * (1.0f - h.u - h.v) + n1 * h.u + n2 * h.v;
					ln = len2(in3) <= 1e-24f ? fn : normalize(in3);
				}
			}
			// local record: both normals face the local ray
			const f3 lShading = dot(local.d, ln) < 0.0f ? ln : -ln;
			const f3 lGeom = dot(local.d, fn) < 0.0f ? fn : -fn;
			const float* m = in.normalToWorld;
			auto toWorld = [&](f3 n) {
				const f3 w = mk(m[0] * n.x + m[1] * n.y + m[2] * n.z, m[3] * n.x + m[4] * n.y + m[5] * n.z,
					m[6] * n.x + m[7] * n.y + m[8] * n.z);
				const float l2 = len2(w);
				return (l2 <= 1e-24f || !isfinite(l2)) ? mk(0.0f, 0.0f, 0.0f) : w * (1.0f / sqrtf(l2));
			};
			f3 wn = toWorld(lShading);
			f3 wg = toWorld(lGeom);
			if (dot(ray.d, wn) > 0.0f) wn = -wn;
			if (dot(ray.d, wg) > 0.0f) wg = -wg;
			sf.normal = wn;
			sf.geometricNormal = wg;
			// MeshInstance.cpp:298: recomputed after the flip, so true unless perpendicular
			sf.frontFace = dot(ray.d, wn) < 0.0f;
		}
		return sf;
	}

	// ---- BSDF helpers (BSDF.cpp) -------------------------------------------------------

	__device__ __forceinline__ f3	safeNormalize(f3 v, f3 fallback)
	{
		const float l2 = len2(v);
		if (!isfinite(l2) || l2 <= 0.0f) return fallback;
		return v * (1.0f / sqrtf(l2));
	}

	__device__ __forceinline__ f3	clampColor01(f3 c) { return mk(clamp01(c.x), clamp01(c.y), clamp01(c.z)); }

	__device__ __forceinline__ f3	schlick(f3 f0, float cosTheta)
	{
		cosTheta = clamp01(cosTheta);
		const float a = 1.0f - cosTheta;
		const float f = a * a * a * a * a;
		f0 = clampColor01(f0);
		return mk(f0.x + (1.0f - f0.x) * f, f0.y + (1.0f - f0.y) * f, f0.z + (1.0f - f0.z) * f);
	}

	__device__ __forceinline__ float	dielectricFresnel(float cosTheta, float etaI, float etaT)
	{
		cosTheta = clamp01(fabsf(cosTheta));
		if (!isfinite(etaI) || !isfinite(etaT) || etaI <= 0.0f || etaT <= 0.0f) return 1.0f;
		const float sinI = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
		const float sinT = etaI / etaT * sinI;
		if (sinT >= 1.0f) return 1.0f;
		const float cosT = sqrtf(fmaxf(0.0f, 1.0f - sinT * sinT));
		const float rPar = ((etaT * cosTheta) - (etaI * cosT)) / ((etaT * cosTheta) + (etaI * cosT));
		const float rPer = ((etaI * cosTheta) - (etaT * cosT)) / ((etaI * cosTheta) + (etaT * cosT));
		return clamp01(0.5f * (rPar * rPar + rPer * rPer));
	}

	__device__ __forceinline__ float	conductorChannel(float eta, float k, float c)
	{
		if (!isfinite(eta) || !isfinite(k) || eta <= 0.0f || k < 0.0f) return 0.0f;
		const float c2 = c * c, etaK = eta * eta + k * k, two = 2.0f * eta * c;
		const float pd = etaK * c2 + two + 1.0f, sd = etaK + two + c2;
		if (fabsf(pd) <= 1e-10f || fabsf(sd) <= 1e-10f) return 0.0f;
		return clamp01(0.5f * ((etaK * c2 - two + 1.0f) / pd + (etaK - two + c2) / sd));
	}

	__device__ __forceinline__ f3	conductorFresnel(const GpuMaterial& m, float c)
	{
		c = clamp01(c);
		return mk(conductorChannel(m.eta[0], m.k[0], c), conductorChannel(m.eta[1], m.k[1], c), conductorChannel(m.eta[2], m.k[2], c));
	}

	__device__ __forceinline__ float	ggxD(float noH, float alpha)
	{
		noH = clamp01(noH);
		alpha = fmaxf(1e-4f, alpha);
		const float a2 = alpha * alpha;
		const float den = noH * noH * (a2 - 1.0f) + 1.0f;
		const float v = a2 / (kPiF * den * den);
		return isfinite(v) ? v : 0.0f;
	}

	__device__ __forceinline__ float	smithG1(float noW, float alpha)
	{
		noW = clamp01(noW);
		alpha = fmaxf(1e-4f, alpha);
		if (noW <= 0.0f) return 0.0f;
		const float a2 = alpha * alpha;
		const float den = noW + sqrtf(a2 + (1.0f - a2) * noW * noW);
		if (den <= 1e-10f || !isfinite(den)) return 0.0f;
		return 2.0f * noW / den;
	}

	// d ln(D(noH) G1(noV) G1(noL)) / d alpha of the GGX reflection lobe (the path-replay
	// adjoint's roughness derivative; D and G1 as ggxD / smithG1 above).
	__device__ __forceinline__ float	ggxDlogAlpha(f3 n, f3 view, f3 l, float alpha)
	{
		alpha = fmaxf(1e-4f, alpha);
		const float noV = clamp01(dot(n, view)), noL = clamp01(dot(n, l));
		const f3 h = safeNormalize(view + l, n);
		const float noH = clamp01(dot(n, h));
		const float a2 = alpha * alpha;
		const float den = noH * noH * (a2 - 1.0f) + 1.0f;
		float d = 2.0f / alpha - 4.0f * alpha * noH * noH / den;
		auto g1 = [&](float c) {
			const float sq = sqrtf(a2 + (1.0f - a2) * c * c);
			return -(alpha * (1.0f - c * c) / sq) / (c + sq);
		};
		d += g1(noV) + g1(noL);
		return isfinite(d) ? d : 0.0f;
	}

	__device__ __forceinline__ f3	sampleGGXHalf(f3 n, float alpha, float sx, float sy)
	{
		alpha = fmaxf(1e-4f, alpha);
		const float a2 = alpha * alpha;
		const float phi = 2.0f * kPiF * clamp01(sx);
		const float u = fminf(1.0f - 1e-7f, clamp01(sy));
		const float cosT = sqrtf((1.0f - u) / (1.0f + (a2 - 1.0f) * u));
		const float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
		float sp, cp;
		sincosf(phi, &sp, &cp);
		return Onb(n).local(sinT * cp, sinT * sp, cosT);
	}

	__device__ __forceinline__ float	ggxReflectionPdf(f3 n, f3 view, f3 l, float alpha)
	{
		const float noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0f || noL <= 0.0f) return 0.0f;
		const f3 h = safeNormalize(view + l, n);
		const float voH = dot(view, h), noH = dot(n, h);
		if (voH <= 1e-10f || noH <= 0.0f) return 0.0f;
		const float pdf = ggxD(noH, alpha) * noH / (4.0f * voH);
		return (isfinite(pdf) && pdf > 0.0f) ? pdf : 0.0f;
	}

	__device__ __forceinline__ f3	ggxReflectionBsdfCos(f3 fresnel, f3 n, f3 view, f3 l, float alpha)
	{
		const float noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0f || noL <= 0.0f) return mk(0, 0, 0);
		const f3 h = safeNormalize(view + l, n);
		const float noH = dot(n, h);
		const float scale = ggxD(noH, alpha) * (smithG1(noV, alpha) * smithG1(noL, alpha)) / (4.0f * noV);
		if (!isfinite(scale) || scale <= 0.0f) return mk(0, 0, 0);
		return clampColor01(fresnel) * scale;
	}

	__device__ __forceinline__ f3	metalBsdfCos(const GpuMaterial& m, f3 n, f3 view, f3 l)
	{
		const f3 h = safeNormalize(view + l, n);
		const float voH = fmaxf(0.0f, dot(view, h));
		const f3 F = (m.flags & GPU_MATF_CONDUCTOR) ? conductorFresnel(m, voH) : schlick(clampColor01(ld3(m.color)), voH);
		return ggxReflectionBsdfCos(F, n, view, l, m.alpha);
	}

	__device__ __forceinline__ bool	refractDir(f3 i, f3 mn, float eta, f3& out)
	{
		const float c = fminf(1.0f, dot(-i, mn));
		const f3 perp = (i + mn * c) * eta;
		const float p2 = len2(perp);
		if (p2 > 1.0f) return false;
		const f3 par = mn * -sqrtf(fmaxf(0.0f, 1.0f - p2));
		out = safeNormalize(perp + par, reflect(i, mn));
		return len2(out) > 0.0f;
	}

	__device__ __forceinline__ float	dielectricPdf(f3 n, f3 view, f3 l, float alpha, float etaI, float etaT)
	{
		const float noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0f || noL == 0.0f || etaI <= 0.0f || etaT <= 0.0f) return 0.0f;
		if (noL > 0.0f)
		{
			const f3 h = safeNormalize(view + l, n);
			return dielectricFresnel(dot(view, h), etaI, etaT) * ggxReflectionPdf(n, view, l, alpha);
		}
		const float eta = etaI / etaT;
		f3 h = safeNormalize(view + l * eta, n);
		if (dot(n, h) < 0.0f) h = -h;
		const float noH = dot(n, h), voH = dot(view, h), loH = dot(l, h);
		if (noH <= 0.0f || voH <= 0.0f || loH >= 0.0f) return 0.0f;
		const float den = voH + eta * loH;
		if (fabsf(den) <= 1e-10f) return 0.0f;
		const float F = dielectricFresnel(voH, etaI, etaT);
		const float dHdL = fabsf((eta * eta * loH) / (den * den));
		const float pdf = (1.0f - F) * ggxD(noH, alpha) * noH * dHdL;
		return (isfinite(pdf) && pdf > 0.0f) ? pdf : 0.0f;
	}

	__device__ __forceinline__ f3	dielectricBsdfCos(f3 color, f3 n, f3 view, f3 l, float alpha, float etaI, float etaT)
	{
		const float noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0f || noL == 0.0f || etaI <= 0.0f || etaT <= 0.0f) return mk(0, 0, 0);
		color = clampColor01(color);
		if (noL > 0.0f)
		{
			const f3 h = safeNormalize(view + l, n);
			const float F = dielectricFresnel(dot(view, h), etaI, etaT);
			return ggxReflectionBsdfCos(mk(F, F, F), n, view, l, alpha);
		}
		const float eta = etaI / etaT;
		f3 h = safeNormalize(view + l * eta, n);
		if (dot(n, h) < 0.0f) h = -h;
		const float noH = dot(n, h), voH = dot(view, h), loH = dot(l, h);
		if (noH <= 0.0f || voH <= 0.0f || loH >= 0.0f) return mk(0, 0, 0);
		const float den = voH + eta * loH;
		if (fabsf(den) <= 1e-10f) return mk(0, 0, 0);
		const float F = dielectricFresnel(voH, etaI, etaT);
		const float scale = (1.0f - F) * ggxD(noH, alpha) * (smithG1(noV, alpha) * smithG1(fabsf(noL), alpha)) * eta * eta
			* fabsf(voH * loH) / (noV * den * den);
		if (!isfinite(scale) || scale <= 0.0f) return mk(0, 0, 0);
		return color * scale;
	}

	// ---- scatter record (Material.hpp:23-41) ----------------------------------------------

	enum : std::uint32_t
	{
		PDF_NONE = 0,
		PDF_COSINE = 1,
		PDF_SPHERE = 2,
		PDF_HG = 3,
		PDF_BSDF = 4,
	};

	struct	Scatter
	{
		std::uint32_t	pdfType;
		bool	specular;
		f3	attenuation;
		f3	direction;		// specular ray direction or sampled BSDF direction
		f3	incident;		// incoming ray direction (BSDF evaluations)
		f3	cosineW;		// cosine basis axis; Henyey-Greenstein: the phase direction
		float	g;			// Henyey-Greenstein anisotropy
		bool	absorbing;	// dielectric medium absorption
		bool	mirrorFallback;	// rough metal: the GGX sample reflected below the horizon
		bool	subsurfaceExit;	// principled Burley subsurface lobe sampled: the exit is relocated
	};

	// materialCanTransmitThroughGeometry (ColorHelper.cpp:327-355)
	__device__ __forceinline__ bool	canTransmit(const GpuMaterial& m)
	{
		return m.kind == GPU_MAT_DIELECTRIC || m.kind == GPU_MAT_ISOTROPIC || m.kind == GPU_MAT_HG
			|| (m.kind == GPU_MAT_PRINCIPLED && (m.transmission > 0.0f || (m.flags & GPU_MATF_THIN_SUBSURFACE) != 0));
	}

	// Henyey-Greenstein pdf of scatterPDFValue (ColorHelper.cpp:56-77)
	__device__ __forceinline__ float	hgPdf(f3 phaseDir, float gIn, f3 dir)
	{
		const float dl2 = len2(dir), pl2 = len2(phaseDir);
		if (dl2 <= 0.0f || pl2 <= 0.0f || !isfinite(dl2) || !isfinite(pl2)) return 0.0f;
		const f3 nd = dir * (1.0f / sqrtf(dl2));
		const f3 pd = phaseDir * (1.0f / sqrtf(pl2));
		const float c = fmaxf(-1.0f, fminf(1.0f, dot(nd, pd)));
		const float g = fmaxf(-0.99f, fminf(0.99f, gIn));
		const float g2 = g * g;
		const float den = fmaxf(1e-12f, 1.0f + g2 - (2.0f * g * c));
		return (1.0f - g2) / (4.0f * kPiF * den * sqrtf(den));
	}

	// leavesOpaqueGeometricSurface (ColorHelper.cpp:357-374)
	__device__ __forceinline__ bool	leavesOpaque(const Surface& sf, const GpuMaterial& m, f3 dir)
	{
		if (canTransmit(m)) return true;
		f3 g = sf.geometricNormal;
		if (len2(g) <= 1e-12f) g = sf.normal;
		if (len2(g) <= 1e-12f) return true;
		return dot(g, dir) > 1e-7f;
	}

	// GGX-Schlick reflection lobe (BSDF.cpp:312-330)
	__device__ __forceinline__ f3	schlickBsdfCos(f3 f0, f3 n, f3 view, f3 l, float alpha)
	{
		const f3 h = safeNormalize(view + l, n);
		const float voH = fmaxf(0.0f, dot(view, h));
		return ggxReflectionBsdfCos(schlick(f0, voH), n, view, l, alpha);
	}

	__device__ __forceinline__ float	cosinePdf(f3 n, f3 l)
	{
		const float c = dot(n, l);
		return c <= 0.0f ? 0.0f : c / kPiF;
	}

	// DiffuseGlossy::evaluateBSDFCos / scatteringPDF (DiffuseGlossy.cpp:151-186)
	__device__ __forceinline__ f3	diffuseGlossyEval(const GpuMaterial& m, f3 n, f3 view, f3 l)
	{
		const float noL = dot(n, l);
		const f3 diffuse = noL <= 0.0f ? mk(0, 0, 0) : clampColor01(ld3(m.color)) * (noL / kPiF);
		const f3 glossy = ggxReflectionBsdfCos(clampColor01(ld3(m.glossyColor)), n, view, l, m.alpha);
		return diffuse * (1.0f - m.glossyWeight) + glossy * m.glossyWeight;
	}

	__device__ __forceinline__ float	diffuseGlossyPdf(const GpuMaterial& m, f3 n, f3 view, f3 l)
	{
		return cosinePdf(n, l) * (1.0f - m.glossyWeight) + ggxReflectionPdf(n, view, l, m.alpha) * m.glossyWeight;
	}

	__device__ __forceinline__ f3	scaledColor(f3 c, float f)
	{
		return (!isfinite(f) || f <= 0.0f) ? mk(0, 0, 0) : c * f;
	}

	// Principled::evaluateBSDFCos (Principled.cpp:763-853)
	__device__ __forceinline__ f3	principledEval(const GpuMaterial& m, const Surface& sf, f3 view, f3 l)
	{
		const f3 base = clampColor01(ld3(m.color));
		const f3 n = sf.normal;
		const float noL = dot(n, l);
		const float diffuse = m.diffuseScale;
		const float surfaceDiffuse = diffuse * (1.0f - m.subsurface);
		f3 r = (noL <= 0.0f || surfaceDiffuse <= 0.0f) ? mk(0, 0, 0) : base * (surfaceDiffuse * noL / kPiF);
		// subsurfaceBSDFCos (:251-271): base * subsurface colour, two-sided when thin
		const float ssScale = diffuse * m.subsurface;
		const float ssCos = (m.flags & GPU_MATF_THIN_SUBSURFACE) ? fabsf(noL) : noL;
		if (ssCos > 0.0f && ssScale > 0.0f) r = r + ld3(m.subsurfaceAlbedo) * (ssScale * ssCos / kPiF);
		const float sheenScale = diffuse * m.sheen;
		if (noL > 0.0f && sheenScale > 0.0f)
		{
			const f3 h = safeNormalize(view + l, n);
			const float x = 1.0f - clamp01(dot(l, h));
			const float fall = x * x * x * x * x;
			const f3 sheenColor = mk(1, 1, 1) * 0.5f + base * 0.5f;
			r = r + sheenColor * (sheenScale * fall * noL / kPiF);
		}
		if (noL > 0.0f)
		{
			const f3 h = safeNormalize(view + l, n);
			const float F = dielectricFresnel(dot(view, h), 1.0f, m.ior);
			r = r + scaledColor(ggxReflectionBsdfCos(mk(F, F, F), n, view, l, m.alpha), 1.0f - m.metallic);
		}
		r = r + scaledColor(schlickBsdfCos(base, n, view, l, m.alpha), m.metallic);
		const float transmission = m.transmission * (1.0f - m.metallic);
		if (noL < 0.0f && transmission > 0.0f)
		{
			const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
			r = r + scaledColor(dielectricBsdfCos(base, n, view, l, m.alpha, etaI, etaT), transmission);
		}
		r = r + scaledColor(schlickBsdfCos(mk(0.04f, 0.04f, 0.04f), n, view, l, m.clearcoatAlpha), m.clearcoat * 0.25f);
		return r;
	}

	// Principled::scatteringPDF (Principled.cpp:855-921): lobe-weighted mixture
	__device__ __forceinline__ float	principledPdf(const GpuMaterial& m, const Surface& sf, f3 view, f3 l)
	{
		const f3 n = sf.normal;
		const float cosP = cosinePdf(n, l);
		const float spec = ggxReflectionPdf(n, view, l, m.alpha);
		float trans = 0.0f;
		if (dot(n, l) < 0.0f)
		{
			const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
			trans = dielectricPdf(n, view, l, m.alpha, etaI, etaT);
		}
		const float coat = ggxReflectionPdf(n, view, l, m.clearcoatAlpha);
		// subsurfaceCosinePDF (:236-249)
		const float ssP = (m.flags & GPU_MATF_THIN_SUBSURFACE) ? fabsf(dot(n, l)) / (2.0f * kPiF) : cosP;
		const float pdfs[7] = {cosP, ssP, cosP, spec, spec, trans, coat};
		float pdf = 0.0f;
		for (int i = 0; i < 7; i++)
		{
			if (m.lobe[i] > 0.0f && pdfs[i] > 0.0f) pdf += m.lobe[i] * pdfs[i];
		}
		return (isfinite(pdf) && pdf > 0.0f) ? pdf : 0.0f;
	}

	template <int Feat = kFeatAll>
	__device__ __forceinline__ float	bsdfPdf(const GpuMaterial& m, const Surface& sf, const Scatter& sc, f3 dir)
	{
		const f3 view = safeNormalize(-sc.incident, sf.normal);
		if (m.kind == GPU_MAT_METAL || ((Feat & kFeatRich) && m.kind == GPU_MAT_GLOSSY))
		{
			return ggxReflectionPdf(sf.normal, view, dir, m.alpha);
		}
		if ((Feat & kFeatRich) && m.kind == GPU_MAT_DIFFUSE_GLOSSY) return diffuseGlossyPdf(m, sf.normal, view, dir);
		if ((Feat & kFeatRich) && m.kind == GPU_MAT_PRINCIPLED) return principledPdf(m, sf, view, dir);
		const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
		return dielectricPdf(sf.normal, view, dir, m.alpha, etaI, etaT);
	}

	template <int Feat = kFeatAll>
	__device__ __forceinline__ f3	bsdfEval(const GpuMaterial& m, const Surface& sf, const Scatter& sc, f3 dir)
	{
		const f3 view = safeNormalize(-sc.incident, sf.normal);
		if (m.kind == GPU_MAT_METAL)
		{
			return metalBsdfCos(m, sf.normal, view, dir);
		}
		if ((Feat & kFeatRich) && m.kind == GPU_MAT_GLOSSY) return ggxReflectionBsdfCos(clampColor01(ld3(m.color)), sf.normal, view, dir, m.alpha);
		if ((Feat & kFeatRich) && m.kind == GPU_MAT_DIFFUSE_GLOSSY) return diffuseGlossyEval(m, sf.normal, view, dir);
		if ((Feat & kFeatRich) && m.kind == GPU_MAT_PRINCIPLED) return principledEval(m, sf, view, dir);
		const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
		return dielectricBsdfCos(ld3(m.color), sf.normal, view, dir, m.alpha, etaI, etaT);
	}

	// scatterPDFValue (ColorHelper.cpp:39-93)
	template <int Feat = kFeatAll>
	__device__ __forceinline__ float	scatterPdf(const GpuMaterial& m, const Surface& sf, const Scatter& sc, f3 dir)
	{
		if (sc.pdfType == PDF_COSINE)
		{
			const float c = dot(dir, sc.cosineW);
			return c <= 0.0f ? 0.0f : c / kPiF;
		}
		if ((Feat & kFeatRich) && sc.pdfType == PDF_SPHERE) return 1.0f / (4.0f * kPiF);
		if ((Feat & kFeatRich) && sc.pdfType == PDF_HG) return hgPdf(sc.cosineW, sc.g, dir);
		if (sc.pdfType == PDF_BSDF) return bsdfPdf<Feat>(m, sf, sc, dir);
		return 0.0f;
	}

	// scatterBSDFCos (ColorHelper.cpp:95-113)
	template <int Feat = kFeatAll>
	__device__ __forceinline__ f3	scatterBsdfCos(const GpuMaterial& m, const Surface& sf, const Scatter& sc, f3 dir)
	{
		if (sc.pdfType == PDF_BSDF) return bsdfEval<Feat>(m, sf, sc, dir);
		return sc.attenuation * scatterPdf<Feat>(m, sf, sc, dir);
	}

	__device__ __forceinline__ f3	normalizeOrZero(f3 d)
	{
		// Ray::_normalize
		const float l = sqrtf(len2(d));
		if (!isfinite(l) || l <= 0.0f) return mk(0, 0, 0);
		return d * (1.0f / l);
	}

	// Sampler::cosineHemisphere (Sampler.cpp:239-246) in the ONB of `n`; `flip` mirrors it
	// below the surface (the thin subsurface lobe's far side, Principled.cpp:156-167)
	__device__ __forceinline__ f3	cosineHemisphereDir(f3 n, const PixelSampler& ps, bool flip = false)
	{
		float sx, sy;
		ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
		const float phi = (2.0f * kPiF) * sx;
		const float r = sqrtf(sy);
		const float z = sqrtf(fmaxf(0.0f, 1.0f - sy));
		float sp, cp;
		sincosf(phi, &sp, &cp);
		return Onb(n).local(r * cp, r * sp, flip ? -z : z);
	}

	// GGX half-vector reflection sample (Glossy.cpp:22-40, Principled.cpp:334-357)
	__device__ __forceinline__ f3	sampleGGXReflection(f3 inDir, f3 n, float alpha, const PixelSampler& ps)
	{
		const f3 view = safeNormalize(-inDir, n);
		float sx, sy;
		ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
		f3 h = sampleGGXHalf(n, alpha, sx, sy);
		if (dot(view, h) <= 0.0f) h = n;
		f3 dir = reflect(inDir, h);
		if (dot(dir, n) <= 0.0f) dir = reflect(inDir, n);
		return dir;
	}

	// Finishes a BSDF-sampled scatter: attenuation = bsdfCos / pdf, rejected when either is 0.
	__device__ __forceinline__ bool	acceptBsdfSample(const GpuMaterial& m, const Surface& sf, f3 dir, Scatter& sc)
	{
		const float pdf = bsdfPdf(m, sf, sc, dir);
		const f3 f = bsdfEval(m, sf, sc, dir);
		if (pdf <= 0.0f || !isfinite(pdf) || maxc(f) <= 0.0f) return false;
		sc.direction = dir;
		sc.attenuation = f * (1.0f / pdf);
		sc.pdfType = PDF_BSDF;
		return true;
	}

	// Material::scatter for the supported materials. Returns false when the material
	// does not scatter (emissive) or the BSDF sample is rejected.
	template <int Feat = kFeatAll>
	__device__ __forceinline__ bool	scatter(const GpuMaterial& m, const Surface& sf, f3 inDir, const PixelSampler& ps, Scatter& sc)
	{
		sc.incident = inDir;
		sc.absorbing = false;
		sc.specular = false;
		sc.mirrorFallback = false;
		sc.subsurfaceExit = false;
		sc.pdfType = PDF_NONE;
		if (m.kind == GPU_MAT_LAMBERTIAN)
		{
			sc.attenuation = ld3(m.color);
			sc.pdfType = PDF_COSINE;
			sc.cosineW = normalize(sf.normal);
			return true;
		}
		if constexpr ((Feat & kFeatRich) != 0)
		{
		if (m.kind == GPU_MAT_ISOTROPIC)
		{
			sc.attenuation = ld3(m.color);
			sc.pdfType = PDF_SPHERE;
			return true;
		}
		if (m.kind == GPU_MAT_GLOSSY)
		{
			return acceptBsdfSample(m, sf, sampleGGXReflection(inDir, sf.normal, m.alpha, ps), sc);
		}
		if (m.kind == GPU_MAT_DIFFUSE_GLOSSY)
		{
			const f3 dir = ps.sample1D(DIM_MATERIAL_DECISION) < static_cast<double>(m.glossyWeight)
				? sampleGGXReflection(inDir, sf.normal, m.alpha, ps) : cosineHemisphereDir(sf.normal, ps);
			return acceptBsdfSample(m, sf, dir, sc);
		}
		if (m.kind == GPU_MAT_PRINCIPLED)
		{
			// setMediumAbsorption (Principled.cpp:129-140): transmissive with a positive coefficient
			sc.absorbing = (m.flags & GPU_MATF_ABSORPTION) != 0 && m.transmission > 0.0f;
			const f3 base = clampColor01(ld3(m.color));
			if (m.flags & GPU_MATF_SMOOTH_TRANSMISSION)
			{
				// scatterSmoothTransmission (Principled.cpp:401-440)
				const f3 view = safeNormalize(-inDir, sf.normal);
				const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
				const float ratio = etaI / etaT;
				const float cosT = fminf(1.0f, dot(view, sf.normal));
				const float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
				f3 dir;
				if (ratio * sinT > 1.0f || static_cast<double>(dielectricFresnel(cosT, etaI, etaT)) > ps.sample1D(DIM_MATERIAL_DECISION))
				{
					dir = reflect(inDir, sf.normal);
					sc.attenuation = mk(1, 1, 1);
				}
				else
				{
					const float c = fminf(dot(-inDir, sf.normal), 1.0f);
					const f3 perp = (inDir + sf.normal * c) * ratio;
					dir = perp + sf.normal * -sqrtf(fabsf(1.0f - len2(perp)));
					sc.attenuation = base;
				}
				sc.direction = normalizeOrZero(dir);
				sc.specular = true;
				return true;
			}
			// selectComponent (Principled.cpp:204-216)
			double target = ps.sample1D(DIM_MATERIAL_DECISION);
			int comp = 0;
			for (int i = 0; i < 7; i++)
			{
				target -= static_cast<double>(m.lobe[i]);
				if (target <= 0.0)
				{
					comp = i;
					break;
				}
			}
			f3 dir;
			if (comp <= 2)
			{
				const bool thin = comp == 1 && (m.flags & GPU_MATF_THIN_SUBSURFACE) != 0;
				dir = cosineHemisphereDir(sf.normal, ps, thin && ps.sample1D(DIM_SUBSURFACE_THIN_SIDE) < 0.5);
				sc.subsurfaceExit = comp == 1 && (m.flags & GPU_MATF_BURLEY_SUBSURFACE) != 0;
			}
			else if (comp == 5)
			{
				// sampleGGXTransmission (Principled.cpp:359-399)
				const f3 view = safeNormalize(-inDir, sf.normal);
				const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
				float sx, sy;
				ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
				f3 h = sampleGGXHalf(sf.normal, m.alpha, sx, sy);
				if (dot(view, h) <= 0.0f) h = sf.normal;
				if (!refractDir(inDir, h, etaI / etaT, dir)) dir = reflect(inDir, h);
			}
			else
			{
				dir = sampleGGXReflection(inDir, sf.normal, comp == 6 ? m.clearcoatAlpha : m.alpha, ps);
			}
			return acceptBsdfSample(m, sf, dir, sc);
		}
		if (m.kind == GPU_MAT_HG)
		{
			sc.attenuation = ld3(m.color);
			sc.pdfType = PDF_HG;
			sc.cosineW = len2(inDir) <= 0.0f ? mk(0.0f, 0.0f, 1.0f) : normalize(inDir);
			sc.g = m.anisotropy;
			return true;
		}
		}
		if (m.kind == GPU_MAT_METAL)
		{
			const f3 view = safeNormalize(-inDir, sf.normal);
			const float cosTheta = fminf(1.0f, fmaxf(0.0f, dot(view, sf.normal)));
			if (m.roughness <= 1e-4f)
			{
				sc.direction = normalizeOrZero(reflect(inDir, sf.normal));
				sc.attenuation = (m.flags & GPU_MATF_CONDUCTOR) ? conductorFresnel(m, cosTheta) : schlick(clampColor01(ld3(m.color)), cosTheta);
				sc.specular = true;
				return true;
			}
			float sx, sy;
			ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
			f3 h = sampleGGXHalf(sf.normal, m.alpha, sx, sy);
			if (dot(view, h) <= 0.0f)
			{
				h = sf.normal;
				sc.mirrorFallback = true;
			}
			f3 dir = reflect(inDir, h);
			if (dot(dir, sf.normal) <= 0.0f)
			{
				dir = reflect(inDir, sf.normal);
				sc.mirrorFallback = true;
			}
			const float pdf = ggxReflectionPdf(sf.normal, view, dir, m.alpha);
			const f3 f = metalBsdfCos(m, sf.normal, view, dir);
			if (pdf <= 0.0f || !isfinite(pdf) || maxc(f) <= 0.0f) return false;
			sc.direction = dir;
			sc.attenuation = f * (1.0f / pdf);
			sc.pdfType = PDF_BSDF;
			return true;
		}
		if (m.kind == GPU_MAT_DIELECTRIC)
		{
			sc.absorbing = (m.flags & GPU_MATF_ABSORPTION) != 0;
			const f3 view = safeNormalize(-inDir, sf.normal);
			const float etaI = sf.frontFace ? 1.0f : m.ior, etaT = sf.frontFace ? m.ior : 1.0f;
			const float ratio = etaI / etaT;
			if (m.roughness <= 1e-4f)
			{
				const float cosT = fminf(dot(view, sf.normal), 1.0f);
				const float sinT = sqrtf(1.0f - cosT * cosT);
				const bool cannot = ratio * sinT > 1.0f;
				f3 dir;
				if (cannot || static_cast<double>(dielectricFresnel(cosT, etaI, etaT)) > ps.sample1D(DIM_MATERIAL_DECISION))
				{
					dir = reflect(inDir, sf.normal);
					sc.attenuation = mk(1, 1, 1);
				}
				else
				{
					// Utilities::refract (Utilities.hpp:59-66)
					const float c = fminf(dot(-inDir, sf.normal), 1.0f);
					const f3 perp = (inDir + sf.normal * c) * ratio;
					const f3 par = sf.normal * -sqrtf(fabsf(1.0f - len2(perp)));
					dir = perp + par;
					sc.attenuation = ld3(m.color);
				}
				sc.direction = normalizeOrZero(dir);
				sc.specular = true;
				return true;
			}
			float sx, sy;
			ps.sample2D(DIM_BSDF_DIRECTION, sx, sy);
			f3 h = sampleGGXHalf(sf.normal, m.alpha, sx, sy);
			if (dot(view, h) <= 0.0f) h = sf.normal;
			const float F = dielectricFresnel(dot(view, h), etaI, etaT);
			f3 refr;
			const bool canRefr = refractDir(inDir, h, ratio, refr);
			const f3 dir = (!canRefr || ps.sample1D(DIM_MATERIAL_DECISION) < static_cast<double>(F)) ? reflect(inDir, h) : refr;
			const float pdf = dielectricPdf(sf.normal, view, dir, m.alpha, etaI, etaT);
			const f3 f = dielectricBsdfCos(ld3(m.color), sf.normal, view, dir, m.alpha, etaI, etaT);
			if (pdf <= 0.0f || !isfinite(pdf) || maxc(f) <= 0.0f) return false;
			sc.direction = dir;
			sc.attenuation = f * (1.0f / pdf);
			sc.pdfType = PDF_BSDF;
			return true;
		}
		return false;
	}

	// ---- path-level helpers (ColorHelper.cpp) -------------------------------------------------

	__device__ __forceinline__ f3	clampRay(f3 c)
	{
		const float l = luminance(c);
		if (!isfinite(l)) return mk(0, 0, 0);
		if (l > 1.0e12f) return c * (1.0e12f / l);
		return c;
	}

	__device__ __forceinline__ bool	terminated(f3 t)
	{
		const float m = maxc(t);
		return !isfinite(m) || m <= 1e-6f;
	}

	__device__ __forceinline__ float	powerHeuristic(float a, float b)
	{
		const float a2 = a * a, b2 = b * b, s = a2 + b2;
		if (s <= 0.0f || !isfinite(s)) return 0.0f;
		return a2 / s;
	}

	__device__ __forceinline__ bool	russianRoulette(f3& thr, int bounce, const PixelSampler& ps)
	{
		if (bounce < 3) return true;
		const float surv = fmaxf(0.05f, fminf(0.95f, maxc(thr)));
		if (ps.sample1D(DIM_RUSSIAN_ROULETTE) > static_cast<double>(surv)) return false;
		thr = thr * (1.0f / surv);
		return true;
	}

	// ---- finite lights ----------------------------------------------------------------------

	struct	LightSample
	{
		f3	direction;
		f3	emitted;
		float	pdf;
		float	tMax;
		bool	valid;
		std::uint32_t	material;	// the light's emissive material (finite lights)
	};

	// Sphere::sampleLight (Sphere.cpp:382-442): cone sampling toward the sphere, FP64 like
	// the reference except the fp32 cone direction of Sphere::randomToSphere.
	// IESProfile::candelaAt (IESProfile.cpp:245-300): bilinear in the vertical and horizontal
	// angle (degrees) over the profile's grid, zero outside its vertical range; past the last
	// horizontal angle it interpolates towards the first one a turn later.
	__device__ inline double	iesCandelaAt(const double* V, int nv, const double* H, int nh, const double* C, double vdeg, double hdeg)
	{
		if (nv <= 0 || nh <= 0 || !isfinite(vdeg) || vdeg < V[0] || vdeg > V[nv - 1]) return 0.0;
		int v0 = 0;
		while (v0 + 1 < nv && V[v0 + 1] < vdeg) v0++;
		const int v1 = min(v0 + 1, nv - 1);
		const double vSpan = V[v1] - V[v0];
		const double vt = vSpan > 0.0 ? (vdeg - V[v0]) / vSpan : 0.0;
		int h0 = 0, h1 = 0;
		double ht = 0.0;
		if (nh > 1)
		{
			double h = 0.0;
			if (isfinite(hdeg))
			{
				h = fmod(hdeg, 360.0);
				if (h < 0.0) h += 360.0;
			}
			if (h < H[0]) h += 360.0;
			if (h > H[nh - 1] && H[nh - 1] >= 360.0) h = fmin(h, H[nh - 1]);
			while (h0 + 1 < nh && H[h0 + 1] < h) h0++;
			if (h0 + 1 < nh)
			{
				h1 = h0 + 1;
				const double span = H[h1] - H[h0];
				ht = span > 0.0 ? (h - H[h0]) / span : 0.0;
			}
			else
			{
				h1 = 0;
				const double span = (H[0] + 360.0) - H[h0];
				ht = span > 0.0 ? (h - H[h0]) / span : 0.0;
			}
		}
		const double c00 = C[h0 * nv + v0], c01 = C[h0 * nv + v1], c10 = C[h1 * nv + v0], c11 = C[h1 * nv + v1];
		const double a = c00 * (1.0 - vt) + c01 * vt, b = c10 * (1.0 - vt) + c11 * vt;
		return a * (1.0 - ht) + b * ht;
	}

	// IESProfile::relativeIntensity (IESProfile.cpp:302-323) in FP64 for an emission direction
	// (from the light towards the shaded point): the vertical angle from the profile's axis,
	// the horizontal angle in a frame built from the axis and turned by the rotation, and the
	// candela there over the profile's mean. `t` is the profile's table (GpuLight::ies).
	__device__ inline double	iesRelativeIntensity(const double* t, f3 emission)
	{
		const double mean = t[0];
		if (!(mean > 0.0) || !isfinite(mean)) return 0.0;
		const double ax = t[1], ay = t[2], az = t[3], rotation = t[4];
		const int nv = static_cast<int>(t[5]), nh = static_cast<int>(t[6]);
		const double* V = t + 7;
		const double* H = V + nv;
		const double* C = H + nh;
		double dx = emission.x, dy = emission.y, dz = emission.z;
		const double l2 = dx * dx + dy * dy + dz * dz;
		if (!isfinite(l2) || l2 <= 0.0)
		{
			dx = ax; dy = ay; dz = az;
		}
		else
		{
			const double l = sqrt(l2);
			dx = dx / l; dy = dy / l; dz = dz / l;
		}
		const double cosT = fmin(fmax(dx * ax + dy * ay + dz * az, -1.0), 1.0);
		constexpr double kDeg = 180.0 / 3.14159265358979323846;
		const double vdeg = acos(cosT) * kDeg;
		const bool yHelper = fabs(ay) < 0.95;
		const double hx = yHelper ? 0.0 : 1.0, hy = yHelper ? 1.0 : 0.0, hz = 0.0;
		double tx = hy * az - hz * ay, ty = hz * ax - hx * az, tz = hx * ay - hy * ax;
		const double tl2 = tx * tx + ty * ty + tz * tz;
		if (!isfinite(tl2) || tl2 <= 0.0)
		{
			tx = 1.0; ty = 0.0; tz = 0.0;
		}
		else
		{
			const double tl = sqrt(tl2);
			tx = tx / tl; ty = ty / tl; tz = tz / tl;
		}
		const double bx = ay * tz - az * ty, by = az * tx - ax * tz, bz = ax * ty - ay * tx;
		double hdeg = atan2(dx * bx + dy * by + dz * bz, dx * tx + dy * ty + dz * tz) * kDeg + rotation;
		if (!isfinite(hdeg)) hdeg = 0.0;
		else
		{
			hdeg = fmod(hdeg, 360.0);
			if (hdeg < 0.0) hdeg += 360.0;
		}
		return iesCandelaAt(V, nv, H, nh, C, vdeg, hdeg) / mean;
	}

	__device__ __forceinline__ LightSample	sampleSphereLight(const SceneView& s, const GpuLight& L, f3 origin, const PixelSampler& ps)
	{
		LightSample ls;
		ls.valid = false;
		const GpuSphere sp = s.spheres[L.index];
		const double cx = static_cast<double>(sp.center[0]) - origin.x, cy = static_cast<double>(sp.center[1]) - origin.y,
			cz = static_cast<double>(sp.center[2]) - origin.z;
		const double r = sp.radius;
		const double d2 = cx * cx + cy * cy + cz * cz;
		if (r <= 0.0 || !isfinite(d2) || d2 <= r * r) return ls;
		float sx, sy;
		ps.sample2D(DIM_LIGHT_SURFACE_POINT, sx, sy);
		const float rr = static_cast<float>(r * r / d2);
		const float z = 1.0f + sy * (sqrtf(fmaxf(0.0f, 1.0f - rr)) - 1.0f);
		const float phi = (2.0f * kPiF) * sx;
		const float rad = sqrtf(fmaxf(0.0f, 1.0f - z * z));
		const float lx = cosf(phi) * rad, ly = sinf(phi) * rad;
		f3 dir = Onb(mk(static_cast<float>(cx), static_cast<float>(cy), static_cast<float>(cz))).local(lx, ly, z);
		const float dl2 = len2(dir);
		if (!isfinite(dl2) || dl2 <= 0.0f) return ls;
		dir = dir * (1.0f / sqrtf(dl2));
		const double ox = -cx, oy = -cy, oz = -cz;	// origin - center
		const double a = static_cast<double>(dir.x) * dir.x + static_cast<double>(dir.y) * dir.y + static_cast<double>(dir.z) * dir.z;
		const double b = ox * dir.x + oy * dir.y + oz * dir.z;
		const double c = d2 - r * r;
		const double disc = b * b - a * c;
		if (disc < 0.0) return ls;
		const double sq = sqrt(disc);
		double root = (-b - sq) / a;
		if (root < 0.001 || root > 1.7976931348623157e308) root = (-b + sq) / a;
		if (root < 0.001 || root > 1.7976931348623157e308) return ls;
		const double cosMax = sqrt(1.0 - r * r / d2);
		const double solid = 2.0 * 3.14159265358979323846 * (1.0 - cosMax);
		if (solid <= 0.0 || !isfinite(solid)) return ls;
		ls.direction = dir;
		ls.pdf = static_cast<float>(1.0 / solid) * L.selectionProbability;
		ls.tMax = static_cast<float>(root);
		if (ls.pdf <= 0.0f || ls.tMax <= kTMinF || !isfinite(ls.pdf) || !isfinite(ls.tMax)) return ls;
		ls.emitted = ld3(s.materials[L.material].color);
		if (L.ies)
		{
			// Sphere::sampleLight (Sphere.cpp:451-461): an IES profile scales the sample's
			// radiance (the light's radiance seen by a ray that hits it is not scaled)
			ls.emitted = ls.emitted * static_cast<float>(iesRelativeIntensity(s.iesData + (L.ies - 1u), -dir));
		}
		ls.valid = maxc(ls.emitted) > 0.0f;
		return ls;
	}

	template <int Feat = kFeatAll>
	__device__ __forceinline__ LightSample	sampleLight(const SceneView& s, const GpuParams& p, f3 origin, const PixelSampler& ps)
	{
		LightSample ls;
		ls.valid = false;
		std::uint32_t idx = 0;
		const std::uint32_t n = p.lightCount;
		if (n > 1)
		{
			const double u = ps.sample1D(DIM_LIGHT_SELECTION);
			if (p.lightWeighted)
			{
				const double* cdf = s.lightCdf + p.lightBase;
				const double target = u * cdf[n - 1];
				// std::upper_bound: first element > target
				std::uint32_t lo = 0, hi = n;
				while (lo < hi)
				{
					const std::uint32_t mid = (lo + hi) >> 1;
					if (cdf[mid] > target) hi = mid;
					else lo = mid + 1;
				}
				idx = lo < n - 1 ? lo : n - 1;
			}
			else
			{
				const std::uint32_t r = static_cast<std::uint32_t>(u * static_cast<double>(n));
				idx = r < n - 1 ? r : n - 1;
			}
		}
		const GpuLight& L = s.lights[p.lightBase + idx];
		ls.material = L.material;
		if (!(L.selectionProbability > 0.0f)) return ls;
		if ((Feat & kFeatAnalytic) && L.kind == GPU_HIT_DIRECTIONAL)
		{
			// DirectionalLight::sampleLight: delta direction, pdf 1, tMax = T_MAX
			ls.direction = ld3(L.direction);
			ls.pdf = L.selectionProbability;
			ls.tMax = INFINITY;
			ls.emitted = ld3(s.materials[L.material].color);
			ls.valid = len2(ls.direction) > 0.0f && ls.pdf > 0.0f && maxc(ls.emitted) > 0.0f;
			return ls;
		}
		if ((Feat & kFeatAnalytic) && L.kind == GPU_HIT_SPHERE)
		{
			LightSample sl = sampleSphereLight(s, L, origin, ps);
			sl.material = L.material;
			return sl;
		}
		f3 point, normal;
		if ((Feat & kFeatAnalytic) && L.kind == GPU_HIT_INSTANCE)
		{
			// MeshInstance::sampleLight / _sampleSurface (MeshInstance.cpp:409-466,510-546): a
			// triangle by area (std::lower_bound over the FP64 prefix sums), a point on it
			// (Triangle::sampleSurface, Triangle.cpp:320-341), mapped to world space
			if (L.triCount == 0u || !(L.area > 0.0f)) return ls;
			const double* cdf = s.lightTriCdf + L.triBase;
			const double target = ps.sample1D(DIM_LIGHT_SURFACE_SELECTION) * cdf[L.triCount - 1];
			std::uint32_t lo = 0, hi = L.triCount;
			while (lo < hi)
			{
				const std::uint32_t mid = (lo + hi) >> 1;
				if (cdf[mid] < target) lo = mid + 1;
				else hi = mid;
			}
			const std::uint32_t k = lo < L.triCount - 1 ? lo : L.triCount - 1;
			const double prev = k == 0 ? 0.0 : cdf[k - 1];
			if (!(cdf[k] - prev > 0.0)) return ls;
			const GpuTri& t = s.meshTris[s.lightTriSlot[L.triBase + k]];
			float sx, sy;
			ps.sample2D(DIM_LIGHT_SURFACE_POINT, sx, sy);
			float u = sx, v = sy;
			if (u + v > 1.0f) { u = 1.0f - u; v = 1.0f - v; }
			const f3 lp = ld3(t.v0) + ld3(t.e1) * u + ld3(t.e2) * v;
			const float* M = L.localToWorld;
			point = mk(M[0] * lp.x + M[1] * lp.y + M[2] * lp.z + M[3], M[4] * lp.x + M[5] * lp.y + M[6] * lp.z + M[7],
				M[8] * lp.x + M[9] * lp.y + M[10] * lp.z + M[11]);
			const float* N = s.instances[L.index].normalToWorld;
			const f3 fn = faceNormalOf(t);
			normal = normalizeOrZero(mk(N[0] * fn.x + N[1] * fn.y + N[2] * fn.z, N[3] * fn.x + N[4] * fn.y + N[5] * fn.z,
				N[6] * fn.x + N[7] * fn.y + N[8] * fn.z));
			if (len2(normal) <= 0.0f) return ls;
		}
		else
		{
			if (L.kind != GPU_HIT_RECTANGLE && L.kind != GPU_HIT_TRIANGLE)
			{
				return ls;	// Hittable::sampleLight default: never valid (pdfValue 0)
			}
			float sx, sy;
			ps.sample2D(DIM_LIGHT_SURFACE_POINT, sx, sy);
			if (L.kind == GPU_HIT_RECTANGLE)
			{
				const GpuRect& r = s.rects[L.index];
				point = ld3(r.position) + ld3(r.widthAxis) * ((sx - 0.5f) * r.width) + ld3(r.heightAxis) * ((sy - 0.5f) * r.height);
				normal = ld3(r.normal);
			}
			else
			{
				const GpuTri& t = s.topTris[L.index];
				float u = sx, v = sy;
				if (u + v > 1.0f) { u = 1.0f - u; v = 1.0f - v; }
				point = ld3(t.v0) + ld3(t.e1) * u + ld3(t.e2) * v;
				normal = faceNormalOf(t);
			}
		}
		const f3 d = point - origin;
		const float d2 = len2(d);
		if (!isfinite(d2) || d2 <= 0.0f) return ls;
		const float dist = sqrtf(d2);
		ls.direction = d * (1.0f / dist);
		const float c = fabsf(dot(ls.direction, normal));
		if (L.area <= 0.0f || c <= 0.0f) return ls;
		const float pdf = d2 / (c * L.area);
		if (!isfinite(pdf) || pdf <= 0.0f) return ls;
		ls.pdf = pdf * L.selectionProbability;
		ls.tMax = dist;
		if (ls.pdf <= 0.0f || ls.tMax <= kTMinF || !isfinite(ls.pdf) || !isfinite(ls.tMax)) return ls;
		ls.emitted = ld3(s.materials[L.material].color);
		if (maxc(ls.emitted) <= 0.0f) return ls;
		ls.valid = true;
		return ls;
	}

	// Hittable::pdfValue of one light for a direction from `origin` (re-intersects it).
	template <int Feat = kFeatAll>
	__device__ __forceinline__ float	lightPdfValue(const SceneView& s, const GpuLight& L, f3 origin, f3 vec)
	{
		const Ray r = makeRay(origin, normalizeOrZero(vec));
		const float vl2 = len2(vec);
		float t;
		if ((Feat & kFeatAnalytic) && L.kind == GPU_HIT_SPHERE)
		{
			// Sphere::pdfValue (Sphere.cpp:362-375)
			const GpuSphere sp = s.spheres[L.index];
			if (!sphereHit(sp, r, kTMinF, INFINITY, t)) return 0.0f;
			const double cx = static_cast<double>(sp.center[0]) - origin.x, cy = static_cast<double>(sp.center[1]) - origin.y,
				cz = static_cast<double>(sp.center[2]) - origin.z;
			const double rr = static_cast<double>(sp.radius) * sp.radius;
			const double cosMax = sqrt(1.0 - rr / (cx * cx + cy * cy + cz * cz));
			return static_cast<float>(1.0 / (2.0 * 3.14159265358979323846 * (1.0 - cosMax)));
		}
		if ((Feat & kFeatAnalytic) && L.kind == GPU_HIT_INSTANCE)
		{
			// MeshInstance::pdfValue (MeshInstance.cpp:468-497): this instance alone
			if (!(L.area > 0.0f) || !(vl2 > 0.0f)) return 0.0f;
			HitInfo h;
			if (!instanceClosest(s, L.index, L.hittable, r, INFINITY, h)) return 0.0f;
			const float* N = s.instances[L.index].normalToWorld;
			const f3 fn = faceNormalOf(s.meshTris[h.tri]);
			const f3 gn = normalizeOrZero(mk(N[0] * fn.x + N[1] * fn.y + N[2] * fn.z, N[3] * fn.x + N[4] * fn.y + N[5] * fn.z,
				N[6] * fn.x + N[7] * fn.y + N[8] * fn.z));
			const float c = fabsf(dot(vec, gn) / sqrtf(vl2));
			if (!(c > 0.0f)) return 0.0f;
			return (h.t * h.t * vl2) / (c * L.area);
		}
		if (L.kind != GPU_HIT_RECTANGLE && L.kind != GPU_HIT_TRIANGLE)
		{
			return 0.0f;	// directional lights and the Hittable::pdfValue default
		}
		if (L.kind == GPU_HIT_RECTANGLE)
		{
			const GpuRect& rc = s.rects[L.index];
			if (!rectHit(rc, r, kTMinF, INFINITY, t)) return 0.0f;
			const float c = fabsf(dot(vec, ld3(rc.normal)) / sqrtf(vl2));
			return (t * t * vl2) / (c * L.area);
		}
		const GpuTri& tri = s.topTris[L.index];
		float u, v;
		if (!triHit(tri, r, kTMinF, INFINITY, t, u, v) || L.area <= 0.0f) return 0.0f;
		const float c = fabsf(dot(vec, faceNormalOf(tri)) / sqrtf(vl2));
		if (c <= 0.0f) return 0.0f;
		return (t * t * vl2) / (c * L.area);
	}

	// lightPDFValue (ColorHelper.cpp:233-257)
	template <int Feat = kFeatAll>
	__device__ __forceinline__ float	lightsPdf(const SceneView& s, const GpuParams& p, f3 origin, f3 dir)
	{
		if (p.lightCount == 0) return 0.0f;
		const GpuLight* lights = s.lights + p.lightBase;
		if (p.lightCount == 1) return lightPdfValue<Feat>(s, lights[0], origin, dir);
		float sum = 0.0f;
		for (std::uint32_t i = 0; i < p.lightCount; i++)
		{
			sum += lights[i].selectionProbability * lightPdfValue<Feat>(s, lights[i], origin, dir);
		}
		return sum;
	}

	// ---- environment (EnvironmentMap.cpp) ------------------------------------------------------

	__device__ __forceinline__ float	wrap01(float v)
	{
		v = v - floorf(v);
		if (v < 0.0f) v += 1.0f;
		return v;
	}

	__device__ __forceinline__ void	dirToUV(const GpuParams& p, f3 d, float& u, float& v)
	{
		const float l2 = len2(d);
		if (!isfinite(l2) || l2 <= 0.0f) { u = 0.0f; v = 0.5f; return; }
		const f3 n = d * (1.0f / sqrtf(l2));
		const float phi = atan2f(n.z, n.x);
		const float y = fminf(1.0f, fmaxf(-1.0f, n.y));
		u = wrap01(0.5f + (phi / (2.0f * kPiF)) + (p.envRotation / 360.0f));
		v = clamp01(0.5f + (asinf(y) / kPiF));
	}

	__device__ __forceinline__ f3	envLookup(const SceneView& s, const GpuParams& p, f3 d)
	{
		float u, v;
		dirToUV(p, d, u, v);
		const int W = static_cast<int>(p.envWidth), H = static_cast<int>(p.envHeight);
		const float x = u * static_cast<float>(W) - 0.5f;
		const float y = (1.0f - v) * static_cast<float>(H) - 0.5f;
		const float fx = floorf(x), fy = floorf(y);
		int x0 = static_cast<int>(fx) % W; if (x0 < 0) x0 += W;
		int x1 = (static_cast<int>(fx) + 1) % W; if (x1 < 0) x1 += W;
		const int y0 = min(max(static_cast<int>(fy), 0), H - 1);
		const int y1 = min(max(static_cast<int>(fy) + 1, 0), H - 1);
		const float tx = x - fx, ty = y - fy;
		const float4* px = s.envPixels + p.envBase;
		const float4 c00 = px[y0 * W + x0], c10 = px[y0 * W + x1];
		const float4 c01 = px[y1 * W + x0], c11 = px[y1 * W + x1];
		const f3 top = mk(c00.x, c00.y, c00.z) * (1.0f - tx) + mk(c10.x, c10.y, c10.z) * tx;
		const f3 bot = mk(c01.x, c01.y, c01.z) * (1.0f - tx) + mk(c11.x, c11.y, c11.z) * tx;
		const f3 c = top * (1.0f - ty) + bot * ty;
		return mk(isfinite(c.x) ? fmaxf(0.0f, c.x) : 0.0f, isfinite(c.y) ? fmaxf(0.0f, c.y) : 0.0f, isfinite(c.z) ? fmaxf(0.0f, c.z) : 0.0f);
	}

	__device__ __forceinline__ float	envPdf(const SceneView& s, const GpuParams& p, f3 d)
	{
		if (p.envTotalWeight <= 0.0) return 1.0f / (4.0f * kPiF);
		float u, v;
		dirToUV(p, d, u, v);
		const int x = min(static_cast<int>(floorf(wrap01(u) * static_cast<float>(p.envWidth))), static_cast<int>(p.envWidth) - 1);
		const int y = min(static_cast<int>(floorf((1.0f - clamp01(v)) * static_cast<float>(p.envHeight))), static_cast<int>(p.envHeight) - 1);
		return s.envPdf[p.envBase + y * static_cast<int>(p.envWidth) + x];
	}

	__device__ __forceinline__ double	clampSample(double v)
	{
		if (!isfinite(v) || v <= 0.0) return 0.0;
		if (v >= 1.0) return 0.99999999999999989;	// nextafter(1.0, 0.0)
		return v;
	}

	__device__ __forceinline__ bool	hasEnvLight(const GpuParams& p)
	{
		return p.hasEnvironment && p.envLighting && p.envStrength > 0.0f;
	}

	__device__ __forceinline__ bool	hasConstantBackgroundLight(const GpuParams& p)
	{
		return !hasEnvLight(p) && p.sky == GPU_SKY_NONE && fmaxf(p.background[0], fmaxf(p.background[1], p.background[2])) > 0.0f;
	}

	// EnvironmentMap::sample with the ColorHelper.cpp:858-894 wrapper (minus visibility)
	__device__ __forceinline__ LightSample	sampleInfinite(const SceneView& s, const GpuParams& p, const PixelSampler& ps)
	{
		LightSample ls;
		ls.valid = false;
		ls.tMax = INFINITY;
		if (hasEnvLight(p))
		{
			const double sel = ps.sample1D(DIM_ENVIRONMENT_SELECTION);
			float jx, jy;
			ps.sample2D(DIM_ENVIRONMENT_POINT, jx, jy);
			jx = static_cast<float>(clampSample(jx));
			jy = static_cast<float>(clampSample(jy));
			f3 dir;
			if (p.envTotalWeight <= 0.0)
			{
				const float z = 1.0f - 2.0f * jy;
				const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
				float sp, cp;
				sincosf(2.0f * kPiF * jx, &sp, &cp);
				dir = mk(r * cp, z, r * sp);
				ls.pdf = 1.0f / (4.0f * kPiF);
			}
			else
			{
				const double su = clampSample(sel);
				const double target = su * p.envTotalWeight;
				const std::uint32_t n = p.envWidth * p.envHeight;
				std::uint32_t lo = 0, hi = n;
				if (s.envGuide && n >= kEnvGuideMinTexels)
				{
					// the guide bucket of the sample bounds the search (its product with G is
					// exact: G is a power of two)
					const std::uint32_t buckets = envGuideBuckets(n);
					const std::uint32_t g = min(static_cast<std::uint32_t>(su * static_cast<double>(buckets)), buckets - 1u);
					const std::uint32_t* guide = s.envGuide + p.envBase;
					lo = guide[g];
					hi = min(guide[g + 1] + 1u, n);
				}
				while (lo < hi)	// std::lower_bound: first element >= target
				{
					const std::uint32_t mid = (lo + hi) >> 1;
					if (s.envCdf[p.envBase + mid] < target) lo = mid + 1;
					else hi = mid;
				}
				const std::uint32_t index = lo < n ? lo : n - 1;
				const std::uint32_t y = index / p.envWidth, x = index % p.envWidth;
				const float u = (static_cast<float>(x) + jx) / static_cast<float>(p.envWidth);
				const float v = 1.0f - ((static_cast<float>(y) + jy) / static_cast<float>(p.envHeight));
				const float phi = (wrap01(u - (p.envRotation / 360.0f)) - 0.5f) * 2.0f * kPiF;
				const float yy = sinf((clamp01(v) - 0.5f) * kPiF);
				const float r = sqrtf(fmaxf(0.0f, 1.0f - yy * yy));
				float sp, cp;
				sincosf(phi, &sp, &cp);
				dir = mk(r * cp, yy, r * sp);
				ls.pdf = envPdf(s, p, dir);
			}
			if (!(ls.pdf > 0.0f) || !isfinite(ls.pdf)) return ls;
			ls.direction = dir;
			ls.emitted = envLookup(s, p, dir) * p.envStrength;
			if (maxc(ls.emitted) <= 0.0f) return ls;
			ls.valid = true;
			return ls;
		}
		if (hasConstantBackgroundLight(p))
		{
			float sx, sy;
			ps.sample2D(DIM_ENVIRONMENT_POINT, sx, sy);
			const float z = 1.0f - 2.0f * sy;
			const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
			float sp, cp;
			sincosf(2.0f * kPiF * sx, &sp, &cp);
			ls.direction = mk(r * cp, r * sp, z);
			ls.emitted = mk(p.background[0], p.background[1], p.background[2]);
			ls.pdf = 1.0f / (4.0f * kPiF);
			ls.valid = maxc(ls.emitted) > 0.0f;
		}
		return ls;
	}

	__device__ __forceinline__ float	infiniteLightPdf(const SceneView& s, const GpuParams& p, f3 d)
	{
		if (hasEnvLight(p)) return envPdf(s, p, d);
		if (hasConstantBackgroundLight(p)) return 1.0f / (4.0f * kPiF);
		return 0.0f;
	}

	__device__ __forceinline__ f3	clampConstantBackground(const GpuParams& p, f3 c)
	{
		const float bl = p.backgroundLuminance;
		if (!isfinite(bl) || bl <= 0.0f) return c;
		const float l = luminance(c), cap = bl * 4.0f;
		if (!isfinite(l)) return mk(0, 0, 0);
		if (cap > 0.0f && l > cap) return c * (cap / l);
		return c;
	}
}
