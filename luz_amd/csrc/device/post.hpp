// Display post-processing of a rendered image (applyPostProcessing, Threads.cpp:55-99), in
// FP64 like the reference: exposure, bloom (bright-pass with isolated-firefly rejection, a
// separable 9-tap Gaussian, added back at 0.25), the view transform to display sRGB
// (standard, AgX or ACES fit, ColorManagement.cpp:162-300) with the optional contrast curve,
// and the display-space firefly suppression (Image.cpp:140-268, 506-526).
//
// Each stage is a pixel-parallel kernel over W x H x 3 doubles; the windowed stages (bloom
// bright-pass, blur, firefly suppression) read a 5 x 5 or 9-tap neighbourhood from L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace luzgpu
{
	enum	PostView : int
	{
		POST_VIEW_STANDARD = 0,
		POST_VIEW_AGX = 1,
		POST_VIEW_ACES = 2,
		POST_VIEW_RAW = 3
	};

	struct	PostParams
	{
		std::uint32_t	width, height;
		double	exposureScale;		// 2^exposure, formed on the host (Image.cpp:467-484)
		double	contrast;
		double	blur[9];			// createSeparableKernel(9, 2.0) (Gaussian.cpp:22-43), host-formed
		int	view;
	};

	// Bloom constants (Defaults.hpp:47-49)
	constexpr double	kBloomThreshold = 1.0;
	constexpr double	kBloomSoftKnee = 0.5;
	constexpr double	kBloomIntensity = 0.25;

	__device__ __forceinline__ double	postStdMax(double a, double b) { return a < b ? b : a; }
	__device__ __forceinline__ double	postStdMin(double a, double b) { return b < a ? b : a; }

	// clampNonNegative / clampUnit (ColorManagement.cpp:9-31, Image.cpp:14-26)
	__device__ __forceinline__ double	postNonNegative(double v) { return (!isfinite(v) || v <= 0.0) ? 0.0 : v; }
	__device__ __forceinline__ double	postUnit(double v)
	{
		if (!isfinite(v) || v <= 0.0) return 0.0;
		return v >= 1.0 ? 1.0 : v;
	}

	struct	PostColor
	{
		double	r, g, b;
	};

	__device__ __forceinline__ PostColor	postLoad(const double* img, size_t i) { return PostColor{img[3 * i], img[3 * i + 1], img[3 * i + 2]}; }
	__device__ __forceinline__ void	postStore(double* img, size_t i, PostColor c)
	{
		img[3 * i] = c.r;
		img[3 * i + 1] = c.g;
		img[3 * i + 2] = c.b;
	}

	// sanitizeSceneLinear (ColorManagement.cpp:302-309)
	__device__ __forceinline__ PostColor	postSanitize(PostColor c)
	{
		return PostColor{postNonNegative(c.r), postNonNegative(c.g), postNonNegative(c.b)};
	}

	// ColorManagement::luminance (ColorManagement.cpp:313-320)
	__device__ __forceinline__ double	postLuminance(PostColor c)
	{
		#pragma clang fp contract(off)
		return c.r * 0.2722287168 + c.g * 0.6740817658 + c.b * 0.0536895174;
	}

	// sceneLuminance (Image.cpp:37-40)
	__device__ __forceinline__ double	postSceneLuminance(PostColor c) { return postLuminance(postSanitize(c)); }

	__device__ __forceinline__ PostColor	postScale(PostColor c, double s) { return PostColor{c.r * s, c.g * s, c.b * s}; }

	// multiply (ColorManagement.cpp:33-40): row-major 3 x 3 times colour
	__device__ __forceinline__ PostColor	postMul(const double (&m)[3][3], PostColor c)
	{
		#pragma clang fp contract(off)
		return PostColor{m[0][0] * c.r + m[0][1] * c.g + m[0][2] * c.b, m[1][0] * c.r + m[1][1] * c.g + m[1][2] * c.b,
			m[2][0] * c.r + m[2][1] * c.g + m[2][2] * c.b};
	}

	__device__ __forceinline__ PostColor	postClamp(PostColor c) { return PostColor{postUnit(c.r), postUnit(c.g), postUnit(c.b)}; }

	// linearToSRGB / encodeSRGB (ColorManagement.cpp:203-221)
	__device__ __forceinline__ double	postEncode(double v)
	{
		#pragma clang fp contract(off)
		v = postUnit(v);
		if (v <= 0.0031308) return 12.92 * v;
		return (1.055 * pow(v, 1.0 / 2.4)) - 0.055;
	}

	__device__ __forceinline__ PostColor	postEncode(PostColor c) { return PostColor{postEncode(c.r), postEncode(c.g), postEncode(c.b)}; }

	// rrtAndOdtFit (ColorManagement.cpp:51-70)
	__device__ __forceinline__ double	postAcesFit(double v)
	{
		#pragma clang fp contract(off)
		if (!isfinite(v) || v <= 0.0) return 0.0;
		const double a = v * (v + 0.0245786) - 0.000090537;
		const double b = v * (0.983729 * v + 0.4329510) + 0.238081;
		if (b <= 0.0 || !isfinite(b)) return 0.0;
		return a / b;
	}

	// agxDefaultContrast / agxCurve (ColorManagement.cpp:72-110)
	__device__ __forceinline__ double	postAgxContrast(double v)
	{
		#pragma clang fp contract(off)
		v = postStdMin(postStdMax(v, 0.0), 1.0);
		const double v2 = v * v, v3 = v2 * v, v4 = v2 * v2, v5 = v4 * v, v6 = v4 * v2;
		return postUnit((15.5 * v6) - (40.14 * v5) + (31.96 * v4) - (6.868 * v3) + (0.4298 * v2) + (0.1191 * v) - 0.00232);
	}

	__device__ __forceinline__ double	postAgxCurve(double v)
	{
		#pragma clang fp contract(off)
		constexpr double minEv = -12.47393, maxEv = 4.026069;
		if (!isfinite(v) || v <= 0.0) v = minEv;
		else v = postStdMin(postStdMax(log2(v), minEv), maxEv);
		return postAgxContrast((v - minEv) / (maxEv - minEv));
	}

	// viewTransformToLinearSRGB (ColorManagement.cpp:272-300) of a sanitized scene-linear
	// ACEScg colour. The matrices are the reference's colour-science constants
	// (ColorManagement.cpp:120-160).
	__device__ inline PostColor	postViewToLinearSRGB(PostColor c, int view)
	{
		constexpr double toSrgb[3][3] = {{1.7048586763, -0.6217160219, -0.0832993717}, {-0.1300768242, 1.1407357748, -0.0105598017},
			{-0.0239640729, -0.1289755083, 1.1530140189}};
		constexpr double acesIn[3][3] = {{0.59719, 0.35458, 0.04823}, {0.07600, 0.90834, 0.01566}, {0.02840, 0.13383, 0.83777}};
		constexpr double acesOut[3][3] = {{1.60475, -0.53108, -0.07367}, {-0.10208, 1.10813, -0.00605}, {-0.00327, -0.07276, 1.07602}};
		constexpr double agxIn[3][3] = {{0.842479062253094, 0.0784335999999992, 0.0792237451477643},
			{0.0423282422610123, 0.878468636469772, 0.0791661274605434}, {0.0423756549057051, 0.0784336, 0.879142973793104}};
		constexpr double agxOut[3][3] = {{1.19687900512017, -0.0980208811401368, -0.0990297440797205},
			{-0.0528968517574562, 1.15190312990417, -0.0989611768448433}, {-0.0529716355144438, -0.0980434501171241, 1.15107367264116}};
		const PostColor lin = postSanitize(postMul(toSrgb, c));
		switch (view)
		{
			case POST_VIEW_AGX:
			{
				const PostColor a = postMul(agxIn, lin);
				return postClamp(postMul(agxOut, PostColor{postAgxCurve(a.r), postAgxCurve(a.g), postAgxCurve(a.b)}));
			}
			case POST_VIEW_ACES:
			{
				const PostColor a = postMul(acesIn, lin);
				return postClamp(postMul(acesOut, PostColor{postAcesFit(a.r), postAcesFit(a.g), postAcesFit(a.b)}));
			}
			case POST_VIEW_RAW:
				return lin;
			default:
				return postClamp(lin);
		}
	}

	// applyExposure (Image.cpp:467-484)
	__global__ void	post_exposure(const double* __restrict__ in, double* __restrict__ out, PostParams P)
	{
		#pragma clang fp contract(off)
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= static_cast<size_t>(P.width) * P.height) return;
		postStore(out, i, postScale(postSanitize(postLoad(in, i)), P.exposureScale));
	}

	// extractBloom (Image.cpp:599-640) with suppressIsolatedBloomFirefly (Image.cpp:42-103)
	__global__ void	post_bloom_extract(const double* __restrict__ img, double* __restrict__ bright, PostParams P)
	{
		#pragma clang fp contract(off)
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= static_cast<size_t>(P.width) * P.height) return;
		const long long x = static_cast<long long>(i % P.width), y = static_cast<long long>(i / P.width);
		const double threshold = kBloomThreshold;
		PostColor pixel = postSanitize(postLoad(img, i));
		const double lum0 = postSceneLuminance(pixel);
		if (!(threshold <= 0.0 || lum0 <= threshold * 16.0))
		{
			double neighborMax = 0.0;
			unsigned similar = 0, brightN = 0;
			for (int oy = -2; oy <= 2; oy++)
			{
				for (int ox = -2; ox <= 2; ox++)
				{
					if (ox == 0 && oy == 0) continue;
					const long long sx = x + ox, sy = y + oy;
					if (sx < 0 || sy < 0 || sx >= static_cast<long long>(P.width) || sy >= static_cast<long long>(P.height)) continue;
					const double nl = postSceneLuminance(postLoad(img, static_cast<size_t>(sy) * P.width + static_cast<size_t>(sx)));
					neighborMax = postStdMax(neighborMax, nl);
					if (nl >= lum0 / 16.0) similar++;
					if (nl > threshold) brightN++;
				}
			}
			if (!(similar > 0 || brightN >= 3))
			{
				const double ceiling = postStdMax(threshold, neighborMax * 4.0);
				if (lum0 > ceiling) pixel = postScale(pixel, ceiling / lum0);
			}
		}
		const double luminance = postLuminance(pixel);
		double contribution = postStdMax(luminance - threshold, 0.0);
		const double knee = threshold * kBloomSoftKnee;
		const double soft = postStdMin(postStdMax(luminance - threshold + knee, 0.0), 2.0 * knee);	// std::clamp
		contribution = postStdMax(contribution, (soft * soft) / ((4.0 * knee) + 0.00001));
		const double weight = luminance > 0.0 ? contribution / luminance : 0.0;
		postStore(bright, i, postScale(pixel, weight));
	}

	// Gaussian::blur (Gaussian.cpp:80-150): the horizontal pass, clamped at the borders
	__global__ void	post_blur_h(const double* __restrict__ in, double* __restrict__ out, PostParams P)
	{
		#pragma clang fp contract(off)
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= static_cast<size_t>(P.width) * P.height) return;
		const long long x = static_cast<long long>(i % P.width);
		const size_t row = (i / P.width) * P.width;
		PostColor acc{0.0, 0.0, 0.0};
		for (int k = -4; k <= 4; k++)
		{
			const long long sx = x + k < 0 ? 0 : (x + k >= static_cast<long long>(P.width) ? P.width - 1 : x + k);
			const PostColor c = postLoad(in, row + static_cast<size_t>(sx));
			const double w = P.blur[k + 4];
			acc = PostColor{acc.r + c.r * w, acc.g + c.g * w, acc.b + c.b * w};
		}
		postStore(out, i, acc);
	}

	// the vertical pass, and the bloom added back (applyBloom, Threads.cpp:55-75)
	__global__ void	post_blur_v_add(const double* __restrict__ in, double* __restrict__ img, PostParams P)
	{
		#pragma clang fp contract(off)
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= static_cast<size_t>(P.width) * P.height) return;
		const long long y = static_cast<long long>(i / P.width);
		const size_t x = i % P.width;
		PostColor acc{0.0, 0.0, 0.0};
		for (int k = -4; k <= 4; k++)
		{
			const long long sy = y + k < 0 ? 0 : (y + k >= static_cast<long long>(P.height) ? P.height - 1 : y + k);
			const PostColor c = postLoad(in, static_cast<size_t>(sy) * P.width + x);
			const double w = P.blur[k + 4];
			acc = PostColor{acc.r + c.r * w, acc.g + c.g * w, acc.b + c.b * w};
		}
		const PostColor base = postLoad(img, i);
		postStore(img, i, PostColor{base.r + acc.r * kBloomIntensity, base.g + acc.g * kBloomIntensity, base.b + acc.b * kBloomIntensity});
	}

	// The view transform to display-encoded sRGB: applyViewTransformAndEncodeSRGB
	// (Image.cpp:571-597) at the default contrast, else applyViewTransform, applyContrast and
	// gammaCorrect (Image.cpp:486-569). Not launched for the raw view.
	__global__ void	post_view(const double* __restrict__ in, double* __restrict__ out, PostParams P)
	{
		#pragma clang fp contract(off)
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= static_cast<size_t>(P.width) * P.height) return;
		PostColor c = postViewToLinearSRGB(postSanitize(postLoad(in, i)), P.view);
		if (P.contrast != 1.0)
		{
			c = PostColor{postUnit(((c.r - 0.5) * P.contrast) + 0.5), postUnit(((c.g - 0.5) * P.contrast) + 0.5),
				postUnit(((c.b - 0.5) * P.contrast) + 0.5)};
		}
		postStore(out, i, postEncode(c));
	}

	__device__ __forceinline__ double	postMaxChannel(PostColor c) { return postStdMax(c.r, postStdMax(c.g, c.b)); }
	__device__ __forceinline__ double	postMinChannel(PostColor c) { return postStdMin(c.r, postStdMin(c.g, c.b)); }

	// isDisplayFireflyCandidate / isBrightDisplayOutlierCandidate /
	// isUnsupportedSaturatedDisplayPixel (Image.cpp:115-139)
	__device__ __forceinline__ bool	postFireflyCandidate(PostColor c)
	{
		return postMaxChannel(c) >= 0.99 && postMinChannel(c) >= 0.85 && postSceneLuminance(c) >= 0.95;
	}
	__device__ __forceinline__ bool	postBrightOutlier(PostColor c) { return postMaxChannel(c) >= 0.50 && postSceneLuminance(c) >= 0.35; }
	__device__ __forceinline__ bool	postSaturated(PostColor c)
	{
		return postMaxChannel(c) >= 0.995 && postMinChannel(c) >= 0.96 && postSceneLuminance(c) >= 0.98;
	}

	// medianNeighborColor (Image.cpp:141-152): the colour of middle luminance rank. The
	// candidates are ordered by an insertion sort on luminance, which equals the reference's
	// ordering whenever luminances differ or tied candidates are the same colour.
	__device__ inline PostColor	postMedianColor(PostColor* c, int n)
	{
		double key[24];
		for (int k = 0; k < n; k++) key[k] = postSceneLuminance(c[k]);
		for (int a = 1; a < n; a++)
		{
			const double kv = key[a];
			const PostColor cv = c[a];
			int b = a - 1;
			while (b >= 0 && kv < key[b])
			{
				key[b + 1] = key[b];
				c[b + 1] = c[b];
				b--;
			}
			key[b + 1] = kv;
			c[b + 1] = cv;
		}
		return c[n / 2];
	}

	// suppressIsolatedFireflies (Image.cpp:506-526) with suppressDisplayFirefly
	// (Image.cpp:154-268), over the display-encoded image.
	__global__ void	post_fireflies(const double* __restrict__ src, double* __restrict__ dst, PostParams P)
	{
		#pragma clang fp contract(off)
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= static_cast<size_t>(P.width) * P.height) return;
		const PostColor pixel = postLoad(src, i);
		auto keep = [&]() { postStore(dst, i, pixel); };
		if (P.width < 3 || P.height < 3)
		{
			keep();
			return;
		}
		const long long x = static_cast<long long>(i % P.width), y = static_cast<long long>(i / P.width);
		const double luminance = postSceneLuminance(pixel);
		const bool saturatedCandidate = postSaturated(pixel);
		const bool displayCandidate = postFireflyCandidate(pixel);
		if (!displayCandidate && !postBrightOutlier(pixel))
		{
			keep();
			return;
		}
		unsigned similar = 0, brightN = 0, saturated = 0;
		double immediate[8];
		int nImmediate = 0;
		PostColor candidates[24];
		int nCandidates = 0;
		for (int oy = -2; oy <= 2; oy++)
		{
			for (int ox = -2; ox <= 2; ox++)
			{
				if (ox == 0 && oy == 0) continue;
				const long long sx = x + ox, sy = y + oy;
				if (sx < 0 || sy < 0 || sx >= static_cast<long long>(P.width) || sy >= static_cast<long long>(P.height)) continue;
				const PostColor nb = postLoad(src, static_cast<size_t>(sy) * P.width + static_cast<size_t>(sx));
				const double nl = postSceneLuminance(nb);
				if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1) immediate[nImmediate++] = nl;
				if (nl >= luminance * 0.65) similar++;
				if (nl >= 0.80) brightN++;
				if (postSaturated(nb)) saturated++;
				if (!postFireflyCandidate(nb)) candidates[nCandidates++] = nb;
			}
		}
		if (!displayCandidate)
		{
			if (nImmediate == 0)
			{
				keep();
				return;
			}
			for (int a = 1; a < nImmediate; a++)
			{
				const double v = immediate[a];
				int b = a - 1;
				while (b >= 0 && v < immediate[b])
				{
					immediate[b + 1] = immediate[b];
					b--;
				}
				immediate[b + 1] = v;
			}
			const double localMedian = immediate[nImmediate / 2];
			if (luminance < localMedian * 1.65 || luminance < localMedian + 0.08)
			{
				keep();
				return;
			}
			if (nCandidates > 0)
			{
				postStore(dst, i, postMedianColor(candidates, nCandidates));
				return;
			}
		}
		if (saturatedCandidate && saturated < 3u && nCandidates > 0)
		{
			postStore(dst, i, postMedianColor(candidates, nCandidates));
			return;
		}
		if (similar >= 4u || brightN >= 6u || nCandidates == 0)
		{
			keep();
			return;
		}
		postStore(dst, i, postMedianColor(candidates, nCandidates));
	}

	__global__ void	post_to_float(const double* __restrict__ in, float* __restrict__ out, size_t n)
	{
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i < n) out[i] = static_cast<float>(in[i]);
	}
}
