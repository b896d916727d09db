// Adaptive per-pixel stopping: the running FP64 statistics of a pixel's samples
// (AdaptiveAccumulator, Threads.cpp:31-53) and the stopping rule checked after a sample
// (adaptiveSampleConverged, Threads.cpp:194-262).
//
// On the device a render proceeds in rounds over a list of still-active pixels: the first
// round renders the minimum sample count of every pixel, each later round `checkInterval`
// more samples (capped at the scene's total) of the pixels left active. The rule can only
// stop a pixel at min + k * interval samples or at the total, which are exactly the round
// boundaries, so checking once per round stops every pixel after the same number of
// samples as the per-sample loop of _threadRender (Threads.cpp:551-583).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace luzgpu
{
	// AdaptiveAccumulator fields, in doubles per pixel
	constexpr int	kAdaptiveColorSum = 0;		// 3
	constexpr int	kAdaptiveColorSquare = 3;	// 3
	constexpr int	kAdaptiveLumSum = 6;
	constexpr int	kAdaptiveLumSquare = 7;
	constexpr int	kAdaptiveMaxLum = 8;
	constexpr int	kAdaptiveMoments = 9;

	// The scene's stopping rule with its derived counts resolved on the host:
	// minSamples = min(adaptiveminsamples, samples), darkMinimum = darkMinimumSamples
	// (Threads.cpp:183-192).
	struct	AdaptiveRule
	{
		std::uint32_t	maxSamples;
		std::uint32_t	minSamples;
		std::uint32_t	checkInterval;
		std::uint32_t	darkMinimum;
		double	threshold;
	};

	// std::max's choice (the first argument unless it compares below the second), which
	// differs from fmax when an argument is NaN.
	__device__ __forceinline__ double	stdMax(double a, double b) { return a < b ? b : a; }

	// AdaptiveAccumulator::add (Threads.cpp:39-52) with one cleaned sample colour. The
	// products and sums are rounded separately, as the reference compiles them.
	__device__ __forceinline__ void	adaptiveAdd(double* m, double r, double g, double b)
	{
		#pragma clang fp contract(off)
		// sampleLuminance (Threads.cpp:139-148) = ColorManagement::luminance (ColorManagement.cpp:313-320)
		double lum = r * 0.2722287168 + g * 0.6740817658 + b * 0.0536895174;
		if (!isfinite(lum)) lum = 0.0;
		m[kAdaptiveColorSum] += r;
		m[kAdaptiveColorSum + 1] += g;
		m[kAdaptiveColorSum + 2] += b;
		m[kAdaptiveColorSquare] += r * r;
		m[kAdaptiveColorSquare + 1] += g * g;
		m[kAdaptiveColorSquare + 2] += b * b;
		m[kAdaptiveLumSum] += lum;
		m[kAdaptiveLumSquare] += lum * lum;
		m[kAdaptiveMaxLum] = stdMax(m[kAdaptiveMaxLum], lum);
	}

	// confidenceInterval95 over sampleMeanVariance (Threads.cpp:150-172)
	__device__ __forceinline__ double	confidenceInterval95(double sum, double square, std::uint32_t count)
	{
		#pragma clang fp contract(off)
		double meanVariance = 0.0;
		if (count > 1u)
		{
			const double n = static_cast<double>(count);
			const double variance = (square - (sum * sum / n)) / (n - 1.0);
			meanVariance = stdMax(0.0, variance / n);
		}
		if (!isfinite(meanVariance)) return __builtin_huge_val();
		return 1.96 * sqrt(meanVariance);
	}

	// channelConverged (Threads.cpp:174-180)
	__device__ __forceinline__ bool	channelConverged(double mean, double ci, double threshold)
	{
		#pragma clang fp contract(off)
		const double target = threshold * stdMax(fabs(mean), 0.02);
		return ci <= target;
	}

	// adaptiveSampleConverged (Threads.cpp:194-262) after `used` samples of a pixel.
	__device__ __forceinline__ bool	adaptiveConverged(const AdaptiveRule& r, std::uint32_t used, const double* m)
	{
		#pragma clang fp contract(off)
		if (used >= r.maxSamples) return true;
		if (used < r.minSamples) return false;
		if ((used - r.minSamples) % r.checkInterval != 0u) return false;
		if (used <= 1u) return false;
		const double n = static_cast<double>(used);
		const double lumMean = m[kAdaptiveLumSum] / n;
		if (m[kAdaptiveMaxLum] < 0.02 && used < r.darkMinimum) return false;
		const double lumCI = confidenceInterval95(m[kAdaptiveLumSum], m[kAdaptiveLumSquare], used);
		const double rCI = confidenceInterval95(m[kAdaptiveColorSum], m[kAdaptiveColorSquare], used);
		const double gCI = confidenceInterval95(m[kAdaptiveColorSum + 1], m[kAdaptiveColorSquare + 1], used);
		const double bCI = confidenceInterval95(m[kAdaptiveColorSum + 2], m[kAdaptiveColorSquare + 2], used);
		return channelConverged(lumMean, lumCI, r.threshold)
			&& channelConverged(m[kAdaptiveColorSum] / n, rCI, r.threshold)
			&& channelConverged(m[kAdaptiveColorSum + 1] / n, gCI, r.threshold)
			&& channelConverged(m[kAdaptiveColorSum + 2] / n, bCI, r.threshold);
	}

	// After a round that brought every listed pixel to `used` samples: records `used` for
	// each, and appends the pixels the rule keeps active to `next` (order unspecified; one
	// atomic per wave). `list` == nullptr lists every pixel 0..count-1.
	__global__ void	adaptive_check(AdaptiveRule rule, const std::uint32_t* __restrict__ list, std::uint32_t count, std::uint32_t used,
		const double* __restrict__ moments, std::uint32_t* __restrict__ usedOut, std::uint32_t* __restrict__ next,
		std::uint32_t* __restrict__ nextCount)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		bool active = false;
		std::uint32_t pixel = 0;
		if (i < count)
		{
			pixel = list ? list[i] : i;
			usedOut[pixel] = used;
			active = !adaptiveConverged(rule, used, moments + static_cast<size_t>(kAdaptiveMoments) * pixel);
		}
		const unsigned long long mask = __ballot(active);
		const std::uint32_t lane = threadIdx.x & 63u;
		const int leader = __ffsll(static_cast<long long>(mask)) - 1;
		std::uint32_t base = 0;
		if (mask && static_cast<int>(lane) == leader) base = atomicAdd(nextCount, static_cast<std::uint32_t>(__popcll(mask)));
		base = __shfl(base, leader < 0 ? 0 : leader);
		if (active) next[base + static_cast<std::uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)))] = pixel;
	}
}
