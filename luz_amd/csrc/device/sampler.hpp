// Bit-exact device restatement of the reference's stateless sampler
// (src/core/Sampler.cpp:28-237, dimensions include/luz/Sampler.hpp:15-37).
// A sample is a pure function of (render seed, x, y, stream, sample index, bounce,
// dimension): SplitMix64 hashing for most dimensions and a per-dimension rank-1 shift
// `fract(0.5 + (i + 1) * step + shift)` in FP64 for the seven "progressive" ones.
// FP64 adds/multiplies are IEEE-exact on CDNA; contraction is disabled below so the
// FP64 expression rounds exactly as the reference's separate mul and adds do.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace luzgpu
{
	enum SamplerDim : std::uint32_t
	{
		DIM_CAMERA = 0,
		DIM_LENS = 1,
		DIM_LIGHT_STRATEGY = 2,
		DIM_LIGHT_SELECTION = 3,
		DIM_LIGHT_SURFACE_SELECTION = 4,
		DIM_LIGHT_SURFACE_POINT = 5,
		DIM_BSDF_DIRECTION = 6,
		DIM_MATERIAL_DECISION = 7,
		DIM_MATERIAL_FUZZ = 8,
		DIM_RUSSIAN_ROULETTE = 9,
		DIM_VOLUME_DISTANCE = 10,
		DIM_ATMOSPHERE = 11,
		DIM_ENVIRONMENT_SELECTION = 12,
		DIM_ENVIRONMENT_POINT = 13,
		DIM_ENVIRONMENT_STRATEGY = 14,
		DIM_LIGHT_EMISSION_SIDE = 15,
		DIM_LIGHT_EMISSION_DIRECTION = 16,
		DIM_SUBSURFACE_SPATIAL = 17,
		DIM_SUBSURFACE_PROFILE_COMPONENT = 18,
		DIM_SUBSURFACE_THIN_SIDE = 19,
	};

	__host__ __device__ constexpr std::uint64_t	splitMix64(std::uint64_t v)
	{
		v += 0x9e3779b97f4a7c15ull;
		v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
		v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
		return (v ^ (v >> 31));
	}

	__host__ __device__ constexpr double	unitFromHash32(std::uint32_t h)
	{
		return (static_cast<double>(h >> 8) * (1.0 / 16777216.0));
	}

	__host__ __device__ inline double	unitFromHash64(std::uint64_t h)
	{
		return (static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0));
	}

	__host__ __device__ constexpr double	dimensionStep(std::uint32_t dim, std::uint32_t component)
	{
#pragma clang fp contract(off)
		const std::uint64_t key = (static_cast<std::uint64_t>(dim) + 1ull) * 0x9e3779b97f4a7c15ull
			^ (static_cast<std::uint64_t>(component) + 1ull) * 0xda942042e4dd58b5ull;
		double step = unitFromHash32(static_cast<std::uint32_t>(splitMix64(key) >> 32));
		if (step < 0.05 || step > 0.95)
		{
			// x - floor(x) for x = step + 0.618...: x lies in [0.618, 0.668) or (1.568, 1.618),
			// so its floor is 0 or 1 and x - 1 is exact
			const double x = step + 0.6180339887498948;
			step = x >= 1.0 ? x - 1.0 : x;
		}
		return (step);
	}

	// dimensionStep depends on the dimension key and the component alone: the steps of the
	// first 64 bounces (keys below 2048) are evaluated at compile time, the rest when used.
	constexpr std::uint32_t	kStepTableKeys = 32u * 64u;
	struct	DimensionSteps
	{
		double	v[2 * kStepTableKeys];
	};
	constexpr DimensionSteps	makeDimensionSteps()
	{
		DimensionSteps t{};
		for (std::uint32_t k = 0; k < kStepTableKeys; k++)
		{
			t.v[2 * k] = dimensionStep(k, 0);
			t.v[2 * k + 1] = dimensionStep(k, 1);
		}
		return t;
	}
	static __constant__ DimensionSteps	kDimensionSteps = makeDimensionSteps();

	__device__ __forceinline__ double	stepOf(std::uint32_t dimKey, std::uint32_t component)
	{
		return dimKey < kStepTableKeys ? kDimensionSteps.v[2 * dimKey + component] : dimensionStep(dimKey, component);
	}

	struct	PixelSampler
	{
		std::uint64_t	baseKey;
		std::uint32_t	sampleIndex;
		std::uint32_t	bounce;

		__host__ __device__ void	begin(std::uint32_t seed, std::uint32_t x, std::uint32_t y, std::uint32_t sample, std::uint32_t stream)
		{
			baseKey = seed;
			baseKey ^= (static_cast<std::uint64_t>(x) + 1ull) * 0xbf58476d1ce4e5b9ull;
			baseKey ^= (static_cast<std::uint64_t>(y) + 1ull) * 0x94d049bb133111ebull;
			baseKey ^= (static_cast<std::uint64_t>(stream) + 1ull) * 0xda942042e4dd58b5ull;
			sampleIndex = sample;
			bounce = 0;
		}

		__host__ __device__ static bool	progressive(std::uint32_t dim)
		{
			const std::uint32_t b = dim % 32u;
			return (b == 0 || b == 1 || b == 4 || b == 5 || b == 12 || b == 13 || b == 16);
		}

		__host__ __device__ std::uint64_t	contextKey(std::uint32_t dimKey) const
		{
			return (baseKey ^ ((static_cast<std::uint64_t>(dimKey) + 1ull) * 0x9e3779b97f4a7c15ull));
		}

		__host__ __device__ double	hashed(std::uint32_t dimKey, std::uint32_t component) const
		{
			std::uint64_t key = contextKey(dimKey);
			key ^= (static_cast<std::uint64_t>(sampleIndex) + 1ull) * 0x94d049bb133111ebull;
			key ^= (static_cast<std::uint64_t>(component) + 1ull) * 0xbf58476d1ce4e5b9ull;
			return (unitFromHash64(splitMix64(key)));
		}

		__device__ double	sample1D(std::uint32_t dim) const
		{
#pragma clang fp contract(off)
			const std::uint32_t dk = dim + bounce * 32u;
			if (!progressive(dim))
			{
				return (hashed(dk, 0));
			}
			const double shift = unitFromHash32(static_cast<std::uint32_t>(splitMix64(contextKey(dk)) >> 32));
			const double step = stepOf(dk, 0);
			const double s = static_cast<double>(sampleIndex) + 1.0;
			const double prod = s * step;
			const double x = (0.5 + prod) + shift;
			return (x - floor(x));
		}

		__device__ void	sample2D(std::uint32_t dim, float& sx, float& sy) const
		{
#pragma clang fp contract(off)
			const std::uint32_t dk = dim + bounce * 32u;
			if (!progressive(dim))
			{
				sx = static_cast<float>(hashed(dk, 0));
				sy = static_cast<float>(hashed(dk, 1));
				return;
			}
			const std::uint64_t key = contextKey(dk);
			const double shiftX = unitFromHash32(static_cast<std::uint32_t>(splitMix64(key) >> 32));
			const double shiftY = unitFromHash32(static_cast<std::uint32_t>(splitMix64(key ^ 0xda942042e4dd58b5ull) >> 32));
			const double stepX = stepOf(dk, 0);
			const double stepY = stepOf(dk, 1);
			const double s = static_cast<double>(sampleIndex) + 1.0;
			const double px = s * stepX;
			const double py = s * stepY;
			const double x = (0.5 + px) + shiftX;
			const double y = (0.5 + py) + shiftY;
			sx = static_cast<float>(x - floor(x));
			sy = static_cast<float>(y - floor(y));
		}
	};
}
