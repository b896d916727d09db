// NFOR denoiser on the device (Denoise::applyNFOR, src/core/Denoise/NFOR.cpp:977-1102), in
// FP64 like the reference.
//
// Inputs are the half-buffer statistics the integrator accumulates with denoising on
// (Threads.cpp:274-372): mean colours of the even and odd samples (A / B), the per-pixel
// colour variance, the mean primary features of each half (11 per pixel) and their variance.
// Stages, one kernel each:
//   firefly suppression of A, B and of their average (NFOR.cpp:145-211);
//   feature prefiltering, each half guided by the other (NL-means weights, :403-464);
//   first-order feature regression of A with B's features and of B with A's, at two
//   bandwidths (k = 0.5, 1) (:739-832);
//   MSE estimation, its and the bandwidth selection's cross-bilateral filtering (:834-936);
//   per-pixel blend, then a final regression of the blend (:1074-1101).
//
// The regression solves a 12 x 12 weighted least-squares system per pixel. Each lane owns
// one pixel: it walks the window in the reference's raster order, so every sum is formed in
// the reference's order, keeps the system in registers and solves it by the reference's
// Gauss-Jordan elimination with partial pivoting.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace luzgpu
{
	constexpr int	kNforFeatures = 11;
	constexpr int	kNforBasis = kNforFeatures + 1;			// REGRESSION_SIZE
	constexpr int	kNforUpper = kNforBasis * (kNforBasis + 1) / 2;	// 78 upper-triangle entries
	constexpr int	kNforSystem = kNforUpper + kNforBasis * 3;		// + the RGB right-hand side

	// Per-pixel FP64 statistics of a denoised render, in sample order (_threadRender with
	// denoise buffers, Threads.cpp:585-627): [0,3) half A colour sum, [3,6) half B, [6,9) all
	// colours, [9,12) colour squares, then kNforFeatures each of half A, half B, all features
	// and feature squares.
	constexpr int	kDenoiseStats = 12 + 4 * kNforFeatures;

	// Adds one sample (cleaned colour, primary features) to a pixel's statistics `a`;
	// `odd`: the sample index is odd (half B, sampler stream 2).
	__device__ __forceinline__ void	denoiseAddSample(double* a, const float* color, const float* feat, bool odd)
	{
		#pragma clang fp contract(off)
		const int half = odd ? 3 : 0;
		for (int j = 0; j < 3; j++)
		{
			const double c = color[j];
			a[half + j] += c;
			a[6 + j] += c;
			a[9 + j] += c * c;
		}
		const int fh = odd ? 12 + kNforFeatures : 12;
		for (int j = 0; j < kNforFeatures; j++)
		{
			const double f = feat[j];
			a[fh + j] += f;
			a[12 + 2 * kNforFeatures + j] += f;
			a[12 + 3 * kNforFeatures + j] += f * f;
		}
	}

	struct	NforParams
	{
		std::uint32_t	width, height;
		std::uint32_t	prefilterRadius, regressionRadius, patchRadius;
		double	minVariance;
	};

	__device__ __forceinline__ std::uint32_t	nforClamp(long long v, std::uint32_t limit)
	{
		return v < 0 ? 0u : (v >= static_cast<long long>(limit) ? limit - 1u : static_cast<std::uint32_t>(v));
	}

	// colorDistanceSquared, colorLuminance (NFOR.cpp:84-103)
	__device__ __forceinline__ double	nforDist2(const double* a, const double* b)
	{
		const double r = a[0] - b[0], g = a[1] - b[1], bl = a[2] - b[2];
		return ((r * r) + (g * g) + (bl * bl)) / 3.0;
	}

	__device__ __forceinline__ double	nforLuminance(const double* c)
	{
		const double l = c[0] * 0.2722287168 + c[1] * 0.6740817658 + c[2] * 0.0536895174;
		return (!isfinite(l) || l <= 0.0) ? 0.0 : l;
	}

	// patchDistance (NFOR.cpp:236-311); border patches clamp their coordinates
	__device__ inline double	nforPatchDistance(const double* guide, const double* variance, const NforParams& P, std::uint32_t x1,
		std::uint32_t y1, std::uint32_t x2, std::uint32_t y2, std::uint32_t patchRadius)
	{
		if (x1 == x2 && y1 == y2) return 0.0;
		const long long r = patchRadius;
		double distance = 0.0;
		unsigned samples = 0;
		for (long long oy = -r; oy <= r; oy++)
		{
			for (long long ox = -r; ox <= r; ox++)
			{
				const std::uint32_t ia = nforClamp(static_cast<long long>(y1) + oy, P.height) * P.width + nforClamp(static_cast<long long>(x1) + ox, P.width);
				const std::uint32_t ib = nforClamp(static_cast<long long>(y2) + oy, P.height) * P.width + nforClamp(static_cast<long long>(x2) + ox, P.width);
				const double lv = variance[ia] + variance[ib] + P.minVariance;
				distance += nforDist2(guide + 3 * ia, guide + 3 * ib) / lv;
				samples++;
			}
		}
		return distance / static_cast<double>(samples);
	}

	// colorWeightFromDistance (NFOR.cpp:313-318)
	__device__ __forceinline__ double	nforWeight(double distance, double bandwidth, double minVariance)
	{
		return exp(-distance / fmax(minVariance, bandwidth * bandwidth));
	}

	// suppressDenoiseFireflies (NFOR.cpp:145-211): a bright pixel with fewer than two similar
	// neighbours in its 5 x 5 window that stands 2.5x and 0.18 above the window's median
	// luminance takes the luminance-median of its darker neighbours. `average` != 0: the
	// input is the mean of a and b (averageColorBuffers) instead of a.
	__global__ void	nfor_fireflies(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
		NforParams P, int average)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= P.width * P.height) return;
		auto colorAt = [&](std::uint32_t k, double* c) {
			if (average)
			{
				for (int j = 0; j < 3; j++) c[j] = (a[3 * k + j] + b[3 * k + j]) * 0.5;
			}
			else
			{
				for (int j = 0; j < 3; j++) c[j] = a[3 * k + j];
			}
		};
		double center[3];
		colorAt(i, center);
		for (int j = 0; j < 3; j++) out[3 * i + j] = center[j];
		if (P.width < 3 || P.height < 3) return;
		const std::uint32_t x = i % P.width, y = i / P.width;
		const double lum = nforLuminance(center);
		double neighbor[24];
		std::uint32_t candidate[24];
		double candidateLum[24];
		int n = 0, nc = 0;
		unsigned similar = 0;
		for (int oy = -2; oy <= 2; oy++)
		{
			for (int ox = -2; ox <= 2; ox++)
			{
				if (ox == 0 && oy == 0) continue;
				const long long sx = static_cast<long long>(x) + ox, sy = static_cast<long long>(y) + oy;
				if (sx < 0 || sy < 0 || sx >= P.width || sy >= P.height) continue;
				const std::uint32_t k = static_cast<std::uint32_t>(sy) * P.width + static_cast<std::uint32_t>(sx);
				double c[3];
				colorAt(k, c);
				const double l = nforLuminance(c);
				neighbor[n++] = l;
				if (l >= lum * 0.65) similar++;
				else
				{
					candidate[nc] = k;
					candidateLum[nc] = l;
					nc++;
				}
			}
		}
		// isDenoiseFirefly (NFOR.cpp:115-143)
		if (nc == 0 || lum < 0.30 || n == 0 || similar >= 2u) return;
		for (int p = 1; p < n; p++)
		{
			const double v = neighbor[p];
			int q = p - 1;
			while (q >= 0 && neighbor[q] > v)
			{
				neighbor[q + 1] = neighbor[q];
				q--;
			}
			neighbor[q + 1] = v;
		}
		const double median = neighbor[n / 2];
		if (!(lum >= median * 2.5 && lum >= median + 0.18)) return;
		// medianColorByLuminance (NFOR.cpp:105-113)
		for (int p = 1; p < nc; p++)
		{
			const double v = candidateLum[p];
			const std::uint32_t k = candidate[p];
			int q = p - 1;
			while (q >= 0 && candidateLum[q] > v)
			{
				candidateLum[q + 1] = candidateLum[q];
				candidate[q + 1] = candidate[q];
				q--;
			}
			candidateLum[q + 1] = v;
			candidate[q + 1] = k;
		}
		double c[3];
		colorAt(candidate[nc / 2], c);
		for (int j = 0; j < 3; j++) out[3 * i + j] = c[j];
	}

	// doubledVariance and averageFeatureVarianceMap (NFOR.cpp:331-353)
	__global__ void	nfor_variances(const double* __restrict__ colorVariance, const double* __restrict__ featureVariance,
		double* __restrict__ regressionVariance, double* __restrict__ featureVarianceAverage, std::uint32_t n)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= n) return;
		regressionVariance[i] = colorVariance[i] * 2.0;
		double v = 0.0;
		for (int k = 0; k < kNforFeatures; k++) v += featureVariance[kNforFeatures * i + k];
		featureVarianceAverage[i] = v / static_cast<double>(kNforFeatures);
	}

	// prefilterFeatures (NFOR.cpp:403-464): `image`'s features averaged with weights from
	// `guide`'s feature distances and a spatial Gaussian.
	__global__ void	nfor_prefilter(const double* __restrict__ image, const double* __restrict__ guide,
		const double* __restrict__ varianceAverage, double* __restrict__ out, NforParams P)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= P.width * P.height) return;
		const std::uint32_t x = i % P.width, y = i / P.width;
		const long long r = P.prefilterRadius;
		const double sigma = fmax(1.0, static_cast<double>(P.prefilterRadius));
		double sum[kNforFeatures] = {};
		double total = 0.0;
		for (long long oy = -r; oy <= r; oy++)
		{
			for (long long ox = -r; ox <= r; ox++)
			{
				const std::uint32_t k = nforClamp(static_cast<long long>(y) + oy, P.height) * P.width + nforClamp(static_cast<long long>(x) + ox, P.width);
				const double lv = varianceAverage[i] + varianceAverage[k] + P.minVariance;
				double d = 0.0;
				for (int f = 0; f < kNforFeatures; f++)
				{
					const double diff = guide[kNforFeatures * i + f] - guide[kNforFeatures * k + f];
					d += diff * diff;
				}
				const double distance = (d / static_cast<double>(kNforFeatures)) / lv;
				const double spatial = static_cast<double>(ox * ox + oy * oy);
				const double w = exp(-distance / 0.25) * exp(-spatial / (2.0 * sigma * sigma));
				for (int f = 0; f < kNforFeatures; f++) sum[f] += image[kNforFeatures * k + f] * w;
				total += w;
			}
		}
		for (int f = 0; f < kNforFeatures; f++)
		{
			out[kNforFeatures * i + f] = total <= 2.220446049250313e-16 ? image[kNforFeatures * i + f] : sum[f] / total;
		}
	}

	// weightedAverageFallback (NFOR.cpp:518-553), when a system is singular
	__device__ inline void	nforFallback(const double* image, const double* variance, const NforParams& P, std::uint32_t x, std::uint32_t y,
		double* out)
	{
		const long long r = P.regressionRadius;
		double acc[3] = {0.0, 0.0, 0.0}, total = 0.0;
		for (long long oy = -r; oy <= r; oy++)
		{
			for (long long ox = -r; ox <= r; ox++)
			{
				const std::uint32_t k = nforClamp(static_cast<long long>(y) + oy, P.height) * P.width + nforClamp(static_cast<long long>(x) + ox, P.width);
				const double spatial = static_cast<double>(ox * ox + oy * oy);
				const double w = (1.0 / (variance[k] + P.minVariance)) * exp(-spatial / 8.0);
				for (int j = 0; j < 3; j++) acc[j] += image[3 * k + j] * w;
				total += w;
			}
		}
		const std::uint32_t c = y * P.width + x;
		for (int j = 0; j < 3; j++) out[j] = total <= 2.220446049250313e-16 ? image[3 * c + j] : acc[j] / total;
	}

	constexpr int	kNforTile = 16;						// pixels per tile side (one lane each)
	constexpr int	kNforMaxPatchRadius = 8;
	constexpr int	kNforApronMax = kNforTile + 2 * kNforMaxPatchRadius;

	// regressionPass (NFOR.cpp:679-737) with one lane per pixel, a 16 x 16 tile per block.
	// The window is visited in the reference's raster order and every sum is formed in its
	// order. The system lives in registers (78 upper-triangle entries + 36 right-hand-side
	// entries while accumulating, the completed 12 x 15 system while solving, every index
	// static after unrolling) and is solved by the reference's Gauss-Jordan elimination with
	// partial pivoting (solveLinearSystemRGB, NFOR.cpp:466-516). A pair pass
	// (regressionPairPass, :739-832) is two launches, one per bandwidth.
	//
	// Patch distances: for a window displacement d, the term of patch offset o between pixel
	// p and its sample p + d is D(p + o, p + o + d) (patchDistance, NFOR.cpp:240-311), which
	// depends on p + o only. The block evaluates D(q, q + d) once for every q of its tile and
	// apron into LDS, and each lane sums its (2r+1)^2 terms from there in the reference's
	// order. Lanes whose patch or sample would be clamped at the image border take the
	// direct, clamped form.
	__global__ __launch_bounds__(kNforTile * kNforTile) void	nfor_regression_tile(const double* __restrict__ image,
		const double* __restrict__ features, const double* __restrict__ variance, double* __restrict__ out, NforParams P,
		double bandwidth)
	{
		#pragma clang fp contract(off)
		__shared__ double pairTerms[2][kNforApronMax * kNforApronMax];
		const int tx = static_cast<int>(threadIdx.x) % kNforTile, ty = static_cast<int>(threadIdx.x) / kNforTile;
		const std::uint32_t tilesX = (P.width + kNforTile - 1) / kNforTile;
		const long long x0 = static_cast<long long>(blockIdx.x % tilesX) * kNforTile;
		const long long y0 = static_cast<long long>(blockIdx.x / tilesX) * kNforTile;
		const std::uint32_t x = static_cast<std::uint32_t>(x0) + tx, y = static_cast<std::uint32_t>(y0) + ty;
		const bool active = x < P.width && y < P.height;
		const std::uint32_t c = active ? y * P.width + x : 0u;
		const long long pr = P.patchRadius;
		const int apron = kNforTile + 2 * static_cast<int>(pr);
		const double patchTerms = static_cast<double>((2 * pr + 1) * (2 * pr + 1));
		const bool patchInside = active && static_cast<long long>(x) >= pr && static_cast<long long>(y) >= pr
			&& static_cast<long long>(x) + pr < static_cast<long long>(P.width) && static_cast<long long>(y) + pr < static_cast<long long>(P.height);
		double upper[kNforUpper];
		double rhs[kNforBasis][3];
		{
			int k = 0;
			#pragma unroll
			for (int row = 0; row < kNforBasis; row++)
			{
				#pragma unroll
				for (int col = row; col < kNforBasis; col++) upper[k++] = (col == row) ? 1e-4 : 0.0;	// ridge
				rhs[row][0] = rhs[row][1] = rhs[row][2] = 0.0;
			}
		}
		double center[kNforFeatures];
		#pragma unroll
		for (int j = 0; j < kNforFeatures; j++) center[j] = features[kNforFeatures * static_cast<size_t>(c) + j];
		const double bw2 = fmax(P.minVariance, bandwidth * bandwidth);
		const long long r = P.regressionRadius;
		int buf = 0;
		for (long long oy = -r; oy <= r; oy++)
		{
			for (long long ox = -r; ox <= r; ox++)
			{
				// D(q, q + d) over the tile and its apron (entries off the image are clamped
				// reads whose values no lane uses); double-buffered, so one barrier per step
				for (int i = static_cast<int>(threadIdx.x); i < apron * apron; i += kNforTile * kNforTile)
				{
					const long long qx = x0 - pr + i % apron, qy = y0 - pr + i / apron;
					const std::uint32_t qa = nforClamp(qy, P.height) * P.width + nforClamp(qx, P.width);
					const std::uint32_t qb = nforClamp(qy + oy, P.height) * P.width + nforClamp(qx + ox, P.width);
					pairTerms[buf][i] = nforDist2(image + 3 * static_cast<size_t>(qa), image + 3 * static_cast<size_t>(qb))
						/ (variance[qa] + variance[qb] + P.minVariance);
				}
				__syncthreads();
				const double* terms = pairTerms[buf];
				buf ^= 1;
				if (!active) continue;
				const long long ux = static_cast<long long>(x) + ox, uy = static_cast<long long>(y) + oy;
				const std::uint32_t sx = nforClamp(ux, P.width), sy = nforClamp(uy, P.height);
				const size_t s = static_cast<size_t>(sy) * P.width + sx;
				double distance = 0.0;
				if (sx == x && sy == y)
				{
					distance = 0.0;
				}
				else if (patchInside && ux == sx && uy == sy && ux >= pr && uy >= pr && ux + pr < static_cast<long long>(P.width)
					&& uy + pr < static_cast<long long>(P.height))
				{
					for (long long py = -pr; py <= pr; py++)
					{
						for (long long px = -pr; px <= pr; px++)
						{
							distance += terms[(ty + pr + py) * apron + (tx + pr + px)];
						}
					}
					distance = distance / patchTerms;
				}
				else
				{
					distance = nforPatchDistance(image, variance, P, x, y, sx, sy, P.patchRadius);
				}
				const double w = exp(-distance / bw2);	// colorWeightFromDistance (NFOR.cpp:313-318)
				double basis[kNforBasis];
				basis[0] = 1.0;
				#pragma unroll
				for (int j = 0; j < kNforFeatures; j++) basis[j + 1] = features[kNforFeatures * s + j] - center[j];
				const double cr = image[3 * s], cg = image[3 * s + 1], cb = image[3 * s + 2];
				// addRegressionSample (NFOR.cpp:582-600)
				int k = 0;
				#pragma unroll
				for (int row = 0; row < kNforBasis; row++)
				{
					const double wb = w * basis[row];
					#pragma unroll
					for (int col = row; col < kNforBasis; col++) upper[k++] += wb * basis[col];
					rhs[row][0] += wb * cr;
					rhs[row][1] += wb * cg;
					rhs[row][2] += wb * cb;
				}
			}
		}
		if (!active) return;
		// completeRegressionMatrix (NFOR.cpp:628-637), then solveLinearSystemRGB
		double m[kNforBasis][kNforBasis];
		{
			int k = 0;
			#pragma unroll
			for (int row = 0; row < kNforBasis; row++)
			{
				#pragma unroll
				for (int col = row; col < kNforBasis; col++)
				{
					m[row][col] = upper[k];
					m[col][row] = upper[k];
					k++;
				}
			}
		}
		bool solved = true;
		#pragma unroll
		for (int pivot = 0; pivot < kNforBasis; pivot++)
		{
			int best = pivot;
			double bestValue = fabs(m[pivot][pivot]);
			#pragma unroll
			for (int row = pivot + 1; row < kNforBasis; row++)
			{
				const double v = fabs(m[row][pivot]);
				if (v > bestValue)
				{
					bestValue = v;
					best = row;
				}
			}
			if (bestValue < 1e-10)
			{
				solved = false;
				break;
			}
			// row swap (columns below the pivot are zero in every row already)
			#pragma unroll
			for (int row = pivot + 1; row < kNforBasis; row++)
			{
				if (row == best)
				{
					#pragma unroll
					for (int col = pivot; col < kNforBasis; col++)
					{
						const double t = m[pivot][col];
						m[pivot][col] = m[row][col];
						m[row][col] = t;
					}
					#pragma unroll
					for (int j = 0; j < 3; j++)
					{
						const double t = rhs[pivot][j];
						rhs[pivot][j] = rhs[row][j];
						rhs[row][j] = t;
					}
				}
			}
			const double pivotValue = m[pivot][pivot];
			#pragma unroll
			for (int col = pivot; col < kNforBasis; col++) m[pivot][col] /= pivotValue;
			#pragma unroll
			for (int j = 0; j < 3; j++) rhs[pivot][j] /= pivotValue;
			#pragma unroll
			for (int row = 0; row < kNforBasis; row++)
			{
				if (row == pivot) continue;
				const double factor = m[row][pivot];
				#pragma unroll
				for (int col = pivot; col < kNforBasis; col++) m[row][col] -= factor * m[pivot][col];
				#pragma unroll
				for (int j = 0; j < 3; j++) rhs[row][j] -= factor * rhs[pivot][j];
			}
		}
		double* o = out + 3 * static_cast<size_t>(c);
		if (solved)
		{
			o[0] = rhs[0][0];
			o[1] = rhs[0][1];
			o[2] = rhs[0][2];
		}
		else
		{
			nforFallback(image, variance, P, x, y, o);
		}
	}

	// estimateMSE (NFOR.cpp:875-893) of the first and the second bandwidth
	__global__ void	nfor_mse(const double* __restrict__ colorA, const double* __restrict__ colorB, const double* __restrict__ a0,
		const double* __restrict__ b0, const double* __restrict__ a1, const double* __restrict__ b1,
		const double* __restrict__ variance, double* __restrict__ mse0, double* __restrict__ mse1, std::uint32_t n)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= n) return;
		auto mse = [&](const double* fa, const double* fb) {
			const double mseA = nforDist2(colorB + 3 * i, fa + 3 * i) - 2.0 * variance[i];
			const double mseB = nforDist2(colorA + 3 * i, fb + 3 * i) - 2.0 * variance[i];
			const double residual = nforDist2(fa + 3 * i, fb + 3 * i) * 0.25;
			return fmax(0.0, ((mseA + mseB) * 0.5) - residual);
		};
		mse0[i] = mse(a0, b0);
		mse1[i] = mse(a1, b1);
	}

	// filterScalarMapPair (NFOR.cpp:834-873): radius-3 cross-bilateral filter of two maps,
	// weights from 3 x 3 colour patches of `guide` (bandwidth 1). `select` != 0: the maps are
	// first turned into bandwidthSelectionMap(first, second) and (second, first) (:907-919).
	__global__ void	nfor_filter_pair(const double* __restrict__ first, const double* __restrict__ second,
		const double* __restrict__ guide, const double* __restrict__ guideVariance, double* __restrict__ out0,
		double* __restrict__ out1, NforParams P, int select)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= P.width * P.height) return;
		const std::uint32_t x = i % P.width, y = i / P.width;
		auto value = [&](std::uint32_t k, double& v0, double& v1) {
			if (select)
			{
				v0 = first[k] <= second[k] ? 1.0 : 0.0;
				v1 = second[k] <= first[k] ? 1.0 : 0.0;
			}
			else
			{
				v0 = first[k];
				v1 = second[k];
			}
		};
		double f0 = 0.0, f1 = 0.0, total = 0.0;
		for (long long oy = -3; oy <= 3; oy++)
		{
			for (long long ox = -3; ox <= 3; ox++)
			{
				const std::uint32_t sx = nforClamp(static_cast<long long>(x) + ox, P.width);
				const std::uint32_t sy = nforClamp(static_cast<long long>(y) + oy, P.height);
				const std::uint32_t k = sy * P.width + sx;
				const double w = nforWeight(nforPatchDistance(guide, guideVariance, P, x, y, sx, sy, 1u), 1.0, P.minVariance);
				double v0, v1;
				value(k, v0, v1);
				f0 += v0 * w;
				f1 += v1 * w;
				total += w;
			}
		}
		if (total > 2.220446049250313e-16)
		{
			out0[i] = f0 / total;
			out1[i] = f1 / total;
		}
		else
		{
			value(i, out0[i], out1[i]);
		}
	}

	// The per-pixel blend of applyNFOR (NFOR.cpp:1072-1087) and combinedResultVariance
	// (:895-905): combined colour, blended features and the final regression's variance.
	__global__ void	nfor_combine(const double* __restrict__ sel0, const double* __restrict__ sel1, const double* __restrict__ mse0,
		const double* __restrict__ mse1, const double* __restrict__ a0, const double* __restrict__ a1, const double* __restrict__ b0,
		const double* __restrict__ b1, const double* __restrict__ featA, const double* __restrict__ featB,
		double* __restrict__ combined, double* __restrict__ features, double* __restrict__ variance, NforParams P)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= P.width * P.height) return;
		const double total = sel0[i] + sel1[i];
		const double w0 = total > P.minVariance ? sel0[i] / total : (mse0[i] <= mse1[i] ? 1.0 : 0.0);
		for (int j = 0; j < 3; j++)
		{
			const double ra = a0[3 * i + j] * w0 + a1[3 * i + j] * (1.0 - w0);
			const double rb = b0[3 * i + j] * w0 + b1[3 * i + j] * (1.0 - w0);
			combined[3 * i + j] = (ra + rb) * 0.5;
		}
		for (int f = 0; f < kNforFeatures; f++)
		{
			features[kNforFeatures * i + f] = featA[kNforFeatures * i + f] * 0.5 + featB[kNforFeatures * i + f] * (1.0 - 0.5);
		}
		variance[i] = nforDist2(a1 + 3 * i, b1 + 3 * i) * 0.25;
	}
}
