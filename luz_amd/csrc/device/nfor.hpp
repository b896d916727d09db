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
// The regression solves a 12 x 12 weighted least-squares system per pixel. Two lanes own a
// pixel, each half of its system in registers: they walk the window in the reference's
// raster order, so every entry is summed in the reference's order (with a fused
// multiply-add per term), and a second kernel solves each system by the reference's
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

	__device__ __forceinline__ int	nforClampInt(int v, int limit)
	{
		return v < 0 ? 0 : (v >= limit ? limit - 1 : v);
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

	constexpr int	kNforTile = 16;						// pixels per tile side
	constexpr int	kNforTilePixels = kNforTile * kNforTile;
	constexpr int	kNforMaxPatchRadius = 8;
	constexpr int	kNforChunk = 4;						// window displacements per pair of barriers
	constexpr int	kNforHalfSlots = kNforSystem / 2;	// 57 system entries per half

	// The two halves of a pixel's system, by rows of the upper triangle (each row's entries
	// and its three right-hand sides): 39 matrix + 18 right-hand-side entries each.
	__host__ __device__ constexpr int	nforHalfRow(int half, int i)
	{
		constexpr int rows[2][6] = {{0, 3, 4, 7, 8, 11}, {1, 2, 5, 6, 9, 10}};
		return rows[half][i];
	}

	// index of upper-triangle entry (row, col >= row) in the row-major packing
	__host__ __device__ constexpr int	nforUpperIndex(int row, int col)
	{
		return row * kNforBasis - row * (row - 1) / 2 + (col - row);
	}

	// Dynamic LDS of nfor_accumulate, in doubles. Staged: the colours (three planes) and
	// variances of the tile grown by the window and patch radii, and the 11 feature planes of
	// the tile grown by the window radius, each slot holding the pixel of its clamped
	// coordinate. Both forms: the patch terms and the weights of kNforChunk displacements.
	struct	NforLdsLayout
	{
		int	colorWidth, featureWidth, apron;
		std::uint32_t	colors, variances, features, pairs, weights, doubles;
	};

	__host__ __device__ inline NforLdsLayout	nforLdsLayout(std::uint32_t regressionRadius, std::uint32_t patchRadius, bool staged)
	{
		NforLdsLayout L;
		L.colorWidth = staged ? kNforTile + 2 * static_cast<int>(regressionRadius + patchRadius) : 0;
		L.featureWidth = staged ? kNforTile + 2 * static_cast<int>(regressionRadius) : 0;
		L.apron = kNforTile + 2 * static_cast<int>(patchRadius);
		const std::uint32_t cc = static_cast<std::uint32_t>(L.colorWidth * L.colorWidth);
		const std::uint32_t ff = static_cast<std::uint32_t>(L.featureWidth * L.featureWidth);
		L.colors = 0;
		L.variances = 3 * cc;
		L.features = L.variances + cc;
		L.pairs = L.features + kNforFeatures * ff;
		L.weights = L.pairs + kNforChunk * static_cast<std::uint32_t>(L.apron * L.apron);
		L.doubles = L.weights + kNforChunk * kNforTilePixels;
		return L;
	}

	// a half of a fresh system: the ridge 1e-4 on the diagonal, zero elsewhere
	template <int Half>
	__device__ __forceinline__ void	nforClearHalf(double (&acc)[kNforHalfSlots])
	{
		int e = 0;
		#pragma unroll
		for (int i = 0; i < 6; i++)
		{
			const int row = nforHalfRow(Half, i);
			#pragma unroll
			for (int col = row; col < kNforBasis; col++) acc[e++] = col == row ? 1e-4 : 0.0;
			acc[e++] = 0.0;
			acc[e++] = 0.0;
			acc[e++] = 0.0;
		}
	}

	// addRegressionSample (NFOR.cpp:582-600) for one half of the system. Each entry adds
	// (w * basis[row]) * basis[col] with one fused multiply-add: one rounding per sample where
	// the reference rounds the product and the sum separately, so the systems differ from
	// the reference's by FP64 rounding only (the same entries, summed in the same order).
	template <int Half>
	__device__ __forceinline__ void	nforAddSample(double (&acc)[kNforHalfSlots], const double (&basis)[kNforBasis], double w, double cr,
		double cg, double cb)
	{
		#pragma clang fp contract(off)
		int e = 0;
		#pragma unroll
		for (int i = 0; i < 6; i++)
		{
			const int row = nforHalfRow(Half, i);
			const double wb = w * basis[row];
			#pragma unroll
			for (int col = row; col < kNforBasis; col++)
			{
				acc[e] = fma(wb, basis[col], acc[e]);
				e++;
			}
			acc[e] = fma(wb, cr, acc[e]);
			acc[e + 1] = fma(wb, cg, acc[e + 1]);
			acc[e + 2] = fma(wb, cb, acc[e + 2]);
			e += 3;
		}
	}

	// The half's entries into the pixel's system (slot-major: entry k of pixel p at k * n + p;
	// entries 0..77 the upper triangle, 78 + 3 * row + channel the right-hand sides).
	template <int Half>
	__device__ __forceinline__ void	nforStoreHalf(const double (&acc)[kNforHalfSlots], double* __restrict__ system, size_t n, size_t pixel)
	{
		int e = 0;
		#pragma unroll
		for (int i = 0; i < 6; i++)
		{
			const int row = nforHalfRow(Half, i);
			#pragma unroll
			for (int col = row; col < kNforBasis; col++) system[static_cast<size_t>(nforUpperIndex(row, col)) * n + pixel] = acc[e++];
			#pragma unroll
			for (int j = 0; j < 3; j++) system[static_cast<size_t>(kNforUpper + 3 * row + j) * n + pixel] = acc[e++];
		}
	}

	// regressionPass (NFOR.cpp:679-737), accumulation: a 16 x 16 tile per block of 512 lanes,
	// two per pixel, each owning half of the pixel's 12 x 12 system and its right-hand sides
	// (nforHalfRow). The halves are whole waves (lanes 0-255 and 256-511), so every branch on
	// the half is wave-uniform and every register index static; both halves form the pixel's
	// weights. The window is visited in the reference's raster order, so every entry is summed
	// in the reference's order; nfor_solve then solves the systems.
	//
	// Patch distances: for a window displacement d, the term of patch offset o between pixel
	// p and its sample p + d is D(p + o, p + o + d) (patchDistance, NFOR.cpp:240-311), which
	// depends on p + o only. For kNforChunk displacements at a time the block evaluates
	// D(q, q + d) once for every q of its tile and apron into LDS, the two halves form the
	// chunk's weights between them (two displacements each) into LDS, and both then add the
	// chunk's samples; three barriers fence the steps. Lanes whose patch or sample would be
	// clamped at the image border take the direct, clamped form.
	//
	// Staged: the block first copies every colour, variance and feature its window reads into
	// LDS, so the per-step reads are LDS reads (radius 9, patch radius 1: 158 KiB); the
	// unstaged form reads them from global memory, for windows too wide for LDS.
	template <bool Staged>
	__global__ __launch_bounds__(2 * kNforTilePixels) void	nfor_accumulate(const double* __restrict__ image,
		const double* __restrict__ features, const double* __restrict__ variance, double* __restrict__ system, NforParams P,
		double bandwidth)
	{
		#pragma clang fp contract(off)
		extern __shared__ double lds[];
		const NforLdsLayout L = nforLdsLayout(P.regressionRadius, P.patchRadius, Staged);
		double* const colors = lds + L.colors;
		double* const vars = lds + L.variances;
		double* const feats = lds + L.features;
		double* const pairs = lds + L.pairs;
		double* const weights = lds + L.weights;
		const int cw = L.colorWidth, fw = L.featureWidth, apron = L.apron;
		const int cc = cw * cw, ff = fw * fw, aa = apron * apron;
		const int lane = static_cast<int>(threadIdx.x) % kNforTilePixels;
		const int half = static_cast<int>(threadIdx.x) / kNforTilePixels;	// wave-uniform
		// each 32-lane half of a wave takes tile rows h and h + 8: 8 rows of the staged feature
		// (34 wide) and patch-term (18 wide) planes apart, their 16 doubles fall on disjoint
		// LDS banks
		const int tx = lane % kNforTile, ty = lane / 32 + ((lane % 32) >= kNforTile ? 8 : 0);
		const std::uint32_t tilesX = (P.width + kNforTile - 1) / kNforTile;
		const long long x0 = static_cast<long long>(blockIdx.x % tilesX) * kNforTile;
		const long long y0 = static_cast<long long>(blockIdx.x / tilesX) * kNforTile;
		const std::uint32_t x = static_cast<std::uint32_t>(x0) + tx, y = static_cast<std::uint32_t>(y0) + ty;
		const bool active = x < P.width && y < P.height;
		const std::uint32_t c = active ? y * P.width + x : 0u;
		const long long r = P.regressionRadius, pr = P.patchRadius;
		const long long cx0 = x0 - r - pr, cy0 = y0 - r - pr;	// staged colour region origin
		const long long fx0 = x0 - r, fy0 = y0 - r;				// staged feature region origin
		if constexpr (Staged)
		{
			for (int i = static_cast<int>(threadIdx.x); i < cc; i += 2 * kNforTilePixels)
			{
				const std::uint32_t g = nforClamp(cy0 + i / cw, P.height) * P.width + nforClamp(cx0 + i % cw, P.width);
				colors[i] = image[3 * static_cast<size_t>(g)];
				colors[cc + i] = image[3 * static_cast<size_t>(g) + 1];
				colors[2 * cc + i] = image[3 * static_cast<size_t>(g) + 2];
				vars[i] = variance[g];
			}
			for (int i = static_cast<int>(threadIdx.x); i < ff; i += 2 * kNforTilePixels)
			{
				const std::uint32_t g = nforClamp(fy0 + i / fw, P.height) * P.width + nforClamp(fx0 + i % fw, P.width);
				#pragma unroll
				for (int j = 0; j < kNforFeatures; j++) feats[j * ff + i] = features[kNforFeatures * static_cast<size_t>(g) + j];
			}
			__syncthreads();
		}
		const double patchTerms = static_cast<double>((2 * pr + 1) * (2 * pr + 1));
		const bool patchInside = active && static_cast<long long>(x) >= pr && static_cast<long long>(y) >= pr
			&& static_cast<long long>(x) + pr < static_cast<long long>(P.width) && static_cast<long long>(y) + pr < static_cast<long long>(P.height);
		double acc[kNforHalfSlots];
		if (half == 0) nforClearHalf<0>(acc);
		else nforClearHalf<1>(acc);
		// 32-bit index arithmetic from here on (the image and window sizes are far below 2^31)
		const int ix = static_cast<int>(x), iy = static_cast<int>(y);
		const int iw = static_cast<int>(P.width), ih = static_cast<int>(P.height);
		const int ir = static_cast<int>(r), ipr = static_cast<int>(pr);
		const int fBase = (ty + ir) * fw + (tx + ir);				// the pixel's staged feature slot
		const int cBase = (ty + ir + ipr) * cw + (tx + ir + ipr);	// the pixel's staged colour slot
		const double bw2 = fmax(P.minVariance, bandwidth * bandwidth);
		const int side = 2 * ir + 1;
		const int steps = side * side;
		// the chunk's displacements, raster order: (stepY[k], stepX[k]) = (oy, ox)
		int stepY[kNforChunk], stepX[kNforChunk];
		// one window sample: its weight w, its basis (1, features - the pixel's) and colour
		auto sampleAt = [&]<int Half>(int k, double w) {
			const int oy = stepY[k], ox = stepX[k];
			double basis[kNforBasis];
			double cr, cg, cb;
			basis[0] = 1.0;
			if constexpr (Staged)
			{
				const int fi = fBase + oy * fw + ox;
				const int ci = cBase + oy * cw + ox;
				#pragma unroll
				for (int j = 0; j < kNforFeatures; j++) basis[j + 1] = feats[j * ff + fi] - feats[j * ff + fBase];
				cr = colors[ci];
				cg = colors[cc + ci];
				cb = colors[2 * cc + ci];
			}
			else
			{
				const size_t sp = static_cast<size_t>(nforClampInt(iy + oy, ih)) * P.width + nforClampInt(ix + ox, iw);
				#pragma unroll
				for (int j = 0; j < kNforFeatures; j++)
				{
					basis[j + 1] = features[kNforFeatures * sp + j] - features[kNforFeatures * static_cast<size_t>(c) + j];
				}
				cr = image[3 * sp];
				cg = image[3 * sp + 1];
				cb = image[3 * sp + 2];
			}
			nforAddSample<Half>(acc, basis, w, cr, cg, cb);
		};
		// colorWeightFromDistance (NFOR.cpp:313-318) of one window displacement
		auto weightAt = [&](int k) {
			const int ux = ix + stepX[k], uy = iy + stepY[k];
			const int sx = nforClampInt(ux, iw), sy = nforClampInt(uy, ih);
			double distance = 0.0;
			if (sx == ix && sy == iy)
			{
				distance = 0.0;
			}
			else if (patchInside && ux == sx && uy == sy && ux >= ipr && uy >= ipr && ux + ipr < iw && uy + ipr < ih)
			{
				const double* terms = pairs + k * aa;
				if (pr == 1)
				{
					// the common 3 x 3 patch: the nine reads issued together, then summed in order
					double t[9];
					#pragma unroll
					for (int i = 0; i < 9; i++) t[i] = terms[(ty + i / 3) * apron + (tx + i % 3)];
					#pragma unroll
					for (int i = 0; i < 9; i++) distance += t[i];
				}
				else
				{
					for (long long py = -pr; py <= pr; py++)
					{
						for (long long px = -pr; px <= pr; px++)
						{
							distance += terms[(ty + pr + py) * apron + (tx + pr + px)];
						}
					}
				}
				distance = distance / patchTerms;
			}
			else
			{
				distance = nforPatchDistance(image, variance, P, x, y, static_cast<std::uint32_t>(sx), static_cast<std::uint32_t>(sy),
					P.patchRadius);
			}
			return exp(-distance / bw2);
		};
		int nextY = -ir, nextX = -ir;
		for (int base = 0; base < steps; base += kNforChunk)
		{
			const int count = min(kNforChunk, steps - base);
			#pragma unroll
			for (int k = 0; k < kNforChunk; k++)
			{
				stepY[k] = nextY;
				stepX[k] = nextX;
				if (k < count && ++nextX > ir)
				{
					nextX = -ir;
					nextY++;
				}
			}
			// D(q, q + d) over the tile and its apron for the chunk's displacements (entries off
			// the image are clamped reads whose values no lane uses)
			for (int i = static_cast<int>(threadIdx.x); i < count * aa; i += 2 * kNforTilePixels)
			{
				const int k = i / aa, j = i % aa;
				const int oy = stepY[k], ox = stepX[k];
				const int jx = j % apron, jy = j / apron;
				if constexpr (Staged)
				{
					const int a = (jy + ir) * cw + (jx + ir);	// q = tile origin - pr + (jx, jy)
					const int b = a + oy * cw + ox;
					const double dr = colors[a] - colors[b], dg = colors[cc + a] - colors[cc + b], db = colors[2 * cc + a] - colors[2 * cc + b];
					pairs[i] = (((dr * dr) + (dg * dg) + (db * db)) / 3.0) / (vars[a] + vars[b] + P.minVariance);
				}
				else
				{
					const long long qx = x0 - pr + jx, qy = y0 - pr + jy;
					const std::uint32_t qa = nforClamp(qy, P.height) * P.width + nforClamp(qx, P.width);
					const std::uint32_t qb = nforClamp(qy + oy, P.height) * P.width + nforClamp(qx + ox, P.width);
					pairs[i] = nforDist2(image + 3 * static_cast<size_t>(qa), image + 3 * static_cast<size_t>(qb))
						/ (variance[qa] + variance[qb] + P.minVariance);
				}
			}
			__syncthreads();
			// the halves split the chunk's weights (steps k = half, half + 2), then both add
			// every sample of the chunk
			if (active)
			{
				if (count == kNforChunk)
				{
					// two independent exp chains in flight
					const double w0 = weightAt(half), w1 = weightAt(half + 2);
					weights[half * kNforTilePixels + lane] = w0;
					weights[(half + 2) * kNforTilePixels + lane] = w1;
				}
				else
				{
					for (int k = half; k < count; k += 2) weights[k * kNforTilePixels + lane] = weightAt(k);
				}
			}
			__syncthreads();
			if (active)
			{
				if (half == 0)
				{
					for (int k = 0; k < count; k++) sampleAt.template operator()<0>(k, weights[k * kNforTilePixels + lane]);
				}
				else
				{
					for (int k = 0; k < count; k++) sampleAt.template operator()<1>(k, weights[k * kNforTilePixels + lane]);
				}
			}
			__syncthreads();
		}
		if (!active) return;
		const size_t n = static_cast<size_t>(P.width) * P.height;
		if (half == 0) nforStoreHalf<0>(acc, system, n, c);
		else nforStoreHalf<1>(acc, system, n, c);
	}

	// completeRegressionMatrix (NFOR.cpp:628-637), then solveLinearSystemRGB (:466-516): the
	// reference's Gauss-Jordan elimination with partial pivoting on each pixel's system from
	// nfor_accumulate; a singular system takes weightedAverageFallback.
	__global__ __launch_bounds__(256) void	nfor_solve(const double* __restrict__ system, const double* __restrict__ image,
		const double* __restrict__ variance, double* __restrict__ out, NforParams P)
	{
		#pragma clang fp contract(off)
		const size_t n = static_cast<size_t>(P.width) * P.height;
		const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (c >= n) return;
		const std::uint32_t x = static_cast<std::uint32_t>(c % P.width), y = static_cast<std::uint32_t>(c / P.width);
		double upper[kNforUpper];
		double rhs[kNforBasis][3];
		#pragma unroll
		for (int k = 0; k < kNforUpper; k++) upper[k] = system[static_cast<size_t>(k) * n + c];
		#pragma unroll
		for (int row = 0; row < kNforBasis; row++)
		{
			#pragma unroll
			for (int j = 0; j < 3; j++) rhs[row][j] = system[static_cast<size_t>(kNforUpper + 3 * row + j) * n + c];
		}
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
		double* o = out + 3 * c;
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
