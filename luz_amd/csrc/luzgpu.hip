// luzgpu: device memory management, kernels and the C ABI (include/luzgpu.h).
//
// Kernels
//   render_mega      one thread per (pixel, sample): the whole path in registers, the
//                    traversal stack in LDS. The plain single-scene kernel.
//   render_wavefront stages (raygen / extend / shade / shadow) over SoA path state in
//                    HBM with wave-level ballot compaction; see wavefront.hpp.
//   resolve          per pixel: sum of the per-sample colours in sample order, divided
//                    by the sample count, cleaned (Threads.cpp:563-583).
#include <hip/hip_runtime.h>

#include "../../include/luzgpu.h"
#include "common/layout.h"
#include "device/adjoint.hpp"
#include "device/integrator.hpp"
#include "device/nfor.hpp"
#include "device/post.hpp"
#include "device/wavefront.hpp"
#include "host/accel.hpp"

#include <atomic>
#include <unordered_set>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace luzgpu
{
	namespace
	{
		thread_local std::string	g_lastError;

		// Device addresses of `scratch` (luzgpu_render_noalloc), one per device, looked up by
		// luzgpu_create with that device current: in host code the array's name is only a
		// stand-in, not a device address.
		constexpr int	kScratchDevices = 64;
		std::mutex	g_scratchMutex;
		void*	g_scratchPtr[kScratchDevices] = {};

		struct	Error
		{
			int	code;
			std::string	message;
		};

		void	check(hipError_t e, const char* what)
		{
			if (e != hipSuccess)
			{
				throw Error{e == hipErrorOutOfMemory ? LUZGPU_E_OOM : LUZGPU_E_DEVICE,
					std::string(what) + ": " + hipGetErrorString(e)};
			}
		}

		struct	DeviceBuffer
		{
			void*	ptr = nullptr;
			size_t	bytes = 0;

			DeviceBuffer() = default;
			DeviceBuffer(const DeviceBuffer&) = delete;
			DeviceBuffer& operator=(const DeviceBuffer&) = delete;
			~DeviceBuffer() { release(); }

			void	release()
			{
				if (ptr) (void)hipFree(ptr);
				ptr = nullptr;
				bytes = 0;
			}

			void	reserve(size_t n)
			{
				if (n <= bytes) return;
				release();
				check(hipMalloc(&ptr, n), "hipMalloc");
				bytes = n;
			}

			template <typename T>
			void	upload(const std::vector<T>& v, hipStream_t stream)
			{
				if (v.empty()) return;
				reserve(v.size() * sizeof(T));
				check(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
			}

			template <typename T> T*	as() const { return static_cast<T*>(ptr); }
		};

		// A HIP event destroyed with its scope, so refused or failing calls leak none.
		struct	Event
		{
			hipEvent_t	e = nullptr;

			Event() { check(hipEventCreate(&e), "hipEventCreate"); }
			Event(const Event&) = delete;
			Event& operator=(const Event&) = delete;
			~Event() { if (e) (void)hipEventDestroy(e); }

			void	record(hipStream_t st) { check(hipEventRecord(e, st), "hipEventRecord"); }
			operator hipEvent_t() const { return e; }
		};

		float	elapsedMs(const Event& a, const Event& b)
		{
			float ms = 0.0f;
			check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
			return ms;
		}
	}

	// Scalars the kernels need beyond GpuParams.
	struct	LaunchCounters
	{
		unsigned long long	closestRays;
		unsigned long long	shadowRays;
	};

	constexpr int	kMegaBlock = 128;
	constexpr std::uint32_t	kMaxImageSide = 65536;
	constexpr std::uint64_t	kMaxImagePixels = 1ull << 28;
	constexpr int	kStackDepth = 48;

	// Device memory of luzgpu_render_noalloc, which allocates nothing: one fixed array per
	// device, held until the process ends. Layout: 8 ray / visit counters, the FP64 pixel
	// sums of at most kNoAllocMaxPixels pixels, then the per-sample colours of one plain-kernel
	// launch (kNoAllocStagingSamples of them), which also hold the resolved fp32 image before
	// a readback to host memory. Every call on a device shares that device's copy.
	constexpr std::uint32_t	kNoAllocMaxPixels = 1u << 20;
	constexpr std::uint32_t	kNoAllocStagingSamples = 2u << 20;
	constexpr size_t	kNoAllocCountersBytes = 256;
	constexpr size_t	kNoAllocSumsBytes = static_cast<size_t>(kNoAllocMaxPixels) * 3 * sizeof(double);
	constexpr size_t	kNoAllocStagingBytes = static_cast<size_t>(kNoAllocStagingSamples) * 3 * sizeof(float);
	constexpr size_t	kNoAllocScratchBytes = kNoAllocCountersBytes + kNoAllocSumsBytes + kNoAllocStagingBytes;
	static_assert(kNoAllocStagingSamples >= kNoAllocMaxPixels, "the staging area must hold one fp32 image");

	static __device__ __align__(256) unsigned char	scratch[kNoAllocScratchBytes];

	// counters: [0] closest-hit rays, [1] shadow rays, [2] nodes visited, [3] primitives tested.
	// Renders samples [sampleBegin, sampleBegin + sampleCount) of `pixels` pixels: entries of
	// `list` (adaptive rounds), or the whole image when `list` is nullptr.
	template <bool Count, int Feat>
	__global__ __launch_bounds__(kMegaBlock) void	render_mega(SceneView s, GpuParams p, const std::uint32_t* __restrict__ list,
		std::uint32_t pixels, float* __restrict__ sampleColors, unsigned long long* __restrict__ counters)
	{
		__shared__ std::uint32_t stackMem[kStackDepth * kMegaBlock];
		const std::uint64_t gid = static_cast<std::uint64_t>(blockIdx.x) * kMegaBlock + threadIdx.x;
		std::uint32_t closest = 0, shadow = 0;
		VisitCounter visits{0u, 0u};
		if (gid < static_cast<std::uint64_t>(pixels) * p.sampleCount)
		{
			const std::uint32_t slot = static_cast<std::uint32_t>(gid / pixels);
			const std::uint32_t entry = static_cast<std::uint32_t>(gid % pixels);
			const std::uint32_t pixel = list ? list[entry] : entry;
			LdsStack<kStackDepth, kMegaBlock> stack{stackMem + threadIdx.x, 0};
			const f3 c = cleanColor(tracePath<Count, Feat>(s, p, pixel % p.width, pixel / p.width, p.sampleBegin + slot, stack, closest, shadow, visits));
			float* out = sampleColors + (static_cast<size_t>(slot) * pixels + entry) * 3;
			out[0] = c.x;
			out[1] = c.y;
			out[2] = c.z;
		}
		// one atomic per wave for the ray counters
		for (int off = 32; off > 0; off >>= 1)
		{
			closest += __shfl_down(closest, off);
			shadow += __shfl_down(shadow, off);
			if (Count)
			{
				visits.nodes += __shfl_down(visits.nodes, off);
				visits.prims += __shfl_down(visits.prims, off);
			}
		}
		if ((threadIdx.x & 63) == 0)
		{
			atomicAdd(counters + 0, static_cast<unsigned long long>(closest));
			atomicAdd(counters + 1, static_cast<unsigned long long>(shadow));
			if (Count)
			{
				atomicAdd(counters + 2, static_cast<unsigned long long>(visits.nodes));
				atomicAdd(counters + 3, static_cast<unsigned long long>(visits.prims));
			}
		}
	}

	template <bool Count>
	void	launchMegaC(int features, size_t grid, hipStream_t st, const SceneView& s, const GpuParams& p, const std::uint32_t* list,
		std::uint32_t pixels, float* colors, unsigned long long* counters)
	{
		switch (features & kFeatEvery)
		{
			case kFeatEvery: render_mega<Count, kFeatEvery><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, counters); break;
			case 0: render_mega<Count, 0><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, counters); break;
			case kFeatAnalytic: render_mega<Count, kFeatAnalytic><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, counters); break;
			case kFeatRich: render_mega<Count, kFeatRich><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, counters); break;
			default: render_mega<Count, kFeatAll><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, counters); break;
		}
	}

	void	launchMega(bool count, int features, size_t grid, hipStream_t st, const SceneView& s, const GpuParams& p,
		const std::uint32_t* list, std::uint32_t pixels, float* colors, unsigned long long* counters)
	{
		if (count) launchMegaC<true>(features, grid, st, s, p, list, pixels, colors, counters);
		else launchMegaC<false>(features, grid, st, s, p, list, pixels, colors, counters);
	}

	// One scene of a batched plain-kernel launch. Its (pixel, sample) threads start on a wave
	// boundary, so every wave renders one scene and reads that scene's parameters with
	// scalar loads.
	struct	MegaJob
	{
		GpuParams	params;			// sampleBegin / sampleCount: this launch's slice
		std::uint32_t	firstWave;	// first wave of the slice in the launch
		std::uint32_t	pixels;
		std::uint32_t	count;		// pixels * sampleCount
		std::uint32_t	pixelBase;	// first pixel of the scene in the launch's pixel numbering
		std::uint64_t	colorBase;	// first sample colour of the slice
		double*	sums;			// the scene's FP64 pixel sums
	};

	// Several scenes in one launch (render_batch on the plain kernel): small scenes that
	// cannot fill the GPU alone run side by side, each path still in one thread.
	template <bool Count, int Feat>
	__global__ __launch_bounds__(kMegaBlock) void	render_mega_batch(SceneView s, const MegaJob* __restrict__ jobs, std::uint32_t jobCount,
		float* __restrict__ sampleColors, unsigned long long* __restrict__ counters)
	{
		__shared__ std::uint32_t stackMem[kStackDepth * kMegaBlock];
		const std::uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * kMegaBlock + threadIdx.x) >> 6);
		std::uint32_t lo = 0, hi = jobCount;
		while (hi - lo > 1u)
		{
			const std::uint32_t mid = (lo + hi) >> 1;
			if (jobs[mid].firstWave <= wave) lo = mid;
			else hi = mid;
		}
		const MegaJob& J = jobs[__builtin_amdgcn_readfirstlane(lo)];
		const std::uint32_t local = (wave - J.firstWave) * 64u + (threadIdx.x & 63u);
		std::uint32_t closest = 0, shadow = 0;
		VisitCounter visits{0u, 0u};
		if (wave >= J.firstWave && local < J.count)
		{
			const GpuParams& p = J.params;
			const std::uint32_t slot = local / J.pixels, pixel = local % J.pixels;
			LdsStack<kStackDepth, kMegaBlock> stack{stackMem + threadIdx.x, 0};
			const f3 c = cleanColor(tracePath<Count, Feat>(s, p, pixel % p.width, pixel / p.width, p.sampleBegin + slot, stack, closest,
				shadow, visits));
			float* out = sampleColors + (J.colorBase + local) * 3;
			out[0] = c.x;
			out[1] = c.y;
			out[2] = c.z;
		}
		for (int off = 32; off > 0; off >>= 1)
		{
			closest += __shfl_down(closest, off);
			shadow += __shfl_down(shadow, off);
			if (Count)
			{
				visits.nodes += __shfl_down(visits.nodes, off);
				visits.prims += __shfl_down(visits.prims, off);
			}
		}
		if ((threadIdx.x & 63) == 0)
		{
			atomicAdd(counters + 0, static_cast<unsigned long long>(closest));
			atomicAdd(counters + 1, static_cast<unsigned long long>(shadow));
			if (Count)
			{
				atomicAdd(counters + 2, static_cast<unsigned long long>(visits.nodes));
				atomicAdd(counters + 3, static_cast<unsigned long long>(visits.prims));
			}
		}
	}

	// Adds a batched launch's sample colours into every scene's FP64 pixel sums, in sample
	// order: one thread per pixel of the launch.
	__global__ void	accumulate_batch(const float* __restrict__ colors, const MegaJob* __restrict__ jobs, std::uint32_t jobCount,
		std::uint32_t totalPixels)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= totalPixels) return;
		std::uint32_t lo = 0, hi = jobCount;
		while (hi - lo > 1u)
		{
			const std::uint32_t mid = (lo + hi) >> 1;
			if (jobs[mid].pixelBase <= i) lo = mid;
			else hi = mid;
		}
		const MegaJob& J = jobs[lo];
		const std::uint32_t px = i - J.pixelBase;
		double r = J.sums[3 * static_cast<size_t>(px)], g = J.sums[3 * static_cast<size_t>(px) + 1], b = J.sums[3 * static_cast<size_t>(px) + 2];
		for (std::uint32_t k = 0; k < J.params.sampleCount; k++)
		{
			const float* c = colors + (J.colorBase + static_cast<std::uint64_t>(k) * J.pixels + px) * 3;
			r += static_cast<double>(c[0]);
			g += static_cast<double>(c[1]);
			b += static_cast<double>(c[2]);
		}
		J.sums[3 * static_cast<size_t>(px)] = r;
		J.sums[3 * static_cast<size_t>(px) + 1] = g;
		J.sums[3 * static_cast<size_t>(px) + 2] = b;
	}

	template <bool Count>
	void	launchMegaBatchC(int features, size_t grid, hipStream_t st, const SceneView& s, const MegaJob* jobs, std::uint32_t n,
		float* colors, unsigned long long* counters)
	{
		switch (features & kFeatEvery)
		{
			case kFeatEvery: render_mega_batch<Count, kFeatEvery><<<grid, kMegaBlock, 0, st>>>(s, jobs, n, colors, counters); break;
			case 0: render_mega_batch<Count, 0><<<grid, kMegaBlock, 0, st>>>(s, jobs, n, colors, counters); break;
			case kFeatAnalytic: render_mega_batch<Count, kFeatAnalytic><<<grid, kMegaBlock, 0, st>>>(s, jobs, n, colors, counters); break;
			case kFeatRich: render_mega_batch<Count, kFeatRich><<<grid, kMegaBlock, 0, st>>>(s, jobs, n, colors, counters); break;
			default: render_mega_batch<Count, kFeatAll><<<grid, kMegaBlock, 0, st>>>(s, jobs, n, colors, counters); break;
		}
	}

	// The plain kernel with denoising on: each (pixel, sample) also writes its primary
	// features (kNforFeatures floats, RenderSample of _calculatePixelSample, ColorHelper.cpp:1119-1126).
	template <int Feat>
	__global__ __launch_bounds__(kMegaBlock) void	render_mega_denoise(SceneView s, GpuParams p, const std::uint32_t* __restrict__ list,
		std::uint32_t pixels, float* __restrict__ sampleColors, float* __restrict__ sampleFeatures, unsigned long long* __restrict__ counters)
	{
		__shared__ std::uint32_t stackMem[kStackDepth * kMegaBlock];
		const std::uint64_t gid = static_cast<std::uint64_t>(blockIdx.x) * kMegaBlock + threadIdx.x;
		std::uint32_t closest = 0, shadow = 0;
		VisitCounter visits{0u, 0u};
		if (gid < static_cast<std::uint64_t>(pixels) * p.sampleCount)
		{
			const std::uint32_t slot = static_cast<std::uint32_t>(gid / pixels);
			const std::uint32_t entry = static_cast<std::uint32_t>(gid % pixels);
			const std::uint32_t pixel = list ? list[entry] : entry;
			LdsStack<kStackDepth, kMegaBlock> stack{stackMem + threadIdx.x, 0};
			float f[kNforFeatures];
			const f3 c = cleanColor(tracePath<false, Feat>(s, p, pixel % p.width, pixel / p.width, p.sampleBegin + slot, stack, closest,
				shadow, visits, static_cast<NoAdjoint*>(nullptr), f));
			const size_t k = static_cast<size_t>(slot) * pixels + entry;
			sampleColors[3 * k] = c.x;
			sampleColors[3 * k + 1] = c.y;
			sampleColors[3 * k + 2] = c.z;
			for (int j = 0; j < kNforFeatures; j++) sampleFeatures[kNforFeatures * k + j] = f[j];
		}
		for (int off = 32; off > 0; off >>= 1)
		{
			closest += __shfl_down(closest, off);
			shadow += __shfl_down(shadow, off);
		}
		if ((threadIdx.x & 63) == 0)
		{
			if (closest) atomicAdd(counters + 0, static_cast<unsigned long long>(closest));
			if (shadow) atomicAdd(counters + 1, static_cast<unsigned long long>(shadow));
		}
	}

	void	launchMegaDenoise(int features, size_t grid, hipStream_t st, const SceneView& s, const GpuParams& p, const std::uint32_t* list,
		std::uint32_t pixels, float* colors, float* feats, unsigned long long* counters)
	{
		switch (features & kFeatEvery)
		{
			case kFeatEvery: render_mega_denoise<kFeatEvery><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, feats, counters); break;
			case 0: render_mega_denoise<0><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, feats, counters); break;
			case kFeatAnalytic: render_mega_denoise<kFeatAnalytic><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, feats, counters); break;
			case kFeatRich: render_mega_denoise<kFeatRich><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, feats, counters); break;
			default: render_mega_denoise<kFeatAll><<<grid, kMegaBlock, 0, st>>>(s, p, list, pixels, colors, feats, counters); break;
		}
	}

	// `list` / `pixels`: as render_mega_denoise; `moments` (adaptive rounds, or nullptr):
	// the pixels' AdaptiveAccumulator.
	__global__ void	denoise_accumulate(const float* __restrict__ colors, const float* __restrict__ feats, double* __restrict__ acc,
		const std::uint32_t* __restrict__ list, std::uint32_t pixels, std::uint32_t count, std::uint32_t firstSample,
		double* __restrict__ moments)
	{
		#pragma clang fp contract(off)
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= pixels) return;
		const size_t px = list ? list[i] : i;
		double a[kDenoiseStats];
		for (int j = 0; j < kDenoiseStats; j++) a[j] = acc[static_cast<size_t>(kDenoiseStats) * px + j];
		double m[kAdaptiveMoments];
		if (moments)
		{
			for (int j = 0; j < kAdaptiveMoments; j++) m[j] = moments[kAdaptiveMoments * px + j];
		}
		for (std::uint32_t k = 0; k < count; k++)
		{
			const size_t idx = static_cast<size_t>(k) * pixels + i;
			denoiseAddSample(a, colors + 3 * idx, feats + kNforFeatures * idx, ((firstSample + k) & 1u) != 0u);
			if (moments) adaptiveAdd(m, colors[3 * idx], colors[3 * idx + 1], colors[3 * idx + 2]);
		}
		for (int j = 0; j < kDenoiseStats; j++) acc[static_cast<size_t>(kDenoiseStats) * px + j] = a[j];
		if (moments)
		{
			for (int j = 0; j < kAdaptiveMoments; j++) moments[kAdaptiveMoments * px + j] = m[j];
		}
	}

	// sampleMeanVariance (Threads.cpp:150-161)
	__device__ __forceinline__ double	sampleMeanVariance(double sum, double square, std::uint32_t n)
	{
		if (n <= 1u) return 0.0;
		const double m = static_cast<double>(n);
		return fmax(0.0, ((square - (sum * sum / m)) / (m - 1.0)) / m);
	}

	// storeDenoisePixel (Threads.cpp:340-372) and the pixel's raw mean (:628-631): the NFOR
	// input buffers (FP64) and the image.
	// `used` (adaptive renders, or nullptr): each pixel's sample count instead of `n`.
	__global__ void	denoise_finalize(const double* __restrict__ acc, std::uint32_t pixels, std::uint32_t uniformCount,
		const std::uint32_t* __restrict__ used, float* __restrict__ image, double* __restrict__ meanImage, double* __restrict__ colorA,
		double* __restrict__ colorB, double* __restrict__ colorVariance, double* __restrict__ featA, double* __restrict__ featB,
		double* __restrict__ featVariance)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= pixels) return;
		const std::uint32_t n = used ? used[i] : uniformCount;
		const double* a = acc + static_cast<size_t>(kDenoiseStats) * i;
		const std::uint32_t nA = (n + 1u) / 2u, nB = n / 2u;	// samples 0, 2, 4... and 1, 3, 5...
		const double dn = static_cast<double>(n);
		double cv = 0.0;
		for (int j = 0; j < 3; j++)
		{
			const double fallback = n == 0u ? 0.0 : a[6 + j] / dn;
			colorA[3 * i + j] = nA == 0u ? fallback : a[j] / static_cast<double>(nA);
			colorB[3 * i + j] = nB == 0u ? fallback : a[3 + j] / static_cast<double>(nB);
			cv += sampleMeanVariance(a[6 + j], a[9 + j], n);
			double m = a[6 + j] / dn;
			if (!isfinite(m)) m = 0.0;
			meanImage[3 * i + j] = m;
			image[3 * i + j] = static_cast<float>(m);
		}
		colorVariance[i] = cv / 3.0;
		for (int j = 0; j < kNforFeatures; j++)
		{
			const double fallback = n == 0u ? 0.0 : a[12 + 2 * kNforFeatures + j] / dn;
			featA[kNforFeatures * i + j] = nA == 0u ? fallback : a[12 + j] / static_cast<double>(nA);
			featB[kNforFeatures * i + j] = nB == 0u ? fallback : a[12 + kNforFeatures + j] / static_cast<double>(nB);
			featVariance[kNforFeatures * i + j] = sampleMeanVariance(a[12 + 2 * kNforFeatures + j], a[12 + 3 * kNforFeatures + j], n);
		}
	}

	// Path-replay adjoint, one thread per (pixel, sample): dLdI is dLoss/dImage (W*H*3),
	// grads receives kGradPerMaterial FP64 sums per material.
	template <int Feat>
	__global__ __launch_bounds__(kMegaBlock) void	adjoint_mega(SceneView s, GpuParams p, const float* __restrict__ dLdI,
		double* __restrict__ grads, std::uint32_t materialCount)
	{
		__shared__ std::uint32_t stackMem[kStackDepth * kMegaBlock];
		extern __shared__ float ldsGrads[];
		const std::uint32_t n = materialCount * kGradPerMaterial;
		for (std::uint32_t i = threadIdx.x; i < n; i += kMegaBlock) ldsGrads[i] = 0.0f;
		__syncthreads();
		const std::uint32_t pixels = p.width * p.height;
		const size_t id = static_cast<size_t>(blockIdx.x) * kMegaBlock + threadIdx.x;
		if (id < static_cast<size_t>(pixels) * p.sampleCount)
		{
			const std::uint32_t pixel = static_cast<std::uint32_t>(id % pixels), slot = static_cast<std::uint32_t>(id / pixels);
			LdsStack<kStackDepth, kMegaBlock> stack{stackMem + threadIdx.x, 0};
			AdjointRecorder rec;
			rec.grads = ldsGrads;
			rec.count = 0;
			rec.pendingMaterial = -1;
			rec.pendingScale = 0.0f;
			rec.materials = s.materials;
			const float inv = 1.0f / static_cast<float>(p.sampleCount);
			rec.weight = mk(dLdI[3 * static_cast<size_t>(pixel)] * inv, dLdI[3 * static_cast<size_t>(pixel) + 1] * inv,
				dLdI[3 * static_cast<size_t>(pixel) + 2] * inv);
			std::uint32_t closest = 0, shadow = 0;
			VisitCounter visits{0u, 0u};
			const f3 L = tracePath<false, Feat>(s, p, pixel % p.width, pixel / p.width, p.sampleBegin + slot, stack, closest, shadow,
				visits, &rec);
			rec.finish(L);
		}
		__syncthreads();
		for (std::uint32_t i = threadIdx.x; i < n; i += kMegaBlock)
		{
			if (ldsGrads[i] != 0.0f) atomicAdd(grads + i, static_cast<double>(ldsGrads[i]));
		}
	}

	void	launchAdjoint(int features, size_t grid, size_t shmem, hipStream_t st, const SceneView& s, const GpuParams& p,
		const float* dLdI, double* grads, std::uint32_t materialCount)
	{
		switch (features & kFeatAll)
		{
			case 0: adjoint_mega<0><<<grid, kMegaBlock, shmem, st>>>(s, p, dLdI, grads, materialCount); break;
			case kFeatAnalytic: adjoint_mega<kFeatAnalytic><<<grid, kMegaBlock, shmem, st>>>(s, p, dLdI, grads, materialCount); break;
			case kFeatRich: adjoint_mega<kFeatRich><<<grid, kMegaBlock, shmem, st>>>(s, p, dLdI, grads, materialCount); break;
			default: adjoint_mega<kFeatAll><<<grid, kMegaBlock, shmem, st>>>(s, p, dLdI, grads, materialCount); break;
		}
	}

	// Adds one launch's samples (render_mega's layout) into the FP64 pixel sums in sample
	// order, and in adaptive rounds into the pixels' AdaptiveAccumulator.
	__global__ void	accumulate(const float* __restrict__ sampleColors, double* __restrict__ sums, const std::uint32_t* __restrict__ list,
		std::uint32_t pixels, std::uint32_t count, double* __restrict__ moments)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= pixels) return;
		const size_t px = list ? list[i] : i;
		double r = sums[3 * px], g = sums[3 * px + 1], b = sums[3 * px + 2];
		double m[kAdaptiveMoments];
		if (moments)
		{
			for (int j = 0; j < kAdaptiveMoments; j++) m[j] = moments[kAdaptiveMoments * px + j];
		}
		for (std::uint32_t k = 0; k < count; k++)
		{
			const float* c = sampleColors + (static_cast<size_t>(k) * pixels + i) * 3;
			const double cr = c[0], cg = c[1], cb = c[2];
			r += cr;
			g += cg;
			b += cb;
			if (moments) adaptiveAdd(m, cr, cg, cb);
		}
		sums[3 * px] = r;
		sums[3 * px + 1] = g;
		sums[3 * px + 2] = b;
		if (moments)
		{
			for (int j = 0; j < kAdaptiveMoments; j++) moments[kAdaptiveMoments * px + j] = m[j];
		}
	}

	// The image of an adaptive render: each pixel's sum over its own sample count,
	// cleanColor'd (Threads.cpp:580-581).
	__global__ void	resolve_used(const double* __restrict__ sums, const std::uint32_t* __restrict__ used, float* __restrict__ out,
		std::uint32_t pixels)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= pixels) return;
		const double n = static_cast<double>(used[i]);
		for (int k = 0; k < 3; k++)
		{
			const double v = sums[3 * i + k] / n;
			out[3 * i + k] = isfinite(v) ? static_cast<float>(v) : 0.0f;
		}
	}

	__global__ void	resolve(const double* __restrict__ sums, float* __restrict__ out, std::uint32_t pixels, double count)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= pixels) return;
		for (int k = 0; k < 3; k++)
		{
			const double v = sums[3 * i + k] / count;
			out[3 * i + k] = isfinite(v) ? static_cast<float>(v) : 0.0f;
		}
	}

	// The FP64 mean image (cleanColor'd, Threads.cpp:580-581) for post-processing: each
	// pixel's sum over `used[i]` samples, or over `count` when used is nullptr.
	__global__ void	resolve_double(const double* __restrict__ sums, const std::uint32_t* __restrict__ used, double count,
		double* __restrict__ out, std::uint32_t pixels)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i >= pixels) return;
		const double n = used ? static_cast<double>(used[i]) : count;
		for (int k = 0; k < 3; k++)
		{
			const double v = sums[3 * i + k] / n;
			out[3 * i + k] = isfinite(v) ? v : 0.0;
		}
	}

	__global__ void	to_float(const double* __restrict__ in, float* __restrict__ out, size_t n)
	{
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i < n) out[i] = static_cast<float>(in[i]);
	}

	__global__ void	zero_doubles(double* p, size_t n)
	{
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i < n) p[i] = 0.0;
	}

	__global__ void	zero_u32(std::uint32_t* p, std::uint32_t n)
	{
		const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
		if (i < n) p[i] = 0u;
	}

	__global__ void	zero_u64(unsigned long long* p, size_t n)
	{
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i < n) p[i] = 0ull;
	}

	__global__ void	primary_hits(SceneView s, GpuParams p, long long* ids, float* ts)
	{
		__shared__ std::uint32_t stackMem[kStackDepth * kMegaBlock];
		const std::uint32_t pixels = p.width * p.height;
		const std::uint32_t i = blockIdx.x * kMegaBlock + threadIdx.x;
		if (i >= pixels) return;
		PixelSampler ps;
		ps.begin(p.renderSeed, i % p.width, i / p.width, 0, 0);
		const Ray ray = cameraRay(p, i % p.width, i / p.width, ps);
		LdsStack<kStackDepth, kMegaBlock> stack{stackMem + threadIdx.x, 0};
		HitInfo h;
		if (!traverse<false>(s, p, ray, kTMinF, INFINITY, h, stack, nullptr, volumeLogU(p, ps)))
		{
			ids[2 * i] = -1;
			ids[2 * i + 1] = -1;
			ts[i] = -1.0f;
			return;
		}
		// the second id is the triangle within the mesh for instances (cubes keep their
		// face index in h.tri, which is not reported)
		const bool instance = (s.hittables[h.hittable].kindIndex >> kKindShift) == GPU_HIT_INSTANCE;
		ids[2 * i] = h.hittable;
		ids[2 * i + 1] = instance ? s.meshTris[h.tri].prim : -1;
		ts[i] = h.t;
	}

	__global__ void	sampler_eval(const std::uint32_t* keys, double* s1, float* s2, size_t n)
	{
		const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i >= n) return;
		const std::uint32_t* k = keys + 7 * i;
		PixelSampler ps;
		ps.begin(k[0], k[1], k[2], k[3], k[4]);
		ps.bounce = k[5];
		s1[i] = ps.sample1D(k[6]);
		ps.sample2D(k[6], s2[2 * i], s2[2 * i + 1]);
	}

	struct	GpuScene
	{
		PackedScene	host;
		DeviceBuffer	tlas, blas, topTris, meshTris, meshNormals, tlasPrims, hittables, rects, instances, materials,
			lights, lightCdf, lightTriCdf, lightTriSlot, envPixels, envPdf, envCdf, envGuide, iesData, unbounded, spheres, cubeFaces, volumes, planes;
		SceneView	view{};
		luzgpu_scene_info	info{};

		void	releaseDevice()
		{
			for (DeviceBuffer* b : {&tlas, &blas, &topTris, &meshTris, &meshNormals, &tlasPrims, &hittables, &rects, &instances,
				&materials, &lights, &lightCdf, &lightTriCdf, &lightTriSlot, &envPixels, &envPdf, &envCdf, &envGuide, &iesData, &unbounded, &spheres,
				&cubeFaces, &volumes, &planes})
			{
				b->release();
			}
			view = SceneView{};
		}
	};
}

struct luzgpu_ctx
{
	int	device = 0;
	hipStream_t	stream = nullptr;
	luzgpu::DeviceBuffer	sampleColors;
	luzgpu::DeviceBuffer	sums;
	luzgpu::DeviceBuffer	counters;
	luzgpu::DeviceBuffer	output;
	luzgpu::DeviceBuffer	scratch;
	luzgpu::WavefrontState	wave;
	// merged arrays of the last multi-scene batch, keyed by the scenes' serials
	std::vector<std::uint64_t>	batchKey;
	std::unique_ptr<luzgpu::GpuScene>	batch;
	std::vector<luzgpu::GpuParams>	batchParams;
	luzgpu::DeviceBuffer	batchSums;
	std::unordered_set<luzgpu_scene*>	scenes;	// loaded on this context and not yet released
	// autotuned launch shapes, keyed by scene serials and per-scene (width, height, spp, depth)
	std::vector<std::pair<std::vector<std::uint64_t>, luzgpu_tuning>>	tuned;
	luzgpu::DeviceBuffer	tuneSums;
	// denoising: per-sample features, per-pixel statistics, NFOR inputs and temporaries
	luzgpu::DeviceBuffer	sampleFeatures, denoiseStats, nforBuffers, nforTemp, nforSystems;
	// adaptive rounds: per-pixel accumulators, sample counts, two active-pixel lists, counters
	luzgpu::DeviceBuffer	adaptiveMem;
	// display post-processing: the FP64 mean image and three working images
	luzgpu::DeviceBuffer	postMem;
	// batched plain-kernel launches: their job tables
	luzgpu::DeviceBuffer	megaJobs;
	// per-pixel sample counts of the last single-scene render: `countsUniform` samples each,
	// or (countsUniform == 0) the adaptive rounds' counts in adaptiveMem
	std::uint32_t	countsPixels = 0;
	std::uint32_t	countsUniform = 0;
	std::uint32_t*	countsDevice = nullptr;
};

struct luzgpu_scene
{
	luzgpu_ctx*	ctx = nullptr;
	luzgpu::GpuScene	gpu;
	std::uint64_t	serial = 0;		// identity for the context's merged-batch cache
};

namespace luzgpu
{
	namespace
	{
		template <typename F>
		int	guarded(F&& f)
		{
			try
			{
				f();
				return LUZGPU_OK;
			}
			catch (const Error& e)
			{
				g_lastError = e.message;
				return e.code;
			}
			catch (const std::bad_alloc&)
			{
				g_lastError = "host allocation failed";
				return LUZGPU_E_OOM;
			}
			catch (const std::runtime_error& e)
			{
				g_lastError = e.what();
				const std::string m = e.what();
				if (m.find("Unsupported") != std::string::npos) return LUZGPU_E_UNSUPPORTED;
				return LUZGPU_E_IO;
			}
			catch (const std::exception& e)
			{
				g_lastError = e.what();
				return LUZGPU_E_INVALID;
			}
		}

		void	uploadScene(luzgpu_ctx* ctx, GpuScene& g)
		{
			PackedScene& h = g.host;
			hipStream_t st = ctx->stream;
			g.tlas.upload(h.tlasWide, st);
			g.blas.upload(h.blasWide, st);
			g.topTris.upload(h.topTris, st);
			g.meshTris.upload(h.meshTris, st);
			g.meshNormals.upload(h.meshNormals, st);
			g.tlasPrims.upload(h.tlasLeaf, st);
			g.hittables.upload(h.hittables, st);
			g.rects.upload(h.rects, st);
			g.instances.upload(h.instances, st);
			g.materials.upload(h.materials, st);
			g.lights.upload(h.lights, st);
			g.lightCdf.upload(h.lightCdf, st);
			g.lightTriCdf.upload(h.lightTriCdf, st);
			g.lightTriSlot.upload(h.lightTriSlot, st);
			g.envPixels.upload(h.envPixels, st);
			g.envPdf.upload(h.envPdf, st);
			g.envCdf.upload(h.envCdf, st);
			g.envGuide.upload(h.envGuide, st);
			g.iesData.upload(h.iesData, st);
			g.unbounded.upload(h.unbounded, st);
			g.spheres.upload(h.spheres, st);
			g.cubeFaces.upload(h.cubeFaces, st);
			g.volumes.upload(h.volumes, st);
			g.planes.upload(h.planes, st);
			check(hipStreamSynchronize(st), "scene upload");
			SceneView& v = g.view;
			v.tlas = g.tlas.as<Bvh4Node>();
			v.blas = g.blas.as<Bvh4Node>();
			v.topTris = g.topTris.as<GpuTri>();
			v.meshTris = g.meshTris.as<GpuTri>();
			v.meshNormals = g.meshNormals.as<float>();
			v.tlasLeaf = g.tlasPrims.as<uint2>();
			v.hittables = g.hittables.as<GpuHittable>();
			v.rects = g.rects.as<GpuRect>();
			v.instances = g.instances.as<GpuInstance>();
			v.materials = g.materials.as<GpuMaterial>();
			v.lights = g.lights.as<GpuLight>();
			v.lightCdf = h.lightCdf.empty() ? nullptr : g.lightCdf.as<double>();
			v.lightTriCdf = h.lightTriCdf.empty() ? nullptr : g.lightTriCdf.as<double>();
			v.lightTriSlot = h.lightTriSlot.empty() ? nullptr : g.lightTriSlot.as<std::uint32_t>();
			v.envPixels = g.envPixels.as<float4>();
			v.envPdf = g.envPdf.as<float>();
			v.envCdf = g.envCdf.as<double>();
			v.envGuide = h.envGuide.empty() ? nullptr : g.envGuide.as<std::uint32_t>();
			v.iesData = h.iesData.empty() ? nullptr : g.iesData.as<double>();
			v.unbounded = g.unbounded.as<std::uint32_t>();
			v.spheres = g.spheres.as<GpuSphere>();
			v.cubeFaces = g.cubeFaces.as<GpuRect>();
			v.volumes = g.volumes.as<GpuVolume>();
			v.planes = g.planes.as<GpuPlane>();
		}

		// The render parameters of one call: the scene's (base) parameters with the caller's
		// settings applied (FlagsParser overrides, FlagsParser.cpp:87-117).
		GpuParams	resolveParams(const GpuParams& base, const PackedScene& host, const luzgpu_settings* s)
		{
			GpuParams p = base;
			if (s)
			{
				p.renderSeed = s->seed;
				if (s->samples >= 0) p.totalSamples = static_cast<std::uint32_t>(s->samples);
				if (s->max_bounces >= 0) p.maxBounces = s->max_bounces;
				p.sampleBegin = s->sample_begin;
				p.sampleCount = s->sample_count >= 0 ? static_cast<std::uint32_t>(s->sample_count)
					: (p.totalSamples > s->sample_begin ? p.totalSamples - s->sample_begin : 0u);
				if (s->width > 0 || s->height > 0)
				{
					p.width = s->width > 0 ? static_cast<std::uint32_t>(s->width) : p.width;
					p.height = s->height > 0 ? static_cast<std::uint32_t>(s->height) : p.height;
					p.camera = makeCamera(host.camera, p.width, p.height, host.metersPerUnit);
				}
			}
			else
			{
				p.renderSeed = 0x9e3779b9u;	// Sampler.cpp:28 default
				p.sampleBegin = 0;
				p.sampleCount = p.totalSamples;
			}
			// 64-bit checks: a sample range or an image size that wraps in 32 bits would
			// otherwise size buffers and sampler indices wrongly
			if (static_cast<std::uint64_t>(p.sampleBegin) + p.sampleCount > p.totalSamples)
			{
				throw Error{LUZGPU_E_INVALID, "sample range [sample_begin, sample_begin + sample_count) outside [0, samples)"};
			}
			if (p.width == 0 || p.height == 0 || p.width > kMaxImageSide || p.height > kMaxImageSide
				|| static_cast<std::uint64_t>(p.width) * p.height > kMaxImagePixels)
			{
				throw Error{LUZGPU_E_INVALID, "image size outside 1..65536 per side and 2^28 pixels"};
			}
			return p;
		}

		GpuParams	resolveParams(const GpuScene& g, const luzgpu_settings* s)
		{
			return resolveParams(g.host.params, g.host, s);
		}

		unsigned	blocks(size_t n, unsigned b) { return static_cast<unsigned>((n + b - 1) / b); }

		WaveTuning	tuningOf(const luzgpu_settings* s)
		{
			WaveTuning t;
			if (s)
			{
				t.grid = s->traversal_grid;
				t.extendRefill = s->extend_refill;
				t.shadowRefill = s->shadow_refill;
				t.chunk = s->fetch_chunk;
				t.ldsStack = static_cast<int>(s->lds_stack);
			}
			return t;
		}

		WaveTuning	tuningOf(const luzgpu_tuning& r)
		{
			WaveTuning t;
			t.grid = r.traversal_grid;
			t.extendRefill = r.extend_refill;
			t.shadowRefill = r.shadow_refill;
			t.chunk = r.fetch_chunk;
			t.ldsStack = static_cast<int>(r.lds_stack);
			return t.resolved();
		}

		std::vector<std::uint64_t>	tuneKey(const std::vector<std::uint64_t>& serials, const std::vector<BatchJob>& jobs)
		{
			std::vector<std::uint64_t> k = serials;
			for (const BatchJob& j : jobs)
			{
				k.push_back((static_cast<std::uint64_t>(j.params.width) << 32) | j.params.height);
				k.push_back((static_cast<std::uint64_t>(j.params.sampleCount) << 32) | static_cast<std::uint32_t>(j.params.maxBounces));
			}
			return k;
		}

		// Picks the traversal launch shape (LDS stack depth, rays claimed per queue atomic,
		// lane-refill thresholds, persistent grid) by timing a calibration render of the same
		// jobs: each scene's samples cut so the whole calibration is one full pass (at most
		// kMaxPassPaths), coordinate descent over the knobs, two sweeps, each candidate timed on
		// its second run (the first captures its graph). The image does not depend on any knob.
		luzgpu_tuning	autotuneJobs(luzgpu_ctx* ctx, const SceneView& view, int features, const std::vector<BatchJob>& jobs)
		{
			const auto t0 = std::chrono::steady_clock::now();
			hipStream_t st = ctx->stream;
			std::uint64_t total = 0, sumLen = 0;
			for (const BatchJob& j : jobs)
			{
				total += static_cast<std::uint64_t>(j.params.width) * j.params.height * j.params.sampleCount;
				sumLen += static_cast<std::uint64_t>(j.params.width) * j.params.height * 3;
			}
			const double scale = std::min(1.0, static_cast<double>(kMaxPassPaths) / static_cast<double>(std::max<std::uint64_t>(total, 1)));
			ctx->tuneSums.reserve(sumLen * sizeof(double));
			std::vector<BatchJob> cal = jobs;
			std::uint64_t off = 0;
			for (BatchJob& j : cal)
			{
				j.params.sampleCount = std::max<std::uint32_t>(1u, static_cast<std::uint32_t>(j.params.sampleCount * scale));
				j.sums = ctx->tuneSums.as<double>() + off;
				off += static_cast<std::uint64_t>(j.params.width) * j.params.height * 3;
			}
			ctx->counters.reserve(8 * sizeof(unsigned long long));
			Event e0, e1;
			std::uint32_t trials = 0;
			auto timeOf = [&](WaveTuning t) {
				float ms = 0.0f;
				for (int rep = 0; rep < 2; rep++)
				{
					e0.record(st);
					renderWavefront(ctx->wave, view, features, cal, ctx->counters.as<unsigned long long>(), st, false, nullptr, t);
					e1.record(st);
					check(hipEventSynchronize(e1), "autotune");
					ms = elapsedMs(e0, e1);
				}
				trials++;
				return static_cast<double>(ms);
			};
			WaveTuning best = WaveTuning{}.resolved();
			const double defaultMs = timeOf(best);
			double bestMs = defaultMs;
			for (int sweep = 0; sweep < 2; sweep++)
			{
				auto tryValues = [&](auto set, std::initializer_list<int> values) {
					for (int v : values)
					{
						WaveTuning t = best;
						set(t, v);
						if (t == best) continue;
						const double ms = timeOf(t.resolved());
						if (ms < bestMs)
						{
							bestMs = ms;
							best = t;
						}
					}
				};
				tryValues([](WaveTuning& t, int v) { t.ldsStack = v; }, {8, 16, 24});
				tryValues([](WaveTuning& t, int v) { t.chunk = static_cast<unsigned>(v); }, {64, 256, 1024});
				tryValues([](WaveTuning& t, int v) { t.extendRefill = v; }, {1, 16, 48, 64});
				tryValues([](WaveTuning& t, int v) { t.shadowRefill = v; }, {1, 16, 48, 64});
				tryValues([](WaveTuning& t, int v) { t.grid = static_cast<unsigned>(v); }, {1024, 1280, 2048});
			}
			luzgpu_tuning r{};
			r.traversal_grid = best.grid;
			r.extend_refill = best.extendRefill;
			r.shadow_refill = best.shadowRefill;
			r.fetch_chunk = best.chunk;
			r.lds_stack = static_cast<std::uint32_t>(best.ldsStack);
			r.trials = trials;
			r.default_ms = defaultMs;
			r.tuned_ms = bestMs;
			r.calibration_paths = static_cast<std::uint64_t>(0);
			for (const BatchJob& j : cal) r.calibration_paths += static_cast<std::uint64_t>(j.params.width) * j.params.height * j.params.sampleCount;
			r.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			return r;
		}

		// NFOR input buffers of W x H pixels in one FP64 allocation (field offsets in doubles).
		struct	NforLayout
		{
			size_t	mean, colorA, colorB, colorVariance, featA, featB, featVariance, total;
			explicit NforLayout(size_t n)
			{
				mean = 0;
				colorA = mean + 3 * n;
				colorB = colorA + 3 * n;
				colorVariance = colorB + 3 * n;
				featA = colorVariance + n;
				featB = featA + kNforFeatures * n;
				featVariance = featB + kNforFeatures * n;
				total = featVariance + kNforFeatures * n;
			}
		};

		NforParams	nforParams(std::uint32_t w, std::uint32_t h, const luzgpu_nfor_settings* ns)
		{
			NforParams P{w, h, 1u, 2u, 1u, 1e-5};	// NFORSettings defaults (NFOR.hpp:39-48)
			if (ns)
			{
				P.prefilterRadius = ns->feature_prefilter_radius;
				P.regressionRadius = ns->regression_radius;
				P.patchRadius = ns->patch_radius;
				P.minVariance = ns->min_variance;
			}
			if (P.regressionRadius == 0 || P.patchRadius == 0)
			{
				throw Error{LUZGPU_E_INVALID, "NFOR radii must be positive."};	// NFOR.cpp:985-988
			}
			if (P.regressionRadius > 32 || P.patchRadius > static_cast<std::uint32_t>(kNforMaxPatchRadius) || P.prefilterRadius > 32
				|| !(P.minVariance >= 0.0))
			{
				throw Error{LUZGPU_E_INVALID, "NFOR settings out of range: regression radius <= 32, patch radius <= 8, prefilter radius <= 32"};
			}
			return P;
		}

		// Denoise::applyNFOR (NFOR.cpp:977-1102) on device buffers, into `out` (W x H x 3, FP64).
		void	runNfor(luzgpu_ctx* ctx, const NforParams& P, const double* colorA, const double* colorB, const double* colorVariance,
			const double* featA, const double* featB, const double* featVariance, double* out)
		{
			hipStream_t st = ctx->stream;
			const size_t n = static_cast<size_t>(P.width) * P.height;
			// temporaries, in doubles per pixel: 3 x 3 suppressed colours, 2 variances, 2 x 11
			// filtered features, 4 x 3 regressions, 6 scalar maps, 3 + 11 + 1 for the blend
			const size_t per = 9 + 2 + 2 * kNforFeatures + 12 + 6 + 3 + kNforFeatures + 1;
			ctx->nforTemp.reserve(per * n * sizeof(double));
			double* t = ctx->nforTemp.as<double>();
			double* supA = t; t += 3 * n;
			double* supB = t; t += 3 * n;
			double* color = t; t += 3 * n;
			double* regVar = t; t += n;
			double* featVarAvg = t; t += n;
			double* ffA = t; t += kNforFeatures * n;
			double* ffB = t; t += kNforFeatures * n;
			double* a0 = t; t += 3 * n;
			double* a1 = t; t += 3 * n;
			double* b0 = t; t += 3 * n;
			double* b1 = t; t += 3 * n;
			double* m0 = t; t += n;
			double* m1 = t; t += n;
			double* f0 = t; t += n;
			double* f1 = t; t += n;
			double* s0 = t; t += n;
			double* s1 = t; t += n;
			double* combined = t; t += 3 * n;
			double* finalFeat = t; t += kNforFeatures * n;
			double* finalVar = t; t += n;
			const unsigned g = blocks(n, 256);
			nfor_fireflies<<<g, 256, 0, st>>>(colorA, colorA, supA, P, 0);
			nfor_fireflies<<<g, 256, 0, st>>>(colorB, colorB, supB, P, 0);
			nfor_fireflies<<<g, 256, 0, st>>>(supA, supB, color, P, 1);
			nfor_variances<<<g, 256, 0, st>>>(colorVariance, featVariance, regVar, featVarAvg, static_cast<std::uint32_t>(n));
			nfor_prefilter<<<g, 256, 0, st>>>(featA, featB, featVarAvg, ffA, P);
			nfor_prefilter<<<g, 256, 0, st>>>(featB, featA, featVarAvg, ffB, P);
			const unsigned tiles = blocks(P.width, kNforTile) * blocks(P.height, kNforTile);
			constexpr unsigned tb = 2 * kNforTilePixels;
			// each regression's per-pixel systems (nfor_accumulate -> nfor_solve)
			ctx->nforSystems.reserve(static_cast<size_t>(kNforSystem) * n * sizeof(double));
			double* systems = ctx->nforSystems.as<double>();
			// the LDS-staged accumulation when its window fits the device's LDS per block
			// (radius 9, patch radius 1: 158 KiB of the MI355X's 160 KiB), else the unstaged form
			const size_t stagedBytes = static_cast<size_t>(nforLdsLayout(P.regressionRadius, P.patchRadius, true).doubles) * sizeof(double);
			const size_t plainBytes = static_cast<size_t>(nforLdsLayout(P.regressionRadius, P.patchRadius, false).doubles) * sizeof(double);
			int device = 0, ldsLimit = 0;
			check(hipGetDevice(&device), "hipGetDevice");
			check(hipDeviceGetAttribute(&ldsLimit, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "LDS per block");
			const bool staged = stagedBytes <= static_cast<size_t>(ldsLimit)
				&& hipFuncSetAttribute(reinterpret_cast<const void*>(&nfor_accumulate<true>), hipFuncAttributeMaxDynamicSharedMemorySize,
					static_cast<int>(stagedBytes)) == hipSuccess;
			if (!staged && (plainBytes > static_cast<size_t>(ldsLimit)
				|| hipFuncSetAttribute(reinterpret_cast<const void*>(&nfor_accumulate<false>), hipFuncAttributeMaxDynamicSharedMemorySize,
					static_cast<int>(plainBytes)) != hipSuccess))
			{
				throw Error{LUZGPU_E_UNSUPPORTED, "NFOR patch radius too large for the device's LDS"};
			}
			auto regress = [&](const double* img, const double* feat, const double* var, double* dst, double bandwidth) {
				if (staged) nfor_accumulate<true><<<tiles, tb, stagedBytes, st>>>(img, feat, var, systems, P, bandwidth);
				else nfor_accumulate<false><<<tiles, tb, plainBytes, st>>>(img, feat, var, systems, P, bandwidth);
				nfor_solve<<<g, 256, 0, st>>>(systems, img, var, dst, P);
			};
			regress(supA, ffB, regVar, a0, 0.5);
			regress(supA, ffB, regVar, a1, 1.0);
			regress(supB, ffA, regVar, b0, 0.5);
			regress(supB, ffA, regVar, b1, 1.0);
			nfor_mse<<<g, 256, 0, st>>>(supA, supB, a0, b0, a1, b1, regVar, m0, m1, static_cast<std::uint32_t>(n));
			nfor_filter_pair<<<g, 256, 0, st>>>(m0, m1, color, colorVariance, f0, f1, P, 0);
			nfor_filter_pair<<<g, 256, 0, st>>>(f0, f1, color, colorVariance, s0, s1, P, 1);
			nfor_combine<<<g, 256, 0, st>>>(s0, s1, f0, f1, a0, a1, b0, b1, ffA, ffB, combined, finalFeat, finalVar, P);
			regress(combined, finalFeat, finalVar, out, 1.0);
			check(hipGetLastError(), "NFOR launch");
		}

		// The context's tuning for these jobs, tuned now if there is none yet.
		WaveTuning	tunedFor(luzgpu_ctx* ctx, const std::vector<std::uint64_t>& serials, const SceneView& view, int features,
			const std::vector<BatchJob>& jobs, luzgpu_tuning* report)
		{
			const std::vector<std::uint64_t> key = tuneKey(serials, jobs);
			for (const auto& e : ctx->tuned)
			{
				if (e.first == key)
				{
					if (report) *report = e.second;
					return tuningOf(e.second);
				}
			}
			const luzgpu_tuning r = autotuneJobs(ctx, view, features, jobs);
			ctx->tuned.emplace_back(key, r);
			if (report) *report = r;
			return tuningOf(r);
		}

		// The integrator a call runs: LUZGPU_INTEGRATOR_AUTO picks the plain kernel for at most
		// 4M paths. A batch of 32 small scenes measured the crossover between 2.1M paths (the
		// plain kernel 1.3x faster) and 8.4M (the wavefront 1.2x faster).
		std::uint32_t	resolveIntegrator(std::uint32_t requested, std::uint64_t paths, const char* who)
		{
			if (requested == LUZGPU_INTEGRATOR_AUTO)
			{
				return paths <= (4ull << 20) ? LUZGPU_INTEGRATOR_MEGAKERNEL : LUZGPU_INTEGRATOR_WAVEFRONT;
			}
			if (requested > LUZGPU_INTEGRATOR_MEGAKERNEL)
			{
				throw Error{LUZGPU_E_INVALID, std::string(who) + ": unknown integrator"};
			}
			return requested;
		}

		luzgpu_post_settings	postOf(const PackedScene& h)
		{
			luzgpu_post_settings r{};
			r.exposure = h.post.exposure;
			r.contrast = h.post.contrast;
			r.bloom = h.post.bloom ? 1u : 0u;
			r.view_transform = static_cast<std::uint32_t>(h.post.viewTransform);
			return r;
		}

		// Kernel parameters of a post-processing run, validated as the reference's setters do
		// (Scene.cpp:261-300); the exposure scale and the blur taps are formed on the host
		// exactly as Image::applyExposure and createSeparableKernel form them.
		PostParams	postParams(const luzgpu_post_settings& s, std::uint32_t width, std::uint32_t height)
		{
			PostParams P{};
			P.width = width;
			P.height = height;
			if (!std::isfinite(s.exposure) || !std::isfinite(std::pow(2.0, s.exposure)))
			{
				throw Error{LUZGPU_E_INVALID, "Exposure must be finite."};
			}
			if (!std::isfinite(s.contrast) || s.contrast < 0.0)
			{
				throw Error{LUZGPU_E_INVALID, "Contrast must be non-negative."};
			}
			if (s.view_transform > LUZGPU_VIEW_RAW || s.bloom > 1u)
			{
				throw Error{LUZGPU_E_INVALID, "post-processing: unknown view transform or bloom flag"};
			}
			P.exposureScale = std::pow(2.0, s.exposure);
			P.contrast = s.contrast;
			P.view = static_cast<int>(s.view_transform);
			double sum = 0.0;
			for (int x = -4; x <= 4; x++)
			{
				P.blur[x + 4] = std::exp(-(static_cast<double>(x * x)) / (2.0 * 2.0 * 2.0));
				sum += P.blur[x + 4];
			}
			for (double& v : P.blur) v /= sum;
			return P;
		}

		// applyPostProcessing (Threads.cpp:77-99) from the FP64 image `in` into `out` (device
		// pointers, W x H x 3; they may be equal).
		void	runPost(luzgpu_ctx* ctx, const luzgpu_post_settings& s, std::uint32_t width, std::uint32_t height, const double* in,
			double* out)
		{
			const PostParams P = postParams(s, width, height);
			const size_t n = static_cast<size_t>(width) * height;
			hipStream_t st = ctx->stream;
			ctx->postMem.reserve(4 * n * 3 * sizeof(double));
			double* img = ctx->postMem.as<double>() + n * 3;
			double* a = img + n * 3;
			double* b = a + n * 3;
			const unsigned g = blocks(n, 256);
			post_exposure<<<g, 256, 0, st>>>(in, img, P);
			if (s.bloom)
			{
				post_bloom_extract<<<g, 256, 0, st>>>(img, a, P);
				post_blur_h<<<g, 256, 0, st>>>(a, b, P);
				post_blur_v_add<<<g, 256, 0, st>>>(b, img, P);
			}
			if (P.view == POST_VIEW_RAW)
			{
				check(hipMemcpyAsync(out, img, n * 3 * sizeof(double), hipMemcpyDeviceToDevice, st), "post copy");
			}
			else
			{
				post_view<<<g, 256, 0, st>>>(img, a, P);
				post_fireflies<<<g, 256, 0, st>>>(a, out, P);
			}
			check(hipGetLastError(), "post-processing launch");
		}

		// The stopping rule of a render, or false when every pixel takes every sample: the
		// scene's adaptive settings with the caller's overrides (FlagsParser's --adaptive
		// options), in effect when enabled and min samples < samples (adaptiveSamplingCanStop,
		// Threads.cpp:265-271).
		bool	adaptiveRule(const PackedScene& h, const GpuParams& p, const luzgpu_settings* s, AdaptiveRule& rule)
		{
			AdaptiveSettings a = h.adaptive;
			if (s)
			{
				if (s->adaptive == LUZGPU_ADAPTIVE_OFF) a.enabled = false;
				else if (s->adaptive == LUZGPU_ADAPTIVE_ON) a.enabled = true;
				else if (s->adaptive != LUZGPU_ADAPTIVE_SCENE) throw Error{LUZGPU_E_INVALID, "unknown adaptive mode"};
				if (s->adaptive_min_samples) a.minSamples = s->adaptive_min_samples;
				if (s->adaptive_check_interval) a.checkInterval = s->adaptive_check_interval;
				if (s->adaptive_threshold > 0.0)
				{
					if (!std::isfinite(s->adaptive_threshold)) throw Error{LUZGPU_E_INVALID, "Adaptive threshold must be positive."};
					a.threshold = s->adaptive_threshold;
				}
			}
			const std::uint32_t total = p.totalSamples;
			if (!a.enabled || a.minSamples >= total) return false;
			if (p.sampleBegin != 0u || p.sampleCount != total)
			{
				throw Error{LUZGPU_E_INVALID, "adaptive sampling renders the whole sample range [0, samples) in one call "
					"(turn it off to split a render by sample range)"};
			}
			rule.maxSamples = total;
			rule.minSamples = a.minSamples;
			rule.checkInterval = a.checkInterval;
			// darkMinimumSamples (Threads.cpp:183-192)
			const std::uint64_t dark = std::max<std::uint64_t>(256u, static_cast<std::uint64_t>(a.minSamples) + 3ull * a.checkInterval);
			rule.darkMinimum = static_cast<std::uint32_t>(std::min<std::uint64_t>(total, std::max<std::uint64_t>(a.minSamples, dark)));
			rule.threshold = a.threshold;
			return true;
		}

		// One image of an adaptive render and its round state. A job whose rule cannot stop
		// early (minSamples == maxSamples) renders every sample in its first round.
		struct	AdaptiveJob
		{
			AdaptiveRule	rule;
			std::uint32_t	pixels;
			size_t	offset;			// first pixel in the context's per-pixel arrays
			std::uint32_t	first = 0;	// sample index of the job's first sample
			std::uint32_t	used = 0;	// samples every active pixel has
			std::uint32_t	active = 0;
			std::uint32_t	round = 0;	// samples of the coming round
			int	list = -1;			// active-pixel list in use (-1: every pixel)
		};

		// One round's work for one job: samples [begin, begin + count) of `active` pixels,
		// `list` (or every pixel when nullptr), accumulated into the job's sums and `moments`.
		struct	AdaptiveRound
		{
			std::uint32_t	job;
			const std::uint32_t*	list;
			std::uint32_t	active;
			std::uint32_t	begin;
			std::uint32_t	count;
			double*	moments;
		};

		struct	AdaptiveArrays
		{
			double*	moments;
			std::uint32_t*	used;
			std::uint32_t*	lists[2];
			std::uint32_t*	counters;
		};

		AdaptiveArrays	adaptiveArrays(luzgpu_ctx* ctx, size_t pixels, size_t jobs)
		{
			const size_t momentBytes = ((pixels * kAdaptiveMoments * sizeof(double)) + 255) & ~size_t(255);
			const size_t listBytes = ((pixels * sizeof(std::uint32_t)) + 255) & ~size_t(255);
			ctx->adaptiveMem.reserve(momentBytes + 3 * listBytes + jobs * sizeof(std::uint32_t));
			char* b = static_cast<char*>(ctx->adaptiveMem.ptr);
			AdaptiveArrays a;
			a.moments = reinterpret_cast<double*>(b);
			a.used = reinterpret_cast<std::uint32_t*>(b + momentBytes);
			a.lists[0] = reinterpret_cast<std::uint32_t*>(b + momentBytes + listBytes);
			a.lists[1] = reinterpret_cast<std::uint32_t*>(b + momentBytes + 2 * listBytes);
			a.counters = reinterpret_cast<std::uint32_t*>(b + momentBytes + 3 * listBytes);
			return a;
		}

		// Runs the rounds of every job until each pixel has stopped (adaptive.hpp); `render`
		// enqueues one round of every job that still has active pixels. Returns the samples
		// rendered (the sum of the per-pixel counts).
		template <typename Render>
		std::uint64_t	runAdaptive(luzgpu_ctx* ctx, std::vector<AdaptiveJob>& jobs, const AdaptiveArrays& A, Render&& render)
		{
			hipStream_t st = ctx->stream;
			size_t pixels = 0;
			for (AdaptiveJob& j : jobs)
			{
				j.used = 0;
				j.active = j.pixels;
				j.round = j.rule.minSamples;
				j.list = -1;
				pixels += j.pixels;
			}
			zero_doubles<<<blocks(pixels * kAdaptiveMoments, 256), 256, 0, st>>>(A.moments, pixels * kAdaptiveMoments);
			std::uint64_t rendered = 0;
			std::vector<std::uint32_t> next(jobs.size());
			for (;;)
			{
				std::vector<AdaptiveRound> work;
				for (std::uint32_t k = 0; k < jobs.size(); k++)
				{
					const AdaptiveJob& j = jobs[k];
					if (j.active == 0u) continue;
					work.push_back(AdaptiveRound{k, j.list < 0 ? nullptr : A.lists[j.list] + j.offset, j.active, j.first + j.used, j.round,
						A.moments + j.offset * kAdaptiveMoments});
				}
				if (work.empty()) break;
				render(work);
				zero_u32<<<1, 256, 0, st>>>(A.counters, static_cast<std::uint32_t>(jobs.size()));
				for (const AdaptiveRound& w : work)
				{
					const AdaptiveJob& j = jobs[w.job];
					const int to = j.list < 0 ? 0 : 1 - j.list;
					adaptive_check<<<blocks(w.active, 256), 256, 0, st>>>(j.rule, w.list, w.active, j.used + j.round, w.moments,
						A.used + j.offset, A.lists[to] + j.offset, A.counters + w.job);
				}
				check(hipGetLastError(), "adaptive check launch");
				check(hipMemcpyAsync(next.data(), A.counters, next.size() * sizeof(std::uint32_t), hipMemcpyDeviceToHost, st),
					"adaptive counters");
				check(hipStreamSynchronize(st), "adaptive round");
				for (const AdaptiveRound& w : work)
				{
					AdaptiveJob& j = jobs[w.job];
					rendered += static_cast<std::uint64_t>(j.active) * j.round;
					j.used += j.round;
					j.active = next[w.job];
					j.list = j.list < 0 ? 0 : 1 - j.list;
					j.round = std::min(j.rule.checkInterval, j.rule.maxSamples - j.used);
					if (j.active > 0u && j.round == 0u)
					{
						throw Error{LUZGPU_E_DEVICE, "adaptive rounds: pixels left active at the sample total"};
					}
				}
			}
			return rendered;
		}
	}
}

using namespace luzgpu;

namespace
{
	// Merged device arrays of a multi-scene batch, rebuilt only when the scene list changes;
	// returns the list's key (the scenes' serials).
	std::vector<std::uint64_t>	ensureBatch(luzgpu_ctx* ctx, uint32_t n, luzgpu_scene* const* scenes)
	{
		std::vector<std::uint64_t> key(n);
		for (uint32_t i = 0; i < n; i++) key[i] = scenes[i]->serial;
		if (!ctx->batch || ctx->batchKey != key)
		{
			ctx->batch.reset();
			ctx->wave.dropGraph();
			std::vector<const PackedScene*> parts(n);
			for (uint32_t i = 0; i < n; i++) parts[i] = &scenes[i]->gpu.host;
			auto g = std::make_unique<GpuScene>();
			g->host = mergeScenes(parts, ctx->batchParams);
			if (g->host.tlasStackBound + 2 + g->host.blasStackBound > kWaveStackCapacity)
			{
				throw std::runtime_error("Unsupported for the device path: traversal stack bound of the batch");
			}
			uploadScene(ctx, *g);
			ctx->batch = std::move(g);
			ctx->batchKey = key;
		}
		return key;
	}
}

namespace
{
	// The pre-render sync of Renderer::render (Renderer.cpp:18-20: updateLights,
	// updateAccelerationStructure) for a parsed or handed-over scene, then the upload.
	luzgpu_scene*	finishScene(luzgpu_ctx* ctx, const Scene& scene)
	{
		auto sc = std::make_unique<luzgpu_scene>();
		sc->ctx = ctx;
		static std::atomic<std::uint64_t> serials{0};
		sc->serial = ++serials;
		sc->gpu.host = packScene(scene);
		const PackedScene& h = sc->gpu.host;
		// Worst-case stack use: deferred TLAS entries, a leaf-resume entry and the instance
		// exit marker, then deferred mesh-BVH entries. Scenes that could exceed the smaller
		// of the two kernels' stacks are refused rather than traversed incompletely.
		const int tlasDepth = h.tlasDeviceDepth;
		const int blasDepth = h.blasDeviceDepth;
		const int bound = h.tlasStackBound + 2 + h.blasStackBound;
		if (bound > kStackDepth)
		{
			throw std::runtime_error("Unsupported for the device path: traversal stack bound " + std::to_string(bound)
				+ " exceeds " + std::to_string(kStackDepth));
		}
		if (ctx)
		{
			uploadScene(ctx, sc->gpu);
		}
		luzgpu_scene_info& info = sc->gpu.info;
		info.width = h.params.width;
		info.height = h.params.height;
		info.samples = h.params.totalSamples;
		info.max_bounces = h.params.maxBounces;
		info.hittables = static_cast<std::uint32_t>(h.hittables.size());
		info.triangles = static_cast<std::uint32_t>(h.triangleCount);
		info.meshes = static_cast<std::uint32_t>(h.blas.size());
		info.instances = static_cast<std::uint32_t>(h.instances.size());
		info.lights = static_cast<std::uint32_t>(h.lights.size());
		info.tlas_nodes = static_cast<std::uint32_t>(h.tlasWide.size());
		info.blas_nodes = static_cast<std::uint32_t>(h.blasWide.size());
		info.env_width = h.params.envWidth;
		info.env_height = h.params.envHeight;
		info.materials = static_cast<std::uint32_t>(h.materials.size());
		info.tlas_depth = static_cast<std::uint32_t>(tlasDepth);
		info.blas_depth = static_cast<std::uint32_t>(blasDepth);
		if (ctx) ctx->scenes.insert(sc.get());
		return sc.release();
	}
}

extern "C" {

const char*	luzgpu_last_error(void)
{
	return g_lastError.c_str();
}

int	luzgpu_create(luzgpu_ctx** out, int device)
{
	if (!out)
	{
		g_lastError = "luzgpu_create: out is NULL";
		return LUZGPU_E_INVALID;
	}
	*out = nullptr;
	return guarded([&] {
		int count = 0;
		check(hipGetDeviceCount(&count), "hipGetDeviceCount");
		if (device < 0 || device >= count)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_create: device " + std::to_string(device) + " not present"};
		}
		check(hipSetDevice(device), "hipSetDevice");
		if (device < kScratchDevices)
		{
			// once per device, with the device current: the device address of its `scratch`
			std::lock_guard<std::mutex> lock(g_scratchMutex);
			if (!g_scratchPtr[device])
			{
				void* ptr = nullptr;
				check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
				g_scratchPtr[device] = ptr;
			}
		}
		auto ctx = std::make_unique<luzgpu_ctx>();
		ctx->device = device;
		check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
		*out = ctx.release();
	});
}

void	luzgpu_destroy(luzgpu_ctx* ctx)
{
	if (!ctx) return;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	// scenes still loaded lose their device arrays now, while the device is current, and
	// remain valid handles for host-side queries and luzgpu_scene_release
	for (luzgpu_scene* sc : ctx->scenes)
	{
		sc->gpu.releaseDevice();
		sc->ctx = nullptr;
	}
	ctx->scenes.clear();
	ctx->wave.release();
	ctx->batch.reset();
	ctx->batchSums.release();
	ctx->sampleColors.release();
	ctx->sums.release();
	ctx->counters.release();
	ctx->output.release();
	ctx->scratch.release();
	(void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

int	luzgpu_scene_load(luzgpu_ctx* ctx, const char* path, luzgpu_scene** out)
{
	if (!path || !out)
	{
		g_lastError = "luzgpu_scene_load: NULL argument";
		return LUZGPU_E_INVALID;
	}
	*out = nullptr;
	return guarded([&] {
		if (ctx)
		{
			check(hipSetDevice(ctx->device), "hipSetDevice");
		}
		*out = finishScene(ctx, readSceneFile(path));
	});
}

int	luzgpu_scene_create(luzgpu_ctx* ctx, const luzgpu_scene_desc* desc, luzgpu_scene** out)
{
	if (!desc || !out)
	{
		g_lastError = "luzgpu_scene_create: NULL argument";
		return LUZGPU_E_INVALID;
	}
	*out = nullptr;
	return guarded([&] {
		if (ctx)
		{
			check(hipSetDevice(ctx->device), "hipSetDevice");
		}
		*out = finishScene(ctx, sceneFromDesc(*desc));
	});
}

void	luzgpu_scene_release(luzgpu_scene* scene)
{
	if (!scene) return;
	if (scene->ctx)
	{
		(void)hipSetDevice(scene->ctx->device);
		(void)hipStreamSynchronize(scene->ctx->stream);
		scene->ctx->scenes.erase(scene);
	}
	delete scene;
}

int	luzgpu_scene_get_info(const luzgpu_scene* scene, luzgpu_scene_info* out)
{
	if (!scene || !out)
	{
		g_lastError = "luzgpu_scene_get_info: NULL argument";
		return LUZGPU_E_INVALID;
	}
	*out = scene->gpu.info;
	return LUZGPU_OK;
}

int	luzgpu_render(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings, float* rgb_out, luzgpu_stats* stats)
{
	if (!ctx || !scene || !rgb_out || scene->ctx != ctx)
	{
		g_lastError = "luzgpu_render: NULL argument or scene from another context";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		const GpuParams p = resolveParams(scene->gpu, settings);
		if (p.sampleBegin + p.sampleCount > p.totalSamples || p.sampleCount == 0)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render: sample range outside [0, samples) or empty"};
		}
		if (p.maxBounces < 0)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render: negative max_bounces"};
		}
		const std::uint32_t integrator = resolveIntegrator(settings ? settings->integrator : LUZGPU_INTEGRATOR_WAVEFRONT,
			static_cast<std::uint64_t>(p.width) * p.height * p.sampleCount, "luzgpu_render");
		WaveTuning tune = tuningOf(settings).resolved();
		double tuneMs = 0.0;
		if (settings && settings->autotune && integrator == LUZGPU_INTEGRATOR_WAVEFRONT)
		{
			luzgpu_tuning r{};
			const std::vector<BatchJob> cal{BatchJob{p, nullptr}};
			const size_t before = ctx->tuned.size();
			tune = tunedFor(ctx, {scene->serial}, scene->gpu.view, scene->gpu.host.features, cal, &r);
			if (ctx->tuned.size() != before) tuneMs = r.tune_ms;
		}
		AdaptiveRule rule{};
		const bool adaptive = adaptiveRule(scene->gpu.host, p, settings, rule);
		const std::uint32_t pixels = p.width * p.height;
		hipStream_t st = ctx->stream;
		ctx->sums.reserve(static_cast<size_t>(pixels) * 3 * sizeof(double));
		ctx->counters.reserve(8 * sizeof(unsigned long long));
		zero_doubles<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(ctx->sums.as<double>(), static_cast<size_t>(pixels) * 3);
		zero_u64<<<1, 64, 0, st>>>(ctx->counters.as<unsigned long long>(), 8);
		ctx->countsPixels = 0;
		Event e0, e1;
		std::deque<Event> ks, ke;
		StageTimes stages;
		auto markStart = [&]() {
			ks.emplace_back();
			ke.emplace_back();
			ks.back().record(st);
		};
		// samples [q.sampleBegin, + q.sampleCount) of `active` pixels (`list`, or all) into the sums
		auto renderRange = [&](const GpuParams& q, const std::uint32_t* list, std::uint32_t active, double* moments) {
			if (integrator == LUZGPU_INTEGRATOR_MEGAKERNEL)
			{
				// samples per launch: bounded so the per-sample colour buffer stays within 4 GiB
				const std::uint32_t perLaunch = std::max<std::uint32_t>(1u, std::min<std::uint32_t>(q.sampleCount,
					static_cast<std::uint32_t>((4ull << 30) / (static_cast<unsigned long long>(active) * 12ull))));
				ctx->sampleColors.reserve(static_cast<size_t>(active) * perLaunch * 3 * sizeof(float));
				for (std::uint32_t done = 0; done < q.sampleCount; done += perLaunch)
				{
					markStart();
					GpuParams r = q;
					r.sampleBegin = q.sampleBegin + done;
					r.sampleCount = std::min(perLaunch, q.sampleCount - done);
					const size_t threads = static_cast<size_t>(active) * r.sampleCount;
					launchMega(settings && settings->count_visits, scene->gpu.host.features, blocks(threads, kMegaBlock), st, scene->gpu.view, r,
						list, active, ctx->sampleColors.as<float>(), ctx->counters.as<unsigned long long>());
					check(hipGetLastError(), "render launch");
					ke.back().record(st);
					accumulate<<<blocks(active, 256), 256, 0, st>>>(ctx->sampleColors.as<float>(), ctx->sums.as<double>(), list, active,
						r.sampleCount, moments);
				}
			}
			else
			{
				markStart();
				BatchJob job{q, ctx->sums.as<double>()};
				job.pixelList = list;
				job.listCount = active;
				job.moments = moments;
				renderWavefront(ctx->wave, scene->gpu.view, scene->gpu.host.features, {job}, ctx->counters.as<unsigned long long>(), st,
					settings && settings->count_visits, settings && settings->stage_times ? &stages : nullptr, tune);
				check(hipGetLastError(), "render launch");
				ke.back().record(st);
			}
		};
		e0.record(st);
		std::uint64_t rendered = static_cast<std::uint64_t>(pixels) * p.sampleCount;
		AdaptiveArrays A{};
		if (adaptive)
		{
			A = adaptiveArrays(ctx, pixels, 1);
			std::vector<AdaptiveJob> aj(1);
			aj[0].rule = rule;
			aj[0].pixels = pixels;
			aj[0].offset = 0;
			rendered = runAdaptive(ctx, aj, A, [&](const std::vector<AdaptiveRound>& work) {
				const AdaptiveRound& w = work[0];
				GpuParams q = p;
				q.sampleBegin = w.begin;
				q.sampleCount = w.count;
				renderRange(q, w.list, w.active, w.moments);
			});
		}
		else
		{
			renderRange(p, nullptr, pixels, nullptr);
		}
		e1.record(st);
		float* dst = rgb_out;
		if (!(settings && settings->output_on_device))
		{
			ctx->output.reserve(static_cast<size_t>(pixels) * 3 * sizeof(float));
			dst = ctx->output.as<float>();
		}
		if (settings && settings->postprocess)
		{
			ctx->postMem.reserve(4 * static_cast<size_t>(pixels) * 3 * sizeof(double));
			double* mean = ctx->postMem.as<double>();
			resolve_double<<<blocks(pixels, 256), 256, 0, st>>>(ctx->sums.as<double>(), adaptive ? A.used : nullptr,
				static_cast<double>(p.sampleCount), mean, pixels);
			runPost(ctx, postOf(scene->gpu.host), p.width, p.height, mean, mean);
			post_to_float<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(mean, dst, static_cast<size_t>(pixels) * 3);
		}
		else if (adaptive) resolve_used<<<blocks(pixels, 256), 256, 0, st>>>(ctx->sums.as<double>(), A.used, dst, pixels);
		else resolve<<<blocks(pixels, 256), 256, 0, st>>>(ctx->sums.as<double>(), dst, pixels, static_cast<double>(p.sampleCount));
		check(hipGetLastError(), "resolve launch");
		unsigned long long counters[6] = {0, 0, 0, 0, 0, 0};
		check(hipMemcpyAsync(counters, ctx->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost, st), "counter readback");
		if (dst != rgb_out)
		{
			check(hipMemcpyAsync(rgb_out, dst, static_cast<size_t>(pixels) * 3 * sizeof(float), hipMemcpyDeviceToHost, st), "image readback");
		}
		check(hipStreamSynchronize(st), "render");
		const float ms = elapsedMs(e0, e1);
		double kernelMs = 0.0;
		for (size_t i = 0; i < ks.size(); i++) kernelMs += elapsedMs(ks[i], ke[i]);
		ctx->countsPixels = pixels;
		ctx->countsUniform = adaptive ? 0u : p.sampleCount;
		ctx->countsDevice = adaptive ? A.used : nullptr;
		if (stats)
		{
			stats->rendered_samples = rendered;
			stats->closest_hit_rays = counters[0];
			stats->shadow_rays = counters[1];
			stats->render_ms = ms;
			stats->kernel_ms = kernelMs;
			stats->kernel_launches = static_cast<std::uint32_t>(ks.size());
			stats->nodes_visited = counters[2] + counters[4];
			stats->prims_tested = counters[3] + counters[5];
			stats->closest_nodes = counters[2];
			stats->closest_prims = counters[3];
			stats->extend_ms = stages.extendMs;
			stats->shade_ms = stages.shadeMs;
			stats->shadow_ms = stages.shadowMs;
			stats->extend_launches = stages.extendLaunches;
			stats->tune_ms = tuneMs;
		}
	});
}

int	luzgpu_render_noalloc(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings, float* rgb_out, luzgpu_stats* stats)
{
	if (!ctx || !scene || !rgb_out || scene->ctx != ctx)
	{
		g_lastError = "luzgpu_render_noalloc: NULL argument or scene from another context";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		void* base = ctx->device < kScratchDevices ? g_scratchPtr[ctx->device] : nullptr;
		if (!base)
		{
			throw Error{LUZGPU_E_DEVICE, "luzgpu_render_noalloc: no scratch address for device " + std::to_string(ctx->device)};
		}
		const GpuParams p = resolveParams(scene->gpu, settings);
		if (p.sampleCount == 0)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render_noalloc: empty sample range"};
		}
		if (p.maxBounces < 0)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render_noalloc: negative max_bounces"};
		}
		const std::uint32_t requested = settings ? settings->integrator : LUZGPU_INTEGRATOR_MEGAKERNEL;
		if (requested != LUZGPU_INTEGRATOR_MEGAKERNEL && requested != LUZGPU_INTEGRATOR_AUTO)
		{
			throw Error{LUZGPU_E_UNSUPPORTED, "luzgpu_render_noalloc: only the plain kernel (LUZGPU_INTEGRATOR_MEGAKERNEL) renders without allocating"};
		}
		const std::uint32_t pixels = p.width * p.height;
		if (pixels > kNoAllocMaxPixels)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render_noalloc: " + std::to_string(pixels) + " pixels exceed the fixed scratch's limit of "
				+ std::to_string(kNoAllocMaxPixels) + " (1024 x 1024)"};
		}
		AdaptiveRule rule{};
		if (adaptiveRule(scene->gpu.host, p, settings, rule))
		{
			throw Error{LUZGPU_E_UNSUPPORTED, "luzgpu_render_noalloc: adaptive sampling needs per-pixel state; pass adaptive = LUZGPU_ADAPTIVE_OFF"};
		}
		if (settings && settings->postprocess)
		{
			throw Error{LUZGPU_E_UNSUPPORTED, "luzgpu_render_noalloc: display post-processing is not available without allocation"};
		}
		unsigned char* bytes = static_cast<unsigned char*>(base);
		unsigned long long* counters = reinterpret_cast<unsigned long long*>(bytes);
		double* sums = reinterpret_cast<double*>(bytes + kNoAllocCountersBytes);
		float* staging = reinterpret_cast<float*>(bytes + kNoAllocCountersBytes + kNoAllocSumsBytes);
		hipStream_t st = ctx->stream;
		const auto t0 = std::chrono::steady_clock::now();
		// the previous call's contents stay in the array: clear what starts at zero
		zero_doubles<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(sums, static_cast<size_t>(pixels) * 3);
		zero_u64<<<1, 64, 0, st>>>(counters, 8);
		// samples per launch: the per-sample colours of one launch fit the staging area
		const std::uint32_t perLaunch = std::max<std::uint32_t>(1u, std::min<std::uint32_t>(p.sampleCount, kNoAllocStagingSamples / pixels));
		std::uint32_t launches = 0;
		for (std::uint32_t done = 0; done < p.sampleCount; done += perLaunch)
		{
			GpuParams r = p;
			r.sampleBegin = p.sampleBegin + done;
			r.sampleCount = std::min(perLaunch, p.sampleCount - done);
			const size_t threads = static_cast<size_t>(pixels) * r.sampleCount;
			launchMega(settings && settings->count_visits, scene->gpu.host.features, blocks(threads, kMegaBlock), st, scene->gpu.view, r,
				nullptr, pixels, staging, counters);
			check(hipGetLastError(), "render launch");
			accumulate<<<blocks(pixels, 256), 256, 0, st>>>(staging, sums, nullptr, pixels, r.sampleCount, nullptr);
			launches++;
		}
		const bool onDevice = settings && settings->output_on_device;
		float* dst = onDevice ? rgb_out : staging;
		resolve<<<blocks(pixels, 256), 256, 0, st>>>(sums, dst, pixels, static_cast<double>(p.sampleCount));
		check(hipGetLastError(), "resolve launch");
		unsigned long long c[6] = {0, 0, 0, 0, 0, 0};
		check(hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, st), "counter readback");
		if (!onDevice)
		{
			check(hipMemcpyAsync(rgb_out, staging, static_cast<size_t>(pixels) * 3 * sizeof(float), hipMemcpyDeviceToHost, st), "image readback");
		}
		check(hipStreamSynchronize(st), "render");
		const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		// the per-pixel counts of this render are uniform; they no longer point at adaptive state
		ctx->countsPixels = pixels;
		ctx->countsUniform = p.sampleCount;
		ctx->countsDevice = nullptr;
		if (stats)
		{
			*stats = luzgpu_stats{};
			stats->rendered_samples = static_cast<std::uint64_t>(pixels) * p.sampleCount;
			stats->closest_hit_rays = c[0];
			stats->shadow_rays = c[1];
			stats->render_ms = ms;
			stats->kernel_ms = ms;
			stats->kernel_launches = launches;
			stats->nodes_visited = c[2] + c[4];
			stats->prims_tested = c[3] + c[5];
			stats->closest_nodes = c[2];
			stats->closest_prims = c[3];
		}
	});
}

int	luzgpu_render_denoised(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings, const luzgpu_nfor_settings* nfor,
	float* rgb_out, float* denoised_out, luzgpu_nfor_buffers* buffers_out, luzgpu_stats* stats)
{
	if (!ctx || !scene || !rgb_out || !denoised_out || scene->ctx != ctx)
	{
		g_lastError = "luzgpu_render_denoised: NULL argument or scene from another context";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		GpuParams p = resolveParams(scene->gpu, settings);
		p.denoiseStreams = 1u;
		if (p.sampleBegin != 0u || p.sampleCount != p.totalSamples || p.sampleCount == 0u)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render_denoised: the whole sample range [0, samples) is rendered at once"};
		}
		if (p.maxBounces < 0)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render_denoised: negative max_bounces"};
		}
		const NforParams P = nforParams(p.width, p.height, nfor);
		AdaptiveRule rule{};
		const bool adaptive = adaptiveRule(scene->gpu.host, p, settings, rule);
		hipStream_t st = ctx->stream;
		const std::uint32_t pixels = p.width * p.height;
		ctx->countsPixels = 0;
		const NforLayout L(pixels);
		ctx->denoiseStats.reserve(static_cast<size_t>(pixels) * kDenoiseStats * sizeof(double));
		ctx->nforBuffers.reserve((L.total + 3 * static_cast<size_t>(pixels)) * sizeof(double));
		ctx->counters.reserve(8 * sizeof(unsigned long long));
		zero_doubles<<<blocks(static_cast<size_t>(pixels) * kDenoiseStats, 256), 256, 0, st>>>(ctx->denoiseStats.as<double>(),
			static_cast<size_t>(pixels) * kDenoiseStats);
		zero_u64<<<1, 64, 0, st>>>(ctx->counters.as<unsigned long long>(), 8);
		Event e0, e1, e2;
		e0.record(st);
		const std::uint32_t integrator = resolveIntegrator(settings ? settings->integrator : LUZGPU_INTEGRATOR_WAVEFRONT,
			static_cast<std::uint64_t>(p.width) * p.height * p.sampleCount, "luzgpu_render_denoised");
		const WaveTuning tune = tuningOf(settings).resolved();
		// samples [begin, begin + count) of `active` pixels (`list`, or all) into the statistics
		auto renderRange = [&](std::uint32_t begin, std::uint32_t count, const std::uint32_t* list, std::uint32_t active, double* moments) {
			if (integrator == LUZGPU_INTEGRATOR_WAVEFRONT)
			{
				GpuParams q = p;
				q.sampleBegin = begin;
				q.sampleCount = count;
				BatchJob job{q, nullptr};
				job.pixelList = list;
				job.listCount = active;
				job.moments = moments;
				job.denoise = ctx->denoiseStats.as<double>();
				renderWavefront(ctx->wave, scene->gpu.view, scene->gpu.host.features, {job}, ctx->counters.as<unsigned long long>(), st,
					false, nullptr, tune);
				check(hipGetLastError(), "denoise render launch");
				return;
			}
			// samples per launch: per-sample colour + feature buffers within 4 GiB
			const std::uint32_t perLaunch = std::max<std::uint32_t>(1u, std::min<std::uint32_t>(count,
				static_cast<std::uint32_t>((4ull << 30) / (static_cast<unsigned long long>(active) * (12ull + 4ull * kNforFeatures)))));
			ctx->sampleColors.reserve(static_cast<size_t>(active) * perLaunch * 3 * sizeof(float));
			ctx->sampleFeatures.reserve(static_cast<size_t>(active) * perLaunch * kNforFeatures * sizeof(float));
			for (std::uint32_t done = 0; done < count; done += perLaunch)
			{
				GpuParams q = p;
				q.sampleBegin = begin + done;
				q.sampleCount = std::min(perLaunch, count - done);
				const size_t threads = static_cast<size_t>(active) * q.sampleCount;
				launchMegaDenoise(scene->gpu.host.features, blocks(threads, kMegaBlock), st, scene->gpu.view, q, list, active,
					ctx->sampleColors.as<float>(), ctx->sampleFeatures.as<float>(), ctx->counters.as<unsigned long long>());
				check(hipGetLastError(), "denoise render launch");
				denoise_accumulate<<<blocks(active, 256), 256, 0, st>>>(ctx->sampleColors.as<float>(), ctx->sampleFeatures.as<float>(),
					ctx->denoiseStats.as<double>(), list, active, q.sampleCount, q.sampleBegin, moments);
			}
		};
		std::uint64_t rendered = static_cast<std::uint64_t>(pixels) * p.sampleCount;
		AdaptiveArrays A{};
		if (adaptive)
		{
			A = adaptiveArrays(ctx, pixels, 1);
			std::vector<AdaptiveJob> aj(1);
			aj[0].rule = rule;
			aj[0].pixels = pixels;
			aj[0].offset = 0;
			rendered = runAdaptive(ctx, aj, A, [&](const std::vector<AdaptiveRound>& work) {
				const AdaptiveRound& w = work[0];
				renderRange(w.begin, w.count, w.list, w.active, w.moments);
			});
		}
		else
		{
			renderRange(0u, p.sampleCount, nullptr, pixels, nullptr);
		}
		double* nb = ctx->nforBuffers.as<double>();
		ctx->output.reserve(static_cast<size_t>(pixels) * 3 * sizeof(float) * 2);
		float* image = ctx->output.as<float>();
		float* denoised = image + static_cast<size_t>(pixels) * 3;
		denoise_finalize<<<blocks(pixels, 256), 256, 0, st>>>(ctx->denoiseStats.as<double>(), pixels, p.sampleCount,
			adaptive ? A.used : nullptr, image, nb + L.mean, nb + L.colorA, nb + L.colorB, nb + L.colorVariance, nb + L.featA, nb + L.featB,
			nb + L.featVariance);
		e1.record(st);
		double* out = nb + L.total;
		runNfor(ctx, P, nb + L.colorA, nb + L.colorB, nb + L.colorVariance, nb + L.featA, nb + L.featB, nb + L.featVariance, out);
		if (settings && settings->postprocess)
		{
			// the reference post-processes both images (Threads.cpp:530-546)
			ctx->postMem.reserve(4 * static_cast<size_t>(pixels) * 3 * sizeof(double));
			double* tmp = ctx->postMem.as<double>();
			runPost(ctx, postOf(scene->gpu.host), p.width, p.height, nb + L.mean, tmp);
			to_float<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(tmp, image, static_cast<size_t>(pixels) * 3);
			runPost(ctx, postOf(scene->gpu.host), p.width, p.height, out, tmp);
			to_float<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(tmp, denoised, static_cast<size_t>(pixels) * 3);
		}
		else
		{
			to_float<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(out, denoised, static_cast<size_t>(pixels) * 3);
		}
		e2.record(st);
		const size_t imageBytes = static_cast<size_t>(pixels) * 3 * sizeof(float);
		const hipMemcpyKind kind = (settings && settings->output_on_device) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
		check(hipMemcpyAsync(rgb_out, image, imageBytes, kind, st), "image copy");
		check(hipMemcpyAsync(denoised_out, denoised, imageBytes, kind, st), "denoised copy");
		if (buffers_out)
		{
			buffers_out->width = p.width;
			buffers_out->height = p.height;
			const std::pair<double*, size_t> fields[] = {{buffers_out->image, L.mean}, {buffers_out->color_a, L.colorA},
				{buffers_out->color_b, L.colorB}, {buffers_out->color_variance, L.colorVariance}, {buffers_out->features_a, L.featA},
				{buffers_out->features_b, L.featB}, {buffers_out->feature_variance, L.featVariance}};
			const size_t sizes[] = {3, 3, 3, 1, kNforFeatures, kNforFeatures, kNforFeatures};
			for (int k = 0; k < 7; k++)
			{
				if (!fields[k].first) continue;
				check(hipMemcpyAsync(fields[k].first, nb + fields[k].second, sizes[k] * pixels * sizeof(double), hipMemcpyDeviceToHost, st),
					"buffer readback");
			}
		}
		unsigned long long counters[2] = {0, 0};
		check(hipMemcpyAsync(counters, ctx->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost, st), "counter readback");
		check(hipStreamSynchronize(st), "denoised render");
		float renderMs = 0.0f, denoiseMs = 0.0f;
		check(hipEventElapsedTime(&renderMs, e0, e1), "hipEventElapsedTime");
		check(hipEventElapsedTime(&denoiseMs, e1, e2), "hipEventElapsedTime");
		ctx->countsPixels = pixels;
		ctx->countsUniform = adaptive ? 0u : p.sampleCount;
		ctx->countsDevice = adaptive ? A.used : nullptr;
		if (stats)
		{
			*stats = luzgpu_stats{};
			stats->rendered_samples = rendered;
			stats->closest_hit_rays = counters[0];
			stats->shadow_rays = counters[1];
			stats->render_ms = renderMs;
			stats->kernel_ms = renderMs;
			stats->denoise_ms = denoiseMs;
		}
	});
}

int	luzgpu_nfor(luzgpu_ctx* ctx, const luzgpu_nfor_buffers* in, const luzgpu_nfor_settings* nfor, double* denoised_out, double* elapsed_ms)
{
	if (!ctx || !in || !denoised_out || !in->color_a || !in->color_b || !in->color_variance || !in->features_a || !in->features_b
		|| !in->feature_variance || in->width == 0 || in->height == 0)
	{
		g_lastError = "luzgpu_nfor: NULL argument or empty image";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		const NforParams P = nforParams(in->width, in->height, nfor);
		hipStream_t st = ctx->stream;
		const size_t n = static_cast<size_t>(in->width) * in->height;
		const NforLayout L(n);
		ctx->nforBuffers.reserve((L.total + 3 * n) * sizeof(double));
		double* nb = ctx->nforBuffers.as<double>();
		const std::pair<const double*, size_t> fields[] = {{in->color_a, L.colorA}, {in->color_b, L.colorB},
			{in->color_variance, L.colorVariance}, {in->features_a, L.featA}, {in->features_b, L.featB}, {in->feature_variance, L.featVariance}};
		const size_t sizes[] = {3, 3, 1, kNforFeatures, kNforFeatures, kNforFeatures};
		for (int k = 0; k < 6; k++)
		{
			check(hipMemcpyAsync(nb + fields[k].second, fields[k].first, sizes[k] * n * sizeof(double), hipMemcpyHostToDevice, st), "buffer upload");
		}
		Event e0, e1;
		e0.record(st);
		runNfor(ctx, P, nb + L.colorA, nb + L.colorB, nb + L.colorVariance, nb + L.featA, nb + L.featB, nb + L.featVariance, nb + L.total);
		e1.record(st);
		check(hipMemcpyAsync(denoised_out, nb + L.total, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st), "denoised readback");
		check(hipStreamSynchronize(st), "NFOR");
		float ms = 0.0f;
		check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
		if (elapsed_ms) *elapsed_ms = ms;
	});
}

int	luzgpu_autotune(luzgpu_ctx* ctx, uint32_t n, luzgpu_scene* const* scenes, const luzgpu_settings* settings, luzgpu_tuning* out)
{
	if (!ctx || !scenes || !out || n == 0 || n > kMaxBatchScenes)
	{
		g_lastError = "luzgpu_autotune: NULL argument or scene count outside 1.." + std::to_string(kMaxBatchScenes);
		return LUZGPU_E_INVALID;
	}
	for (uint32_t i = 0; i < n; i++)
	{
		if (!scenes[i] || scenes[i]->ctx != ctx)
		{
			g_lastError = "luzgpu_autotune: scene " + std::to_string(i) + " is NULL or from another context";
			return LUZGPU_E_INVALID;
		}
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		std::vector<BatchJob> jobs(n);
		std::vector<std::uint64_t> serials(n);
		for (uint32_t i = 0; i < n; i++)
		{
			serials[i] = scenes[i]->serial;
			const luzgpu_settings* si = settings ? settings + i : nullptr;
			const GpuParams p = n == 1 ? resolveParams(scenes[i]->gpu, si) : GpuParams{};
			jobs[i].params = p;
		}
		if (n == 1)
		{
			const GpuParams& p = jobs[0].params;
			if (p.sampleBegin + p.sampleCount > p.totalSamples || p.sampleCount == 0 || p.maxBounces < 0)
			{
				throw Error{LUZGPU_E_INVALID, "luzgpu_autotune: empty or out-of-range samples"};
			}
			*out = autotuneJobs(ctx, scenes[0]->gpu.view, scenes[0]->gpu.host.features, jobs);
		}
		else
		{
			const std::vector<std::uint64_t> key = ensureBatch(ctx, n, scenes);
			for (uint32_t i = 0; i < n; i++)
			{
				jobs[i].params = resolveParams(ctx->batchParams[i], scenes[i]->gpu.host, settings ? settings + i : nullptr);
				const GpuParams& p = jobs[i].params;
				if (p.sampleBegin + p.sampleCount > p.totalSamples || p.sampleCount == 0 || p.maxBounces < 0)
				{
					throw Error{LUZGPU_E_INVALID, "luzgpu_autotune: scene " + std::to_string(i) + ": empty or out-of-range samples"};
				}
			}
			*out = autotuneJobs(ctx, ctx->batch->view, ctx->batch->host.features, jobs);
		}
		// later renders of these jobs with autotune = 1 use this result
		const std::vector<std::uint64_t> key = tuneKey(serials, jobs);
		bool replaced = false;
		for (auto& e : ctx->tuned)
		{
			if (e.first == key)
			{
				e.second = *out;
				replaced = true;
			}
		}
		if (!replaced) ctx->tuned.emplace_back(key, *out);
	});
}

int	luzgpu_render_batch(luzgpu_ctx* ctx, uint32_t n, luzgpu_scene* const* scenes, const luzgpu_settings* settings,
	float* const* rgb_out, luzgpu_stats* stats)
{
	if (!ctx || !scenes || !rgb_out || n == 0 || n > kMaxBatchScenes)
	{
		g_lastError = "luzgpu_render_batch: NULL argument or scene count outside 1.." + std::to_string(kMaxBatchScenes);
		return LUZGPU_E_INVALID;
	}
	for (uint32_t i = 0; i < n; i++)
	{
		if (!scenes[i] || scenes[i]->ctx != ctx || !rgb_out[i])
		{
			g_lastError = "luzgpu_render_batch: scene " + std::to_string(i) + " is NULL, from another context, or has no output";
			return LUZGPU_E_INVALID;
		}
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		hipStream_t st = ctx->stream;
		const std::vector<std::uint64_t> key = ensureBatch(ctx, n, scenes);
		std::vector<BatchJob> jobs(n);
		std::vector<size_t> sumOffset(n);
		size_t sumTotal = 0;
		// adaptive rounds for the whole batch when any scene stops pixels early; the others
		// then take all their samples in their first round
		std::vector<AdaptiveJob> aj(n);
		bool adaptive = false;
		size_t pixelTotal = 0;
		for (uint32_t i = 0; i < n; i++)
		{
			const luzgpu_settings* si = settings ? settings + i : nullptr;
			if (si && si->integrator != (settings ? settings[0].integrator : 0u))
			{
				throw Error{LUZGPU_E_INVALID, "luzgpu_render_batch: every scene of a batch uses the same integrator"};
			}
			if (si && si->integrator > LUZGPU_INTEGRATOR_AUTO)
			{
				throw Error{LUZGPU_E_INVALID, "luzgpu_render_batch: unknown integrator"};
			}
			const GpuParams p = resolveParams(ctx->batchParams[i], scenes[i]->gpu.host, si);
			if (p.sampleBegin + p.sampleCount > p.totalSamples || p.sampleCount == 0 || p.maxBounces < 0)
			{
				throw Error{LUZGPU_E_INVALID, "luzgpu_render_batch: scene " + std::to_string(i) + ": empty or out-of-range samples"};
			}
			jobs[i].params = p;
			sumOffset[i] = sumTotal;
			sumTotal += static_cast<size_t>(p.width) * p.height * 3;
			AdaptiveRule rule{};
			if (adaptiveRule(scenes[i]->gpu.host, p, si, rule))
			{
				adaptive = true;
			}
			else
			{
				rule = AdaptiveRule{p.sampleCount, p.sampleCount, 1u, p.sampleCount, 1.0};
				aj[i].first = p.sampleBegin;
			}
			aj[i].rule = rule;
			aj[i].pixels = p.width * p.height;
			aj[i].offset = pixelTotal;
			pixelTotal += aj[i].pixels;
		}
		ctx->batchSums.reserve(sumTotal * sizeof(double));
		ctx->counters.reserve(8 * sizeof(unsigned long long));
		zero_doubles<<<blocks(sumTotal, 256), 256, 0, st>>>(ctx->batchSums.as<double>(), sumTotal);
		zero_u64<<<1, 64, 0, st>>>(ctx->counters.as<unsigned long long>(), 8);
		for (uint32_t i = 0; i < n; i++) jobs[i].sums = ctx->batchSums.as<double>() + sumOffset[i];
		WaveTuning tune = tuningOf(settings).resolved();
		double tuneMs = 0.0;
		if (settings && settings[0].autotune && settings[0].integrator == LUZGPU_INTEGRATOR_WAVEFRONT)
		{
			luzgpu_tuning r{};
			const size_t before = ctx->tuned.size();
			tune = tunedFor(ctx, key, ctx->batch->view, ctx->batch->host.features, jobs, &r);
			if (ctx->tuned.size() != before) tuneMs = r.tune_ms;
			// the calibration used the counters: clear them again
			zero_u64<<<1, 64, 0, st>>>(ctx->counters.as<unsigned long long>(), 8);
		}
		Event e0, e1;
		const bool timed = settings && settings[0].stage_times;
		StageTimes stages;
		e0.record(st);
		std::uint64_t rendered = 0;
		for (const BatchJob& j : jobs) rendered += static_cast<std::uint64_t>(j.params.width) * j.params.height * j.params.sampleCount;
		AdaptiveArrays A{};
		bool plainKernel = settings && settings[0].integrator == LUZGPU_INTEGRATOR_MEGAKERNEL;
		if (settings && settings[0].integrator == LUZGPU_INTEGRATOR_AUTO)
		{
			std::uint64_t paths = 0;
			for (const BatchJob& j : jobs) paths += static_cast<std::uint64_t>(j.params.width) * j.params.height * j.params.sampleCount;
			plainKernel = !adaptive && resolveIntegrator(LUZGPU_INTEGRATOR_AUTO, paths, "luzgpu_render_batch") == LUZGPU_INTEGRATOR_MEGAKERNEL;
		}
		if (plainKernel && adaptive)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_render_batch: adaptive sampling in a batch runs on the wavefront integrator"};
		}
		if (plainKernel)
		{
			// launches of at most 2^26 paths, each giving every scene with samples left a slice
			constexpr std::uint64_t budget = 1ull << 26;
			std::vector<std::uint32_t> done(n, 0);
			std::uint32_t totalPixels = 0;
			for (uint32_t i = 0; i < n; i++) totalPixels += jobs[i].params.width * jobs[i].params.height;
			ctx->megaJobs.reserve(n * sizeof(MegaJob));
			std::uint64_t totalPaths = 0;
			for (uint32_t i = 0; i < n; i++) totalPaths += static_cast<std::uint64_t>(jobs[i].params.width) * jobs[i].params.height * jobs[i].params.sampleCount;
			// no launch holds more than the budget: one allocation before the first launch
			ctx->sampleColors.reserve(static_cast<size_t>(std::min(totalPaths, budget)) * 3 * sizeof(float));
			// every launch's host table stays alive until the call's final synchronisation (the
			// copies are stream-ordered after the previous launch's kernels)
			std::vector<std::vector<MegaJob>> tables;
			for (;;)
			{
				tables.emplace_back();
				std::vector<MegaJob>& table = tables.back();
				std::uint64_t paths = 0;
				std::uint32_t waves = 0, pixelBase = 0;
				for (uint32_t i = 0; i < n; i++)
				{
					const GpuParams& p = jobs[i].params;
					const std::uint32_t pixels = p.width * p.height;
					MegaJob m{};
					m.params = p;
					m.pixels = pixels;
					m.pixelBase = pixelBase;
					m.sums = jobs[i].sums;
					pixelBase += pixels;
					const std::uint64_t room = budget > paths ? (budget - paths) / pixels : 0;
					const std::uint32_t c = static_cast<std::uint32_t>(std::min<std::uint64_t>(p.sampleCount - done[i], room));
					m.params.sampleBegin = p.sampleBegin + done[i];
					m.params.sampleCount = c;
					m.count = pixels * c;
					m.firstWave = waves;
					m.colorBase = paths;
					waves += (m.count + 63u) / 64u;
					paths += m.count;
					done[i] += c;
					table.push_back(m);
				}
				if (paths == 0) break;
				check(hipMemcpyAsync(ctx->megaJobs.ptr, table.data(), table.size() * sizeof(MegaJob), hipMemcpyHostToDevice, st),
					"batch job table");
				const size_t grid = (static_cast<size_t>(waves) * 64 + kMegaBlock - 1) / kMegaBlock;
				if (settings[0].count_visits)
				{
					launchMegaBatchC<true>(ctx->batch->host.features, grid, st, ctx->batch->view, ctx->megaJobs.as<MegaJob>(), n,
						ctx->sampleColors.as<float>(), ctx->counters.as<unsigned long long>());
				}
				else
				{
					launchMegaBatchC<false>(ctx->batch->host.features, grid, st, ctx->batch->view, ctx->megaJobs.as<MegaJob>(), n,
						ctx->sampleColors.as<float>(), ctx->counters.as<unsigned long long>());
				}
				check(hipGetLastError(), "batch render launch");
				accumulate_batch<<<blocks(totalPixels, 256), 256, 0, st>>>(ctx->sampleColors.as<float>(), ctx->megaJobs.as<MegaJob>(), n,
					totalPixels);
			}
		}
		else if (adaptive)
		{
			A = adaptiveArrays(ctx, pixelTotal, n);
			rendered = runAdaptive(ctx, aj, A, [&](const std::vector<AdaptiveRound>& work) {
				std::vector<BatchJob> round;
				for (const AdaptiveRound& w : work)
				{
					BatchJob b = jobs[w.job];
					b.params.sampleBegin = w.begin;
					b.params.sampleCount = w.count;
					b.pixelList = w.list;
					b.listCount = w.active;
					b.moments = w.moments;
					round.push_back(b);
				}
				renderWavefront(ctx->wave, ctx->batch->view, ctx->batch->host.features, round, ctx->counters.as<unsigned long long>(), st,
					settings && settings[0].count_visits, timed ? &stages : nullptr, tune);
			});
		}
		else
		{
			renderWavefront(ctx->wave, ctx->batch->view, ctx->batch->host.features, jobs, ctx->counters.as<unsigned long long>(), st,
				settings && settings[0].count_visits, timed ? &stages : nullptr, tune);
		}
		check(hipGetLastError(), "batch render launch");
		e1.record(st);
		// per-scene resolve, into device outputs directly or through the staging buffer
		size_t hostBytes = 0;
		for (uint32_t i = 0; i < n; i++)
		{
			if (!(settings && settings[i].output_on_device)) hostBytes += static_cast<size_t>(jobs[i].params.width) * jobs[i].params.height * 3 * sizeof(float);
		}
		if (hostBytes) ctx->output.reserve(hostBytes);
		size_t staged = 0;
		std::vector<std::pair<float*, size_t>> copies;
		for (uint32_t i = 0; i < n; i++)
		{
			const std::uint32_t pixels = jobs[i].params.width * jobs[i].params.height;
			float* dst = rgb_out[i];
			if (!(settings && settings[i].output_on_device))
			{
				dst = reinterpret_cast<float*>(static_cast<char*>(ctx->output.ptr) + staged);
				copies.emplace_back(rgb_out[i], staged);
				staged += static_cast<size_t>(pixels) * 3 * sizeof(float);
			}
			if (settings && settings[i].postprocess)
			{
				ctx->postMem.reserve(4 * static_cast<size_t>(pixels) * 3 * sizeof(double));
				double* mean = ctx->postMem.as<double>();
				resolve_double<<<blocks(pixels, 256), 256, 0, st>>>(jobs[i].sums, adaptive ? A.used + aj[i].offset : nullptr,
					static_cast<double>(jobs[i].params.sampleCount), mean, pixels);
				runPost(ctx, postOf(scenes[i]->gpu.host), jobs[i].params.width, jobs[i].params.height, mean, mean);
				post_to_float<<<blocks(static_cast<size_t>(pixels) * 3, 256), 256, 0, st>>>(mean, dst, static_cast<size_t>(pixels) * 3);
			}
			else if (adaptive) resolve_used<<<blocks(pixels, 256), 256, 0, st>>>(jobs[i].sums, A.used + aj[i].offset, dst, pixels);
			else resolve<<<blocks(pixels, 256), 256, 0, st>>>(jobs[i].sums, dst, pixels, static_cast<double>(jobs[i].params.sampleCount));
		}
		check(hipGetLastError(), "resolve launch");
		unsigned long long counters[6] = {0, 0, 0, 0, 0, 0};
		check(hipMemcpyAsync(counters, ctx->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost, st), "counter readback");
		for (uint32_t i = 0, c = 0; i < n; i++)
		{
			if (settings && settings[i].output_on_device) continue;
			const size_t bytes = static_cast<size_t>(jobs[i].params.width) * jobs[i].params.height * 3 * sizeof(float);
			check(hipMemcpyAsync(copies[c].first, static_cast<char*>(ctx->output.ptr) + copies[c].second, bytes, hipMemcpyDeviceToHost, st),
				"image readback");
			c++;
		}
		check(hipStreamSynchronize(st), "batch render");
		float ms = 0.0f;
		check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
		if (adaptive) ctx->countsPixels = 0;	// the per-pixel counts of the last single render are gone
		if (stats)
		{
			*stats = luzgpu_stats{};
			stats->rendered_samples = rendered;
			stats->closest_hit_rays = counters[0];
			stats->shadow_rays = counters[1];
			stats->render_ms = ms;
			stats->kernel_ms = ms;
			stats->nodes_visited = counters[2] + counters[4];
			stats->prims_tested = counters[3] + counters[5];
			stats->closest_nodes = counters[2];
			stats->closest_prims = counters[3];
			stats->extend_ms = stages.extendMs;
			stats->shade_ms = stages.shadeMs;
			stats->shadow_ms = stages.shadowMs;
			stats->extend_launches = stages.extendLaunches;
			stats->tune_ms = tuneMs;
			stats->kernel_launches = stages.passes;
		}
	});
}

int	luzgpu_adjoint(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings, const float* dl_dimage,
	double* d_albedo, double* d_roughness, double* d_emission)
{
	if (!ctx || !scene || !dl_dimage || scene->ctx != ctx)
	{
		g_lastError = "luzgpu_adjoint: NULL argument or scene from another context";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		const GpuParams p = resolveParams(scene->gpu, settings);
		if (p.sampleBegin + p.sampleCount > p.totalSamples || p.sampleCount == 0 || p.maxBounces < 0)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_adjoint: empty or out-of-range sample range"};
		}
		AdaptiveRule rule{};
		if (adaptiveRule(scene->gpu.host, p, settings, rule))
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_adjoint: adaptive sampling is in effect for these settings, but the adjoint replays "
				"every sample of every pixel; render and differentiate with adaptive = LUZGPU_ADAPTIVE_OFF"};
		}
		if (p.maxBounces + 1 > kAdjointRecords)
		{
			throw std::runtime_error("Unsupported for the adjoint: more than " + std::to_string(kAdjointRecords - 1) + " bounces");
		}
		const std::uint32_t materials = static_cast<std::uint32_t>(scene->gpu.host.materials.size());
		const size_t shmem = static_cast<size_t>(materials) * kGradPerMaterial * sizeof(float);
		if (shmem > 96 * 1024)
		{
			throw std::runtime_error("Unsupported for the adjoint: more than " + std::to_string(96 * 1024 / (kGradPerMaterial * 4)) + " materials");
		}
		if (scene->gpu.host.features & kFeatSubsurface)
		{
			throw std::runtime_error("Unsupported for the adjoint: Burley subsurface (its exits move the path vertex)");
		}
		if (!scene->gpu.host.iesData.empty())
		{
			throw std::runtime_error("Unsupported for the adjoint: IES-profiled sphere lights");
		}
		hipStream_t st = ctx->stream;
		const std::uint32_t pixels = p.width * p.height;
		const float* dev = dl_dimage;
		if (!(settings && settings->output_on_device))
		{
			ctx->scratch.reserve(static_cast<size_t>(pixels) * 3 * sizeof(float));
			check(hipMemcpyAsync(ctx->scratch.ptr, dl_dimage, static_cast<size_t>(pixels) * 3 * sizeof(float), hipMemcpyHostToDevice, st),
				"dL/dimage upload");
			dev = ctx->scratch.as<float>();
		}
		const size_t n = static_cast<size_t>(materials) * kGradPerMaterial;
		ctx->sums.reserve(n * sizeof(double));
		zero_doubles<<<blocks(n, 256), 256, 0, st>>>(ctx->sums.as<double>(), n);
		const size_t threads = static_cast<size_t>(pixels) * p.sampleCount;
		launchAdjoint(scene->gpu.host.features, blocks(threads, kMegaBlock), shmem, st, scene->gpu.view, p, dev, ctx->sums.as<double>(),
			materials);
		check(hipGetLastError(), "adjoint launch");
		std::vector<double> g(n);
		check(hipMemcpyAsync(g.data(), ctx->sums.ptr, n * sizeof(double), hipMemcpyDeviceToHost, st), "gradient readback");
		check(hipStreamSynchronize(st), "adjoint");
		for (std::uint32_t m = 0; m < materials; m++)
		{
			for (int c = 0; c < 3; c++)
			{
				if (d_albedo) d_albedo[3 * m + c] = g[m * kGradPerMaterial + c];
				if (d_emission) d_emission[3 * m + c] = g[m * kGradPerMaterial + 4 + c];
			}
			if (d_roughness) d_roughness[m] = g[m * kGradPerMaterial + 3];
		}
	});
}

int	luzgpu_scene_set_materials(luzgpu_scene* scene, uint32_t count, const double* albedo, const double* roughness,
	const double* emission)
{
	if (!scene || !scene->ctx)
	{
		g_lastError = "luzgpu_scene_set_materials: NULL scene or a host-only scene";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		PackedScene& h = scene->gpu.host;
		if (count != h.materials.size())
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_scene_set_materials: count must equal the scene's material count ("
				+ std::to_string(h.materials.size()) + ")"};
		}
		for (uint32_t i = 0; i < count; i++)
		{
			GpuMaterial& m = h.materials[i];
			if (albedo && m.kind == GPU_MAT_LAMBERTIAN)
			{
				for (int c = 0; c < 3; c++)
				{
					h.materialColor[i][c] = albedo[3 * i + c];
					m.color[c] = static_cast<float>(albedo[3 * i + c]);
				}
			}
			if (roughness && m.kind == GPU_MAT_METAL)
			{
				const double r = roughness[i];
				if (!std::isfinite(r) || r < 0.0 || r > 1.0)
				{
					throw Error{LUZGPU_E_INVALID, "Metal roughness must be finite and in [0,1]."};
				}
				m.roughness = static_cast<float>(r);
				m.alpha = static_cast<float>(std::max(1e-4, r * r));
			}
			if (emission && m.kind == GPU_MAT_EMISSIVE)
			{
				for (int c = 0; c < 3; c++)
				{
					if (!std::isfinite(emission[3 * i + c]) || emission[3 * i + c] < 0.0)
					{
						throw Error{LUZGPU_E_INVALID, "Emission must be finite and non-negative."};
					}
					h.materialColor[i][c] = emission[3 * i + c];
					m.color[c] = static_cast<float>(emission[3 * i + c]);
				}
			}
		}
		rebuildLights(h);
		luzgpu_ctx* ctx = scene->ctx;
		check(hipSetDevice(ctx->device), "hipSetDevice");
		hipStream_t st = ctx->stream;
		scene->gpu.materials.upload(h.materials, st);
		scene->gpu.lights.upload(h.lights, st);
		scene->gpu.lightCdf.upload(h.lightCdf, st);
		check(hipStreamSynchronize(st), "material upload");
		scene->gpu.view.materials = scene->gpu.materials.as<GpuMaterial>();
		scene->gpu.view.lights = scene->gpu.lights.as<GpuLight>();
		scene->gpu.view.lightCdf = h.lightCdf.empty() ? nullptr : scene->gpu.lightCdf.as<double>();
		// same-sized uploads keep their buffers, so a captured pass graph (keyed on the
		// scene's device pointers) stays valid; merged batch arrays hold copies: rebuild them
		ctx->batchKey.clear();
	});
}

int	luzgpu_scene_post_settings(const luzgpu_scene* scene, luzgpu_post_settings* out)
{
	if (!scene || !out)
	{
		g_lastError = "luzgpu_scene_post_settings: NULL argument";
		return LUZGPU_E_INVALID;
	}
	*out = postOf(scene->gpu.host);
	return LUZGPU_OK;
}

int	luzgpu_postprocess(luzgpu_ctx* ctx, const luzgpu_post_settings* post, uint32_t width, uint32_t height, const double* image_in,
	double* image_out, uint32_t on_device)
{
	if (!ctx || !post || !image_in || !image_out || width == 0 || height == 0)
	{
		g_lastError = "luzgpu_postprocess: NULL argument or empty image";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		hipStream_t st = ctx->stream;
		const size_t n = static_cast<size_t>(width) * height;
		const size_t bytes = n * 3 * sizeof(double);
		if (on_device)
		{
			runPost(ctx, *post, width, height, image_in, image_out);
			check(hipStreamSynchronize(st), "post-processing");
			return;
		}
		(void)postParams(*post, width, height);	// validate before any copy
		ctx->postMem.reserve(4 * bytes);
		double* buf = ctx->postMem.as<double>();
		check(hipMemcpyAsync(buf, image_in, bytes, hipMemcpyHostToDevice, st), "image upload");
		runPost(ctx, *post, width, height, buf, buf);
		check(hipMemcpyAsync(image_out, buf, bytes, hipMemcpyDeviceToHost, st), "image readback");
		check(hipStreamSynchronize(st), "post-processing");
	});
}

int	luzgpu_sample_counts(luzgpu_ctx* ctx, uint32_t* counts_out, size_t count)
{
	if (!ctx || !counts_out)
	{
		g_lastError = "luzgpu_sample_counts: NULL argument";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		if (ctx->countsPixels == 0u)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_sample_counts: no single-scene render on this context since its last adaptive batch"};
		}
		if (count != ctx->countsPixels)
		{
			throw Error{LUZGPU_E_INVALID, "luzgpu_sample_counts: count must be the last render's W*H = " + std::to_string(ctx->countsPixels)};
		}
		if (ctx->countsUniform != 0u)
		{
			std::fill(counts_out, counts_out + count, ctx->countsUniform);
			return;
		}
		check(hipSetDevice(ctx->device), "hipSetDevice");
		check(hipMemcpyAsync(counts_out, ctx->countsDevice, count * sizeof(std::uint32_t), hipMemcpyDeviceToHost, ctx->stream),
			"sample count readback");
		check(hipStreamSynchronize(ctx->stream), "sample count readback");
	});
}

int	luzgpu_primary_hits(luzgpu_ctx* ctx, luzgpu_scene* scene, const luzgpu_settings* settings, int64_t* ids_out, float* t_out)
{
	if (!ctx || !scene || !ids_out || !t_out || scene->ctx != ctx)
	{
		g_lastError = "luzgpu_primary_hits: NULL argument";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		check(hipSetDevice(ctx->device), "hipSetDevice");
		const GpuParams p = resolveParams(scene->gpu, settings);
		const std::uint32_t pixels = p.width * p.height;
		hipStream_t st = ctx->stream;
		ctx->scratch.reserve(static_cast<size_t>(pixels) * (2 * sizeof(long long) + sizeof(float)));
		long long* ids = ctx->scratch.as<long long>();
		float* ts = reinterpret_cast<float*>(ids + 2 * static_cast<size_t>(pixels));
		primary_hits<<<blocks(pixels, kMegaBlock), kMegaBlock, 0, st>>>(scene->gpu.view, p, ids, ts);
		check(hipGetLastError(), "primary_hits launch");
		check(hipMemcpyAsync(ids_out, ids, static_cast<size_t>(pixels) * 2 * sizeof(long long), hipMemcpyDeviceToHost, st), "readback");
		check(hipMemcpyAsync(t_out, ts, static_cast<size_t>(pixels) * sizeof(float), hipMemcpyDeviceToHost, st), "readback");
		check(hipStreamSynchronize(st), "primary_hits");
	});
}

int	luzgpu_sampler_eval(luzgpu_ctx* ctx, size_t n, const uint32_t* keys, double* s1, float* s2)
{
	if (!ctx || (n && (!keys || !s1 || !s2)))
	{
		g_lastError = "luzgpu_sampler_eval: NULL argument";
		return LUZGPU_E_INVALID;
	}
	return guarded([&] {
		if (n == 0) return;
		check(hipSetDevice(ctx->device), "hipSetDevice");
		hipStream_t st = ctx->stream;
		const size_t keyBytes = (n * 7 * sizeof(std::uint32_t) + 15) & ~size_t(15);
		ctx->scratch.reserve(keyBytes + n * (sizeof(double) + 2 * sizeof(float)));
		std::uint32_t* dk = ctx->scratch.as<std::uint32_t>();
		double* d1 = reinterpret_cast<double*>(ctx->scratch.as<char>() + keyBytes);
		float* d2 = reinterpret_cast<float*>(d1 + n);
		check(hipMemcpyAsync(dk, keys, n * 7 * sizeof(std::uint32_t), hipMemcpyHostToDevice, st), "upload");
		sampler_eval<<<blocks(n, 256), 256, 0, st>>>(dk, d1, d2, n);
		check(hipGetLastError(), "sampler_eval launch");
		check(hipMemcpyAsync(s1, d1, n * sizeof(double), hipMemcpyDeviceToHost, st), "readback");
		check(hipMemcpyAsync(s2, d2, n * 2 * sizeof(float), hipMemcpyDeviceToHost, st), "readback");
		check(hipStreamSynchronize(st), "sampler_eval");
	});
}

int	luzgpu_scene_tlas_topology(const luzgpu_scene* scene, char* buf, size_t cap, size_t* needed)
{
	if (!scene)
	{
		g_lastError = "luzgpu_scene_tlas_topology: NULL scene";
		return LUZGPU_E_INVALID;
	}
	const std::string& t = scene->gpu.host.tlasTopology;
	if (needed) *needed = t.size() + 1;
	if (buf && cap > 0)
	{
		const size_t n = std::min(cap - 1, t.size());
		std::memcpy(buf, t.data(), n);
		buf[n] = '\0';
	}
	return LUZGPU_OK;
}

int	luzgpu_scene_blas(const luzgpu_scene* scene, uint32_t mesh, double* boxes, int64_t* links, int64_t* tri_index,
	uint64_t* node_count, uint64_t* tri_count)
{
	if (!scene || mesh >= scene->gpu.host.blas.size())
	{
		g_lastError = "luzgpu_scene_blas: bad scene or mesh index";
		return LUZGPU_E_INVALID;
	}
	const Blas& b = scene->gpu.host.blas[mesh];
	if (node_count) *node_count = b.nodes.size();
	if (tri_count) *tri_count = b.triangleIndices.size();
	for (size_t i = 0; i < b.nodes.size(); i++)
	{
		const BlasNode& n = b.nodes[i];
		if (boxes)
		{
			for (int k = 0; k < 3; k++) { boxes[6 * i + k] = n.box.lo[k]; boxes[6 * i + 3 + k] = n.box.hi[k]; }
		}
		if (links)
		{
			links[5 * i] = static_cast<int64_t>(n.left);
			links[5 * i + 1] = static_cast<int64_t>(n.right);
			links[5 * i + 2] = static_cast<int64_t>(n.start);
			links[5 * i + 3] = static_cast<int64_t>(n.count);
			links[5 * i + 4] = n.leaf ? 1 : 0;
		}
	}
	if (tri_index)
	{
		for (size_t i = 0; i < b.triangleIndices.size(); i++) tri_index[i] = static_cast<int64_t>(b.triangleIndices[i]);
	}
	return LUZGPU_OK;
}

}
