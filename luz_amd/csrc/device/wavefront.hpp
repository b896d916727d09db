// Wavefront integrator: the bounce loop of calculateLightRaysColor (ColorHelper.cpp:1136-1326)
// split into stages over SoA path state in HBM, so the traversal kernels carry only a ray,
// a stack and a hit record (low VGPR count, high occupancy) and the shading kernel carries
// no traversal state.
//
// It is multi-scene by construction: a launch renders a batch of scenes whose device
// arrays were merged into one SceneView (per-scene bases in GpuParams), and one render of
// one scene is a batch of one. A pass renders a slice of samples of every scene of the
// batch; its paths are numbered globally (scene job j owns [pathBase_j, pathBase_j +
// pixels_j * samples_j), sample slot major), and every queued ray carries its scene id.
//
// Per pass:
//   raygen            camera rays (FP64, bit-exact sampler) -> queue 0, path state init
//   for b = 0..maxBounces:
//     extend          closest hit for every ray of queue b (persistent, lane refill)
//     shade           miss / emission / scatter / NEE and environment connections, one
//                     scene per wave iteration (waterfall over the wave's scene ids); the
//                     next ray and up to two shadow rays are appended with one atomic per
//                     wave (ballot + popcount prefix)
//     shadow          any-hit test of the appended shadow rays -> visibility bytes
//     resolve         adds visible connection terms in the reference's order (light, then
//                     environment) and writes the colour of every path that ended
//   accumulate        adds the pass's sample colours into each scene's FP64 pixel sums
// Queue lengths and the pass's job table live in device memory and every stage is a
// grid-stride or work-claiming loop over them, so a pass is a fixed launch sequence that
// the host captures once as a hipGraph and replays for every pass.
#pragma once

#include "adaptive.hpp"
#include "integrator.hpp"
#include "nfor.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace luzgpu
{
	// path flags
	constexpr std::uint32_t	PF_PREV_SPECULAR = 1u;
	constexpr std::uint32_t	PF_PREV_INFINITE = 2u;
	constexpr std::uint32_t	PF_DONE = 4u;		// the path ended in this bounce
	constexpr std::uint32_t	PF_CONN_LIGHT = 8u;	// a light connection is pending
	constexpr std::uint32_t	PF_CONN_ENV = 16u;	// an environment connection is pending
	constexpr std::uint32_t	PF_CONN_FREE_LIGHT = 32u;	// light connection needs no shadow test

	struct	WaveBuffers
	{
		// queue-ordered (double-buffered by bounce parity)
		float4*	qOrigin[2];
		float4*	qDir[2];		// xyz, w = hittable the ray leaves from (bits)
		std::int32_t*	qFrom[2];	// its triangle slot / cube face (RayOrigin::tri)
		std::uint32_t*	qPath[2];
		float4*	qHit;		// t, u, v, hittable (bits)
		std::int32_t*	qTri;
		// path-ordered
		float4*	throughput;	// xyz, w = previous scatter pdf
		float4*	radiance;	// xyz, w = flags (bits)
		float4*	prevOrigin;
		float4*	connection;	// [2 * path + slot]: contribution xyz
		std::uint8_t*	visible;	// [2 * path + slot]
		// shadow queue (rebuilt every bounce)
		float4*	sOrigin;	// xyz, w = tMax
		float4*	sDir;		// xyz, w = scene << 25 | (2 * path + slot) (bits)
		float*	sLogU;		// volume free-flight sample of the shadow ray (batches with volumes)
		int2*	sFrom;		// RayOrigin of the shadow ray (hittable bits, tri)
		// counters: [b] rays in queue b, [kMaxBounceSlots + b] shadow rays of bounce b,
		// [2 * kMaxBounceSlots + b] / [3 * ...] extend / shadow work claims
		std::uint32_t*	counts;
		std::uint32_t	capacity;
		std::uint32_t*	spill;		// traversal stack spill, [entry][grid thread]
		// Queued extension rays carry their scene above the global path id in qPath
		// (scene << pathBits | path), shadow rays above the connection slot in their tag
		// (scene << (pathBits + 1) | 2 * path + slot). pathBits = 31 - the bits of the largest
		// scene index of the render: 31 for one scene, 24 for 127.
		std::uint32_t	pathBits;

		__device__ __forceinline__ std::uint32_t	sceneOfPath(std::uint32_t qp) const { return qp >> pathBits; }
		__device__ __forceinline__ std::uint32_t	pathOf(std::uint32_t qp) const { return qp & ((1u << pathBits) - 1u); }
		__device__ __forceinline__ std::uint32_t	queuedPath(std::uint32_t scene, std::uint32_t path) const { return (scene << pathBits) | path; }
		__device__ __forceinline__ std::uint32_t	shadowTag(std::uint32_t scene, std::uint32_t slot) const
		{
			return pathBits < 31u ? (scene << (pathBits + 1u)) | slot : slot;
		}
		__device__ __forceinline__ std::uint32_t	sceneOfTag(std::uint32_t tag) const { return pathBits < 31u ? tag >> (pathBits + 1u) : 0u; }
		__device__ __forceinline__ std::uint32_t	slotOfTag(std::uint32_t tag) const
		{
			return pathBits < 31u ? tag & ((2u << pathBits) - 1u) : tag;
		}
	};

	// Queued extension rays carry their volume sample in qOrigin.w.
	constexpr std::uint32_t	kMaxBatchScenes = 127;
	// Paths per pass: 2^27 (~39 GB of path state, 1920 x 1080 x 64 samples in one pass), fewer
	// when a batch's scene bits leave less room in the 32-bit queue entries (2^24 at 127 scenes).
	constexpr std::uint32_t	kMaxPassPaths = 1u << 27;

	// pathBits of a render of n scenes (see WaveBuffers).
	inline std::uint32_t	pathBitsFor(std::uint32_t n)
	{
		std::uint32_t sceneBits = 0;
		while ((1u << sceneBits) < n) sceneBits++;
		return 31u - sceneBits;
	}

	// One scene of a pass. The table lives in device memory: a pass is replayed with a new
	// table and the same launches.
	struct	PassJob
	{
		GpuParams	params;		// sampleBegin / sampleCount: this pass's slice
		std::uint32_t	pathBase;	// first global path id of the slice
		std::uint32_t	pixels;		// pixels rendered: the image, or the entries of pixelList
		double*	sums;			// the scene's FP64 pixel sums
		const std::uint32_t*	pixelList;	// adaptive rounds: the active pixels (nullptr: every pixel)
		double*	moments;		// adaptive rounds: per-pixel AdaptiveAccumulator (or nullptr)
		float*	features;		// denoising: primary features per global path id (or nullptr)
		double*	denoise;		// denoising: per-pixel statistics (kDenoiseStats) instead of sums
	};

	// Image pixel of entry i of a pass's pixel range.
	__device__ __forceinline__ std::uint32_t	passPixel(const PassJob& J, std::uint32_t i)
	{
		return J.pixelList ? J.pixelList[i] : i;
	}

	constexpr int	kWaveBlock = 256;
	constexpr int	kWaveLdsStack = 16;		// default LDS stack entries per lane (tunable: 8, 16, 24)
	constexpr int	kWaveLdsStackMin = 8;
	constexpr int	kWaveStackCapacity = 64;
	constexpr int	kMaxBounceSlots = 64;
	constexpr unsigned	kWaveGrid = 2048;	// 256 CUs x 8 waves per SIMD x 4 SIMDs / 4 waves per block
	// default traversal launch shape: as many persistent waves as the traversal kernels'
	// registers let every SIMD hold (256 CUs x 4 SIMDs x 5 waves / 4 waves per block; 2 %
	// faster than 1024 blocks on the sphere frame), whole-wave claims of 64 rays
	constexpr unsigned	kDefaultGrid = 1280;
	constexpr int	kDefaultRefill = 64;
	// Lanes whose next entry is a leaf (primitive tests, an instance entry) wait while other
	// lanes of the wave still have internal nodes to open, until this many are waiting or no
	// lane has a node left; then every live lane takes one step. Each lane's entries are
	// processed in the same order as without the wait, so results do not depend on it; it
	// only lets the wave run the node code and the leaf code less often side by side. The
	// triangle-only kernels wait for 16 lanes (the sphere frame is slower at 8, 12, 24 and 32);
	// the others for 32 on closest-hit rays and 48 on shadow rays (the hazy terrain: 102.8 ms,
	// against 103.7 at 32 / 32, 104.5 at 24 / 24 and 107.6 at 16 / 16).
	template <int Feat, bool AnyHit>
	constexpr int	kLeafBatch = Feat == 0 ? 16 : (AnyHit ? 48 : 32);

	__device__ __forceinline__ std::uint32_t	laneId() { return threadIdx.x & 63u; }

	// The pass's total path count: written into counts[0] (queue 0 holds every path).
	__global__ void	wf_zero_counts(std::uint32_t* counts, std::uint32_t n, const PassJob* jobs, std::uint32_t jobCount)
	{
		const std::uint32_t i = threadIdx.x;
		if (i < n)
		{
			std::uint32_t first = 0;
			if (i == 0)
			{
				for (std::uint32_t j = 0; j < jobCount; j++) first += jobs[j].pixels * jobs[j].params.sampleCount;
			}
			counts[i] = first;
		}
	}

	// log(max(u, 1e-12)) of DIM_VOLUME_DISTANCE for (pixel, sample) at bounce b.
	__device__ __forceinline__ float	pathVolumeLogU(const GpuParams& p, std::uint32_t x, std::uint32_t y, std::uint32_t sample, int b)
	{
		PixelSampler ps;
		ps.begin(p.renderSeed, x, y, sample, sampleStream(p, sample));
		ps.bounce = static_cast<std::uint32_t>(b);
		return volumeLogU(p, ps);
	}

	// Camera rays of every job of the pass in one launch (queue 0 holds the pass's paths in
	// job order, counts[0] of them).
	__global__ __launch_bounds__(kWaveBlock) void	wf_raygen(const PassJob* jobs, std::uint32_t jobCount, WaveBuffers w)
	{
		const std::uint32_t total = w.counts[0];
		for (std::uint32_t g = blockIdx.x * kWaveBlock + threadIdx.x; g < total; g += gridDim.x * kWaveBlock)
		{
			std::uint32_t lo = 0, hi = jobCount;	// the last job whose slice starts at or before g
			while (hi - lo > 1u)
			{
				const std::uint32_t mid = (lo + hi) >> 1;
				if (jobs[mid].pathBase <= g) lo = mid;
				else hi = mid;
			}
			const std::uint32_t job = lo;
			const PassJob& J = jobs[job];
			const GpuParams& p = J.params;
			const std::uint32_t pixels = J.pixels;
			// queue entry g holds the job's id-th camera ray; the queue visits the image in
			// bands of 8 rows, column by column, so a wave's 64 rays cover an 8 x 8 block
			// (coherent first traversals). The path keeps its row-major id, which the later
			// stages and the accumulation decode.
			const std::uint32_t id = g - J.pathBase;
			const std::uint32_t slot = id / pixels, order = id % pixels;
			std::uint32_t pixel;
			if (J.pixelList)
			{
				pixel = J.pixelList[order];
			}
			else
			{
				const std::uint32_t band = order / (8u * p.width);
				const std::uint32_t bandRows = min(8u, p.height - band * 8u);
				const std::uint32_t within = order - band * 8u * p.width;
				pixel = (band * 8u + within % bandRows) * p.width + within / bandRows;
			}
			const std::uint32_t path = J.pathBase + slot * pixels + (J.pixelList ? order : pixel);
			const std::uint32_t x = pixel % p.width, y = pixel / p.width;
			PixelSampler ps;
			ps.begin(p.renderSeed, x, y, p.sampleBegin + slot, sampleStream(p, p.sampleBegin + slot));
			const Ray r = cameraRay(p, x, y, ps);
			const float logU = p.hasVolumes ? volumeLogU(p, ps) : 0.0f;
			w.qOrigin[0][g] = make_float4(r.o.x, r.o.y, r.o.z, logU);
			w.qDir[0][g] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(kNoOrigin.hittable));
			w.qFrom[0][g] = kNoOrigin.tri;
			w.qPath[0][g] = w.queuedPath(job, path);
			// the path's throughput and radiance start as constants the first shade stage
			// supplies itself
		}
	}

	// Queue slots handed to the lanes of one wave, claimed from the queue's fetch counter in
	// chunks of `chunk` rays (one atomic per chunk; tunable 64..1024). All fields are
	// wave-uniform.
	constexpr std::uint32_t	kFetchChunk = 256;	// the most rays a wave claims per atomic

	struct	WaveFeed
	{
		std::uint32_t	next;
		std::uint32_t	end;
		bool	drained;

		// Gives every lane in `idle` (ballot mask) the next queue slot, in lane order; lanes
		// left without one (queue exhausted) get 0xffffffff.
		__device__ __forceinline__ std::uint32_t	hand(unsigned long long idle, std::uint32_t n, std::uint32_t* fetch, std::uint32_t chunk)
		{
			const std::uint32_t want = static_cast<std::uint32_t>(__popcll(idle));
			const std::uint32_t rank = static_cast<std::uint32_t>(__popcll(idle & ((1ull << laneId()) - 1ull)));
			const std::uint32_t avail = end - next;
			std::uint32_t slot;
			if (want <= avail)
			{
				slot = next + rank;
				next += want;
				return slot;
			}
			std::uint32_t c = 0;
			if (laneId() == static_cast<std::uint32_t>(__ffsll(static_cast<long long>(idle)) - 1))
			{
				c = atomicAdd(fetch, chunk);
			}
			c = __builtin_amdgcn_readfirstlane(__shfl(c, __ffsll(static_cast<long long>(idle)) - 1));
			const std::uint32_t cEnd = c >= n ? c : min(c + chunk, n);
			if (rank < avail) slot = next + rank;
			else
			{
				const std::uint32_t k = c + (rank - avail);
				slot = k < cEnd ? k : 0xffffffffu;
			}
			if (c >= n)
			{
				next = end;
				drained = true;
			}
			else
			{
				next = min(c + (want - avail), cEnd);
				end = cEnd;
			}
			return slot;
		}
	};

	// counters: [0] closest rays, [1] shadow rays, [2]/[3] closest nodes/primitives,
	// [4]/[5] shadow nodes/primitives (the last four only when Count)
	__device__ __forceinline__ void	flushCounters(unsigned long long* counters, int slot, std::uint32_t rays,
		VisitCounter visits, bool count)
	{
		for (int off = 32; off > 0; off >>= 1)
		{
			rays += __shfl_down(rays, off);
			if (count)
			{
				visits.nodes += __shfl_down(visits.nodes, off);
				visits.prims += __shfl_down(visits.prims, off);
			}
		}
		if (laneId() == 0 && rays)
		{
			atomicAdd(counters + slot, static_cast<unsigned long long>(rays));
			if (count)
			{
				atomicAdd(counters + 2 + 2 * slot, static_cast<unsigned long long>(visits.nodes));
				atomicAdd(counters + 3 + 2 * slot, static_cast<unsigned long long>(visits.prims));
			}
		}
	}

	// Persistent traversal over one ray queue with lane refill: a lane whose traversal has
	// ended takes the next ray of the queue once at least `refillAt` lanes of its wave are
	// idle (refillAt = 64: the wave claims 64 rays at a time and every lane traverses its
	// ray to the end before the next claim).
	// Source supplies the queue: count(), load(slot, ray, tMax) and store(slot, state).
	template <bool AnyHit, bool Count, int Feat, int LdsDepth, typename Source>
	__device__ __forceinline__ void	traceQueue(const SceneView& s, const PassJob* jobs, Source& src, std::uint32_t* fetch,
		int refillAt, std::uint32_t maxChunk, unsigned long long* counters, int counterSlot)
	{
		__shared__ std::uint32_t stackMem[LdsDepth * kWaveBlock];
		const std::uint32_t gtid = blockIdx.x * kWaveBlock + threadIdx.x;
		SpillStack<LdsDepth, kWaveBlock> stack{stackMem + threadIdx.x, src.spill + gtid, gridDim.x * kWaveBlock, 0,
			kWaveStackCapacity};
		const std::uint32_t n = src.count();
		// Rays a wave claims per atomic on the queue's counter (shared by every XCD): about an
		// eighth of an even share of the queue, a multiple of 64 between 64 and maxChunk, so a
		// long queue is claimed with few atomics and a short one still spreads over the grid.
		const std::uint32_t waves = gridDim.x * (kWaveBlock / 64);
		const std::uint32_t share = ((n / (8u * waves)) + 63u) & ~63u;
		const std::uint32_t chunk = min(max(share, 64u), maxChunk);
		// Waves beyond what the queue can feed leave before touching the claim counter: a
		// whole grid of claims on one counter costs ~50 us when the queue is short. The waves
		// that stay claim until the queue is exhausted, so every ray is still traced.
		if (static_cast<std::uint64_t>(gtid >> 6) * chunk >= n)
		{
			return;
		}
		WaveFeed feed{0u, 0u, false};
		std::uint32_t rays = 0;
		VisitCounter visits{0u, 0u};
		TraceState st;
		std::uint32_t slot = 0xffffffffu;
		bool active = false;
		// the last job's root and unbounded list: a single-scene pass loads them once per lane
		std::uint32_t lastJob = 0xffffffffu, lastUnbounded = 0u;
		std::int32_t lastRoot = -1;
		if (refillAt >= 64)
		{
			// whole-wave cadence: the wave claims 64 rays, each lane runs its traversal to the
			// end in a tight loop, then the wave claims the next 64
			for (;;)
			{
				const std::uint32_t k = feed.hand(~0ull, n, fetch, chunk);
				if (__ballot(k != 0xffffffffu) == 0ull)
				{
					break;
				}
				bool done = true;
				if (k != 0xffffffffu)
				{
					Ray r;
					float tMax, logU;
					RayOrigin from;
					const std::uint32_t job = src.load(k, r, tMax, logU, from);
					if (job != lastJob)
					{
						const GpuParams& jp = jobs[job].params;
						lastJob = job;
						lastRoot = jp.tlasRoot;
						lastUnbounded = unboundedOf(jp);
					}
					done = !traceBegin(st, s, lastRoot, lastUnbounded, r, tMax, logU, from, stack);
				}
				for (;;)
				{
					const bool wantNode = !done && (st.entry & kLeafBit) == 0u;
					const unsigned long long live = __ballot(!done);
					if (live == 0ull)
					{
						break;
					}
					const unsigned long long nodes = __ballot(wantNode);
					const bool leafTurn = nodes == 0ull || __popcll(live & ~nodes) >= kLeafBatch<Feat, AnyHit>;
					if (leafTurn ? !done : wantNode)
					{
						done = traceStep<AnyHit, Count, Feat>(s, st, kTMinF, stack, &visits);
					}
				}
				if (k != 0xffffffffu)
				{
					src.store(k, st);
					rays++;
				}
			}
			flushCounters(counters, counterSlot, rays, visits, Count);
			return;
		}
		while (true)
		{
			const unsigned long long idle = __ballot(!active);
			if (!feed.drained && (__popcll(idle) >= refillAt || idle == ~0ull))
			{
				const std::uint32_t k = feed.hand(idle, n, fetch, chunk);
				if (!active && k != 0xffffffffu)
				{
					slot = k;
					Ray r;
					float tMax, logU;
					RayOrigin from;
					const std::uint32_t job = src.load(slot, r, tMax, logU, from);
					if (job != lastJob)
					{
						const GpuParams& jp = jobs[job].params;
						lastJob = job;
						lastRoot = jp.tlasRoot;
						lastUnbounded = unboundedOf(jp);
					}
					active = traceBegin(st, s, lastRoot, lastUnbounded, r, tMax, logU, from, stack);
					if (!active)
					{
						src.store(slot, st);
						rays++;
					}
				}
			}
			if (__ballot(active) == 0ull && feed.drained)
			{
				break;
			}
			const bool wantNode = active && (st.entry & kLeafBit) == 0u;
			const unsigned long long nodes = __ballot(wantNode);
			const bool leafTurn = nodes == 0ull || __popcll(__ballot(active) & ~nodes) >= kLeafBatch<Feat, AnyHit>;
			if ((leafTurn ? active : wantNode) && traceStep<AnyHit, Count, Feat>(s, st, kTMinF, stack, &visits))
			{
				src.store(slot, st);
				rays++;
				active = false;
			}
		}
		flushCounters(counters, counterSlot, rays, visits, Count);
	}

	struct	ExtendQueue
	{
		const WaveBuffers&	w;
		int	b;
		std::uint32_t*	spill;
		__device__ __forceinline__ std::uint32_t	count() const { return w.counts[b]; }
		// returns the ray's scene job
		__device__ __forceinline__ std::uint32_t	load(std::uint32_t i, Ray& r, float& tMax, float& logU, RayOrigin& from) const
		{
			const float4 o = w.qOrigin[b & 1][i], d = w.qDir[b & 1][i];
			r = makeRay(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
			tMax = INFINITY;
			logU = o.w;
			from = RayOrigin{__float_as_uint(d.w), w.qFrom[b & 1][i]};
			return w.sceneOfPath(w.qPath[b & 1][i]);
		}
		__device__ __forceinline__ void	store(std::uint32_t i, const TraceState& st) const
		{
			w.qHit[i] = make_float4(st.found ? st.hit.t : -1.0f, st.hit.u, st.hit.v, __uint_as_float(st.hit.hittable));
			w.qTri[i] = st.hit.tri;
		}
	};

	struct	ShadowQueue
	{
		const WaveBuffers&	w;
		int	b;
		std::uint32_t*	spill;
		std::uint32_t	volumes;	// the batch has volumes: shadow rays carry a sample
		std::uint32_t	target;	// connection slot of the ray being traced
		__device__ __forceinline__ std::uint32_t	count() const { return w.counts[kMaxBounceSlots + b]; }
		__device__ __forceinline__ std::uint32_t	load(std::uint32_t i, Ray& r, float& tMax, float& logU, RayOrigin& from)
		{
			const float4 o = w.sOrigin[i], d = w.sDir[i];
			r = makeRay(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
			tMax = o.w;
			const std::uint32_t tag = __float_as_uint(d.w);
			target = w.slotOfTag(tag);
			logU = volumes ? w.sLogU[i] : 0.0f;
			const int2 f = w.sFrom[i];
			from = RayOrigin{static_cast<std::uint32_t>(f.x), f.y};
			return w.sceneOfTag(tag);
		}
		__device__ __forceinline__ void	store(std::uint32_t, const TraceState& st) const
		{
			w.visible[target] = st.found ? 0 : 1;
		}
	};

	// closest hit of every ray of queue b
	template <bool Count, int Feat, int LdsDepth>
	// The traversal kernels are asked for at least 4 waves per SIMD (at most 128 VGPRs): the
	// analytic-primitive specialisation otherwise takes 136 and runs 3 (terrain extend 41.4 ->
	// 36.8 ms); the triangle-only one needs 88 and is unaffected.
	__global__ __launch_bounds__(kWaveBlock, 4) void	wf_extend(SceneView s, const PassJob* jobs, WaveBuffers w, int b,
		unsigned long long* counters, int refillAt, std::uint32_t chunk)
	{
		ExtendQueue src{w, b, w.spill};
		traceQueue<false, Count, Feat, LdsDepth>(s, jobs, src, w.counts + 2 * kMaxBounceSlots + b, refillAt, chunk, counters, 0);
	}

	// The plain specialisation is asked for 4 waves per SIMD (128 VGPRs, a few spills): its
	// loads of per-path state are latency-bound, and 2 waves left them exposed.
	template <int Feat>
	__global__ __launch_bounds__(kWaveBlock, (Feat == 0 ? 4 : 1)) void	wf_shade(SceneView s, const PassJob* jobs, WaveBuffers w, int b,
		std::uint32_t volumes)
	{
		const std::uint32_t n = w.counts[b];
		const int q = b & 1;
		std::uint32_t* nextCount = w.counts + b + 1;
		std::uint32_t* shadowCount = w.counts + kMaxBounceSlots + b;
		// Queue appends: each wave counts its continuing paths and its shadow rays (light
		// connections, then environment ones) in LDS, and one thread claims the block's
		// ranges of the next queue and the shadow queue with one atomic each. The counters
		// are shared by every XCD, so per-wave atomics on them queue up behind each other.
		constexpr int kWaves = kWaveBlock / 64;
		__shared__ std::uint32_t waveNext[kWaves], waveShadow[kWaves], blockNext, blockShadow;
		const int wave = static_cast<int>(threadIdx.x >> 6);
		// grid-stride loop with a block-uniform trip count so every lane reaches the ballots
		// and every wave the barriers
		const std::uint32_t stride = gridDim.x * kWaveBlock;
		for (std::uint32_t blockBase = blockIdx.x * kWaveBlock; blockBase < n; blockBase += stride)
		{
			const std::uint32_t i = blockBase + threadIdx.x;
			const bool valid = i < n;
			bool cont = false, wantLight = false, wantEnv = false;
			Ray next{};
			Connection cl{}, ce{};
			f3 shadePoint = mk(0, 0, 0);
			RayOrigin from = kNoOrigin;
			const std::uint32_t qp = valid ? w.qPath[q][i] : 0u;
			const std::uint32_t job = w.sceneOfPath(qp);
			const std::uint32_t path = w.pathOf(qp);
			float nextLogU = 0.0f, shadowLogU = 0.0f;
			// one scene per iteration, so the scene's parameters are wave-uniform (scalar loads)
			unsigned long long pending = __ballot(valid);
			while (pending)
			{
				const int leader = __ffsll(static_cast<long long>(pending)) - 1;
				const std::uint32_t cur = __builtin_amdgcn_readfirstlane(__shfl(job, leader));
				const bool mine = valid && job == cur;
				if (mine)
				{
				const PassJob& J = jobs[cur];
				const GpuParams& p = J.params;
				const bool sampledInfinite = hasEnvLight(p) || hasConstantBackgroundLight(p);
				const std::uint32_t local = path - J.pathBase;
				const std::uint32_t pixel = passPixel(J, local % J.pixels), slot = local / J.pixels;
				const std::uint32_t px = pixel % p.width, py = pixel / p.width;
				PixelSampler ps;
				ps.begin(p.renderSeed, px, py, p.sampleBegin + slot, sampleStream(p, p.sampleBegin + slot));
				ps.bounce = static_cast<std::uint32_t>(b);
				const float4 hv = w.qHit[i];
				const float4 o = w.qOrigin[q][i], d = w.qDir[q][i];
				const Ray ray = makeRay(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
				shadowLogU = o.w;
				// a camera path starts with throughput 1, no radiance, and counts as leaving a
				// specular vertex (the camera): those are not stored by ray generation
				float4 t4 = b == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : w.throughput[path];
				float4 r4 = b == 0 ? make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(PF_PREV_SPECULAR)) : w.radiance[path];
				f3 thr = mk(t4.x, t4.y, t4.z), acc = mk(r4.x, r4.y, r4.z);
				std::uint32_t flags = __float_as_uint(r4.w) & (PF_PREV_SPECULAR | PF_PREV_INFINITE);
				const bool prevSpecular = flags & PF_PREV_SPECULAR;
				const bool prevInfinite = flags & PF_PREV_INFINITE;
				const float prevPdf = t4.w;
				bool done = true;
				if (hv.x < 0.0f)
				{
					float wgt = 1.0f;
					if (b > 0 && !prevSpecular && prevInfinite)
					{
						wgt = powerHeuristic(prevPdf, infiniteLightPdf(s, p, ray.d));
					}
					f3 c = thr * missRadiance<Feat>(s, p, ray, ps, wgt);
					if (b > 0 && hasConstantBackgroundLight(p)) c = clampConstantBackground(p, c);
					acc = acc + clampRay(c);
					if (b == 0 && J.features)
					{
						primaryFeatures(J.features + static_cast<size_t>(kNforFeatures) * path, p, px, py, false, 0.0f, mk(0, 0, 0),
							missRadiance<Feat>(s, p, ray, ps, 1.0f), 0u);
					}
				}
				else
				{
					HitInfo h;
					h.t = hv.x;
					h.u = hv.y;
					h.v = hv.z;
					h.hittable = __float_as_uint(hv.w);
					h.tri = w.qTri[i];
					Surface sf = surfaceAt<Feat>(s, ray, h);
					shadePoint = sf.position;
					from = originOf(s, h);
					const GpuMaterial& m = s.materials[sf.material];
					if (b == 0 && J.features)
					{
						primaryFeatures(J.features + static_cast<size_t>(kNforFeatures) * path, p, px, py, true, h.t, sf.normal,
							ld3(m.color), m.kind);
					}
					bool alive = true;
					if constexpr ((Feat & kFeatRich) != 0)
					{
						if (b == 0 && p.sky == GPU_SKY_ATMOSPHERE)
						{
							atmospherePrimary(p.atmosphere, ray.o, ray.d, h.t, acc, thr);
							alive = !terminated(thr);
						}
					}
					const f3 emitted = m.kind == GPU_MAT_EMISSIVE ? ld3(m.color) : mk(0, 0, 0);
					Scatter sc;
					if (!alive)
					{
					}
					else if (!scatter<Feat>(m, sf, ray.d, ps, sc))
					{
						float lp = 0.0f;
						if (!prevSpecular && p.lightCount > 0)
						{
							const float4 po = w.prevOrigin[path];
							lp = lightsPdf<Feat>(s, p, mk(po.x, po.y, po.z), ray.d);
						}
						const float wgt = prevSpecular ? 1.0f : powerHeuristic(prevPdf, lp);
						acc = acc + clampRay((thr * emitted) * wgt);
					}
					else
					{
						if constexpr ((Feat & kFeatSubsurface) != 0)
						{
							if (sc.subsurfaceExit)
							{
								LocalStack probeStack;
								probeStack.reset();
								std::uint32_t probes = 0;
								relocateSubsurface<Feat>(s, p, m, h, sf, sc, ps, probeStack, shadowLogU, probes);
								from = originOf(s, h);
								shadePoint = sf.position;
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
							if (leavesOpaque(sf, m, sc.direction))
							{
								thr = clampRay(thr * att);
								if (!terminated(thr) && russianRoulette(thr, b, ps))
								{
									next = makeRay(sf.position, sc.direction);
									flags = PF_PREV_SPECULAR;
									t4.w = 0.0f;
									done = false;
								}
							}
						}
						else
						{
							acc = acc + thr * emitted;
							if (!terminated(thr * att))
							{
								if (p.lightCount > 0)
								{
									cl = connectLight<Feat>(s, p, sf, m, sc, b, ps);
									if (cl.active)
									{
										cl.contribution = thr * cl.contribution;	// clampRay applied at resolve
										wantLight = cl.tMax > 0.0f;
										flags |= PF_CONN_LIGHT | (wantLight ? 0u : PF_CONN_FREE_LIGHT);
									}
								}
								if (sampledInfinite)
								{
									ce = connectInfinite<Feat>(s, p, sf, m, sc, b, ps);
									if (ce.active)
									{
										ce.contribution = thr * ce.contribution;
										wantEnv = true;
										flags |= PF_CONN_ENV;
									}
								}
								const f3 dir = scatterDirection<Feat>(sf, sc, ps);
								if (leavesOpaque(sf, m, dir))
								{
									const float pdf = scatterPdf<Feat>(m, sf, sc, dir);
									if (pdf > 0.0f && isfinite(pdf))
									{
										thr = clampRay(thr * att);
										if (!terminated(thr) && russianRoulette(thr, b, ps))
										{
											next = makeRay(sf.position, dir);
											flags = (flags & (PF_CONN_LIGHT | PF_CONN_ENV | PF_CONN_FREE_LIGHT))
												| (samplesVisibleInfinite(p, sc) ? PF_PREV_INFINITE : 0u);
											t4.w = pdf;
											// read back only for an emitter's MIS weight
											if (p.lightCount > 0)
											{
												w.prevOrigin[path] = make_float4(sf.position.x, sf.position.y, sf.position.z, 0.0f);
											}
											done = false;
										}
									}
								}
							}
						}
					}
				}
				if (done || b >= p.maxBounces)
				{
					flags |= PF_DONE;
					done = true;
				}
				cont = !done;
				w.throughput[path] = make_float4(thr.x, thr.y, thr.z, t4.w);
				w.radiance[path] = make_float4(acc.x, acc.y, acc.z, __uint_as_float(flags));
				if (flags & PF_CONN_LIGHT)
				{
					w.connection[2 * path] = make_float4(cl.contribution.x, cl.contribution.y, cl.contribution.z, 0.0f);
					w.visible[2 * path] = wantLight ? 0 : 1;
				}
				if (flags & PF_CONN_ENV)
				{
					w.connection[2 * path + 1] = make_float4(ce.contribution.x, ce.contribution.y, ce.contribution.z, 0.0f);
					w.visible[2 * path + 1] = 0;
				}
				if (cont && p.hasVolumes)
				{
					nextLogU = pathVolumeLogU(p, px, py, p.sampleBegin + slot, b + 1);
				}
				}
				pending &= ~__ballot(mine);
			}
			const unsigned long long mNext = __ballot(cont), mLight = __ballot(wantLight), mEnv = __ballot(wantEnv);
			const std::uint32_t nLight = static_cast<std::uint32_t>(__popcll(mLight));
			if (laneId() == 0)
			{
				waveNext[wave] = static_cast<std::uint32_t>(__popcll(mNext));
				waveShadow[wave] = nLight + static_cast<std::uint32_t>(__popcll(mEnv));
			}
			__syncthreads();
			if (threadIdx.x == 0)
			{
				std::uint32_t tNext = 0, tShadow = 0;
				for (int k = 0; k < kWaves; k++)
				{
					tNext += waveNext[k];
					tShadow += waveShadow[k];
				}
				blockNext = tNext > 0 ? atomicAdd(nextCount, tNext) : 0u;
				blockShadow = tShadow > 0 ? atomicAdd(shadowCount, tShadow) : 0u;
			}
			__syncthreads();
			std::uint32_t nextBase = blockNext, shadowBase = blockShadow;
			for (int k = 0; k < wave; k++)
			{
				nextBase += waveNext[k];
				shadowBase += waveShadow[k];
			}
			const unsigned long long below = (1ull << laneId()) - 1ull;
			const std::uint32_t pos = nextBase + static_cast<std::uint32_t>(__popcll(mNext & below));
			if (cont)
			{
				const int nq = (b + 1) & 1;
				w.qOrigin[nq][pos] = make_float4(next.o.x, next.o.y, next.o.z, nextLogU);
				w.qDir[nq][pos] = make_float4(next.d.x, next.d.y, next.d.z, __uint_as_float(from.hittable));
				w.qFrom[nq][pos] = from.tri;
				w.qPath[nq][pos] = qp;
			}
			const std::uint32_t tagScene = job;
			const std::uint32_t sl = shadowBase + static_cast<std::uint32_t>(__popcll(mLight & below));
			if (wantLight)
			{
				w.sOrigin[sl] = make_float4(shadePoint.x, shadePoint.y, shadePoint.z, cl.tMax);
				w.sDir[sl] = make_float4(cl.direction.x, cl.direction.y, cl.direction.z, __uint_as_float(w.shadowTag(tagScene, 2u * path)));
				w.sFrom[sl] = make_int2(static_cast<int>(from.hittable), from.tri);
				if (volumes) w.sLogU[sl] = shadowLogU;
			}
			const std::uint32_t se = shadowBase + nLight + static_cast<std::uint32_t>(__popcll(mEnv & below));
			if (wantEnv)
			{
				w.sOrigin[se] = make_float4(shadePoint.x, shadePoint.y, shadePoint.z, INFINITY);
				w.sDir[se] = make_float4(ce.direction.x, ce.direction.y, ce.direction.z, __uint_as_float(w.shadowTag(tagScene, 2u * path + 1u)));
				w.sFrom[se] = make_int2(static_cast<int>(from.hittable), from.tri);
				if (volumes) w.sLogU[se] = shadowLogU;
			}
			// the next iteration rewrites the per-wave counts
			__syncthreads();
		}
	}

	// any-hit test of the shadow rays appended in bounce b
	template <bool Count, int Feat, int LdsDepth>
	__global__ __launch_bounds__(kWaveBlock, 4) void	wf_shadow(SceneView s, const PassJob* jobs, WaveBuffers w, int b,
		unsigned long long* counters, int refillAt, std::uint32_t chunk, std::uint32_t volumes)
	{
		ShadowQueue src{w, b, w.spill, volumes, 0u};
		traceQueue<true, Count, Feat, LdsDepth>(s, jobs, src, w.counts + 3 * kMaxBounceSlots + b, refillAt, chunk, counters, 1);
	}

	// sampleColors is indexed by global path id
	__global__ __launch_bounds__(kWaveBlock) void	wf_resolve(WaveBuffers w, int b, float* __restrict__ sampleColors)
	{
		const std::uint32_t n = w.counts[b];
		const int q = b & 1;
		for (std::uint32_t i = blockIdx.x * kWaveBlock + threadIdx.x; i < n; i += gridDim.x * kWaveBlock)
		{
			const std::uint32_t path = w.pathOf(w.qPath[q][i]);
			float4 r4 = w.radiance[path];
			std::uint32_t flags = __float_as_uint(r4.w);
			f3 acc = mk(r4.x, r4.y, r4.z);
			if (flags & PF_CONN_LIGHT)
			{
				if (w.visible[2 * path])
				{
					const float4 c = w.connection[2 * path];
					acc = acc + clampRay(mk(c.x, c.y, c.z));
				}
			}
			if (flags & PF_CONN_ENV)
			{
				if (w.visible[2 * path + 1])
				{
					const float4 c = w.connection[2 * path + 1];
					acc = acc + clampRay(mk(c.x, c.y, c.z));
				}
			}
			if (flags & PF_DONE)
			{
				const f3 c = cleanColor(acc);
				float* out = sampleColors + static_cast<size_t>(path) * 3;
				out[0] = c.x;
				out[1] = c.y;
				out[2] = c.z;
			}
			else
			{
				flags &= (PF_PREV_SPECULAR | PF_PREV_INFINITE);
				w.radiance[path] = make_float4(acc.x, acc.y, acc.z, __uint_as_float(flags));
			}
		}
	}

	// Adds a pass's sample colours of one scene into its FP64 pixel sums, in sample order
	// (the per-pixel sum of _threadRender, Threads.cpp:551-583), and in adaptive rounds into
	// the pixel's AdaptiveAccumulator (Threads.cpp:571-576).
	__global__ void	wf_accumulate(const float* __restrict__ colors, const PassJob* jobs, std::uint32_t job)
	{
		const PassJob& J = jobs[job];
		const std::uint32_t count = J.params.sampleCount;
		if (J.denoise)
		{
			// denoising: half-buffer and feature statistics (Threads.cpp:585-627)
			for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < J.pixels; i += gridDim.x * blockDim.x)
			{
				const size_t px = passPixel(J, i);
				double a[kDenoiseStats];
				for (int j = 0; j < kDenoiseStats; j++) a[j] = J.denoise[kDenoiseStats * px + j];
				double m[kAdaptiveMoments];
				if (J.moments)
				{
					for (int k = 0; k < kAdaptiveMoments; k++) m[k] = J.moments[kAdaptiveMoments * px + k];
				}
				for (std::uint32_t k = 0; k < count; k++)
				{
					const size_t g = static_cast<size_t>(J.pathBase) + static_cast<size_t>(k) * J.pixels + i;
					denoiseAddSample(a, colors + 3 * g, J.features + kNforFeatures * g, ((J.params.sampleBegin + k) & 1u) != 0u);
					if (J.moments) adaptiveAdd(m, colors[3 * g], colors[3 * g + 1], colors[3 * g + 2]);
				}
				for (int j = 0; j < kDenoiseStats; j++) J.denoise[kDenoiseStats * px + j] = a[j];
				if (J.moments)
				{
					for (int k = 0; k < kAdaptiveMoments; k++) J.moments[kAdaptiveMoments * px + k] = m[k];
				}
			}
			return;
		}
		for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < J.pixels; i += gridDim.x * blockDim.x)
		{
			const size_t px = passPixel(J, i);
			double r = J.sums[3 * px], g = J.sums[3 * px + 1], b = J.sums[3 * px + 2];
			double m[kAdaptiveMoments];
			if (J.moments)
			{
				for (int k = 0; k < kAdaptiveMoments; k++) m[k] = J.moments[kAdaptiveMoments * px + k];
			}
			for (std::uint32_t k = 0; k < count; k++)
			{
				const float* c = colors + (static_cast<size_t>(J.pathBase) + static_cast<size_t>(k) * J.pixels + i) * 3;
				const double cr = c[0], cg = c[1], cb = c[2];
				r += cr;
				g += cg;
				b += cb;
				if (J.moments) adaptiveAdd(m, cr, cg, cb);
			}
			J.sums[3 * px] = r;
			J.sums[3 * px + 1] = g;
			J.sums[3 * px + 2] = b;
			if (J.moments)
			{
				for (int k = 0; k < kAdaptiveMoments; k++) J.moments[kAdaptiveMoments * px + k] = m[k];
			}
		}
	}

	// Launch shape of the traversal stages: persistent blocks per launch and the number of
	// idle lanes at which a wave takes new rays (1..64).
	// Launch shape of the traversal kernels; 0 = default. None of it changes results: every
	// ray is traversed in the same order whichever lane, chunk or stack level holds it.
	struct	WaveTuning
	{
		unsigned	grid = 0;			// persistent blocks per traversal launch
		int	extendRefill = 0;		// idle lanes at which a wave takes new closest-hit rays
		int	shadowRefill = 0;		// idle lanes at which a wave takes new shadow rays
		unsigned	chunk = 0;			// most rays a wave claims per queue atomic
		int	ldsStack = 0;			// LDS stack entries per lane (the rest spill to HBM)

		WaveTuning	resolved() const
		{
			WaveTuning t = *this;
			if (t.grid == 0) t.grid = kDefaultGrid;
			if (t.extendRefill == 0) t.extendRefill = kDefaultRefill;
			if (t.shadowRefill == 0) t.shadowRefill = kDefaultRefill;
			if (t.chunk == 0) t.chunk = kFetchChunk;
			if (t.ldsStack == 0) t.ldsStack = kWaveLdsStack;
			if (t.grid > kWaveGrid || t.extendRefill < 1 || t.extendRefill > 64 || t.shadowRefill < 1 || t.shadowRefill > 64
				|| t.chunk < 64 || t.chunk > 1024 || (t.ldsStack != 8 && t.ldsStack != 16 && t.ldsStack != 24))
			{
				throw std::invalid_argument("wavefront tuning out of range: grid 1.." + std::to_string(kWaveGrid)
					+ ", refill 1..64, chunk 64..1024, LDS stack 8, 16 or 24");
			}
			return t;
		}
		bool	operator==(const WaveTuning& o) const
		{
			return grid == o.grid && extendRefill == o.extendRefill && shadowRefill == o.shadowRefill && chunk == o.chunk
				&& ldsStack == o.ldsStack;
		}
	};

	struct	StageTimes
	{
		double	extendMs = 0.0, shadeMs = 0.0, shadowMs = 0.0;
		std::uint32_t	extendLaunches = 0;
		std::uint32_t	passes = 0;
		std::uint32_t	graphReplays = 0;
	};

	// One scene of a render call: its parameters (sampleBegin / sampleCount: the whole
	// range of this call) and its FP64 pixel sums.
	struct	BatchJob
	{
		GpuParams	params;
		double*	sums;
		const std::uint32_t*	pixelList = nullptr;	// adaptive round: the active pixels
		std::uint32_t	listCount = 0;				// ... and how many
		double*	moments = nullptr;				// adaptive round: per-pixel accumulators
		double*	denoise = nullptr;				// denoising: per-pixel statistics (sums unused)
	};

	inline std::uint32_t	jobPixels(const BatchJob& j)
	{
		return j.pixelList ? j.listCount : j.params.width * j.params.height;
	}

	// Device state of one context (grown on demand, kept between renders), and the
	// captured pass graph with what it was captured for.
	struct	WavefrontState
	{
		void*	mem = nullptr;
		size_t	bytes = 0;
		WaveBuffers	buf{};
		float*	colors = nullptr;		// [global path][3]
		float*	features = nullptr;		// denoising: [global path][kNforFeatures], on demand
		std::uint32_t	featureCapacity = 0;
		PassJob*	jobs = nullptr;		// the current pass's table
		PassJob*	plans = nullptr;	// every pass's table of a render
		size_t	planCapacity = 0;
		hipGraphExec_t	graph = nullptr;
		std::vector<unsigned char>	graphKey;

		void	dropGraph()
		{
			if (graph) (void)hipGraphExecDestroy(graph);
			graph = nullptr;
			graphKey.clear();
		}

		void	release()
		{
			dropGraph();
			if (mem) (void)hipFree(mem);
			if (plans) (void)hipFree(plans);
			if (features) (void)hipFree(features);
			features = nullptr;
			featureCapacity = 0;
			mem = nullptr;
			plans = nullptr;
			planCapacity = 0;
			bytes = 0;
		}

		// Carves the state arrays for `capacity` paths out of `base` (each 256-byte aligned)
		// and returns the bytes used; with base == nullptr it only measures.
		size_t	carve(char* base, std::uint32_t capacity)
		{
			const size_t n = capacity;
			const size_t spillBytes = static_cast<size_t>(kWaveStackCapacity - kWaveLdsStackMin) * kWaveGrid * kWaveBlock * 4;
			size_t off = 0;
			auto take = [&](size_t b) { char* r = base ? base + off : nullptr; off += (b + 255) & ~size_t(255); return r; };
			WaveBuffers nb;
			std::memset(&nb, 0, sizeof(nb));	// padding too: graph keys compare the raw bytes
			for (int k = 0; k < 2; k++)
			{
				nb.qOrigin[k] = reinterpret_cast<float4*>(take(16 * n));
				nb.qDir[k] = reinterpret_cast<float4*>(take(16 * n));
				nb.qFrom[k] = reinterpret_cast<std::int32_t*>(take(4 * n));
				nb.qPath[k] = reinterpret_cast<std::uint32_t*>(take(4 * n));
			}
			nb.qHit = reinterpret_cast<float4*>(take(16 * n));
			nb.qTri = reinterpret_cast<std::int32_t*>(take(4 * n));
			nb.throughput = reinterpret_cast<float4*>(take(16 * n));
			nb.radiance = reinterpret_cast<float4*>(take(16 * n));
			nb.prevOrigin = reinterpret_cast<float4*>(take(16 * n));
			nb.connection = reinterpret_cast<float4*>(take(32 * n));
			nb.visible = reinterpret_cast<std::uint8_t*>(take(2 * n));
			nb.sOrigin = reinterpret_cast<float4*>(take(32 * n));
			nb.sDir = reinterpret_cast<float4*>(take(32 * n));
			nb.sLogU = reinterpret_cast<float*>(take(8 * n));
			nb.sFrom = reinterpret_cast<int2*>(take(16 * n));
			float* nc = reinterpret_cast<float*>(take(12 * n));
			nb.counts = reinterpret_cast<std::uint32_t*>(take(kMaxBounceSlots * 4 * 4));
			PassJob* nj = reinterpret_cast<PassJob*>(take(sizeof(PassJob) * kMaxBatchScenes));
			nb.spill = reinterpret_cast<std::uint32_t*>(take(spillBytes));
			nb.capacity = capacity;
			if (base)
			{
				buf = nb;
				colors = nc;
				jobs = nj;
			}
			return off;
		}

		// The layout is fixed between reallocations (captured graphs hold its pointers):
		// a smaller request keeps the larger layout.
		void	ensure(std::uint32_t capacity)
		{
			if (mem && capacity <= buf.capacity)
			{
				return;
			}
			const size_t need = carve(nullptr, capacity);
			if (need > bytes)
			{
				dropGraph();
				if (mem) (void)hipFree(mem);
				mem = nullptr;
				bytes = 0;
				if (hipMalloc(&mem, need) != hipSuccess)
				{
					mem = nullptr;
					throw std::runtime_error("wavefront state allocation failed");
				}
				bytes = need;
			}
			carve(static_cast<char*>(mem), capacity);
		}

		// The per-path feature buffer of denoising passes (not part of the captured layout:
		// passes find it through their job table).
		void	ensureFeatures(std::uint32_t capacity)
		{
			if (features && capacity <= featureCapacity) return;
			if (features) (void)hipFree(features);
			features = nullptr;
			featureCapacity = 0;
			if (hipMalloc(reinterpret_cast<void**>(&features), static_cast<size_t>(capacity) * kNforFeatures * sizeof(float)) != hipSuccess)
			{
				features = nullptr;
				throw std::runtime_error("wavefront feature buffer allocation failed");
			}
			featureCapacity = capacity;
		}

		void	ensurePlans(size_t entries)
		{
			if (entries <= planCapacity) return;
			if (plans) (void)hipFree(plans);
			plans = nullptr;
			planCapacity = 0;
			if (hipMalloc(reinterpret_cast<void**>(&plans), entries * sizeof(PassJob)) != hipSuccess)
			{
				plans = nullptr;
				throw std::runtime_error("wavefront pass-table allocation failed");
			}
			planCapacity = entries;
		}
	};

	inline void	checkHip(hipError_t e, const char* what)
	{
		if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
	}

	// Enqueues the launches of one pass (the table is already in st.jobs).
	template <int Feat>
	inline void	enqueuePassF(WavefrontState& st, const SceneView& s, std::uint32_t jobCount, int maxBounces, bool volumes,
		unsigned long long* counters, hipStream_t stream, bool count, WaveTuning tune, std::vector<hipEvent_t>* marks)
	{
		const WaveBuffers& w = st.buf;
		auto mark = [&]() {
			if (!marks) return;
			hipEvent_t e;
			checkHip(hipEventCreate(&e), "hipEventCreate");
			marks->push_back(e);
			checkHip(hipEventRecord(e, stream), "hipEventRecord");
		};
		const std::uint32_t vol = volumes ? 1u : 0u;
		constexpr int T = Feat & kFeatAll;	// extend and shadow differ only by primitive / material class
		wf_zero_counts<<<1, 4 * kMaxBounceSlots, 0, stream>>>(w.counts, 4 * kMaxBounceSlots, st.jobs, jobCount);
		wf_raygen<<<kWaveGrid, kWaveBlock, 0, stream>>>(st.jobs, jobCount, w);
		for (int b = 0; b <= maxBounces; b++)
		{
			mark();
			const unsigned ch = tune.chunk;
			if (count) wf_extend<true, T, kWaveLdsStack><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.extendRefill, ch);
			else if (tune.ldsStack == 8) wf_extend<false, T, 8><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.extendRefill, ch);
			else if (tune.ldsStack == 24) wf_extend<false, T, 24><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.extendRefill, ch);
			else wf_extend<false, T, 16><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.extendRefill, ch);
			mark();
			wf_shade<Feat><<<kWaveGrid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, vol);
			mark();
			if (count) wf_shadow<true, T, kWaveLdsStack><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.shadowRefill, ch, vol);
			else if (tune.ldsStack == 8) wf_shadow<false, T, 8><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.shadowRefill, ch, vol);
			else if (tune.ldsStack == 24) wf_shadow<false, T, 24><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.shadowRefill, ch, vol);
			else wf_shadow<false, T, 16><<<tune.grid, kWaveBlock, 0, stream>>>(s, st.jobs, w, b, counters, tune.shadowRefill, ch, vol);
			mark();
			wf_resolve<<<kWaveGrid, kWaveBlock, 0, stream>>>(w, b, st.colors);
		}
		for (std::uint32_t j = 0; j < jobCount; j++)
		{
			wf_accumulate<<<1024, 256, 0, stream>>>(st.colors, st.jobs, j);
		}
	}

	inline void	enqueuePass(int features, WavefrontState& st, const SceneView& s, std::uint32_t jobCount, int maxBounces, bool volumes,
		unsigned long long* counters, hipStream_t stream, bool count, WaveTuning tune, std::vector<hipEvent_t>* marks)
	{
		switch (features & kFeatEvery)
		{
			case kFeatEvery: enqueuePassF<kFeatEvery>(st, s, jobCount, maxBounces, volumes, counters, stream, count, tune, marks); break;
			case 0: enqueuePassF<0>(st, s, jobCount, maxBounces, volumes, counters, stream, count, tune, marks); break;
			case kFeatAnalytic: enqueuePassF<kFeatAnalytic>(st, s, jobCount, maxBounces, volumes, counters, stream, count, tune, marks); break;
			case kFeatRich: enqueuePassF<kFeatRich>(st, s, jobCount, maxBounces, volumes, counters, stream, count, tune, marks); break;
			default: enqueuePassF<kFeatAll>(st, s, jobCount, maxBounces, volumes, counters, stream, count, tune, marks); break;
		}
	}

	// Renders every job's sample range into its sums: passes of at most kMaxPassPaths paths
	// over all scenes at once. Without stage timing or visit counting, one pass is captured
	// as a hipGraph (its varying inputs are all in device memory) and replayed for every pass
	// and every later render with the same shape.
	inline void	renderWavefront(WavefrontState& st, const SceneView& s, int features, const std::vector<BatchJob>& jobs,
		unsigned long long* counters, hipStream_t stream, bool count, StageTimes* times, WaveTuning tune)
	{
		tune = tune.resolved();
		if (jobs.empty() || jobs.size() > kMaxBatchScenes)
		{
			throw std::invalid_argument("wavefront: 1.." + std::to_string(kMaxBatchScenes) + " scenes per launch");
		}
		int maxBounces = 0;
		bool volumes = false;
		std::uint64_t total = 0;
		for (const BatchJob& j : jobs)
		{
			if (j.params.maxBounces + 2 > kMaxBounceSlots)
			{
				throw std::runtime_error("Unsupported: more than 62 bounces in the wavefront integrator");
			}
			const std::uint64_t pixels = j.pixelList ? j.listCount : static_cast<std::uint64_t>(j.params.width) * j.params.height;
			if (pixels == 0 || pixels > std::min<std::uint64_t>(kMaxPassPaths, 1ull << pathBitsFor(static_cast<std::uint32_t>(jobs.size()))))
			{
				throw std::runtime_error("Unsupported: image of more pixels than a wavefront pass holds (2^27, fewer in large batches)");
			}
			maxBounces = std::max(maxBounces, j.params.maxBounces);
			volumes = volumes || j.params.hasVolumes;
			total += pixels * j.params.sampleCount;
		}
		const std::uint32_t n = static_cast<std::uint32_t>(jobs.size());
		const std::uint32_t pathBits = pathBitsFor(n);
		const std::uint32_t capacity = static_cast<std::uint32_t>(std::min<std::uint64_t>(std::min<std::uint64_t>(kMaxPassPaths, 1ull << pathBits),
			std::max<std::uint64_t>(total, 1)));
		// pass plan: each pass gives every scene with samples left as many as the budget holds
		std::vector<PassJob> plan;
		std::vector<std::uint32_t> done(n, 0);
		for (;;)
		{
			bool any = false;
			for (std::uint32_t j = 0; j < n; j++) any = any || done[j] < jobs[j].params.sampleCount;
			if (!any) break;
			std::uint32_t budget = capacity, base = 0;
			for (std::uint32_t j = 0; j < n; j++)
			{
				const GpuParams& p = jobs[j].params;
				const std::uint32_t pixels = jobPixels(jobs[j]);
				const std::uint32_t c = std::min(p.sampleCount - done[j], budget / pixels);
				PassJob pj{};
				pj.params = p;
				pj.params.sampleBegin = p.sampleBegin + done[j];
				pj.params.sampleCount = c;
				pj.pathBase = base;
				pj.pixels = pixels;
				pj.sums = jobs[j].sums;
				pj.pixelList = jobs[j].pixelList;
				pj.moments = jobs[j].moments;
				pj.denoise = jobs[j].denoise;
				plan.push_back(pj);
				done[j] += c;
				base += pixels * c;
				budget -= pixels * c;
			}
		}
		const std::uint32_t passes = static_cast<std::uint32_t>(plan.size() / n);
		st.ensure(capacity);
		st.buf.pathBits = pathBits;
		bool denoising = false;
		for (const BatchJob& j : jobs) denoising = denoising || j.denoise != nullptr;
		if (denoising)
		{
			st.ensureFeatures(st.buf.capacity);
			for (PassJob& pj : plan) pj.features = pj.denoise ? st.features : nullptr;
		}
		st.ensurePlans(plan.size());
		// the host table dies with this call: copy it before returning (ordered on the stream)
		checkHip(hipMemcpyAsync(st.plans, plan.data(), plan.size() * sizeof(PassJob), hipMemcpyHostToDevice, stream), "pass tables");
		checkHip(hipStreamSynchronize(stream), "pass tables");
		const bool useGraph = !count && !times;
		if (useGraph)
		{
			// what the captured launches depend on besides device memory
			std::vector<unsigned char> key(sizeof(SceneView) + sizeof(WaveTuning) + 4 * sizeof(std::uint32_t) + sizeof(void*) * 2);
			unsigned char* k = key.data();
			std::memcpy(k, &s, sizeof(SceneView)); k += sizeof(SceneView);
			std::memcpy(k, &tune, sizeof(WaveTuning)); k += sizeof(WaveTuning);
			const std::uint32_t shape[4] = {n, static_cast<std::uint32_t>(maxBounces), volumes ? 1u : 0u,
				static_cast<std::uint32_t>(features)};
			std::memcpy(k, shape, sizeof(shape)); k += sizeof(shape);
			void* ptrs[2] = {st.mem, counters};
			std::memcpy(k, ptrs, sizeof(ptrs));
			// the buffer layout the launches were captured with
			const unsigned char* lb = reinterpret_cast<const unsigned char*>(&st.buf);
			key.insert(key.end(), lb, lb + sizeof(WaveBuffers));
			const void* more[2] = {st.jobs, st.colors};
			const unsigned char* mb = reinterpret_cast<const unsigned char*>(more);
			key.insert(key.end(), mb, mb + sizeof(more));
			if (!st.graph || st.graphKey != key)
			{
				st.dropGraph();
				hipGraph_t g = nullptr;
				checkHip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
				enqueuePass(features, st, s, n, maxBounces, volumes, counters, stream, false, tune, nullptr);
				checkHip(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
				const hipError_t e = hipGraphInstantiate(&st.graph, g, nullptr, nullptr, 0);
				(void)hipGraphDestroy(g);
				checkHip(e, "hipGraphInstantiate");
				st.graphKey = key;
			}
		}
		// stage-timing events, destroyed on every exit (an error thrown mid-pass included)
		struct Marks
		{
			std::vector<hipEvent_t>	v;
			~Marks() { for (hipEvent_t e : v) (void)hipEventDestroy(e); }
		} held;
		std::vector<hipEvent_t>& marks = held.v;
		for (std::uint32_t k = 0; k < passes; k++)
		{
			checkHip(hipMemcpyAsync(st.jobs, st.plans + static_cast<size_t>(k) * n, n * sizeof(PassJob), hipMemcpyDeviceToDevice,
				stream), "pass table");
			if (useGraph)
			{
				checkHip(hipGraphLaunch(st.graph, stream), "hipGraphLaunch");
				if (times) times->graphReplays++;
			}
			else
			{
				enqueuePass(features, st, s, n, maxBounces, volumes, counters, stream, count, tune, times ? &marks : nullptr);
			}
		}
		if (times)
		{
			times->passes += passes;
			checkHip(hipStreamSynchronize(stream), "wavefront stage timing");
			for (size_t k = 0; k + 3 < marks.size(); k += 4)
			{
				float a = 0, b2 = 0, c = 0;
				(void)hipEventElapsedTime(&a, marks[k], marks[k + 1]);
				(void)hipEventElapsedTime(&b2, marks[k + 1], marks[k + 2]);
				(void)hipEventElapsedTime(&c, marks[k + 2], marks[k + 3]);
				times->extendMs += a;
				times->shadeMs += b2;
				times->shadowMs += c;
				times->extendLaunches++;
			}
		}
	}
}
