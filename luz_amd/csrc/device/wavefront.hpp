// Wavefront integrator: the bounce loop of calculateLightRaysColor (ColorHelper.cpp:1136-1326)
// split into stages over SoA path state in HBM, so the traversal kernels carry only a ray,
// a stack and a hit record (low VGPR count, high occupancy) and the shading kernel carries
// no traversal state.
//
// Per pass (S samples of every pixel, path id = sample_slot * pixels + pixel):
//   raygen            camera rays (FP64, bit-exact sampler) -> queue 0, path state init
//   for b = 0..maxBounces:
//     extend          closest hit for every ray of queue b (queue order, coalesced)
//     shade           miss / emission / scatter / NEE and environment connections; the
//                     next ray and up to two shadow rays are appended with one atomic per
//                     wave (ballot + popcount prefix)
//     shadow          any-hit test of the appended shadow rays -> visibility bytes
//     resolve         adds visible connection terms in the reference's order (light, then
//                     environment) and writes the colour of every path that ended
// Queue lengths live in device memory and every stage is a grid-stride loop over them, so
// the host enqueues the whole pass without a synchronisation.
#pragma once

#include "integrator.hpp"

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
		float4*	qDir[2];
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
		float4*	sDir;		// xyz, w = 2 * path + slot (bits)
		// counters: [b] rays in queue b, [kMaxBounce + 1 + b] shadow rays of bounce b
		std::uint32_t*	counts;
		std::uint32_t	capacity;
		std::uint32_t*	spill;		// traversal stack spill, [entry][grid thread]
	};

	constexpr int	kWaveBlock = 256;
	constexpr int	kWaveLdsStack = 16;
	constexpr int	kWaveStackCapacity = 64;
	constexpr int	kMaxBounceSlots = 64;
	constexpr unsigned	kWaveGrid = 2048;	// 256 CUs x 8 waves per SIMD x 4 SIMDs / 4 waves per block
	constexpr int	kDefaultRefill = 16;

	__device__ __forceinline__ std::uint32_t	laneId() { return threadIdx.x & 63u; }

	// Appends `want` (0/1) items per lane with a single atomic per wave; returns the slot.
	__device__ __forceinline__ std::uint32_t	waveAppend(bool want, std::uint32_t* counter)
	{
		const unsigned long long mask = __ballot(want);
		const std::uint32_t total = static_cast<std::uint32_t>(__popcll(mask));
		std::uint32_t base = 0;
		const int leader = __ffsll(static_cast<long long>(mask)) - 1;
		if (total > 0 && static_cast<int>(laneId()) == leader)
		{
			base = atomicAdd(counter, total);
		}
		base = __shfl(base, leader < 0 ? 0 : leader);
		const unsigned long long below = mask & ((1ull << laneId()) - 1ull);
		return base + static_cast<std::uint32_t>(__popcll(below));
	}

	__global__ void	wf_zero_counts(std::uint32_t* counts, std::uint32_t n, std::uint32_t first)
	{
		const std::uint32_t i = threadIdx.x;
		if (i < n) counts[i] = (i == 0) ? first : 0u;
	}

	__global__ __launch_bounds__(kWaveBlock) void	wf_raygen(GpuParams p, WaveBuffers w)
	{
		const std::uint32_t pixels = p.width * p.height;
		const std::uint32_t n = pixels * p.sampleCount;
		for (std::uint32_t id = blockIdx.x * kWaveBlock + threadIdx.x; id < n; id += gridDim.x * kWaveBlock)
		{
			const std::uint32_t pixel = id % pixels, slot = id / pixels;
			PixelSampler ps;
			ps.begin(p.renderSeed, pixel % p.width, pixel / p.width, p.sampleBegin + slot, 0);
			const Ray r = cameraRay(p, pixel % p.width, pixel / p.width, ps);
			w.qOrigin[0][id] = make_float4(r.o.x, r.o.y, r.o.z, 0.0f);
			w.qDir[0][id] = make_float4(r.d.x, r.d.y, r.d.z, 0.0f);
			w.qPath[0][id] = id;
			w.throughput[id] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
			w.radiance[id] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(PF_PREV_SPECULAR));
		}
	}

	// Queue slots handed to the lanes of one wave, claimed from the queue's fetch counter in
	// chunks (one atomic per chunk). All fields are wave-uniform.
	constexpr std::uint32_t	kFetchChunk = 256;

	struct	WaveFeed
	{
		std::uint32_t	next;
		std::uint32_t	end;
		bool	drained;

		// Gives every lane in `idle` (ballot mask) the next queue slot, in lane order; lanes
		// left without one (queue exhausted) get 0xffffffff.
		__device__ __forceinline__ std::uint32_t	hand(unsigned long long idle, std::uint32_t n, std::uint32_t* fetch)
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
				c = atomicAdd(fetch, kFetchChunk);
			}
			c = __builtin_amdgcn_readfirstlane(__shfl(c, __ffsll(static_cast<long long>(idle)) - 1));
			const std::uint32_t cEnd = c >= n ? c : min(c + kFetchChunk, n);
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
	// idle (refillAt = 64: the wave refills only when all its lanes are done).
	// Source supplies the queue: count(), load(slot, ray, tMax) and store(slot, state).
	template <bool AnyHit, bool Count, typename Source>
	__device__ __forceinline__ void	traceQueue(const SceneView& s, std::int32_t tlasRoot, Source& src, std::uint32_t* fetch,
		int refillAt, unsigned long long* counters, int counterSlot)
	{
		__shared__ std::uint32_t stackMem[kWaveLdsStack * kWaveBlock];
		const std::uint32_t gtid = blockIdx.x * kWaveBlock + threadIdx.x;
		SpillStack<kWaveLdsStack, kWaveBlock> stack{stackMem + threadIdx.x, src.spill + gtid, gridDim.x * kWaveBlock, 0,
			kWaveStackCapacity};
		const std::uint32_t n = src.count();
		WaveFeed feed{0u, 0u, false};
		std::uint32_t rays = 0;
		VisitCounter visits{0u, 0u};
		TraceState st;
		std::uint32_t slot = 0xffffffffu;
		bool active = false;
		while (true)
		{
			const unsigned long long idle = __ballot(!active);
			if (!feed.drained && (__popcll(idle) >= refillAt || idle == ~0ull))
			{
				const std::uint32_t k = feed.hand(idle, n, fetch);
				if (!active && k != 0xffffffffu)
				{
					slot = k;
					Ray r;
					float tMax, logU;
					src.load(slot, r, tMax, logU);
					active = traceBegin(st, s, tlasRoot, r, tMax, logU, stack);
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
			if (active && traceStep<AnyHit, Count>(s, st, kTMinF, stack, &visits))
			{
				src.store(slot, st);
				rays++;
				active = false;
			}
		}
		flushCounters(counters, counterSlot, rays, visits, Count);
	}

	// The volume free-flight sample of path `path` at bounce b (scenes with volumes only).
	__device__ __forceinline__ float	pathVolumeLogU(const GpuParams& p, std::uint32_t path, int b)
	{
		if (!p.hasVolumes) return 0.0f;
		const std::uint32_t pixels = p.width * p.height;
		const std::uint32_t pixel = path % pixels, slot = path / pixels;
		PixelSampler ps;
		ps.begin(p.renderSeed, pixel % p.width, pixel / p.width, p.sampleBegin + slot, 0);
		ps.bounce = static_cast<std::uint32_t>(b);
		return volumeLogU(p, ps);
	}

	struct	ExtendQueue
	{
		const WaveBuffers&	w;
		const GpuParams&	p;
		int	b;
		std::uint32_t*	spill;
		__device__ __forceinline__ std::uint32_t	count() const { return w.counts[b]; }
		__device__ __forceinline__ void	load(std::uint32_t i, Ray& r, float& tMax, float& logU) const
		{
			const float4 o = w.qOrigin[b & 1][i], d = w.qDir[b & 1][i];
			r = makeRay(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
			tMax = INFINITY;
			logU = p.hasVolumes ? pathVolumeLogU(p, w.qPath[b & 1][i], b) : 0.0f;
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
		const GpuParams&	p;
		int	b;
		std::uint32_t*	spill;
		std::uint32_t	target;	// connection slot of the ray being traced
		__device__ __forceinline__ std::uint32_t	count() const { return w.counts[kMaxBounceSlots + b]; }
		__device__ __forceinline__ void	load(std::uint32_t i, Ray& r, float& tMax, float& logU)
		{
			const float4 o = w.sOrigin[i], d = w.sDir[i];
			r = makeRay(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
			tMax = o.w;
			target = __float_as_uint(d.w);
			logU = p.hasVolumes ? pathVolumeLogU(p, target >> 1, b) : 0.0f;
		}
		__device__ __forceinline__ void	store(std::uint32_t, const TraceState& st) const
		{
			w.visible[target] = st.found ? 0 : 1;
		}
	};

	// closest hit of every ray of queue b
	template <bool Count>
	__global__ __launch_bounds__(kWaveBlock) void	wf_extend(SceneView s, GpuParams p, WaveBuffers w, int b,
		unsigned long long* counters, int refillAt)
	{
		ExtendQueue src{w, p, b, w.spill};
		traceQueue<false, Count>(s, p.tlasRoot, src, w.counts + 2 * kMaxBounceSlots + b, refillAt, counters, 0);
	}

	__global__ __launch_bounds__(kWaveBlock) void	wf_shade(SceneView s, GpuParams p, WaveBuffers w, int b)
	{
		const std::uint32_t n = w.counts[b];
		const int q = b & 1;
		const std::uint32_t pixels = p.width * p.height;
		std::uint32_t* nextCount = w.counts + b + 1;
		std::uint32_t* shadowCount = w.counts + kMaxBounceSlots + b;
		const bool sampledInfinite = hasEnvLight(p) || hasConstantBackgroundLight(p);
		// grid-stride loop with a wave-uniform trip count so every lane reaches the ballots
		const std::uint32_t stride = gridDim.x * kWaveBlock;
		const std::uint32_t first = blockIdx.x * kWaveBlock + (threadIdx.x & ~63u);
		for (std::uint32_t waveBase = first; waveBase < n; waveBase += stride)
		{
			const std::uint32_t i = waveBase + laneId();
			const bool valid = i < n;
			bool cont = false, wantLight = false, wantEnv = false;
			Ray next{};
			Connection cl{}, ce{};
			f3 shadePoint = mk(0, 0, 0);
			std::uint32_t path = 0;
			if (valid)
			{
				path = w.qPath[q][i];
				const std::uint32_t pixel = path % pixels, slot = path / pixels;
				PixelSampler ps;
				ps.begin(p.renderSeed, pixel % p.width, pixel / p.width, p.sampleBegin + slot, 0);
				ps.bounce = static_cast<std::uint32_t>(b);
				const float4 hv = w.qHit[i];
				const float4 o = w.qOrigin[q][i], d = w.qDir[q][i];
				const Ray ray = makeRay(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
				float4 t4 = w.throughput[path];
				float4 r4 = w.radiance[path];
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
					f3 c = thr * missRadiance(s, p, ray.d, wgt);
					if (b > 0 && hasConstantBackgroundLight(p)) c = clampConstantBackground(p, c);
					acc = acc + clampRay(c);
				}
				else
				{
					HitInfo h;
					h.t = hv.x;
					h.u = hv.y;
					h.v = hv.z;
					h.hittable = __float_as_uint(hv.w);
					h.tri = w.qTri[i];
					const Surface sf = surfaceAt(s, ray, h);
					shadePoint = sf.position;
					const GpuMaterial& m = s.materials[sf.material];
					const f3 emitted = m.kind == GPU_MAT_EMISSIVE ? ld3(m.color) : mk(0, 0, 0);
					Scatter sc;
					if (!scatter(m, sf, ray.d, ps, sc))
					{
						const float4 po = w.prevOrigin[path];
						const float lp = (!prevSpecular && p.lightCount > 0) ? lightsPdf(s, p, mk(po.x, po.y, po.z), ray.d) : 0.0f;
						const float wgt = prevSpecular ? 1.0f : powerHeuristic(prevPdf, lp);
						acc = acc + clampRay((thr * emitted) * wgt);
					}
					else
					{
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
									cl = connectLight(s, p, sf, m, sc, b, ps);
									if (cl.active)
									{
										cl.contribution = thr * cl.contribution;	// clampRay applied at resolve
										wantLight = cl.tMax > 0.0f;
										flags |= PF_CONN_LIGHT | (wantLight ? 0u : PF_CONN_FREE_LIGHT);
									}
								}
								if (sampledInfinite)
								{
									ce = connectInfinite(s, p, sf, m, sc, b, ps);
									if (ce.active)
									{
										ce.contribution = thr * ce.contribution;
										wantEnv = true;
										flags |= PF_CONN_ENV;
									}
								}
								const f3 dir = scatterDirection(sf, sc, ps);
								if (leavesOpaque(sf, m, dir))
								{
									const float pdf = scatterPdf(m, sf, sc, dir);
									if (pdf > 0.0f && isfinite(pdf))
									{
										thr = clampRay(thr * att);
										if (!terminated(thr) && russianRoulette(thr, b, ps))
										{
											next = makeRay(sf.position, dir);
											flags = (flags & (PF_CONN_LIGHT | PF_CONN_ENV | PF_CONN_FREE_LIGHT))
												| (samplesVisibleInfinite(p, sc) ? PF_PREV_INFINITE : 0u);
											t4.w = pdf;
											w.prevOrigin[path] = make_float4(sf.position.x, sf.position.y, sf.position.z, 0.0f);
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
			}
			const std::uint32_t pos = waveAppend(cont, nextCount);
			if (cont)
			{
				const int nq = (b + 1) & 1;
				w.qOrigin[nq][pos] = make_float4(next.o.x, next.o.y, next.o.z, 0.0f);
				w.qDir[nq][pos] = make_float4(next.d.x, next.d.y, next.d.z, 0.0f);
				w.qPath[nq][pos] = path;
			}
			const std::uint32_t sl = waveAppend(wantLight, shadowCount);
			if (wantLight)
			{
				w.sOrigin[sl] = make_float4(shadePoint.x, shadePoint.y, shadePoint.z, cl.tMax);
				w.sDir[sl] = make_float4(cl.direction.x, cl.direction.y, cl.direction.z, __uint_as_float(2u * path));
			}
			const std::uint32_t se = waveAppend(wantEnv, shadowCount);
			if (wantEnv)
			{
				w.sOrigin[se] = make_float4(shadePoint.x, shadePoint.y, shadePoint.z, INFINITY);
				w.sDir[se] = make_float4(ce.direction.x, ce.direction.y, ce.direction.z, __uint_as_float(2u * path + 1u));
			}
		}
	}

	// any-hit test of the shadow rays appended in bounce b
	template <bool Count>
	__global__ __launch_bounds__(kWaveBlock) void	wf_shadow(SceneView s, GpuParams p, WaveBuffers w, int b,
		unsigned long long* counters, int refillAt)
	{
		ShadowQueue src{w, p, b, w.spill, 0u};
		traceQueue<true, Count>(s, p.tlasRoot, src, w.counts + 3 * kMaxBounceSlots + b, refillAt, counters, 1);
	}

	__global__ __launch_bounds__(kWaveBlock) void	wf_resolve(GpuParams p, WaveBuffers w, int b, float* __restrict__ sampleColors)
	{
		const std::uint32_t n = w.counts[b];
		const int q = b & 1;
		for (std::uint32_t i = blockIdx.x * kWaveBlock + threadIdx.x; i < n; i += gridDim.x * kWaveBlock)
		{
			const std::uint32_t path = w.qPath[q][i];
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

	// Device buffers for one context (grown on demand, kept between renders).
	struct	WavefrontState
	{
		void*	mem = nullptr;
		size_t	bytes = 0;
		WaveBuffers	buf{};

		void	release()
		{
			if (mem) (void)hipFree(mem);
			mem = nullptr;
			bytes = 0;
		}

		void	ensure(std::uint32_t capacity)
		{
			const size_t n = capacity;
			const size_t spillBytes = static_cast<size_t>(kWaveStackCapacity - kWaveLdsStack) * kWaveGrid * kWaveBlock * 4;
			const size_t need = n * (16 * 2 * 2 + 4 * 2 + 16 + 4 + 16 * 3 + 32 + 2 + 32 * 2) + kMaxBounceSlots * 4 * 4 + 64 * 16
				+ spillBytes + 256;
			if (need > bytes)
			{
				release();
				if (hipMalloc(&mem, need) != hipSuccess)
				{
					mem = nullptr;
					throw std::runtime_error("wavefront state allocation failed");
				}
				bytes = need;
			}
			char* c = static_cast<char*>(mem);
			auto take = [&](size_t b) { char* r = c; c += (b + 255) & ~size_t(255); return r; };
			for (int k = 0; k < 2; k++)
			{
				buf.qOrigin[k] = reinterpret_cast<float4*>(take(16 * n));
				buf.qDir[k] = reinterpret_cast<float4*>(take(16 * n));
				buf.qPath[k] = reinterpret_cast<std::uint32_t*>(take(4 * n));
			}
			buf.qHit = reinterpret_cast<float4*>(take(16 * n));
			buf.qTri = reinterpret_cast<std::int32_t*>(take(4 * n));
			buf.throughput = reinterpret_cast<float4*>(take(16 * n));
			buf.radiance = reinterpret_cast<float4*>(take(16 * n));
			buf.prevOrigin = reinterpret_cast<float4*>(take(16 * n));
			buf.connection = reinterpret_cast<float4*>(take(32 * n));
			buf.visible = reinterpret_cast<std::uint8_t*>(take(2 * n));
			buf.sOrigin = reinterpret_cast<float4*>(take(32 * n));
			buf.sDir = reinterpret_cast<float4*>(take(32 * n));
			buf.counts = reinterpret_cast<std::uint32_t*>(take(kMaxBounceSlots * 4 * 4));
			buf.spill = reinterpret_cast<std::uint32_t*>(take(spillBytes));
			buf.capacity = capacity;
		}
	};

	// Launch shape of the traversal stages: persistent blocks per launch and the number of
	// idle lanes at which a wave takes new rays (1..64).
	struct	WaveTuning
	{
		unsigned	grid = 0;
		int	extendRefill = 0;
		int	shadowRefill = 0;

		WaveTuning	resolved() const
		{
			WaveTuning t = *this;
			if (t.grid == 0) t.grid = kWaveGrid;
			if (t.extendRefill == 0) t.extendRefill = kDefaultRefill;
			if (t.shadowRefill == 0) t.shadowRefill = kDefaultRefill;
			if (t.grid > kWaveGrid || t.extendRefill < 1 || t.extendRefill > 64 || t.shadowRefill < 1 || t.shadowRefill > 64)
			{
				throw std::invalid_argument("wavefront tuning out of range: grid 1.." + std::to_string(kWaveGrid)
					+ ", refill 1..64");
			}
			return t;
		}
	};

	struct	StageTimes
	{
		double	extendMs = 0.0, shadeMs = 0.0, shadowMs = 0.0;
		std::uint32_t	extendLaunches = 0;
	};

	// Enqueues one launch of the wavefront integrator. With `events`, each stage is
	// bracketed by HIP events (on this stream) and the per-stage times are returned.
	inline void	renderWavefront(WavefrontState& st, const SceneView& s, const GpuParams& p, float* sampleColors,
		unsigned long long* counters, hipStream_t stream, bool count, StageTimes* times, WaveTuning tune)
	{
		tune = tune.resolved();
		if (p.maxBounces + 2 > kMaxBounceSlots)
		{
			throw std::runtime_error("Unsupported: more than 62 bounces in the wavefront integrator");
		}
		// passes of at most ~16M paths keep the SoA state near 4 GB
		const std::uint32_t pixels = p.width * p.height;
		const std::uint32_t perPass = std::max<std::uint32_t>(1u, std::min<std::uint32_t>(p.sampleCount, (16u << 20) / pixels));
		st.ensure(pixels * perPass);
		const WaveBuffers& w = st.buf;
		std::vector<hipEvent_t> ev;
		auto mark = [&]() {
			if (!times) return;
			hipEvent_t e;
			if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
			if (hipEventRecord(e, stream) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
			ev.push_back(e);
		};
		for (std::uint32_t s0 = 0; s0 < p.sampleCount; s0 += perPass)
		{
			GpuParams q = p;
			q.sampleBegin = p.sampleBegin + s0;
			q.sampleCount = std::min(perPass, p.sampleCount - s0);
			float* colors = sampleColors + static_cast<size_t>(s0) * pixels * 3;
			wf_zero_counts<<<1, 4 * kMaxBounceSlots, 0, stream>>>(w.counts, 4 * kMaxBounceSlots, pixels * q.sampleCount);
			wf_raygen<<<kWaveGrid, kWaveBlock, 0, stream>>>(q, w);
			for (int b = 0; b <= q.maxBounces; b++)
			{
				mark();
				if (count) wf_extend<true><<<tune.grid, kWaveBlock, 0, stream>>>(s, q, w, b, counters, tune.extendRefill);
				else wf_extend<false><<<tune.grid, kWaveBlock, 0, stream>>>(s, q, w, b, counters, tune.extendRefill);
				mark();
				wf_shade<<<kWaveGrid, kWaveBlock, 0, stream>>>(s, q, w, b);
				mark();
				if (count) wf_shadow<true><<<tune.grid, kWaveBlock, 0, stream>>>(s, q, w, b, counters, tune.shadowRefill);
				else wf_shadow<false><<<tune.grid, kWaveBlock, 0, stream>>>(s, q, w, b, counters, tune.shadowRefill);
				mark();
				wf_resolve<<<kWaveGrid, kWaveBlock, 0, stream>>>(q, w, b, colors);
			}
		}
		if (times)
		{
			if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("wavefront stage timing failed");
			for (size_t k = 0; k + 3 < ev.size(); k += 4)
			{
				float a = 0, b2 = 0, c = 0;
				(void)hipEventElapsedTime(&a, ev[k], ev[k + 1]);
				(void)hipEventElapsedTime(&b2, ev[k + 1], ev[k + 2]);
				(void)hipEventElapsedTime(&c, ev[k + 2], ev[k + 3]);
				times->extendMs += a;
				times->shadeMs += b2;
				times->shadowMs += c;
				times->extendLaunches++;
			}
			for (hipEvent_t e : ev) (void)hipEventDestroy(e);
		}
	}
}
