// Two-level BVH traversal (closest hit and any hit) over the 64-byte node layout.
//
// One stack serves both levels: entering a mesh instance pushes an exit marker, then
// the mesh root; popping the marker restores the world-space ray. The instance's ray
// keeps the world parameterisation (its direction is the affine image of the world
// direction, not renormalised), so a hit distance needs no rescaling between levels.
//
// Semantics follow the reference's intersection routines:
//   slab test from T_MIN, zero direction components tested against the slab,
//   tMax <= tMin -> miss                          BVHNode.cpp:20-61, Mesh.cpp:99-140
//   Möller–Trumbore, |det| < 1e-12, inclusive edges, t in [tMin, tMax]
//                                                  Triangle.cpp:196-240,272-305
//   rectangle plane test and half-extent check      Rectangle.cpp:212-293
// A later equal-distance hit replaces an earlier one, as with the reference's
// `t > t_max` rejection.
#pragma once

#include "vec.hpp"
#include "../common/layout.h"

namespace luzgpu
{
	constexpr float	kTMinF = 0.001f;

	struct	SceneView
	{
		const Bvh4Node*	tlas;
		const Bvh4Node*	blas;
		const GpuTri*	topTris;
		const GpuTri*	meshTris;
		const float*	meshNormals;
		const std::uint32_t*	tlasPrims;
		const GpuHittable*	hittables;
		const GpuRect*	rects;
		const GpuInstance*	instances;
		const GpuMaterial*	materials;
		const GpuLight*	lights;
		const double*	lightCdf;
		const float4*	envPixels;
		const float*	envPdf;
		const double*	envCdf;
		const std::uint32_t*	unbounded;
	};

	struct	Ray
	{
		f3	o, d, inv;
	};

	__device__ __forceinline__ float	safeInv(float c)
	{
		// Ray::_inverseComponent: 0 for zero or non-finite components
		if (c == 0.0f || !isfinite(c))
		{
			return 0.0f;
		}
		const float r = 1.0f / c;
		return isfinite(r) ? r : 0.0f;
	}

	__device__ __forceinline__ Ray	makeRay(f3 o, f3 d)
	{
		return Ray{o, d, mk(safeInv(d.x), safeInv(d.y), safeInv(d.z))};
	}

	// Slab entry distance for one axis; zero inverse components test the origin against
	// the slab instead (lo/hi of +/-inf keep or reject the interval).
	__device__ __forceinline__ void	slab(float o, float inv, float bmin, float bmax, float& tn, float& tf)
	{
		if (inv == 0.0f)
		{
			const bool inside = !(o < bmin || o > bmax);
			tn = inside ? -INFINITY : INFINITY;
			tf = inside ? INFINITY : -INFINITY;
			return;
		}
		float t0 = (bmin - o) * inv;
		float t1 = (bmax - o) * inv;
		tn = fminf(t0, t1);
		tf = fmaxf(t0, t1);
	}

	__device__ __forceinline__ bool	boxHit(const Ray& r, const float* lo, const float* hi, float tMax, float& tNear)
	{
		float n0, f0, n1, f1, n2, f2;
		slab(r.o.x, r.inv.x, lo[0], hi[0], n0, f0);
		slab(r.o.y, r.inv.y, lo[1], hi[1], n1, f1);
		slab(r.o.z, r.inv.z, lo[2], hi[2], n2, f2);
		const float tn = fmaxf(fmaxf(n0, n1), fmaxf(n2, kTMinF));
		const float tf = fminf(fminf(f0, f1), fminf(f2, tMax));
		tNear = tn;
		return tf > tn;
	}

	__device__ __forceinline__ bool	triHit(const GpuTri& tri, const Ray& r, float tMin, float tMax, float& t, float& u, float& v)
	{
		const f3 e1 = ld3(tri.e1), e2 = ld3(tri.e2);
		const f3 p = cross(r.d, e2);
		const float det = dot(e1, p);
		if (fabsf(det) < 1e-12f)
		{
			return false;
		}
		const float inv = 1.0f / det;
		const f3 tv = r.o - ld3(tri.v0);
		u = dot(tv, p) * inv;
		if (u < 0.0f || u > 1.0f)
		{
			return false;
		}
		const f3 q = cross(tv, e1);
		v = dot(r.d, q) * inv;
		if (v < 0.0f || u + v > 1.0f)
		{
			return false;
		}
		t = dot(e2, q) * inv;
		return !(t > tMax || t < tMin);
	}

	__device__ __forceinline__ bool	rectHit(const GpuRect& rc, const Ray& r, float tMin, float tMax, float& t)
	{
		const f3 n = ld3(rc.normal);
		const f3 pos = ld3(rc.position);
		const float a = dot(r.o - pos, n);
		const float b = dot(r.d, n);
		if (b == 0.0f || (a < 0.0f && b < 0.0f) || (a > 0.0f && b > 0.0f))
		{
			return false;
		}
		t = -a / b;
		if (t > tMax || t < tMin)
		{
			return false;
		}
		const f3 d = (r.o + r.d * t) - pos;
		return !(fabsf(dot(d, ld3(rc.widthAxis))) > rc.halfWidth || fabsf(dot(d, ld3(rc.heightAxis))) > rc.halfHeight);
	}

	__device__ __forceinline__ Ray	toLocal(const GpuInstance& in, const Ray& w)
	{
		const float* m = in.worldToLocal;
		const f3 o = mk(m[0] * w.o.x + m[1] * w.o.y + m[2] * w.o.z + m[3],
			m[4] * w.o.x + m[5] * w.o.y + m[6] * w.o.z + m[7],
			m[8] * w.o.x + m[9] * w.o.y + m[10] * w.o.z + m[11]);
		const f3 d = mk(m[0] * w.d.x + m[1] * w.d.y + m[2] * w.d.z,
			m[4] * w.d.x + m[5] * w.d.y + m[6] * w.d.z,
			m[8] * w.d.x + m[9] * w.d.y + m[10] * w.d.z);
		return makeRay(o, d);
	}

	// Stack entries: bit31 set = leaf (bits 27..30 = count-1, bits 0..26 = first prim);
	// otherwise a node index; kExitInstance returns to the world-space ray.
	constexpr std::uint32_t	kLeafBit = 0x80000000u;
	constexpr std::uint32_t	kExitInstance = 0x7fffffffu;

	__device__ __forceinline__ std::uint32_t	leafEntry(std::int32_t start, std::int32_t count)
	{
		return kLeafBit | (static_cast<std::uint32_t>(count - 1) << 27) | static_cast<std::uint32_t>(start);
	}

	struct	HitInfo
	{
		float	t;
		float	u, v;
		std::uint32_t	hittable;	// top-level hittable index, 0xffffffff: none
		std::int32_t	tri;		// mesh triangle slot (global) for instances, -1 otherwise
	};

	// Per-thread traversal stack in LDS, laid out [depth][thread] so the 64 lanes of a
	// wave touch 64 consecutive banks.
	template <int Depth, int Block>
	struct	LdsStack
	{
		std::uint32_t*	base;
		int	sp;

		// The host refuses scenes whose BVH depth could exceed Depth; the clamp only keeps
		// every LDS access inside this thread's column whatever happens.
		__device__ __forceinline__ void	push(std::uint32_t v) { base[min(sp, Depth - 1) * Block] = v; sp++; }
		__device__ __forceinline__ std::uint32_t	pop() { sp--; return base[min(sp, Depth - 1) * Block]; }
		__device__ __forceinline__ bool	empty() const { return sp == 0; }
		__device__ __forceinline__ void	reset() { sp = 0; }
	};

	// Node / primitive visit counts (only maintained when Count is true).
	struct	VisitCounter
	{
		std::uint32_t	nodes;
		std::uint32_t	prims;
	};

	// LDS stack of LdsDepth entries per thread, continued in a per-thread column of global
	// memory (`spill`, stride = threads in the grid) for deeper entries. Keeping the LDS
	// part short lets 8 waves per SIMD share a CU's LDS.
	template <int LdsDepth, int Block>
	struct	SpillStack
	{
		std::uint32_t*	base;
		std::uint32_t*	spill;
		std::uint32_t	stride;
		int	sp;
		int	capacity;

		__device__ __forceinline__ void	push(std::uint32_t v)
		{
			if (sp < LdsDepth) base[sp * Block] = v;
			else spill[static_cast<size_t>(min(sp, capacity - 1) - LdsDepth) * stride] = v;
			sp++;
		}
		__device__ __forceinline__ std::uint32_t	pop()
		{
			sp--;
			if (sp < LdsDepth) return base[sp * Block];
			return spill[static_cast<size_t>(min(sp, capacity - 1) - LdsDepth) * stride];
		}
		__device__ __forceinline__ bool	empty() const { return sp == 0; }
		__device__ __forceinline__ void	reset() { sp = 0; }
	};

	template <bool AnyHit, bool Count = false, typename Stack>
	__device__ __forceinline__ bool	traverse(const SceneView& s, std::int32_t tlasRoot, const Ray& worldRay, float tMin, float tMax,
		HitInfo& hit, Stack& stack, VisitCounter* counter = nullptr)
	{
		hit.t = tMax;
		hit.hittable = 0xffffffffu;
		hit.tri = -1;
		// an any-hit query returns with entries still on the stack: every traversal
		// starts from an empty stack
		stack.reset();
		if (tlasRoot < 0)
		{
			return false;
		}
		Ray ray = worldRay;
		bool inInstance = false;
		std::uint32_t instHittable = 0;
		const Bvh4Node* nodes = s.tlas;
		float closest = tMax;
		bool found = false;
		std::uint32_t entry = static_cast<std::uint32_t>(tlasRoot);
		while (true)
		{
			if (entry == kExitInstance)
			{
				ray = worldRay;
				inInstance = false;
				nodes = s.tlas;
			}
			else if ((entry & kLeafBit) == 0)
			{
				const Bvh4Node& n = nodes[entry];
				if (Count) counter->nodes++;
				// decode the four child boxes (plane = fma(q, 2^exp, origin)) and test them
				float sc[3];
				for (int a = 0; a < 3; a++) sc[a] = __int_as_float((static_cast<int>(n.exp[a]) + 127) << 23);
				std::uint32_t e[4];
				float tn[4];
				int hits = 0;
				for (int k = 0; k < 4; k++)
				{
					if (n.count[k] == 255) continue;
					float lo[3], hi[3];
					for (int a = 0; a < 3; a++)
					{
						lo[a] = fmaf(static_cast<float>(n.qlo[a][k]), sc[a], n.origin[a]);
						hi[a] = fmaf(static_cast<float>(n.qhi[a][k]), sc[a], n.origin[a]);
					}
					float t;
					if (boxHit(ray, lo, hi, closest, t))
					{
						tn[hits] = t;
						e[hits] = n.count[k] > 0 ? leafEntry(n.child[k], n.count[k]) : static_cast<std::uint32_t>(n.child[k]);
						hits++;
					}
				}
				if (hits > 0)
				{
					// order by entry distance (insertion sort of <= 4), push far to near
					for (int i = 1; i < hits; i++)
					{
						const float kt = tn[i];
						const std::uint32_t ke = e[i];
						int j = i - 1;
						while (j >= 0 && tn[j] > kt)
						{
							tn[j + 1] = tn[j];
							e[j + 1] = e[j];
							j--;
						}
						tn[j + 1] = kt;
						e[j + 1] = ke;
					}
					for (int i = hits - 1; i > 0; i--) stack.push(e[i]);
					entry = e[0];
					continue;
				}
			}
			else
			{
				const std::int32_t start = static_cast<std::int32_t>(entry & 0x07ffffffu);
				const std::int32_t count = static_cast<std::int32_t>((entry >> 27) & 0xfu) + 1;
				bool descended = false;
				for (std::int32_t i = 0; i < count; i++)
				{
					if (Count) counter->prims++;
					if (inInstance)
					{
						const GpuTri& tri = s.meshTris[start + i];
						float t, u, v;
						if (triHit(tri, ray, tMin, closest, t, u, v))
						{
							found = true;
							closest = t;
							hit.t = t; hit.u = u; hit.v = v;
							hit.hittable = instHittable;
							hit.tri = start + i;
							if (AnyHit) return true;
						}
						continue;
					}
					const std::uint32_t h = s.tlasPrims[start + i];
					const GpuHittable gh = s.hittables[h];
					const std::uint32_t kind = gh.kindIndex >> 30;
					const std::uint32_t idx = gh.kindIndex & 0x3fffffffu;
					if (kind == GPU_HIT_TRIANGLE)
					{
						float t, u, v;
						if (triHit(s.topTris[idx], ray, tMin, closest, t, u, v))
						{
							found = true;
							closest = t;
							hit.t = t; hit.u = u; hit.v = v;
							hit.hittable = h;
							hit.tri = -1;
							if (AnyHit) return true;
						}
					}
					else if (kind == GPU_HIT_RECTANGLE)
					{
						float t;
						if (rectHit(s.rects[idx], ray, tMin, closest, t))
						{
							found = true;
							closest = t;
							hit.t = t;
							hit.hittable = h;
							hit.tri = -1;
							if (AnyHit) return true;
						}
					}
					else
					{
						const GpuInstance& in = s.instances[idx];
						if (i + 1 < count)
						{
							stack.push(leafEntry(start + i + 1, count - i - 1));
						}
						stack.push(kExitInstance);
						ray = toLocal(in, worldRay);
						inInstance = true;
						instHittable = h;
						nodes = s.blas;
						entry = static_cast<std::uint32_t>(in.blasRoot);
						descended = true;
						break;
					}
				}
				if (descended)
				{
					continue;
				}
			}
			if (stack.empty())
			{
				break;
			}
			entry = stack.pop();
		}
		return found;
	}
}
