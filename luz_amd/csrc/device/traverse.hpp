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
		const uint2*	tlasLeaf;	// top-level leaf items: (hittable index, kind << 28 | index)
		const GpuHittable*	hittables;
		const GpuRect*	rects;
		const GpuInstance*	instances;
		const GpuMaterial*	materials;
		const GpuLight*	lights;
		const double*	lightCdf;
		const double*	lightTriCdf;		// emissive mesh instances: triangle-area prefix sums
		const std::uint32_t*	lightTriSlot;	// ... and the mesh-triangle slot of each entry
		const float4*	envPixels;
		const float*	envPdf;
		const double*	envCdf;
		const std::uint32_t*	envGuide;	// texel-aligned with envCdf
		const double*	iesData;	// IES profiles of sphere lights (GpuLight::ies)
		const std::uint32_t*	unbounded;	// hittable ids of the planes, tested after the BVH
		const GpuSphere*	spheres;
		const GpuRect*	cubeFaces;
		const GpuVolume*	volumes;
		const GpuPlane*	planes;
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
		// hardware reciprocal (1 ulp): the inverse only feeds box tests, whose intervals are
		// widened by far more than that
		const float r = __builtin_amdgcn_rcpf(c);
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
		// widened by 2^-20 relative like the quantised-plane test (the inverse is 1 ulp)
		const float tn = fmaxf(fmaxf(n0, n1) * (1.0f - 0x1p-20f), fmaxf(n2 * (1.0f - 0x1p-20f), kTMinF));
		const float tf = fminf(fminf(f0, f1) * (1.0f + 0x1p-20f), fminf(f2 * (1.0f + 0x1p-20f), tMax));
		tNear = tn;
		return tf > tn;
	}

	// The 48-byte triangle is fetched as three 16-byte loads issued together (one memory
	// round trip), not field by field.
	__device__ __forceinline__ bool	triHit(const GpuTri& tri, const Ray& r, float tMin, float tMax, float& t, float& u, float& v)
	{
		const float4* tp = reinterpret_cast<const float4*>(&tri);
		const float4 A = tp[0], B = tp[1], C = tp[2];
		f3 v0 = mk(A.x, A.y, A.z);
		const f3 e1 = mk(B.x, B.y, B.z), e2 = mk(C.x, C.y, C.z);
		// v0 is needed only after the first exit: pin it here so its load is issued with
		// the edges' instead of after the determinant test
		asm volatile("" : "+v"(v0.x), "+v"(v0.y), "+v"(v0.z));
		const f3 p = cross(r.d, e2);
		const float det = dot(e1, p);
		const f3 tv = r.o - v0;
		if (fabsf(det) < 1e-12f)
		{
			return false;
		}
		const float inv = __builtin_amdgcn_rcpf(det);	// 1 ulp, against the FP64 reference's 1 / det
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

	__device__ __forceinline__ bool	zeroDir(const Ray& r)
	{
		return r.d.x == 0.0f && r.d.y == 0.0f && r.d.z == 0.0f;
	}

	// Sphere::hit / hitAny (Sphere.cpp:226-301), in FP64 from the fp32 ray: the reference
	// solves the quadratic for a unit direction (a = 1).
	__device__ __forceinline__ bool	sphereHit(const GpuSphere& sp, const Ray& r, float tMin, float tMax, float& t)
	{
		if (zeroDir(r)) return false;
		const double ox = static_cast<double>(r.o.x) - sp.center[0], oy = static_cast<double>(r.o.y) - sp.center[1],
			oz = static_cast<double>(r.o.z) - sp.center[2];
		const double dx = r.d.x, dy = r.d.y, dz = r.d.z;
		const double b = ox * dx + oy * dy + oz * dz;
		const double rr = static_cast<double>(sp.radius) * sp.radius;
		const double c = (ox * ox + oy * oy + oz * oz) - rr;
		const double disc = b * b - c;
		if (disc < 0.0) return false;
		const double sq = sqrt(disc);
		double root = -b - sq;
		if (root < tMin || root > tMax)
		{
			root = -b + sq;
			if (root < tMin || root > tMax) return false;
		}
		t = static_cast<float>(root);
		return true;
	}

	// ConstantVolume::sampleScatteringDistance (ConstantVolume.cpp:62-101) with the boundary's
	// hitInterval over (-T_MAX, T_MAX) (Sphere.cpp:303-339, Cube.cpp:120-164). `logU` is
	// log(max(u, 1e-12)) of the path's DIM_VOLUME_DISTANCE sample at this bounce.
	__device__ __forceinline__ bool	volumeHit(const GpuVolume& v, const Ray& r, float tMin, float tMax, float logU, float& t)
	{
		if (zeroDir(r)) return false;
		double entry, exit;
		const double o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
		if (v.sphere)
		{
			const double ox = o[0] - v.center[0], oy = o[1] - v.center[1], oz = o[2] - v.center[2];
			const double b = ox * d[0] + oy * d[1] + oz * d[2];
			const double c = (ox * ox + oy * oy + oz * oz) - static_cast<double>(v.radius) * v.radius;
			const double disc = b * b - c;
			if (disc < 0.0) return false;
			const double sq = sqrt(disc);
			entry = fmin(-b - sq, -b + sq);
			exit = fmax(-b - sq, -b + sq);
			if (entry >= exit) return false;
		}
		else
		{
			entry = -1.7976931348623157e308;
			exit = 1.7976931348623157e308;
			const float lo[3] = {v.lo[0], v.lo[1], v.lo[2]}, hi[3] = {v.hi[0], v.hi[1], v.hi[2]};
			for (int a = 0; a < 3; a++)
			{
				if (d[a] == 0.0)
				{
					if (o[a] < lo[a] || o[a] > hi[a]) return false;
					continue;
				}
				const double inv = 1.0 / d[a];
				double t0 = (lo[a] - o[a]) * inv, t1 = (hi[a] - o[a]) * inv;
				if (t0 > t1)
				{
					const double x = t0; t0 = t1; t1 = x;
				}
				entry = fmax(entry, t0);
				exit = fmin(exit, t1);
				if (entry >= exit) return false;
			}
		}
		if (entry < tMin) entry = tMin;
		if (exit > tMax) exit = tMax;
		if (entry >= exit) return false;
		if (entry < 0.0) entry = 0.0;
		const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
		if (!(len > 0.0) || !isfinite(len)) return false;
		const double inside = (exit - entry) * len;
		const double dist = static_cast<double>(v.negInvDensity) * logU;
		if (dist > inside) return false;
		t = static_cast<float>(entry + dist / len);
		return true;
	}

	// Plane::hit / hitAny (Plane.cpp:51-87)
	__device__ __forceinline__ bool	planeHit(const GpuPlane& pl, const Ray& r, float tMin, float tMax, float& t)
	{
		const f3 o = ld3(pl.orientation);
		const float d = dot(r.d, o);
		if (d == 0.0f) return false;
		t = dot(ld3(pl.point) - r.o, o) / d;
		return !(t > tMax || t < tMin);
	}

	__device__ __forceinline__ Ray	toLocal(const float4 r0, const float4 r1, const float4 r2, const Ray& w)
	{
		const float m[12] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
		const f3 o = mk(m[0] * w.o.x + m[1] * w.o.y + m[2] * w.o.z + m[3],
			m[4] * w.o.x + m[5] * w.o.y + m[6] * w.o.z + m[7],
			m[8] * w.o.x + m[9] * w.o.y + m[10] * w.o.z + m[11]);
		const f3 d = mk(m[0] * w.d.x + m[1] * w.d.y + m[2] * w.d.z,
			m[4] * w.d.x + m[5] * w.d.y + m[6] * w.d.z,
			m[8] * w.d.x + m[9] * w.d.y + m[10] * w.d.z);
		return makeRay(o, d);
	}

	__device__ __forceinline__ Ray	toLocal(const GpuInstance& in, const Ray& w)
	{
		const float4* mp = reinterpret_cast<const float4*>(in.worldToLocal);
		return toLocal(mp[0], mp[1], mp[2], w);
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

	// The primitive a ray leaves from. A planar one (triangle, rectangle, cube face, plane) is
	// left out of the ray's traversal: a ray cannot meet the plane it starts on again, and
	// the FP64 reference never does, but an fp32 hit position may sit a little behind that
	// plane, so a grazing ray would meet it again beyond T_MIN. Curved and volume hits, and
	// camera rays, exclude nothing (hittable 0xffffffff).
	struct	RayOrigin
	{
		std::uint32_t	hittable;
		std::int32_t	tri;		// instance: mesh triangle slot; cube: face; else -1
	};
	constexpr RayOrigin	kNoOrigin{0xffffffffu, -1};

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

	// A private traversal stack (scratch memory) for the rare traversals made outside the
	// traversal kernels, such as re-intersecting one emissive mesh instance for its light pdf.
	// Scenes are admitted only when a BLAS traversal fits (their bound is checked on load).
	constexpr int	kLocalStackDepth = 48;
	struct	LocalStack
	{
		std::uint32_t	v[kLocalStackDepth];
		int	sp;

		__device__ __forceinline__ void	push(std::uint32_t x) { v[min(sp, kLocalStackDepth - 1)] = x; sp++; }
		__device__ __forceinline__ std::uint32_t	pop() { sp--; return v[min(sp, kLocalStackDepth - 1)]; }
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

	// Resumable per-lane traversal state: one call of traceStep processes one stack entry
	// (a node, a leaf or an instance exit), so a persistent kernel can hand a lane a new ray
	// as soon as its traversal ends instead of waiting for the whole wave.
	struct	TraceState
	{
		Ray	world;		// world-space ray, restored when an instance is left
		Ray	ray;		// ray of the current level
		HitInfo	hit;	// hit.t is the current closest distance
		const Bvh4Node*	nodes;	// node array of the current level
		float	logU;	// log(max(u, 1e-12)) of the volume free-flight sample (scenes with volumes)
		std::uint32_t	unbounded;	// this ray's scene: unbounded list base << 16 | count
		std::uint32_t	entry;
		std::uint32_t	instHittable;
		RayOrigin	from;	// planar primitive the ray leaves from
		bool	inInstance;
		bool	found;
	};

	// Starts a traversal; returns false when there is nothing to traverse.
	template <typename Stack>
	__device__ __forceinline__ bool	traceBegin(TraceState& st, const SceneView& s, std::int32_t tlasRoot, std::uint32_t unbounded,
		const Ray& worldRay, float tMax, float logU, RayOrigin from, Stack& stack)
	{
		const Bvh4Node* tlas = s.tlas;
		st.logU = logU;
		st.from = from;
		st.unbounded = unbounded;
		st.world = worldRay;
		st.ray = worldRay;
		st.hit.t = tMax;
		st.hit.hittable = 0xffffffffu;
		st.hit.tri = -1;
		// without bounded hittables only the unbounded ones are tested: start from the
		// instance-exit entry, which leads straight to the end of the traversal
		st.entry = tlasRoot >= 0 ? static_cast<std::uint32_t>(tlasRoot) : kExitInstance;
		st.nodes = tlas;
		st.instHittable = 0;
		st.inInstance = false;
		st.found = false;
		// an any-hit query ends with entries still on the stack: every traversal starts
		// from an empty stack
		stack.reset();
		return tlasRoot >= 0 || (unbounded & 0xffffu) != 0u;
	}

	// Processes the current entry; returns true when the traversal has ended (st.found and
	// st.hit hold the result).
	template <bool AnyHit, bool Count, int Feat, typename Stack>
	__device__ __forceinline__ bool	traceStep(const SceneView& s, TraceState& st, float tMin, Stack& stack,
		VisitCounter* counter)
	{
		const std::uint32_t entry = st.entry;
		if (entry == kExitInstance)
		{
			st.ray = st.world;
			st.inInstance = false;
			st.nodes = s.tlas;
		}
		else if ((entry & kLeafBit) == 0)
		{
			// the whole 64-byte node in four 16-byte loads issued together:
			//   n0 = origin xyz | exp[3]      n1 = qlo[axis][child] x3 | qhi[0][child]
			//   n2 = qhi[1], qhi[2] | child 0, 1      n3 = child 2, 3 | count[4] | pad
			const uint4* np = reinterpret_cast<const uint4*>(st.nodes + entry);
			const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
			if (Count) counter->nodes++;
			// decode the four child boxes (plane = fma(q, 2^exp, origin)) and test them
			const float org[3] = {__uint_as_float(n0.x), __uint_as_float(n0.y), __uint_as_float(n0.z)};
			int ex[3];
			for (int a = 0; a < 3; a++) ex[a] = static_cast<int>(static_cast<std::int8_t>((n0.w >> (8 * a)) & 0xffu));
			const std::uint32_t qlo[3] = {n1.x, n1.y, n1.z};
			const std::uint32_t qhi[3] = {n1.w, n2.x, n2.y};
			const std::uint32_t child[4] = {n2.z, n2.w, n3.x, n3.y};
			// per child: the entry distance as an ordering key (hit distances are >= T_MIN > 0,
			// so their bit patterns order like the floats; a miss sorts last) and the stack entry
			constexpr std::uint32_t kMissKey = 0xffffffffu;
			std::uint32_t key[4], e[4];
			const f3 inv = st.ray.inv;
			if (inv.x != 0.0f && inv.y != 0.0f && inv.z != 0.0f)
			{
				// Slab distances straight from the quantised planes: with s = 2^exp / d and
				// b = (origin - o) / d per axis, a plane at q is at t = fma(q, s, b). Rounding
				// differs from decoding the plane first by a few ulps, so the interval is
				// widened by 2^-20 relative (the quantised boxes already enclose the FP64 ones).
				// 2^exp is a normal power of two (exp in [-100, 127]), so ldexp(1 / d, exp) is
				// exactly 2^exp * (1 / d).
				const float s0 = __builtin_amdgcn_ldexpf(inv.x, ex[0]), s1 = __builtin_amdgcn_ldexpf(inv.y, ex[1]),
					s2 = __builtin_amdgcn_ldexpf(inv.z, ex[2]);
				const float b0 = (org[0] - st.ray.o.x) * inv.x, b1 = (org[1] - st.ray.o.y) * inv.y, b2 = (org[2] - st.ray.o.z) * inv.z;
				// fma(q, s, b) is monotone in q (increasing for s >= 0), so on each axis the
				// q_lo plane is the near one when the direction is positive and the q_hi plane
				// otherwise: choosing the words once per node gives the min / max of each pair
				const std::uint32_t near0 = inv.x > 0.0f ? qlo[0] : qhi[0], far0 = inv.x > 0.0f ? qhi[0] : qlo[0];
				const std::uint32_t near1 = inv.y > 0.0f ? qlo[1] : qhi[1], far1 = inv.y > 0.0f ? qhi[1] : qlo[1];
				const std::uint32_t near2 = inv.z > 0.0f ? qlo[2] : qhi[2], far2 = inv.z > 0.0f ? qhi[2] : qlo[2];
				#pragma unroll
				for (int k = 0; k < 4; k++)
				{
					const std::uint32_t cnt = (n3.z >> (8 * k)) & 0xffu;
					const float l0 = fmaf(static_cast<float>((near0 >> (8 * k)) & 0xffu), s0, b0);
					const float h0 = fmaf(static_cast<float>((far0 >> (8 * k)) & 0xffu), s0, b0);
					const float l1 = fmaf(static_cast<float>((near1 >> (8 * k)) & 0xffu), s1, b1);
					const float h1 = fmaf(static_cast<float>((far1 >> (8 * k)) & 0xffu), s1, b1);
					const float l2 = fmaf(static_cast<float>((near2 >> (8 * k)) & 0xffu), s2, b2);
					const float h2 = fmaf(static_cast<float>((far2 >> (8 * k)) & 0xffu), s2, b2);
					const float near = fmaxf(fmaxf(l0, l1), l2) * (1.0f - 0x1p-20f);
					const float far = fminf(fminf(h0, h1), h2) * (1.0f + 0x1p-20f);
					const float t = fmaxf(near, kTMinF);
					key[k] = cnt != 255u && fminf(far, st.hit.t) > t ? __float_as_uint(t) : kMissKey;
					e[k] = cnt > 0 ? leafEntry(static_cast<std::int32_t>(child[k]), static_cast<std::int32_t>(cnt)) : child[k];
				}
			}
			else
			{
				// an axis-parallel ray: AABB::hit tests the origin against that slab (AABB.cpp:37-80)
				float sc[3];
				for (int a = 0; a < 3; a++) sc[a] = __int_as_float((ex[a] + 127) << 23);
				#pragma unroll
				for (int k = 0; k < 4; k++)
				{
					const std::uint32_t cnt = (n3.z >> (8 * k)) & 0xffu;
					float lo[3], hi[3];
					for (int a = 0; a < 3; a++)
					{
						lo[a] = fmaf(static_cast<float>((qlo[a] >> (8 * k)) & 0xffu), sc[a], org[a]);
						hi[a] = fmaf(static_cast<float>((qhi[a] >> (8 * k)) & 0xffu), sc[a], org[a]);
					}
					float t;
					key[k] = cnt != 255u && boxHit(st.ray, lo, hi, st.hit.t, t) ? __float_as_uint(t) : kMissKey;
					e[k] = cnt > 0 ? leafEntry(static_cast<std::int32_t>(child[k]), static_cast<std::int32_t>(cnt)) : child[k];
				}
			}
			// the hit children by entry distance, in child order among equal distances (a
			// compare-exchange network of adjacent pairs is a stable sort), pushed far to near,
			// the nearest taken next. An any-hit query ends at its first hit whatever the order
			// (BVHNode::hitAny is unordered, BVHNode.cpp:326-380); near first finds occluders sooner.
			auto exchange = [&](int i) {
				const bool swap = key[i] > key[i + 1];
				const std::uint32_t ki = key[i], ei = e[i];
				key[i] = swap ? key[i + 1] : ki;
				e[i] = swap ? e[i + 1] : ei;
				key[i + 1] = swap ? ki : key[i + 1];
				e[i + 1] = swap ? ei : e[i + 1];
			};
			exchange(0); exchange(1); exchange(0); exchange(2); exchange(1); exchange(0);
			if (key[0] != kMissKey)
			{
				if (key[3] != kMissKey) stack.push(e[3]);
				if (key[2] != kMissKey) stack.push(e[2]);
				if (key[1] != kMissKey) stack.push(e[1]);
				st.entry = e[0];
				return false;
			}
		}
		else
		{
			const std::int32_t start = static_cast<std::int32_t>(entry & 0x07ffffffu);
			const std::int32_t count = static_cast<std::int32_t>((entry >> 27) & 0xfu) + 1;
			for (std::int32_t i = 0; i < count; i++)
			{
				if (Count) counter->prims++;
				if (st.inInstance)
				{
					if (start + i == st.from.tri && st.instHittable == st.from.hittable) continue;
					const GpuTri& tri = s.meshTris[start + i];
					float t, u, v;
					if (triHit(tri, st.ray, tMin, st.hit.t, t, u, v))
					{
						st.found = true;
						st.hit.t = t; st.hit.u = u; st.hit.v = v;
						st.hit.hittable = st.instHittable;
						st.hit.tri = start + i;
						if (AnyHit) return true;
					}
					continue;
				}
				const uint2 item = s.tlasLeaf[start + i];
				const std::uint32_t h = item.x;
				const std::uint32_t kind = item.y >> kKindShift;
				const std::uint32_t idx = item.y & kIndexMask;
				if (kind == GPU_HIT_TRIANGLE)
				{
					float t, u, v;
					if (h != st.from.hittable && triHit(s.topTris[idx], st.ray, tMin, st.hit.t, t, u, v))
					{
						st.found = true;
						st.hit.t = t; st.hit.u = u; st.hit.v = v;
						st.hit.hittable = h;
						st.hit.tri = -1;
						if (AnyHit) return true;
					}
				}
				else if (kind == GPU_HIT_RECTANGLE)
				{
					float t;
					if (h != st.from.hittable && rectHit(s.rects[idx], st.ray, tMin, st.hit.t, t))
					{
						st.found = true;
						st.hit.t = t;
						st.hit.hittable = h;
						st.hit.tri = -1;
						if (AnyHit) return true;
					}
				}
				else if ((Feat & kFeatAnalytic) && kind == GPU_HIT_SPHERE)
				{
					float t;
					if (sphereHit(s.spheres[idx], st.ray, tMin, st.hit.t, t))
					{
						st.found = true;
						st.hit.t = t;
						st.hit.hittable = h;
						st.hit.tri = -1;
						if (AnyHit) return true;
					}
				}
				else if ((Feat & kFeatAnalytic) && kind == GPU_HIT_CUBE)
				{
					// Cube::hit (Cube.cpp:97-111): its six faces, closest first-found
					for (int f = 0; f < 6; f++)
					{
						float t;
						if ((h != st.from.hittable || f != st.from.tri) && rectHit(s.cubeFaces[6 * idx + f], st.ray, tMin, st.hit.t, t))
						{
							st.found = true;
							st.hit.t = t;
							st.hit.hittable = h;
							st.hit.tri = f;
							if (AnyHit) return true;
						}
					}
				}
				else if ((Feat & kFeatAnalytic) && kind == GPU_HIT_VOLUME)
				{
					float t;
					if (volumeHit(s.volumes[idx], st.ray, tMin, st.hit.t, st.logU, t))
					{
						st.found = true;
						st.hit.t = t;
						st.hit.hittable = h;
						st.hit.tri = -1;
						if (AnyHit) return true;
					}
				}
				else if ((Feat & kFeatAnalytic) && kind == GPU_HIT_HIDDEN_SPHERE)
				{
					// an invisible sphere (Sphere.cpp:227-233, 271-277): never hit
				}
				else
				{
					// 3x4 map and mesh links: one group of four 16-byte loads
					const float4* ip = reinterpret_cast<const float4*>(s.instances + idx);
					const float4 r0 = ip[0], r1 = ip[1], r2 = ip[2];
					const int4 links = reinterpret_cast<const int4*>(ip)[3];
					if (i + 1 < count)
					{
						stack.push(leafEntry(start + i + 1, count - i - 1));
					}
					stack.push(kExitInstance);
					st.ray = toLocal(r0, r1, r2, st.world);
					st.inInstance = true;
					st.nodes = s.blas;
					st.instHittable = h;
					st.entry = static_cast<std::uint32_t>(links.x);
					return false;
				}
			}
		}
		if (stack.empty())
		{
			// the unbounded hittables after the acceleration structure (HitHelper.cpp:5-32)
			const std::uint32_t ubBase = st.unbounded >> 16, ubCount = (Feat & kFeatAnalytic) ? (st.unbounded & 0xffffu) : 0u;
			for (std::uint32_t i = 0; i < ubCount; i++)
			{
				const std::uint32_t h = s.unbounded[ubBase + i];
				const GpuPlane& pl = s.planes[s.hittables[h].kindIndex & kIndexMask];
				float t;
				if (h != st.from.hittable && planeHit(pl, st.world, tMin, st.hit.t, t))
				{
					st.found = true;
					st.hit.t = t;
					st.hit.hittable = h;
					st.hit.tri = -1;
					if (AnyHit) return true;
				}
			}
			return true;
		}
		st.entry = stack.pop();
		return false;
	}

	__device__ __forceinline__ std::uint32_t	unboundedOf(const GpuParams& p)
	{
		return (p.unboundedBase << 16) | p.unboundedCount;
	}

	// Closest hit of a world ray against one mesh instance alone (MeshInstance::hit,
	// MeshInstance.cpp:303-327): its BLAS traversed in local space, t shared with the world
	// ray (the local direction is not renormalised).
	__device__ inline bool	instanceClosest(const SceneView& s, std::uint32_t instance, std::uint32_t hittable, const Ray& world,
		float tMax, HitInfo& hit)
	{
		const GpuInstance& in = s.instances[instance];
		TraceState st;
		st.world = world;
		st.ray = toLocal(in, world);
		st.hit.t = tMax;
		st.hit.hittable = 0xffffffffu;
		st.hit.tri = -1;
		st.nodes = s.blas;
		st.logU = 0.0f;
		st.unbounded = 0u;
		st.entry = static_cast<std::uint32_t>(in.blasRoot);
		st.instHittable = hittable;
		st.from = kNoOrigin;
		st.inInstance = true;
		st.found = false;
		LocalStack stack;
		stack.reset();
		while (!traceStep<false, false, 0>(s, st, kTMinF, stack, nullptr))
		{
		}
		hit = st.hit;
		return st.found;
	}

	// The exclusion identity of a hit (see RayOrigin).
	__device__ __forceinline__ RayOrigin	originOf(const SceneView& s, const HitInfo& h)
	{
		const std::uint32_t kind = s.hittables[h.hittable].kindIndex >> kKindShift;
		const bool planar = kind == GPU_HIT_TRIANGLE || kind == GPU_HIT_RECTANGLE || kind == GPU_HIT_INSTANCE || kind == GPU_HIT_CUBE
			|| kind == GPU_HIT_PLANE;
		return planar ? RayOrigin{h.hittable, h.tri} : kNoOrigin;
	}

	template <bool AnyHit, bool Count = false, int Feat = kFeatAll, typename Stack>
	__device__ __forceinline__ bool	traverse(const SceneView& s, const GpuParams& p, const Ray& worldRay, float tMin, float tMax,
		HitInfo& hit, Stack& stack, VisitCounter* counter = nullptr, float logU = 0.0f, RayOrigin from = kNoOrigin)
	{
		TraceState st;
		if (traceBegin(st, s, p.tlasRoot, unboundedOf(p), worldRay, tMax, logU, from, stack))
		{
			while (!traceStep<AnyHit, Count, Feat>(s, st, tMin, stack, counter))
			{
			}
		}
		hit = st.hit;
		return st.found;
	}
}
