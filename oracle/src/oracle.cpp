// ORACLE — test infrastructure only. FP64 CPU restatement of the reference's radiance
// path, used by tests/ (as the checker), by __graft_entry__.smoke() and as an alternative
// CPU baseline. Nothing in the product links or calls it.
//
// Each function follows the reference file:line it cites, in the reference's own
// operation order, so that on the same libm it reproduces the reference's FP64 values.
// Scenes are read with the product's .luz reader (luz_amd/csrc/host/scene_file.cpp), whose
// output is pinned to the reference separately (topology fixtures, golden images).
//
// Pinning (tests/test_oracle.py): sampler against the reference's golden vectors
// (bit-exact); primary hits against the reference's brute-force IDs; Cornell and sphere
// images against the reference's FP64 renders (golden .npz); closed forms (Beer–Lambert
// exp(-2) style attenuation is not on this path; GGX/cosine pdf normalisation and the
// conductor Fresnel normal-incidence value are checked by Monte Carlo / closed form).
#include "../../luz_amd/csrc/host/scene.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace oracle
{
	using luzgpu::dvec3;
	const double	T_MIN = 0.001;					// Defaults.hpp
	const double	T_MAX = std::numeric_limits<double>::max();
	const double	PI = 3.14159265358979323846;	// Defaults.hpp D_PI

	// ---- value types (Vector3.hpp, Color.hpp) -------------------------------------------
	struct	V
	{
		double	x = 0, y = 0, z = 0;
	};
	inline V	operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
	inline V	operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
	inline V	operator*(double f, V v) { return {f * v.x, f * v.y, f * v.z}; }		// Vector3.hpp:77-80
	inline V	operator*(V v, double f) { return f * v; }							// Vector3.hpp:82-85
	inline V	operator*(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
	inline V	operator/(V v, double f) { return (1.0 / f) * v; }					// Vector3.hpp:92-95
	inline double	dot(V a, V b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }	// Utilities.hpp:13-19
	inline V	cross(V a, V b)
	{
		return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
	}
	inline double	len2(V a) { return (a.x * a.x) + (a.y * a.y) + (a.z * a.z); }
	inline V	normalize(V a) { return a / std::sqrt(len2(a)); }					// Utilities.hpp:38-41
	inline V	reflect(V v, V n) { return v - 2.0 * dot(v, n) * n; }					// Utilities.hpp:43-46
	inline V	fromD(const dvec3& d) { return {d[0], d[1], d[2]}; }
	inline double	at(V v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

	struct	C	// Color: division divides directly (Color.hpp:68-71)
	{
		double	r = 0, g = 0, b = 0;
	};
	inline C	operator+(C a, C b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
	inline C	operator*(C a, double f) { return {a.r * f, a.g * f, a.b * f}; }
	inline C	operator*(C a, C b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
	inline C	operator/(C a, double f) { return {a.r / f, a.g / f, a.b / f}; }
	inline double	maxChannel(C c) { return std::max(c.r, std::max(c.g, c.b)); }
	inline double	luminance(C c) { return c.r * 0.2722287168 + c.g * 0.6740817658 + c.b * 0.0536895174; }
	inline C	fromC(const dvec3& d) { return {d[0], d[1], d[2]}; }

	// ---- sampler (Sampler.cpp:28-237) --------------------------------------------------------
	inline std::uint64_t	splitMix64(std::uint64_t v)
	{
		v += 0x9e3779b97f4a7c15ull;
		v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
		v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
		return v ^ (v >> 31);
	}
	inline double	unit32(std::uint32_t h) { return static_cast<double>(h >> 8) * (1.0 / 16777216.0); }
	inline double	unit64(std::uint64_t h) { return static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0); }
	inline double	fract(double v) { return v - std::floor(v); }

	double	dimensionStep(std::uint32_t dim, std::uint32_t comp)	// Sampler.cpp:59-98
	{
		const std::uint64_t key = (static_cast<std::uint64_t>(dim) + 1ull) * 0x9e3779b97f4a7c15ull
			^ (static_cast<std::uint64_t>(comp) + 1ull) * 0xda942042e4dd58b5ull;
		double step = unit32(static_cast<std::uint32_t>(splitMix64(key) >> 32));
		if (step < 0.05 || step > 0.95) step = fract(step + 0.6180339887498948);
		return step;
	}

	struct	Sampler
	{
		std::uint64_t	baseKey = 0;
		std::uint32_t	sample = 0, bounce = 0;

		void	begin(std::uint32_t seed, std::size_t x, std::size_t y, std::uint32_t s, std::uint32_t stream)	// :161-173
		{
			baseKey = seed;
			baseKey ^= (static_cast<std::uint64_t>(x) + 1ull) * 0xbf58476d1ce4e5b9ull;
			baseKey ^= (static_cast<std::uint64_t>(y) + 1ull) * 0x94d049bb133111ebull;
			baseKey ^= (static_cast<std::uint64_t>(stream) + 1ull) * 0xda942042e4dd58b5ull;
			sample = s;
			bounce = 0;
		}
		std::uint64_t	ctxKey(std::uint32_t dk) const { return baseKey ^ ((static_cast<std::uint64_t>(dk) + 1ull) * 0x9e3779b97f4a7c15ull); }
		static bool	progressive(std::uint32_t d)	// :112-126
		{
			const std::uint32_t b = d % 32;
			return b == 0 || b == 1 || b == 4 || b == 5 || b == 12 || b == 13 || b == 16;
		}
		double	hashed(std::uint32_t dk, std::uint32_t comp) const	// :128-135
		{
			std::uint64_t key = ctxKey(dk);
			key ^= (static_cast<std::uint64_t>(sample) + 1ull) * 0x94d049bb133111ebull;
			key ^= (static_cast<std::uint64_t>(comp) + 1ull) * 0xbf58476d1ce4e5b9ull;
			return unit64(splitMix64(key));
		}
		double	s1(std::uint32_t dim) const	// :193-209
		{
			const std::uint32_t dk = dim + bounce * 32;
			if (!progressive(dim)) return hashed(dk, 0);
			const double shift = unit32(static_cast<std::uint32_t>(splitMix64(ctxKey(dk)) >> 32));
			const double step = dimensionStep(dk, 0);
			return fract(0.5 + (static_cast<double>(sample) + 1.0) * step + shift);
		}
		void	s2(std::uint32_t dim, float& x, float& y) const	// :211-237
		{
			const std::uint32_t dk = dim + bounce * 32;
			if (!progressive(dim))
			{
				x = static_cast<float>(hashed(dk, 0));
				y = static_cast<float>(hashed(dk, 1));
				return;
			}
			const std::uint64_t key = ctxKey(dk);
			const double sx = unit32(static_cast<std::uint32_t>(splitMix64(key) >> 32));
			const double sy = unit32(static_cast<std::uint32_t>(splitMix64(key ^ 0xda942042e4dd58b5ull) >> 32));
			const double stx = dimensionStep(dk, 0), sty = dimensionStep(dk, 1);
			const double s = static_cast<double>(sample) + 1.0;
			x = static_cast<float>(fract(0.5 + s * stx + sx));
			y = static_cast<float>(fract(0.5 + s * sty + sy));
		}
	};

	enum Dim : std::uint32_t
	{
		D_CAMERA = 0, D_LENS = 1, D_LIGHT_STRATEGY = 2, D_LIGHT_SELECTION = 3, D_LIGHT_SURFACE_SELECTION = 4, D_LIGHT_SURFACE_POINT = 5,
		D_BSDF_DIRECTION = 6, D_MATERIAL_DECISION = 7, D_RR = 9, D_VOLUME_DISTANCE = 10, D_ENV_SELECTION = 12,
		D_ENV_POINT = 13, D_ENV_STRATEGY = 14, D_SUBSURFACE_SPATIAL = 17, D_SUBSURFACE_PROFILE_COMPONENT = 18,
		D_SUBSURFACE_THIN_SIDE = 19
	};

	// ---- rays (Ray.hpp) -----------------------------------------------------------------------
	inline double	invComp(double c)
	{
		if (c == 0.0 || !std::isfinite(c)) return 0.0;
		const double i = 1.0 / c;
		return std::isfinite(i) ? i : 0.0;
	}
	struct	Ray
	{
		V	o, d, inv;
		static Ray	normalized(V o, V d) { return Ray{o, d, {invComp(d.x), invComp(d.y), invComp(d.z)}}; }	// :23-26
		static Ray	make(V o, V d)	// :18-21, _normalize :53-65
		{
			const double l = std::sqrt((d.x * d.x) + (d.y * d.y) + (d.z * d.z));
			const V n = (!std::isfinite(l) || l <= 0.0) ? V{} : d / l;
			return normalized(o, n);
		}
		V	at(double t) const { return o + (d * t); }
	};

	// ---- hit record (Hittable.hpp:12-36) ----------------------------------------------------
	struct	Hit
	{
		double	t = 0;
		V	p, n, gn;
		bool	front = true;
		std::uint32_t	material = 0;
		std::int64_t	top = -1, tri = -1;
	};
	inline void	setFace(Hit&, const Ray& r, V outward, V& n, bool& front)
	{
		front = ((r.d.x * outward.x) + (r.d.y * outward.y) + (r.d.z * outward.z)) < 0.0;
		n = front ? outward : outward * -1.0;
	}

	// ---- geometry ---------------------------------------------------------------------------------
	struct	Tri	// Triangle.cpp:94-126
	{
		V	v0, v1, v2, e1, e2, face;
		V	n0, n1, n2;
		bool	hasN = false, degenerate = false;
		double	area = 0;
		V	lo, hi;
		std::uint32_t	material = 0;

		void	init()
		{
			e1 = v1 - v0;
			e2 = v2 - v0;
			const V g = cross(e1, e2);
			const double ls = len2(g);
			degenerate = ls <= 1e-24;
			if (degenerate) { face = V{}; area = 0.0; }
			else { const double l = std::sqrt(ls); face = g / l; area = l / 2.0; }
			lo = {std::min(v0.x, std::min(v1.x, v2.x)) - T_MIN, std::min(v0.y, std::min(v1.y, v2.y)) - T_MIN,
				std::min(v0.z, std::min(v1.z, v2.z)) - T_MIN};
			hi = {std::max(v0.x, std::max(v1.x, v2.x)) + T_MIN, std::max(v0.y, std::max(v1.y, v2.y)) + T_MIN,
				std::max(v0.z, std::max(v1.z, v2.z)) + T_MIN};
		}

		// Triangle::hit (Triangle.cpp:196-270): returns t, u, v; record built by caller
		bool	intersect(const Ray& r, double tMin, double tMax, double& t, double& u, double& v) const
		{
			if (degenerate) return false;
			const V p = cross(r.d, e2);
			const double det = dot(e1, p);
			if (std::fabs(det) < 1e-12) return false;
			const double invDet = 1.0 / det;
			const V tv = r.o - v0;
			u = dot(tv, p) * invDet;
			if (u < 0.0 || u > 1.0) return false;
			const V q = cross(tv, e1);
			v = dot(r.d, q) * invDet;
			if (v < 0.0 || u + v > 1.0) return false;
			t = dot(e2, q) * invDet;
			return !(t > tMax || t < tMin);
		}

		void	record(const Ray& r, double t, double u, double v, Hit& h) const	// :241-270
		{
			h.t = t;
			V n = face;
			if (hasN)
			{
				const V in = (n0 * (1.0 - u - v)) + (n1 * u) + (n2 * v);
				n = len2(in) <= 1e-24 ? face : normalize(in);
			}
			setFace(h, r, n, h.n, h.front);
			bool f2;
			setFace(h, r, face, h.gn, f2);
			h.p = r.at(t);
		}
	};

	struct	Rect	// Rectangle.cpp:14-50,212-259
	{
		V	pos, n, wa, ha;
		double	w = 0, h = 0;
		bool	ok = false;
		std::uint32_t	material = 0;

		bool	intersect(const Ray& r, double tMin, double tMax, double& t, double& wd, double& hd) const
		{
			if (!ok || w <= 0.0 || h <= 0.0) return false;
			const double a = dot(r.o - pos, n);
			const double b = dot(r.d, n);
			if (b == 0.0 || (a < 0.0 && b < 0.0) || (a > 0.0 && b > 0.0)) return false;
			t = -a / b;
			if (t > tMax || t < tMin) return false;
			const V d = r.at(t) - pos;
			wd = dot(d, wa);
			hd = dot(d, ha);
			return !(std::fabs(wd) > (w / 2.0) || std::fabs(hd) > (h / 2.0));
		}
	};

	// Rectangle basis from the orientation (Rectangle.cpp:14-50, 127-152)
	Rect	makeRect(V pos, V ori, double w, double h, std::uint32_t material)
	{
		Rect o;
		o.pos = pos;
		o.w = w;
		o.h = h;
		o.ok = len2(ori) > 0.0;
		if (o.ok)
		{
			o.n = normalize(ori);
			if (std::fabs(o.n.y) > 0.999) { o.wa = {1, 0, 0}; o.ha = {0, 0, 1}; }
			else if (std::fabs(o.n.z) > 0.999) { o.wa = {1, 0, 0}; o.ha = {0, 1, 0}; }
			else if (std::fabs(o.n.x) > 0.999) { o.wa = {0, 0, 1}; o.ha = {0, 1, 0}; }
			else
			{
				V helper = {0, 1, 0};
				if (std::fabs(dot(helper, o.n)) > 0.9) helper = {1, 0, 0};
				o.wa = normalize(cross(helper, o.n));
				o.ha = normalize(cross(o.n, o.wa));
			}
		}
		o.material = material;
		return o;
	}

	inline bool	zeroDir(const Ray& r) { return r.d.x == 0.0 && r.d.y == 0.0 && r.d.z == 0.0; }

	struct	Sphere	// Sphere.cpp
	{
		V	c;
		double	r = 1.0;
		bool	visible = true;	// an invisible sphere is never hit, yet stays a light (:227-233, 271-277)
		// Sphere::setIESProfile (:188-202): the profile (index into the scene's), its vertical
		// axis and horizontal rotation (degrees)
		int	ies = -1;
		V	iesDirection{0.0, -1.0, 0.0};
		double	iesRotation = 0.0;

		// Sphere::hit / hitAny (:226-301): unit direction, a = 1
		bool	intersect(const Ray& ray, double tMin, double tMax, double& t) const
		{
			if (!visible || zeroDir(ray)) return false;
			const V oc = ray.o - c;
			const double b = dot(oc, ray.d);
			const double cc = dot(oc, oc) - (r * r);
			const double disc = (b * b) - cc;
			if (disc < 0.0) return false;
			const double sq = std::sqrt(disc);
			double root = -b - sq;
			if (root < tMin || root > tMax)
			{
				root = -b + sq;
				if (root < tMin || root > tMax) return false;
			}
			t = root;
			return true;
		}
		// Sphere::hitInterval (:303-339)
		bool	interval(const Ray& ray, double tMin, double tMax, double& t0, double& t1) const
		{
			if (zeroDir(ray)) return false;
			const V oc = ray.o - c;
			const double b = dot(oc, ray.d);
			const double cc = dot(oc, oc) - (r * r);
			const double disc = (b * b) - cc;
			if (disc < 0.0) return false;
			const double sq = std::sqrt(disc);
			double r0 = -b - sq, r1 = -b + sq;
			if (r0 > r1) std::swap(r0, r1);
			if (r0 < tMin) r0 = tMin;
			if (r1 > tMax) r1 = tMax;
			if (r0 >= r1) return false;
			t0 = r0;
			t1 = r1;
			return true;
		}
	};

	// Cube::hitInterval over its bounding box (Cube.cpp:120-164)
	inline bool	boxInterval(V lo, V hi, const Ray& r, double tMin, double tMax, double& t0, double& t1)
	{
		double a0 = tMin, a1 = tMax;
		for (int a = 0; a < 3; a++)
		{
			const double d = at(r.d, a), o = at(r.o, a);
			if (d == 0.0)
			{
				if (o < at(lo, a) || o > at(hi, a)) return false;
				continue;
			}
			const double inv = 1.0 / d;
			double x0 = (at(lo, a) - o) * inv, x1 = (at(hi, a) - o) * inv;
			if (x0 > x1) std::swap(x0, x1);
			a0 = std::max(a0, x0);
			a1 = std::min(a1, x1);
			if (a0 >= a1) return false;
		}
		t0 = a0;
		t1 = a1;
		return true;
	}

	struct	Cube	// Cube.cpp:72-118
	{
		Rect	faces[6];
		V	lo, hi;
	};

	struct	Plane	// Plane.cpp:51-87
	{
		double	y = 0.0;
		V	o;
		bool	intersect(const Ray& r, double tMin, double tMax, double& t) const
		{
			const V pos = {0.0, y, 0.0};
			const double d = dot(r.d, o);
			if (d == 0.0) return false;
			t = dot(pos - r.o, o) / d;
			return !(t > tMax || t < tMin);
		}
	};

	struct	Volume	// ConstantVolume.cpp
	{
		bool	isSphere = false;
		Sphere	sphere;
		V	lo, hi;
		double	negInv = -1.0;

		// sampleScatteringDistance (:62-101) with u = DIM_VOLUME_DISTANCE of this bounce
		bool	sample(const Ray& r, double tMin, double tMax, double u, double& t) const
		{
			double e0, e1;
			if (isSphere ? !sphere.interval(r, -T_MAX, T_MAX, e0, e1) : (zeroDir(r) || !boxInterval(lo, hi, r, -T_MAX, T_MAX, e0, e1)))
			{
				return false;
			}
			if (e0 < tMin) e0 = tMin;
			if (e1 > tMax) e1 = tMax;
			if (e0 >= e1) return false;
			if (e0 < 0) e0 = 0;
			const double len = std::sqrt(len2(r.d));
			if (len <= 0.0 || !std::isfinite(len)) return false;
			const double inside = (e1 - e0) * len;
			const double dist = negInv * std::log(std::max(u, 1e-12));
			if (dist > inside) return false;
			t = e0 + dist / len;
			return true;
		}
	};

	// AABB::hit / boxHitDistance (AABB.cpp:37-80, BVHNode.cpp:20-61)
	inline bool	boxHit(V lo, V hi, const Ray& r, double tMax, double& tNear)
	{
		double tMin = T_MIN;
		if (r.d.x == 0.0 && r.d.y == 0.0 && r.d.z == 0.0) return false;
		for (int a = 0; a < 3; a++)
		{
			const double inv = at(r.inv, a), o = at(r.o, a);
			if (inv == 0.0)
			{
				if (o < at(lo, a) || o > at(hi, a)) return false;
				continue;
			}
			double t0 = (at(lo, a) - o) * inv, t1 = (at(hi, a) - o) * inv;
			if (inv < 0.0) std::swap(t0, t1);
			tMin = t0 > tMin ? t0 : tMin;
			tMax = t1 < tMax ? t1 : tMax;
			if (tMax <= tMin) return false;
		}
		tNear = tMin;
		return true;
	}

	// A plain median-split BVH over a mesh's triangles. Closest-hit results do not depend
	// on the tree shape (only exact-distance ties could), so this oracle uses its own.
	struct	MeshBvh
	{
		struct	Node
		{
			V	lo, hi;
			std::uint32_t	left = 0, right = 0, start = 0, count = 0;
		};
		std::vector<Node>	nodes;
		std::vector<std::uint32_t>	order;

		void	build(const std::vector<Tri>& tris)
		{
			order.resize(tris.size());
			std::iota(order.begin(), order.end(), 0u);
			nodes.clear();
			if (!tris.empty()) make(tris, 0, static_cast<std::uint32_t>(tris.size()));
		}

		std::uint32_t	make(const std::vector<Tri>& tris, std::uint32_t s, std::uint32_t e)
		{
			Node n;
			n.lo = {INFINITY, INFINITY, INFINITY};
			n.hi = {-INFINITY, -INFINITY, -INFINITY};
			for (std::uint32_t i = s; i < e; i++)
			{
				const Tri& t = tris[order[i]];
				n.lo = {std::min(n.lo.x, t.lo.x), std::min(n.lo.y, t.lo.y), std::min(n.lo.z, t.lo.z)};
				n.hi = {std::max(n.hi.x, t.hi.x), std::max(n.hi.y, t.hi.y), std::max(n.hi.z, t.hi.z)};
			}
			const std::uint32_t idx = static_cast<std::uint32_t>(nodes.size());
			nodes.push_back(n);
			if (e - s <= 4)
			{
				nodes[idx].start = s;
				nodes[idx].count = e - s;
				return idx;
			}
			const V ext = n.hi - n.lo;
			const int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
			const std::uint32_t mid = s + (e - s) / 2;
			std::nth_element(order.begin() + s, order.begin() + mid, order.begin() + e, [&](std::uint32_t a, std::uint32_t b) {
				return at(tris[a].lo, axis) + at(tris[a].hi, axis) < at(tris[b].lo, axis) + at(tris[b].hi, axis);
			});
			const std::uint32_t l = make(tris, s, mid);
			const std::uint32_t r = make(tris, mid, e);
			nodes[idx].left = l;
			nodes[idx].right = r;
			return idx;
		}

		// closest (any = false) or any hit in [tMin, tMax]
		bool	trace(const std::vector<Tri>& tris, const Ray& r, double tMin, double tMax, bool any,
			std::uint32_t& triOut, double& tOut, double& uOut, double& vOut) const
		{
			if (nodes.empty()) return false;
			std::uint32_t stack[128];
			int sp = 0;
			stack[sp++] = 0;
			bool found = false;
			double closest = tMax;
			while (sp > 0)
			{
				const Node& n = nodes[stack[--sp]];
				double tn;
				if (!boxHit(n.lo, n.hi, r, closest, tn)) continue;
				if (n.count > 0)
				{
					for (std::uint32_t i = n.start; i < n.start + n.count; i++)
					{
						double t, u, v;
						if (tris[order[i]].intersect(r, tMin, closest, t, u, v))
						{
							found = true;
							closest = t;
							triOut = order[i];
							tOut = t; uOut = u; vOut = v;
							if (any) return true;
						}
					}
					continue;
				}
				stack[sp++] = n.right;
				stack[sp++] = n.left;
			}
			return found;
		}
	};

	struct	Mesh
	{
		std::vector<Tri>	tris;
		MeshBvh	bvh;
		V	lo, hi;	// root box of the reference's mesh BVH (union of padded triangle boxes)
	};

	struct	Instance	// MeshInstance.cpp:48-167
	{
		std::uint32_t	mesh = 0, material = 0;
		V	position, scale;
		double	cx = 1, sx = 0, cy = 1, sy = 0, cz = 1, sz = 0;
		V	lo, hi;

		V	rotate(V v) const
		{
			double x = v.x, y = v.y, z = v.z;
			double ny = (y * cx) - (z * sx), nz = (y * sx) + (z * cx);
			y = ny; z = nz;
			double nx = (x * cy) + (z * sy);
			nz = (-x * sy) + (z * cy);
			x = nx; z = nz;
			nx = (x * cz) - (y * sz);
			ny = (x * sz) + (y * cz);
			return {nx, ny, z};
		}
		V	inverseRotate(V v) const
		{
			double x = v.x, y = v.y, z = v.z;
			double nx = (x * cz) + (y * sz), ny = (-x * sz) + (y * cz);
			x = nx; y = ny;
			nx = (x * cy) - (z * sy);
			double nz = (x * sy) + (z * cy);
			x = nx; z = nz;
			ny = (y * cx) + (z * sx);
			nz = (-y * sx) + (z * cx);
			return {x, ny, nz};
		}
		V	transformPoint(V p) const	// :123-126
		{
			return rotate({p.x * scale.x, p.y * scale.y, p.z * scale.z}) + position;
		}
		V	transformNormal(V n) const	// :151-164
		{
			const V s = rotate({n.x / scale.x, n.y / scale.y, n.z / scale.z});
			const double ls = len2(s);
			if (!std::isfinite(ls) || ls <= 1e-24) return V{};
			return s / std::sqrt(ls);
		}
		bool	localRay(const Ray& r, Ray& local, double& scaleOut) const	// :251-272
		{
			const V ur = inverseRotate(r.d);
			const V ld = {ur.x / scale.x, ur.y / scale.y, ur.z / scale.z};
			scaleOut = std::sqrt(len2(ld));
			if (!std::isfinite(scaleOut) || scaleOut <= 0.0) return false;
			const V uo = inverseRotate(r.o - position);
			const V lo2 = {uo.x / scale.x, uo.y / scale.y, uo.z / scale.z};
			local = Ray::normalized(lo2, ld / scaleOut);
			return true;
		}
	};

	enum Kind { K_TRI, K_RECT, K_INST, K_SPHERE, K_CUBE, K_VOLUME, K_PLANE, K_DIR };

	struct	Top
	{
		Kind	kind;
		std::uint32_t	index;
		std::uint32_t	material;
	};

	struct	Mat
	{
		luzgpu::MaterialKind	kind;
		C	color;
		double	roughness = 0, alpha = 1e-4, ior = 1.5, g = 0.0;
		C	glossyColor = {1, 1, 1};
		double	glossyWeight = 0.5, metallic = 0, transmission = 0, clearcoat = 0, clearcoatRoughness = 0.03, sheen = 0;
		bool	conductor = false;	// Metal::_useConductorFresnel (Metal.cpp:118-125)
		C	eta = {1, 1, 1}, k = {0, 0, 0};
		C	absorption = {0, 0, 0};	// dielectric / principled interior absorption, 1/m
		// principled subsurface (Principled.cpp:39-88): weight, tint, Burley diffusion
		// length (max channel of radius * scale, metres), thin or Burley
		double	subsurface = 0.0;
		C	subsurfaceColor = {1, 1, 1};
		double	subsurfaceRadius = 0.0;
		bool	subsurfaceThin = false;
	};

	struct	Light
	{
		Kind	kind;
		std::uint32_t	index;
		C	emitted;
		double	area;
		V	dir;	// directional: -direction
		std::uint32_t	material = 0;
		// emissive mesh instance: the triangle-area prefix sums its sampler searches (the
		// mesh's own when the instance scales area uniformly, else the transformed ones)
		std::vector<double>	triCdf;
	};

	struct	Env
	{
		std::size_t	W = 0, H = 0;
		std::vector<C>	px;
		std::vector<double>	weights, solid, cdf;
		double	total = 0, strength = 1, rotation = 0;
		bool	lighting = true;
	};

	struct	Scene
	{
		std::vector<Top>	top;
		std::vector<Tri>	topTris;	// indexed by Top.index for K_TRI
		std::vector<Rect>	rects;
		std::vector<Sphere>	spheres;
		std::vector<luzgpu::IesProfile>	iesProfiles;
		std::vector<Cube>	cubes;
		std::vector<Plane>	planes;
		std::vector<Volume>	volumes;
		std::vector<V>	directionals;
		std::vector<Mesh>	meshes;
		std::vector<Instance>	instances;
		std::vector<Mat>	mats;
		std::vector<Light>	lights;
		std::vector<double>	lightCdf;
		double	lightTotal = 0;
		bool	hasEnv = false;
		Env	env;
		luzgpu::SkyKind	sky = luzgpu::SKY_NONE;
		C	background;
		int	maxBounces = 8;
		std::uint32_t	width = 0, height = 0, samples = 1;
		double	mpu = 1.0;
		// RenderCamera (RendererInternal.hpp:13-26)
		V	camPos, camU, camV, horizontal, vertical, llc;
		double	lensRadius = 0, invW = 0, invH = 0;
		luzgpu::Camera	camera;
	};

	void	setupCamera(Scene& s)	// RayHelper.cpp:50-82
	{
		const luzgpu::Camera& c = s.camera;
		const double W = s.width, H = s.height;
		s.invW = W > 1.0 ? 1.0 / (W - 1.0) : 0.0;
		s.invH = H > 1.0 ? 1.0 / (H - 1.0) : 0.0;
		s.camPos = fromD(c.position);
		const double focus = c.focusDistance / s.mpu;
		double vw = focus * c.sensorWidth / c.focalLength;
		double vh = focus * c.sensorHeight / c.focalLength;
		const double ia = W / H, va = vw / vh;
		if (ia > va) vh = vw / ia;
		else if (ia < va) vw = vh * ia;
		const V w = normalize(fromD(c.direction));
		V up = fromD(c.up);
		if (!std::isfinite(up.x) || !std::isfinite(up.y) || !std::isfinite(up.z) || len2(up) <= 1e-12) up = {0, 1, 0};
		up = normalize(up);
		if (len2(cross(up, w)) <= 1e-12) up = {0, 0, 1};
		if (len2(cross(up, w)) <= 1e-12) up = {1, 0, 0};
		s.camU = normalize(cross(up, w));
		s.camV = cross(w, s.camU);
		s.lensRadius = (c.pinhole ? 0.0 : c.focalLength / c.fNumber) / (2.0 * s.mpu);
		s.horizontal = s.camU * vw;
		s.vertical = s.camV * vh;
		s.llc = s.camPos + (s.horizontal / 2.0) + (s.vertical / 2.0) + (w * focus);
	}

	Scene	build(const luzgpu::Scene& in, std::uint32_t width, std::uint32_t height)
	{
		Scene s;
		for (const auto& m : in.materials)
		{
			Mat o;
			o.kind = m.kind;
			o.color = fromC(m.color);
			o.roughness = m.roughness;
			const double r = std::isfinite(m.roughness) ? std::max(0.0, std::min(1.0, m.roughness)) : 0.0;
			o.alpha = std::max(1e-4, r * r);	// BSDF.cpp:54-58
			o.ior = m.ior;
			o.g = m.anisotropy;
			o.glossyColor = fromC(m.glossyColor);
			o.glossyWeight = m.glossyWeight;
			o.metallic = m.metallic;
			o.transmission = m.transmission;
			o.clearcoat = m.clearcoat;
			o.clearcoatRoughness = m.clearcoatRoughness;
			o.sheen = m.sheen;
			o.conductor = m.conductor;
			o.eta = fromC(m.eta);
			o.k = fromC(m.k);
			o.absorption = fromC(m.absorption);
			o.subsurface = m.subsurface;
			o.subsurfaceColor = fromC(m.subsurfaceColor);
			o.subsurfaceRadius = std::max(m.subsurfaceRadius[0] * m.subsurfaceScale,
				std::max(m.subsurfaceRadius[1] * m.subsurfaceScale, m.subsurfaceRadius[2] * m.subsurfaceScale));
			o.subsurfaceThin = m.subsurfaceThin;
			s.mats.push_back(o);
		}
		for (const auto& mesh : in.meshes)
		{
			Mesh m;
			for (const auto& t : mesh.triangles)
			{
				Tri o;
				o.v0 = fromD(t.v0); o.v1 = fromD(t.v1); o.v2 = fromD(t.v2);
				o.n0 = fromD(t.n0); o.n1 = fromD(t.n1); o.n2 = fromD(t.n2);
				o.hasN = t.hasNormals;
				o.init();
				m.tris.push_back(o);
			}
			m.bvh.build(m.tris);
			m.lo = {INFINITY, INFINITY, INFINITY};
			m.hi = {-INFINITY, -INFINITY, -INFINITY};
			for (const Tri& t : m.tris)
			{
				m.lo = {std::min(m.lo.x, t.lo.x), std::min(m.lo.y, t.lo.y), std::min(m.lo.z, t.lo.z)};
				m.hi = {std::max(m.hi.x, t.hi.x), std::max(m.hi.y, t.hi.y), std::max(m.hi.z, t.hi.z)};
			}
			s.meshes.push_back(std::move(m));
		}
		s.iesProfiles = in.iesProfiles;
		for (const auto& h : in.hittables)
		{
			Top t{K_TRI, 0, h.material};
			if (h.kind == luzgpu::HIT_TRIANGLE)
			{
				const auto& tr = in.triangles[h.index];
				Tri o;
				o.v0 = fromD(tr.v0); o.v1 = fromD(tr.v1); o.v2 = fromD(tr.v2);
				o.init();
				o.material = h.material;
				t.kind = K_TRI;
				t.index = static_cast<std::uint32_t>(s.topTris.size());
				s.topTris.push_back(o);
			}
			else if (h.kind == luzgpu::HIT_RECTANGLE)
			{
				const auto& r = in.rectangles[h.index];
				t.kind = K_RECT;
				t.index = static_cast<std::uint32_t>(s.rects.size());
				s.rects.push_back(makeRect(fromD(r.position), fromD(r.orientation), r.width, r.height, h.material));
			}
			else if (h.kind == luzgpu::HIT_SPHERE)
			{
				const auto& sp = in.spheres[h.index];
				t.kind = K_SPHERE;
				t.index = static_cast<std::uint32_t>(s.spheres.size());
				s.spheres.push_back(Sphere{fromD(sp.position), sp.radius, sp.visible, sp.ies, fromD(sp.iesDirection), sp.iesRotation});
			}
			else if (h.kind == luzgpu::HIT_CUBE)
			{
				// Cube::_generateFaces (Cube.cpp:72-95)
				const auto& c = in.cubes[h.index];
				const V p = fromD(c.position);
				Cube o;
				o.faces[0] = makeRect(p - V{0.0, 0.0, c.depth / 2.0}, {0, 0, -1}, c.width, c.height, h.material);
				o.faces[1] = makeRect(p + V{0.0, 0.0, c.depth / 2.0}, {0, 0, 1}, c.width, c.height, h.material);
				o.faces[2] = makeRect(p + V{0.0, c.height / 2.0, 0.0}, {0, 1, 0}, c.width, c.depth, h.material);
				o.faces[3] = makeRect(p - V{0.0, c.height / 2.0, 0.0}, {0, -1, 0}, c.width, c.depth, h.material);
				o.faces[4] = makeRect(p + V{c.width / 2.0, 0.0, 0.0}, {1, 0, 0}, c.depth, c.height, h.material);
				o.faces[5] = makeRect(p - V{c.width / 2.0, 0.0, 0.0}, {-1, 0, 0}, c.depth, c.height, h.material);
				o.lo = {p.x - c.width / 2.0, p.y - c.height / 2.0, p.z - c.depth / 2.0};
				o.hi = {p.x + c.width / 2.0, p.y + c.height / 2.0, p.z + c.depth / 2.0};
				t.kind = K_CUBE;
				t.index = static_cast<std::uint32_t>(s.cubes.size());
				s.cubes.push_back(o);
			}
			else if (h.kind == luzgpu::HIT_VOLUME)
			{
				const auto& v = in.volumes[h.index];
				Volume o;
				o.isSphere = v.sphere;
				o.sphere = Sphere{fromD(v.position), v.radius};
				const V p = fromD(v.position);
				o.lo = {p.x - v.size[0] / 2.0, p.y - v.size[1] / 2.0, p.z - v.size[2] / 2.0};
				o.hi = {p.x + v.size[0] / 2.0, p.y + v.size[1] / 2.0, p.z + v.size[2] / 2.0};
				o.negInv = -1.0 / v.density;
				t.kind = K_VOLUME;
				t.index = static_cast<std::uint32_t>(s.volumes.size());
				s.volumes.push_back(o);
			}
			else if (h.kind == luzgpu::HIT_PLANE)
			{
				const auto& pl = in.planes[h.index];
				t.kind = K_PLANE;
				t.index = static_cast<std::uint32_t>(s.planes.size());
				s.planes.push_back(Plane{pl.y, fromD(pl.orientation)});
			}
			else if (h.kind == luzgpu::HIT_DIRECTIONAL)
			{
				t.kind = K_DIR;
				t.index = static_cast<std::uint32_t>(s.directionals.size());
				s.directionals.push_back(fromD(in.directionals[h.index].direction));
			}
			else
			{
				const auto& inst = in.instances[h.index];
				Instance o;
				o.mesh = inst.mesh;
				o.material = h.material;
				o.position = fromD(inst.position);
				o.scale = fromD(inst.scale);
				const double d2r = 3.14159265358979323846 / 180.0;
				o.cx = std::cos(inst.rotationDegrees[0] * d2r); o.sx = std::sin(inst.rotationDegrees[0] * d2r);
				o.cy = std::cos(inst.rotationDegrees[1] * d2r); o.sy = std::sin(inst.rotationDegrees[1] * d2r);
				o.cz = std::cos(inst.rotationDegrees[2] * d2r); o.sz = std::sin(inst.rotationDegrees[2] * d2r);
				t.kind = K_INST;
				t.index = static_cast<std::uint32_t>(s.instances.size());
				s.instances.push_back(o);
			}
			s.top.push_back(t);
		}
		// lights (Scene.cpp:615-638)
		for (std::size_t i = 0; i < s.top.size(); i++)
		{
			const Top& t = s.top[i];
			if (s.mats[t.material].kind != luzgpu::MAT_EMISSIVE) continue;
			Light l{t.kind, t.index, s.mats[t.material].color, 0.0, V{}};
			l.material = t.material;
			const double lum = luminance(l.emitted);
			double w;
			if (t.kind == K_TRI || t.kind == K_RECT || t.kind == K_SPHERE)
			{
				if (t.kind == K_TRI) l.area = s.topTris[t.index].area;
				else if (t.kind == K_RECT) l.area = s.rects[t.index].h * s.rects[t.index].w;
				else l.area = 4.0 * PI * s.spheres[t.index].r * s.spheres[t.index].r;	// Sphere.cpp:501-515
				w = (l.area <= 0.0 || !std::isfinite(lum) || lum <= 0.0) ? 0.0 : l.area * lum;
			}
			else if (t.kind == K_DIR || t.kind == K_CUBE || t.kind == K_PLANE)
			{
				// Hittable / DirectionalLight::lightSelectionWeight: the luminance
				if (t.kind == K_DIR) l.dir = s.directionals[t.index] * -1.0;
				w = (!std::isfinite(lum) || lum <= 0.0) ? 0.0 : lum;
			}
			else if (t.kind == K_INST)
			{
				// MeshInstance::_surfaceArea / _transformedAreaPrefixSums / lightSelectionWeight
				// (MeshInstance.cpp:345-407,589-603); Mesh prefix sums (Mesh.cpp:525-532)
				const Instance& inst = s.instances[t.index];
				const Mesh& m = s.meshes[inst.mesh];
				const double ax = std::fabs(inst.scale.x), ay = std::fabs(inst.scale.y), az = std::fabs(inst.scale.z);
				const bool uniform = std::fabs(ax - ay) <= 1e-12 && std::fabs(ax - az) <= 1e-12;
				double total = 0.0;
				if (uniform)
				{
					for (const Tri& tr : m.tris)
					{
						total += tr.area;
						l.triCdf.push_back(total);
					}
					const double a = total * (ax * ax);
					l.area = (std::isfinite(a) && a > 0.0) ? a : 0.0;
					if (!(std::isfinite(total) && total > 0.0)) l.triCdf.clear();
				}
				else
				{
					for (const Tri& tr : m.tris)
					{
						const V w0 = inst.transformPoint(tr.v0), w1 = inst.transformPoint(tr.v1), w2 = inst.transformPoint(tr.v2);
						const double a = std::sqrt(len2(cross(w1 - w0, w2 - w0))) / 2.0;
						total += (std::isfinite(a) && a > 0.0) ? a : 0.0;
						l.triCdf.push_back(total);
					}
					if (!(std::isfinite(total) && total > 0.0)) l.triCdf.clear();
					l.area = l.triCdf.empty() ? 0.0 : total;
				}
				w = (!std::isfinite(lum) || lum <= 0.0 || l.area <= 0.0) ? 0.0 : l.area * lum;
			}
			else throw std::runtime_error("oracle: emissive volumes are not restated");
			if (!std::isfinite(w) || w <= 0.0) w = 0.0;
			s.lightTotal += w;
			s.lightCdf.push_back(s.lightTotal);
			s.lights.push_back(l);
		}
		if (s.lightTotal <= 0.0) s.lightCdf.clear();
		// environment (EnvironmentMap.cpp:627-671)
		s.hasEnv = in.hasEnvironment;
		if (s.hasEnv)
		{
			Env& e = s.env;
			e.W = in.environment.width;
			e.H = in.environment.height;
			e.strength = in.environment.strength;
			e.rotation = in.environment.rotation;
			e.lighting = in.environment.lighting;
			for (const auto& p : in.environment.pixels) e.px.push_back(fromC(p));
			e.weights.assign(e.W * e.H, 0.0);
			e.solid.assign(e.W * e.H, 0.0);
			const double dphi = 2.0 * PI / static_cast<double>(e.W);
			for (std::size_t y = 0; y < e.H; y++)
			{
				const double t0 = PI * static_cast<double>(y) / static_cast<double>(e.H);
				const double t1 = PI * static_cast<double>(y + 1) / static_cast<double>(e.H);
				const double sa = dphi * std::max(0.0, std::cos(t0) - std::cos(t1));
				for (std::size_t x = 0; x < e.W; x++)
				{
					const std::size_t i = y * e.W + x;
					const C c = e.px[i];
					const C sc = {std::isfinite(c.r) ? std::max(0.0, c.r) : 0.0, std::isfinite(c.g) ? std::max(0.0, c.g) : 0.0,
						std::isfinite(c.b) ? std::max(0.0, c.b) : 0.0};
					const double lum = luminance(sc);
					const double w = (std::isfinite(lum) && lum > 0.0) ? lum * sa : 0.0;
					e.solid[i] = sa;
					e.weights[i] = w;
					e.total += w;
					e.cdf.push_back(e.total);
				}
			}
			if (!std::isfinite(e.total) || e.total <= 0.0)
			{
				e.total = 0.0;
				e.cdf.clear();
				e.weights.assign(e.W * e.H, 0.0);
			}
		}
		if (in.settings.sky == luzgpu::SKY_ATMOSPHERE) throw std::runtime_error("the atmosphere sky is outside the oracle");
		s.sky = in.settings.sky;
		s.background = fromC(in.settings.background);
		s.maxBounces = in.settings.maxBounces;
		s.width = width ? width : in.settings.width;
		s.height = height ? height : in.settings.height;
		s.samples = in.settings.samples;
		s.mpu = in.settings.metersPerUnit;
		s.camera = in.camera;
		setupCamera(s);
		for (Instance& inst : s.instances)	// MeshInstance.cpp:212-249
		{
			const Mesh& m = s.meshes[inst.mesh];
			inst.lo = {INFINITY, INFINITY, INFINITY};
			inst.hi = {-INFINITY, -INFINITY, -INFINITY};
			for (int c = 0; c < 8; c++)
			{
				const V corner = {(c & 1) ? m.hi.x : m.lo.x, (c & 2) ? m.hi.y : m.lo.y, (c & 4) ? m.hi.z : m.lo.z};
				const V w = inst.rotate(corner * inst.scale) + inst.position;
				inst.lo = {std::min(inst.lo.x, w.x), std::min(inst.lo.y, w.y), std::min(inst.lo.z, w.z)};
				inst.hi = {std::max(inst.hi.x, w.x), std::max(inst.hi.y, w.y), std::max(inst.hi.z, w.z)};
			}
		}
		return s;
	}

	// ---- scene queries (HitHelper.cpp:5-32, ColorHelper.cpp:801-829) ----------------------------
	bool	closestHit(const Scene& s, const Ray& r, Hit& h, const Sampler* ps, bool any = false, double tMax = T_MAX)
	{
		bool found = false;
		double closest = tMax;
		// bounded hittables (the acceleration structure), then the unbounded ones
		for (int pass = 0; pass < 2; pass++)
		for (std::size_t i = 0; i < s.top.size(); i++)
		{
			const Top& t = s.top[i];
			const bool unbounded = t.kind == K_PLANE || t.kind == K_DIR;
			if (unbounded != (pass == 1)) continue;
			if (t.kind == K_DIR) continue;	// DirectionalLight::hit is always false
			if (t.kind == K_SPHERE || t.kind == K_PLANE || t.kind == K_VOLUME)
			{
				double tt;
				bool ok;
				if (t.kind == K_SPHERE) ok = s.spheres[t.index].intersect(r, T_MIN, closest, tt);
				else if (t.kind == K_PLANE) ok = s.planes[t.index].intersect(r, T_MIN, closest, tt);
				else ok = s.volumes[t.index].sample(r, T_MIN, closest, ps->s1(D_VOLUME_DISTANCE), tt);
				if (!ok) continue;
				if (any) return true;
				h.t = tt;
				h.p = r.at(tt);
				if (t.kind == K_SPHERE)
				{
					const Sphere& sp = s.spheres[t.index];
					setFace(h, r, (h.p - sp.c) / sp.r, h.n, h.front);
				}
				else if (t.kind == K_PLANE)
				{
					setFace(h, r, s.planes[t.index].o, h.n, h.front);
				}
				else
				{
					h.n = {1.0, 0.0, 0.0};	// ConstantVolume::hit: arbitrary normal
				}
				h.gn = h.n;
				h.material = t.material;
				h.top = static_cast<std::int64_t>(i);
				h.tri = -1;
				closest = tt;
				found = true;
				continue;
			}
			if (t.kind == K_CUBE)
			{
				for (const Rect& rc : s.cubes[t.index].faces)
				{
					double tt, wd, hd;
					if (!rc.intersect(r, T_MIN, closest, tt, wd, hd)) continue;
					if (any) return true;
					h.t = tt;
					setFace(h, r, rc.n, h.n, h.front);
					h.gn = h.n;
					h.p = r.at(tt);
					h.material = t.material;
					h.top = static_cast<std::int64_t>(i);
					h.tri = -1;
					closest = tt;
					found = true;
				}
				continue;
			}
			if (t.kind == K_TRI)
			{
				double tt, u, v;
				if (s.topTris[t.index].intersect(r, T_MIN, closest, tt, u, v))
				{
					if (any) return true;
					s.topTris[t.index].record(r, tt, u, v, h);
					h.material = t.material;
					h.top = static_cast<std::int64_t>(i);
					h.tri = -1;
					closest = tt;
					found = true;
				}
			}
			else if (t.kind == K_RECT)
			{
				double tt, wd, hd;
				const Rect& rc = s.rects[t.index];
				if (rc.intersect(r, T_MIN, closest, tt, wd, hd))
				{
					if (any) return true;
					h.t = tt;
					setFace(h, r, rc.n, h.n, h.front);
					h.gn = h.n;
					h.p = r.at(tt);
					h.material = t.material;
					h.top = static_cast<std::int64_t>(i);
					h.tri = -1;
					closest = tt;
					found = true;
				}
			}
			else
			{
				const Instance& inst = s.instances[t.index];
				double tn;
				if (!boxHit(inst.lo, inst.hi, r, closest, tn)) continue;
				Ray local;
				double scale;
				if (!inst.localRay(r, local, scale)) continue;
				const Mesh& m = s.meshes[inst.mesh];
				std::uint32_t tri;
				double lt, u, v;
				const double lMax = closest == T_MAX ? T_MAX : closest * scale;
				if (!m.bvh.trace(m.tris, local, T_MIN * scale, lMax, any, tri, lt, u, v)) continue;
				const double wt = lt / scale;
				if (!std::isfinite(wt) || wt < T_MIN || wt > closest) continue;
				if (any) return true;
				// local record, then MeshInstance::_transformHitRecord (:274-301)
				Hit lh;
				m.tris[tri].record(local, lt, u, v, lh);
				V wn = inst.transformNormal(lh.n);
				V wg = inst.transformNormal(lh.gn);
				if (dot(r.d, wn) > 0.0) wn = wn * -1.0;
				if (dot(r.d, wg) > 0.0) wg = wg * -1.0;
				h.t = wt;
				h.p = r.at(wt);
				h.n = wn;
				h.gn = wg;
				h.front = dot(r.d, wn) < 0.0;
				h.material = t.material;
				h.top = static_cast<std::int64_t>(i);
				h.tri = tri;
				closest = wt;
				found = true;
			}
		}
		return found;
	}

	bool	occluded(const Scene& s, const Ray& r, double tMax, const Sampler* ps)	// ColorHelper.cpp:801-829
	{
		if (tMax <= T_MIN) return false;
		Hit h;
		return closestHit(s, r, h, ps, true, tMax);
	}

	// ---- BSDF helpers (BSDF.cpp) -----------------------------------------------------------------
	double	clamp01(double v) { return std::isfinite(v) ? std::max(0.0, std::min(1.0, v)) : 0.0; }
	C	clamp01(C c) { return {clamp01(c.r), clamp01(c.g), clamp01(c.b)}; }
	V	safeNormalize(V v, V fb)
	{
		const double l2 = len2(v);
		if (!std::isfinite(l2) || l2 <= 0.0) return fb;
		return v / std::sqrt(l2);
	}
	C	schlick(C f0, double c)	// :81-92
	{
		c = clamp01(c);
		const double f = std::pow(1.0 - c, 5.0);
		f0 = clamp01(f0);
		return {f0.r + (1.0 - f0.r) * f, f0.g + (1.0 - f0.g) * f, f0.b + (1.0 - f0.b) * f};
	}
	double	dielectricFresnel(double c, double ei, double et)	// :93-125
	{
		c = clamp01(std::fabs(c));
		if (!std::isfinite(ei) || !std::isfinite(et) || ei <= 0.0 || et <= 0.0) return 1.0;
		const double si = std::sqrt(std::max(0.0, 1.0 - c * c));
		const double st = ei / et * si;
		if (st >= 1.0) return 1.0;
		const double ct = std::sqrt(std::max(0.0, 1.0 - st * st));
		const double rp = ((et * c) - (ei * ct)) / ((et * c) + (ei * ct));
		const double rs = ((ei * c) - (et * ct)) / ((ei * c) + (et * ct));
		return clamp01(0.5 * (rp * rp + rs * rs));
	}
	double	ggxD(double noH, double a)	// :209-218
	{
		noH = clamp01(noH);
		a = std::max(1e-4, a);
		const double a2 = a * a, den = noH * noH * (a2 - 1.0) + 1.0, v = a2 / (PI * den * den);
		return std::isfinite(v) ? v : 0.0;
	}
	double	g1(double noW, double a)	// :220-236
	{
		noW = clamp01(noW);
		a = std::max(1e-4, a);
		if (noW <= 0.0) return 0.0;
		const double a2 = a * a, den = noW + std::sqrt(a2 + (1.0 - a2) * noW * noW);
		if (den <= 1e-10 || !std::isfinite(den)) return 0.0;
		return 2.0 * noW / den;
	}
	struct	Onb	// ONB.cpp:5-13
	{
		V	u, v, w;
		explicit Onb(V n)
		{
			w = normalize(n);
			const V a = (std::fabs(w.x) > 0.9) ? V{0, 1, 0} : V{1, 0, 0};
			v = normalize(cross(w, a));
			u = cross(w, v);
		}
		V	local(double a, double b, double c) const { return a * u + b * v + c * w; }
	};
	V	ggxHalf(V n, double a, float sx, float sy)	// :192-207
	{
		a = std::max(1e-4, a);
		const double a2 = a * a, phi = 2.0 * PI * clamp01(sx), u = std::min(1.0 - 1e-7, clamp01(sy));
		const double ct = std::sqrt((1.0 - u) / (1.0 + (a2 - 1.0) * u));
		const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
		return Onb(n).local(st * std::cos(phi), st * std::sin(phi), ct);
	}
	double	ggxReflPdf(V n, V view, V l, double a)	// :243-267
	{
		const double noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0 || noL <= 0.0) return 0.0;
		const V h = safeNormalize(view + l, n);
		const double voH = dot(view, h), noH = dot(n, h);
		if (voH <= 1e-10 || noH <= 0.0) return 0.0;
		const double pdf = ggxD(noH, a) * noH / (4.0 * voH);
		return (std::isfinite(pdf) && pdf > 0.0) ? pdf : 0.0;
	}
	C	ggxReflBsdf(C F, V n, V view, V l, double a)	// :269-287
	{
		const double noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0 || noL <= 0.0) return {};
		const V h = safeNormalize(view + l, n);
		const double sc = ggxD(dot(n, h), a) * (g1(noV, a) * g1(noL, a)) / (4.0 * noV);
		if (!std::isfinite(sc) || sc <= 0.0) return {};
		return clamp01(F) * sc;
	}
	C	conductorFresnel(C eta, C k, double c)	// BSDF.cpp:127-174
	{
		auto channel = [](double e, double kk, double cv) {
			if (!std::isfinite(e) || !std::isfinite(kk) || e <= 0.0 || kk < 0.0) return 0.0;
			const double c2 = cv * cv;
			const double ek = e * e + kk * kk;
			const double tec = 2.0 * e * cv;
			const double dPar = ek * c2 + tec + 1.0;
			const double dPerp = ek + tec + c2;
			if (std::fabs(dPar) <= 1e-10 || std::fabs(dPerp) <= 1e-10) return 0.0;	// MIN_DENOMINATOR (BSDF.cpp:11)
			const double rPar = (ek * c2 - tec + 1.0) / dPar;
			const double rPerp = (ek - tec + c2) / dPerp;
			return clamp01(0.5 * (rPar + rPerp));
		};
		c = clamp01(c);
		return {channel(eta.r, k.r, c), channel(eta.g, k.g, c), channel(eta.b, k.b, c)};
	}
	C	metalBsdf(const Mat& m, V n, V view, V l)	// Metal.cpp:172-200: BSDF.cpp:291-323
	{
		const V h = safeNormalize(view + l, n);
		const double voH = std::max(0.0, dot(view, h));
		return ggxReflBsdf(m.conductor ? conductorFresnel(m.eta, m.k, voH) : schlick(clamp01(m.color), voH), n, view, l, m.alpha);
	}
	bool	refractDir(V i, V mn, double eta, V& out)	// :325-350
	{
		const double c = std::min(1.0, dot(i * -1.0, mn));
		const V perp = eta * (i + c * mn);
		const double p2 = len2(perp);
		if (p2 > 1.0) return false;
		const V par = -std::sqrt(std::max(0.0, 1.0 - p2)) * mn;
		out = safeNormalize(perp + par, reflect(i, mn));
		return len2(out) > 0.0;
	}
	double	dielPdf(V n, V view, V l, double a, double ei, double et)	// :352-392
	{
		const double noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0 || noL == 0.0 || ei <= 0.0 || et <= 0.0) return 0.0;
		if (noL > 0.0)
		{
			const V h = safeNormalize(view + l, n);
			return dielectricFresnel(dot(view, h), ei, et) * ggxReflPdf(n, view, l, a);
		}
		const double eta = ei / et;
		V h = safeNormalize(view + l * eta, n);
		if (dot(n, h) < 0.0) h = h * -1.0;
		const double noH = dot(n, h), voH = dot(view, h), loH = dot(l, h);
		if (noH <= 0.0 || voH <= 0.0 || loH >= 0.0) return 0.0;
		const double den = voH + eta * loH;
		if (std::fabs(den) <= 1e-10) return 0.0;
		const double F = dielectricFresnel(voH, ei, et);
		const double pdf = (1.0 - F) * ggxD(noH, a) * noH * std::fabs((eta * eta * loH) / (den * den));
		return (std::isfinite(pdf) && pdf > 0.0) ? pdf : 0.0;
	}
	C	dielBsdf(C color, V n, V view, V l, double a, double ei, double et)	// :394-458
	{
		const double noV = dot(n, view), noL = dot(n, l);
		if (noV <= 0.0 || noL == 0.0 || ei <= 0.0 || et <= 0.0) return {};
		color = clamp01(color);
		if (noL > 0.0)
		{
			const V h = safeNormalize(view + l, n);
			const double F = dielectricFresnel(dot(view, h), ei, et);
			return ggxReflBsdf({F, F, F}, n, view, l, a);
		}
		const double eta = ei / et;
		V h = safeNormalize(view + l * eta, n);
		if (dot(n, h) < 0.0) h = h * -1.0;
		const double noH = dot(n, h), voH = dot(view, h), loH = dot(l, h);
		if (noH <= 0.0 || voH <= 0.0 || loH >= 0.0) return {};
		const double den = voH + eta * loH;
		if (std::fabs(den) <= 1e-10) return {};
		const double F = dielectricFresnel(voH, ei, et);
		const double sc = (1.0 - F) * ggxD(noH, a) * (g1(noV, a) * g1(std::fabs(noL), a)) * eta * eta * std::fabs(voH * loH)
			/ (noV * den * den);
		if (!std::isfinite(sc) || sc <= 0.0) return {};
		return color * sc;
	}

	// ---- scatter (Material.hpp:23-41 + each material's scatter) -----------------------------------
	enum PdfType { PDF_NONE, PDF_COSINE, PDF_SPHERE, PDF_HG, PDF_BSDF };
	struct	Scatter
	{
		PdfType	type = PDF_NONE;
		bool	specular = false;
		C	att;
		V	dir;		// specular ray direction (normalized) or sampled BSDF direction
		V	incident;
		Onb	basis = Onb(V{0, 0, 1});
		V	phaseDir;	// Henyey-Greenstein
		double	g = 0.0;
		bool	mirrorFallback = false;	// rough metal: the GGX sample reflected below the horizon
		bool	absorbing = false;	// ScatterRecord::hasMediumAbsorption
		C	absorption;			// ScatterRecord::mediumAbsorptionCoefficient
		bool	subsurface = false;	// ScatterRecord::hasSubsurface (the subsurface lobe was sampled)
	};

	// ---- glossy / diffuse-glossy / principled (Glossy.cpp, DiffuseGlossy.cpp, Principled.cpp) -----
	double	roughnessToAlpha(double r) { const double c = clamp01(r); return std::max(1e-4, c * c); }	// BSDF.cpp:54-58
	double	cosPdf(V n, V l) { const double c = dot(n, l); return c <= 0.0 ? 0.0 : c / PI; }
	C	schlickBsdf(C f0, V n, V view, V l, double a)	// BSDF.cpp:312-330
	{
		const V h = safeNormalize(view + l, n);
		return ggxReflBsdf(schlick(f0, std::max(0.0, dot(view, h))), n, view, l, a);
	}
	C	scaledC(C c, double f) { return (!std::isfinite(f) || f <= 0.0) ? C{} : c * f; }	// Principled.cpp:293-300

	struct	Lobes { double w[7] = {}; double total = 0.0; };
	double	lumWeight(C c) { const double l = luminance(c); return (!std::isfinite(l) || l <= 0.0) ? 0.0 : l; }
	double	dielF0(double ior) { const double r0 = (ior - 1.0) / (ior + 1.0); return clamp01(r0 * r0); }
	double	principledDiffuseScale(const Mat& m)	// :141-156
	{
		const double ce = 0.25 * m.clearcoat;
		return (1.0 - m.metallic) * (1.0 - m.transmission) * (1.0 - dielF0(m.ior)) * (1.0 - ce);
	}
	Lobes	principledLobes(const Mat& m, C base)	// componentWeights, :158-202
	{
		Lobes r;
		const double baseLum = std::max(0.05, lumWeight(base));
		const double diffuse = principledDiffuseScale(m);
		const double spec = (1.0 - m.metallic) * std::max(0.03, dielF0(m.ior));
		const double subsurfaceWeight = diffuse * m.subsurface;
		const double surfaceDiffuse = diffuse * (1.0 - m.subsurface);
		r.w[0] = surfaceDiffuse * baseLum;
		r.w[1] = subsurfaceWeight * std::max(0.05, lumWeight(base * clamp01(m.subsurfaceColor)));
		r.w[2] = diffuse * m.sheen * 0.35;
		r.w[3] = spec * (1.0 - 0.5 * m.transmission);
		r.w[4] = m.metallic * std::max(0.05, baseLum);
		r.w[5] = (1.0 - m.metallic) * m.transmission;
		r.w[6] = m.clearcoat * 0.25 * (1.0 - 0.5 * m.roughness);
		for (double x : r.w) if (std::isfinite(x) && x > 0.0) r.total += x;
		if (r.total <= 0.0) { r.w[0] = 1.0; r.total = 1.0; }
		return r;
	}
	C	principledEval(const Mat& m, const Hit& h, V view, V l)	// evaluateBSDFCos, :763-853
	{
		const C base = clamp01(m.color);
		const double a = roughnessToAlpha(m.roughness), ca = roughnessToAlpha(m.clearcoatRoughness);
		const double transmission = m.transmission * (1.0 - m.metallic);
		const double diffuse = principledDiffuseScale(m);
		const V n = h.n;
		C r;
		{	// diffuseBSDFCos, :222-234
			const double noL = dot(n, l), sc = diffuse * (1.0 - m.subsurface);
			if (!(noL <= 0.0 || sc <= 0.0)) r = base * (sc * noL / PI);
		}
		{	// subsurfaceBSDFCos, :251-271 (two-sided when thin)
			double noL = dot(n, l);
			if (m.subsurfaceThin) noL = std::fabs(noL);
			const double sc = diffuse * m.subsurface;
			if (!(noL <= 0.0 || sc <= 0.0)) r = r + (base * clamp01(m.subsurfaceColor)) * (sc * noL / PI);
		}
		{	// sheenBSDFCos, :272-291
			const double noL = dot(n, l), sc = diffuse * m.sheen;
			if (!(noL <= 0.0 || sc <= 0.0))
			{
				const V hv = safeNormalize(view + l, n);
				const double fall = std::pow(1.0 - clamp01(dot(l, hv)), 5.0);
				const C sheenColor = C{1, 1, 1} * (1.0 - 0.5) + base * 0.5;
				r = r + sheenColor * (sc * fall * noL / PI);
			}
		}
		if (dot(n, l) > 0.0)	// dielectricReflectionBSDFCos, :302-327
		{
			const V hv = safeNormalize(view + l, n);
			const double F = dielectricFresnel(dot(view, hv), 1.0, m.ior);
			r = r + scaledC(ggxReflBsdf(C{F, F, F}, n, view, l, a), 1.0 - m.metallic);
		}
		r = r + scaledC(schlickBsdf(base, n, view, l, a), m.metallic);
		if (dot(n, l) < 0.0 && transmission > 0.0)
		{
			const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0;
			r = r + scaledC(dielBsdf(base, n, view, l, a, ei, et), transmission);
		}
		r = r + scaledC(schlickBsdf(C{0.04, 0.04, 0.04}, n, view, l, ca), m.clearcoat * 0.25);
		return r;
	}
	double	principledPdf(const Mat& m, const Hit& h, V view, V l)	// scatteringPDF, :855-921
	{
		const Lobes w = principledLobes(m, clamp01(m.color));
		const double a = roughnessToAlpha(m.roughness), ca = roughnessToAlpha(m.clearcoatRoughness);
		const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0;
		double pdfs[7] = {};
		pdfs[0] = cosPdf(h.n, l);
		// subsurfaceCosinePDF (:236-249)
		pdfs[1] = m.subsurfaceThin ? std::fabs(dot(h.n, l)) / (2.0 * PI) : cosPdf(h.n, l);
		pdfs[2] = pdfs[0];
		pdfs[3] = ggxReflPdf(h.n, view, l, a);
		pdfs[4] = pdfs[3];
		if (dot(h.n, l) < 0.0) pdfs[5] = dielPdf(h.n, view, l, a, ei, et);
		pdfs[6] = ggxReflPdf(h.n, view, l, ca);
		double pdf = 0.0;
		for (int i = 0; i < 7; i++) if (w.w[i] > 0.0 && pdfs[i] > 0.0) pdf += (w.w[i] / w.total) * pdfs[i];
		return (std::isfinite(pdf) && pdf > 0.0) ? pdf : 0.0;
	}
	C	diffuseGlossyEval(const Mat& m, const Hit& h, V view, V l)	// DiffuseGlossy.cpp:151-168
	{
		const double noL = dot(h.n, l);
		const C diffuse = noL <= 0.0 ? C{} : clamp01(m.color) * (noL / PI);
		const C glossy = ggxReflBsdf(clamp01(m.glossyColor), h.n, view, l, roughnessToAlpha(m.roughness));
		return diffuse * (1.0 - m.glossyWeight) + glossy * m.glossyWeight;
	}
	double	diffuseGlossyPdf(const Mat& m, const Hit& h, V view, V l)	// :170-186
	{
		return cosPdf(h.n, l) * (1.0 - m.glossyWeight) + ggxReflPdf(h.n, view, l, roughnessToAlpha(m.roughness)) * m.glossyWeight;
	}

	double	bsdfPdf(const Mat& m, const Hit& h, const Scatter& sc, V l)
	{
		const V view = safeNormalize(sc.incident * -1.0, h.n);
		if (m.kind == luzgpu::MAT_METAL || m.kind == luzgpu::MAT_GLOSSY) return ggxReflPdf(h.n, view, l, m.alpha);	// Metal.cpp:202-216
		if (m.kind == luzgpu::MAT_DIFFUSE_GLOSSY) return diffuseGlossyPdf(m, h, view, l);
		if (m.kind == luzgpu::MAT_PRINCIPLED) return principledPdf(m, h, view, l);
		const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0;		// Dielectric.cpp:230-245
		return dielPdf(h.n, view, l, m.alpha, ei, et);
	}
	C	bsdfEval(const Mat& m, const Hit& h, const Scatter& sc, V l)
	{
		const V view = safeNormalize(sc.incident * -1.0, h.n);
		if (m.kind == luzgpu::MAT_METAL) return metalBsdf(m, h.n, view, l);	// Metal.cpp:172-200
		if (m.kind == luzgpu::MAT_GLOSSY) return ggxReflBsdf(clamp01(m.color), h.n, view, l, m.alpha);	// Glossy.cpp:85-100
		if (m.kind == luzgpu::MAT_DIFFUSE_GLOSSY) return diffuseGlossyEval(m, h, view, l);
		if (m.kind == luzgpu::MAT_PRINCIPLED) return principledEval(m, h, view, l);
		const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0;
		return dielBsdf(m.color, h.n, view, l, m.alpha, ei, et);				// Dielectric.cpp:210-228
	}
	double	scatterPdf(const Mat& m, const Hit& h, const Scatter& sc, V l)	// ColorHelper.cpp:39-93
	{
		if (sc.type == PDF_COSINE)
		{
			const double c = dot(l, sc.basis.w);
			return c <= 0.0 ? 0.0 : c / PI;
		}
		if (sc.type == PDF_SPHERE) return 1.0 / (4.0 * PI);
		if (sc.type == PDF_HG)	// :56-77
		{
			const double dl2 = len2(l), pl2 = len2(sc.phaseDir);
			if (dl2 <= 0.0 || pl2 <= 0.0 || !std::isfinite(dl2) || !std::isfinite(pl2)) return 0.0;
			const V nd = l / std::sqrt(dl2), pd = sc.phaseDir / std::sqrt(pl2);
			const double c = std::max(-1.0, std::min(1.0, dot(nd, pd)));
			const double g = std::max(-0.99, std::min(0.99, sc.g));
			const double g2 = g * g;
			const double den = std::max(1e-12, 1.0 + g2 - (2.0 * g * c));
			return (1.0 - g2) / (4.0 * PI * den * std::sqrt(den));
		}
		if (sc.type == PDF_BSDF) return bsdfPdf(m, h, sc, l);
		return 0.0;
	}
	C	scatterBsdfCos(const Mat& m, const Hit& h, const Scatter& sc, V l)	// ColorHelper.cpp:95-113
	{
		if (sc.type == PDF_BSDF) return bsdfEval(m, h, sc, l);
		return sc.att * scatterPdf(m, h, sc, l);
	}

	bool	scatter(const Mat& m, const Hit& h, const Ray& r, const Sampler& ps, Scatter& sc)
	{
		sc.incident = r.d;
		if (m.kind == luzgpu::MAT_LAMBERTIAN)	// Lambertian.cpp:15-24
		{
			sc.att = m.color;
			sc.type = PDF_COSINE;
			sc.basis = Onb(h.n);
			return true;
		}
		auto ggxReflect = [&](double a) {	// Glossy.cpp:22-40
			const V view = safeNormalize(r.d * -1.0, h.n);
			float sx, sy;
			ps.s2(D_BSDF_DIRECTION, sx, sy);
			V half = ggxHalf(h.n, a, sx, sy);
			if (dot(view, half) <= 0.0) half = h.n;
			V dir = reflect(r.d, half);
			if (dot(dir, h.n) <= 0.0) dir = reflect(r.d, h.n);
			return dir;
		};
		auto cosineDir = [&](bool flip = false) {	// Sampler::cosineHemisphere in ONB(normal)
			float sx, sy;
			ps.s2(D_BSDF_DIRECTION, sx, sy);
			const float phi = static_cast<float>(2.0 * PI) * sx;
			const float rr = std::sqrt(sy);
			const float z = std::sqrt(std::max(0.0f, 1.0f - sy));
			return Onb(h.n).local(static_cast<double>(rr * std::cos(phi)), static_cast<double>(rr * std::sin(phi)),
				flip ? -static_cast<double>(z) : static_cast<double>(z));
		};
		auto accept = [&](V dir) {
			sc.type = PDF_BSDF;
			const double pdf = bsdfPdf(m, h, sc, dir);
			const C f = bsdfEval(m, h, sc, dir);
			if (pdf <= 0.0 || !std::isfinite(pdf) || maxChannel(f) <= 0.0) { sc.type = PDF_NONE; return false; }
			sc.dir = dir;
			sc.att = f / pdf;
			return true;
		};
		if (m.kind == luzgpu::MAT_GLOSSY) return accept(ggxReflect(m.alpha));	// Glossy.cpp:64-83
		if (m.kind == luzgpu::MAT_DIFFUSE_GLOSSY)	// DiffuseGlossy.cpp:122-149
		{
			V dir;
			if (ps.s1(D_MATERIAL_DECISION) < m.glossyWeight) dir = ggxReflect(roughnessToAlpha(m.roughness));
			else dir = cosineDir();
			return accept(dir);
		}
		if (m.kind == luzgpu::MAT_PRINCIPLED)	// Principled.cpp:668-761
		{
			// setMediumAbsorption (:129-140)
			sc.absorbing = m.transmission > 0.0 && (m.absorption.r > 0.0 || m.absorption.g > 0.0 || m.absorption.b > 0.0);
			sc.absorption = m.absorption;
			const C base = clamp01(m.color);
			if (m.metallic <= 1e-6 && m.roughness <= 1e-4 && m.transmission >= 1.0 - 1e-6 && m.clearcoat <= 1e-6 && m.sheen <= 1e-6
				&& m.subsurface <= 1e-6)
			{
				// scatterSmoothTransmission (:401-440)
				const V view = safeNormalize(r.d * -1.0, h.n);
				const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0, ratio = ei / et;
				const double ct = std::min(1.0, dot(view, h.n));
				const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
				V dir;
				if (ratio * st > 1.0 || dielectricFresnel(ct, ei, et) > ps.s1(D_MATERIAL_DECISION))
				{
					dir = reflect(r.d, h.n);
					sc.att = {1, 1, 1};
				}
				else
				{
					const double c = std::fmin(dot(r.d * -1.0, h.n), 1.0);
					const V perp = ratio * (r.d + c * h.n);
					const V par = -std::sqrt(std::fabs(1.0 - len2(perp))) * h.n;
					dir = perp + par;
					sc.att = base;
				}
				sc.dir = Ray::make(h.p, dir).d;
				sc.specular = true;
				return true;
			}
			const Lobes w = principledLobes(m, base);
			double target = ps.s1(D_MATERIAL_DECISION) * w.total;	// selectComponent (:204-216)
			int comp = 0;
			for (int i = 0; i < 7; i++)
			{
				target -= w.w[i];
				if (target <= 0.0) { comp = i; break; }
			}
			V dir;
			// the thin subsurface lobe's far side (:156-167): the local z mirrored
			if (comp <= 2) dir = cosineDir(comp == 1 && m.subsurfaceThin && ps.s1(D_SUBSURFACE_THIN_SIDE) < 0.5);
			else if (comp == 5)	// sampleGGXTransmission (:359-399)
			{
				const V view = safeNormalize(r.d * -1.0, h.n);
				const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0;
				float sx, sy;
				ps.s2(D_BSDF_DIRECTION, sx, sy);
				V half = ggxHalf(h.n, roughnessToAlpha(m.roughness), sx, sy);
				if (dot(view, half) <= 0.0) half = h.n;
				if (!refractDir(r.d, half, ei / et, dir)) dir = reflect(r.d, half);
			}
			else if (comp == 6) dir = ggxReflect(roughnessToAlpha(m.clearcoatRoughness));
			else dir = ggxReflect(roughnessToAlpha(m.roughness));
			sc.subsurface = comp == 1;
			return accept(dir);
		}
		if (m.kind == luzgpu::MAT_ISOTROPIC)	// Isotropic.cpp:15-24
		{
			sc.att = m.color;
			sc.type = PDF_SPHERE;
			return true;
		}
		if (m.kind == luzgpu::MAT_HG)	// HenyeyGreenstein.cpp:54-63
		{
			sc.att = m.color;
			sc.type = PDF_HG;
			sc.phaseDir = len2(r.d) <= 0.0 ? V{0, 0, 1} : normalize(r.d);
			sc.g = m.g;
			return true;
		}
		if (m.kind == luzgpu::MAT_METAL)	// Metal.cpp:126-170
		{
			const V view = safeNormalize(r.d * -1.0, h.n);
			const double ct = std::min(1.0, std::max(0.0, dot(view, h.n)));
			if (m.roughness <= 1e-4)
			{
				sc.dir = Ray::make(h.p, reflect(r.d, h.n)).d;
				sc.att = m.conductor ? conductorFresnel(m.eta, m.k, ct) : schlick(clamp01(m.color), ct);
				sc.specular = true;
				return true;
			}
			float sx, sy;
			ps.s2(D_BSDF_DIRECTION, sx, sy);
			V half = ggxHalf(h.n, m.alpha, sx, sy);
			if (dot(view, half) <= 0.0)
			{
				half = h.n;
				sc.mirrorFallback = true;
			}
			V dir = reflect(r.d, half);
			if (dot(dir, h.n) <= 0.0)
			{
				dir = reflect(r.d, h.n);
				sc.mirrorFallback = true;
			}
			const double pdf = ggxReflPdf(h.n, view, dir, m.alpha);
			const C f = metalBsdf(m, h.n, view, dir);
			if (pdf <= 0.0 || !std::isfinite(pdf) || maxChannel(f) <= 0.0) return false;
			sc.dir = dir;
			sc.att = f / pdf;
			sc.type = PDF_BSDF;
			return true;
		}
		if (m.kind == luzgpu::MAT_DIELECTRIC)	// Dielectric.cpp:127-208
		{
			sc.absorbing = m.absorption.r > 0.0 || m.absorption.g > 0.0 || m.absorption.b > 0.0;	// :129-134
			sc.absorption = m.absorption;
			const V view = safeNormalize(r.d * -1.0, h.n);
			const double ei = h.front ? 1.0 : m.ior, et = h.front ? m.ior : 1.0, ratio = ei / et;
			if (m.roughness <= 1e-4)
			{
				const double ct = std::fmin(dot(view, h.n), 1.0);
				const double st = std::sqrt(1.0 - ct * ct);
				V dir;
				if (ratio * st > 1.0 || dielectricFresnel(ct, ei, et) > ps.s1(D_MATERIAL_DECISION))
				{
					dir = reflect(r.d, h.n);
					sc.att = {1, 1, 1};
				}
				else
				{
					const double c = std::fmin(dot(r.d * -1.0, h.n), 1.0);	// Utilities.hpp:59-66
					const V perp = ratio * (r.d + c * h.n);
					const V par = -std::sqrt(std::fabs(1.0 - len2(perp))) * h.n;
					dir = perp + par;
					sc.att = m.color;
				}
				sc.dir = Ray::make(h.p, dir).d;
				sc.specular = true;
				return true;
			}
			float sx, sy;
			ps.s2(D_BSDF_DIRECTION, sx, sy);
			V half = ggxHalf(h.n, m.alpha, sx, sy);
			if (dot(view, half) <= 0.0) half = h.n;
			const double F = dielectricFresnel(dot(view, half), ei, et);
			V refr;
			const bool can = refractDir(r.d, half, ratio, refr);
			const V dir = (!can || ps.s1(D_MATERIAL_DECISION) < F) ? reflect(r.d, half) : refr;
			const double pdf = dielPdf(h.n, view, dir, m.alpha, ei, et);
			const C f = dielBsdf(m.color, h.n, view, dir, m.alpha, ei, et);
			if (pdf <= 0.0 || !std::isfinite(pdf) || maxChannel(f) <= 0.0) return false;
			sc.dir = dir;
			sc.att = f / pdf;
			sc.type = PDF_BSDF;
			return true;
		}
		return false;	// Emissive / Material::scatter (Material.cpp:40-46)
	}

	// ---- lights and environment -----------------------------------------------------------------------
	bool	leavesOpaque(const Mat& m, const Hit& h, V d)	// ColorHelper.cpp:357-374
	{
		// materialCanTransmitThroughGeometry (:327-355)
		if (m.kind == luzgpu::MAT_DIELECTRIC || m.kind == luzgpu::MAT_ISOTROPIC || m.kind == luzgpu::MAT_HG
			|| (m.kind == luzgpu::MAT_PRINCIPLED && (m.transmission > 0.0 || (m.subsurface > 0.0 && m.subsurfaceThin)))) return true;
		V g = h.gn;
		if (len2(g) <= 1e-12) g = h.n;
		if (len2(g) <= 1e-12) return true;
		return dot(g, d) > 1e-7;
	}

	double	selectionProbability(const Scene& s, std::size_t i)	// ColorHelper.cpp:176-199
	{
		if (s.lights.empty()) return 0.0;
		if (s.lightCdf.empty() || !(s.lightTotal > 0.0)) return 1.0 / static_cast<double>(s.lights.size());
		const double prev = i == 0 ? 0.0 : s.lightCdf[i - 1];
		const double w = s.lightCdf[i] - prev;
		if (w <= 0.0 || s.lightTotal <= 0.0) return 0.0;
		return w / s.lightTotal;
	}

	struct	LS
	{
		V	dir;
		C	emitted;
		double	pdf = 0, tMax = 0;
		bool	valid = false;
		std::int64_t	material = -1;
	};

	// IESProfile::relativeIntensity (IESProfile.cpp:302-323): the candela towards `emission`
	// over the profile's mean, the vertical angle measured from the (normalised) axis, the
	// horizontal one in the frame cross(helper, axis), cross(axis, tangent), turned by
	// `rotation` degrees. The grid lookup is the reader's iesCandelaAt (IESProfile.cpp:245-300).
	double	iesRelativeIntensity(const luzgpu::IesProfile& pr, V emission, V axisIn, double rotation)
	{
		if (pr.meanCandela <= 0.0 || !std::isfinite(pr.meanCandela)) return 0.0;
		auto normalizedOr = [](V v, V fallback) {
			const double l2 = len2(v);
			if (!std::isfinite(l2) || l2 <= 0.0) return fallback;
			return v / std::sqrt(l2);
		};
		const V axis = normalizedOr(axisIn, V{0.0, -1.0, 0.0});
		const V d = normalizedOr(emission, axis);
		const double cosT = std::clamp(dot(d, axis), -1.0, 1.0);
		const double toDegrees = 180.0 / PI;
		const double vdeg = std::acos(cosT) * toDegrees;
		const V helper = std::fabs(axis.y) < 0.95 ? V{0.0, 1.0, 0.0} : V{1.0, 0.0, 0.0};
		const V tangent = normalizedOr(cross(helper, axis), V{1.0, 0.0, 0.0});
		const V bitangent = cross(axis, tangent);
		double h = std::atan2(dot(d, bitangent), dot(d, tangent)) * toDegrees + rotation;
		if (!std::isfinite(h)) h = 0.0;
		else
		{
			h = std::fmod(h, 360.0);
			if (h < 0.0) h += 360.0;
		}
		return luzgpu::iesCandelaAt(pr, vdeg, h) / pr.meanCandela;
	}

	LS	sampleLight(const Scene& s, V origin, const Sampler& ps)	// ColorHelper.cpp:914-971
	{
		LS out;
		if (s.lights.empty()) return out;
		std::size_t idx = 0;
		if (s.lights.size() > 1)	// :201-231
		{
			const double u = ps.s1(D_LIGHT_SELECTION);
			if (!s.lightCdf.empty() && s.lightTotal > 0.0)
			{
				const auto it = std::upper_bound(s.lightCdf.begin(), s.lightCdf.end(), u * s.lightTotal);
				idx = std::min<std::size_t>(static_cast<std::size_t>(it - s.lightCdf.begin()), s.lights.size() - 1);
			}
			else
			{
				idx = std::min<std::size_t>(static_cast<std::size_t>(u * s.lights.size()), s.lights.size() - 1);
			}
		}
		const double sel = selectionProbability(s, idx);
		if (sel <= 0.0 || !std::isfinite(sel)) return out;
		const Light& L = s.lights[idx];
		out.material = L.material;
		if (L.kind == K_DIR)	// DirectionalLight::sampleLight (DirectionalLight.cpp:74-88)
		{
			out.dir = L.dir;
			out.pdf = 1.0 * sel;
			out.tMax = T_MAX;
			out.emitted = L.emitted;
			out.valid = len2(out.dir) > 0.0 && maxChannel(out.emitted) > 0.0;
			return out;
		}
		if (L.kind == K_SPHERE)	// Sphere::sampleLight (Sphere.cpp:382-442)
		{
			const Sphere& sp = s.spheres[L.index];
			const V toC = sp.c - origin;
			const double d2 = len2(toC);
			if (sp.r <= 0.0 || !std::isfinite(d2) || d2 <= sp.r * sp.r) return out;
			float sx, sy;
			ps.s2(D_LIGHT_SURFACE_POINT, sx, sy);
			const float rr = static_cast<float>(sp.r * sp.r / d2);	// :365-380
			const float z = 1.0f + sy * (std::sqrt(std::max(0.0f, 1.0f - rr)) - 1.0f);
			const float phi = static_cast<float>(2.0 * PI) * sx;
			const float rad = std::sqrt(std::max(0.0f, 1.0f - z * z));
			const V local = {static_cast<double>(std::cos(phi) * rad), static_cast<double>(std::sin(phi) * rad), static_cast<double>(z)};
			const Onb uvw(toC);
			V dir = local.x * uvw.u + local.y * uvw.v + local.z * uvw.w;
			const double dl2 = len2(dir);
			if (!std::isfinite(dl2) || dl2 <= 0.0) return out;
			dir = dir / std::sqrt(dl2);
			const V oc = origin - sp.c;
			const double a = len2(dir), b = dot(oc, dir), c = len2(oc) - (sp.r * sp.r);
			const double disc = (b * b) - (a * c);
			if (disc < 0.0) return out;
			const double sq = std::sqrt(disc);
			double root = (-b - sq) / a;
			if (root < T_MIN || root > T_MAX) root = (-b + sq) / a;
			if (root < T_MIN || root > T_MAX) return out;
			const double cosMax = std::sqrt(1.0 - sp.r * sp.r / d2);
			const double solid = 2.0 * PI * (1.0 - cosMax);
			if (solid <= 0.0 || !std::isfinite(solid)) return out;
			out.dir = dir;
			out.pdf = (1.0 / solid) * sel;
			out.tMax = root;
			if (out.pdf <= 0.0 || out.tMax <= T_MIN || !std::isfinite(out.pdf) || !std::isfinite(out.tMax)) return out;
			out.emitted = L.emitted;
			if (sp.ies >= 0)	// :451-461: the profile scales the sample's radiance
			{
				out.emitted = L.emitted * iesRelativeIntensity(s.iesProfiles[static_cast<std::size_t>(sp.ies)], dir * -1.0,
					sp.iesDirection, sp.iesRotation);
			}
			out.valid = maxChannel(out.emitted) > 0.0;
			return out;
		}
		if (L.kind == K_INST)	// MeshInstance::sampleLight / _sampleSurface (MeshInstance.cpp:409-466,510-546)
		{
			const Instance& inst = s.instances[L.index];
			const Mesh& m = s.meshes[inst.mesh];
			if (L.triCdf.empty() || L.area <= 0.0) return out;
			// Mesh::sampleSurface or the transformed prefix sums: lower_bound of u * total
			const double target = ps.s1(D_LIGHT_SURFACE_SELECTION) * L.triCdf.back();
			const auto it = std::lower_bound(L.triCdf.begin(), L.triCdf.end(), target);
			const std::size_t k = std::min<std::size_t>(static_cast<std::size_t>(it - L.triCdf.begin()), L.triCdf.size() - 1);
			const double prev = k == 0 ? 0.0 : L.triCdf[k - 1];
			if (L.triCdf[k] - prev <= 0.0) return out;
			const Tri& tr = m.tris[k];	// Triangle::sampleSurface (Triangle.cpp:320-341)
			if (tr.degenerate || tr.area <= 0.0) return out;
			float sx, sy;
			ps.s2(D_LIGHT_SURFACE_POINT, sx, sy);
			double u = sx, v = sy;
			if (u + v > 1.0) { u = 1.0 - u; v = 1.0 - v; }
			const V local = tr.v0 + ((tr.v1 - tr.v0) * u) + ((tr.v2 - tr.v0) * v);
			const V position = inst.transformPoint(local);
			const V normal = inst.transformNormal(tr.face);
			if (!(std::isfinite(position.x) && std::isfinite(position.y) && std::isfinite(position.z)) || len2(normal) <= 0.0) return out;
			const V d = position - origin;
			const double d2 = len2(d);
			if (!std::isfinite(d2) || d2 <= 0.0) return out;
			const double dist = std::sqrt(d2);
			out.dir = d / dist;
			const double c = std::fabs(dot(out.dir, normal));
			if (!std::isfinite(c) || c <= 0.0) return out;
			const double pdf = d2 / (c * L.area);
			if (!(std::isfinite(pdf) && pdf > 0.0 && std::isfinite(dist) && dist > T_MIN)) return out;
			out.pdf = pdf * sel;
			out.tMax = dist;
			out.emitted = L.emitted;
			out.valid = maxChannel(out.emitted) > 0.0;
			return out;
		}
		if (L.kind != K_RECT && L.kind != K_TRI) return out;	// Hittable::sampleLight default
		float sx, sy;
		ps.s2(D_LIGHT_SURFACE_POINT, sx, sy);
		V point, normal;
		if (L.kind == K_RECT)	// Rectangle.cpp:381-419
		{
			const Rect& r = s.rects[L.index];
			point = r.pos + (r.wa * ((sx - 0.5) * r.w)) + (r.ha * ((sy - 0.5) * r.h));
			normal = r.n;
		}
		else	// Triangle.cpp:391-432
		{
			const Tri& t = s.topTris[L.index];
			double u = sx, v = sy;
			if (u + v > 1.0) { u = 1.0 - u; v = 1.0 - v; }
			point = t.v0 + ((t.v1 - t.v0) * u) + ((t.v2 - t.v0) * v);
			normal = t.face;
		}
		const V d = point - origin;
		const double d2 = len2(d);
		if (!std::isfinite(d2) || d2 <= 0.0) return out;
		const double dist = std::sqrt(d2);
		out.dir = d / dist;
		const double c = std::fabs(dot(out.dir, normal));
		if (L.area <= 0.0 || c <= 0.0) return out;
		const double pdf = d2 / (c * L.area);
		if (!(std::isfinite(pdf) && pdf > 0.0)) return out;
		out.pdf = pdf * sel;
		out.tMax = dist;
		if (out.pdf <= 0.0 || out.tMax <= T_MIN || !std::isfinite(out.pdf) || !std::isfinite(out.tMax)) return out;
		out.emitted = L.emitted;
		if (maxChannel(out.emitted) <= 0.0) return out;
		out.valid = true;
		return out;
	}

	double	lightPdfValue(const Scene& s, const Light& L, V origin, V vec)	// Rectangle.cpp:345-359, Triangle.cpp:343-370
	{
		const Ray r = Ray::make(origin, vec);
		if (L.kind == K_SPHERE)	// Sphere::pdfValue (Sphere.cpp:362-375)
		{
			const Sphere& sp = s.spheres[L.index];
			double t;
			if (!sp.intersect(r, T_MIN, T_MAX, t)) return 0.0;
			const double cosMax = std::sqrt(1.0 - sp.r * sp.r / len2(sp.c - origin));
			return 1.0 / (2.0 * PI * (1.0 - cosMax));
		}
		if (L.kind == K_INST)	// MeshInstance::pdfValue (MeshInstance.cpp:468-497): this instance alone
		{
			if (L.area <= 0.0) return 0.0;
			const double vl = std::sqrt(len2(vec));
			if (!std::isfinite(vl) || vl <= 0.0) return 0.0;
			const Instance& inst = s.instances[L.index];
			Ray local;
			double scale;
			if (!inst.localRay(r, local, scale)) return 0.0;
			const Mesh& m = s.meshes[inst.mesh];
			std::uint32_t tri;
			double lt, u, v;
			if (!m.bvh.trace(m.tris, local, T_MIN * scale, T_MAX * scale, false, tri, lt, u, v)) return 0.0;
			const double wt = lt / scale;
			if (!std::isfinite(wt) || wt < T_MIN || wt > T_MAX) return 0.0;
			const V gn = inst.transformNormal(m.tris[tri].face);
			const double d2 = wt * wt * len2(vec);
			const double c = std::fabs(dot(vec, gn) / vl);
			if (!std::isfinite(d2) || !std::isfinite(c) || c <= 0.0) return 0.0;
			return d2 / (c * L.area);
		}
		if (L.kind != K_RECT && L.kind != K_TRI) return 0.0;
		if (L.kind == K_RECT)
		{
			const Rect& rc = s.rects[L.index];
			double t, wd, hd;
			if (!rc.intersect(r, T_MIN, T_MAX, t, wd, hd)) return 0.0;
			Hit h;
			setFace(h, r, rc.n, h.n, h.front);
			const double d2 = t * t * len2(vec);
			const double c = std::fabs(dot(vec, h.n) / std::sqrt(len2(vec)));
			return d2 / (c * (rc.h * rc.w));
		}
		const Tri& tr = s.topTris[L.index];
		double t, u, v;
		if (!tr.intersect(r, T_MIN, T_MAX, t, u, v) || tr.area <= 0.0) return 0.0;
		Hit g;
		setFace(g, r, tr.face, g.n, g.front);
		const double d2 = t * t * len2(vec);
		const double c = std::fabs(dot(vec, g.n) / std::sqrt(len2(vec)));
		if (c <= 0.0) return 0.0;
		return d2 / (c * tr.area);
	}

	double	lightsPdf(const Scene& s, V origin, V dir)	// ColorHelper.cpp:233-257
	{
		if (s.lights.empty()) return 0.0;
		if (s.lights.size() == 1) return lightPdfValue(s, s.lights[0], origin, dir);
		double sum = 0.0;
		for (std::size_t i = 0; i < s.lights.size(); i++) sum += selectionProbability(s, i) * lightPdfValue(s, s.lights[i], origin, dir);
		return sum;
	}

	double	wrap01(double v) { v = v - std::floor(v); if (v < 0.0) v += 1.0; return v; }
	void	dirToUV(const Env& e, V d, double& u, double& v)	// EnvironmentMap.cpp:600-616
	{
		const double l2 = len2(d);
		if (!std::isfinite(l2) || l2 <= 0.0) { u = 0.0; v = 0.5; return; }
		const V n = d / std::sqrt(l2);
		const double phi = std::atan2(n.z, n.x);
		const double y = std::clamp(n.y, -1.0, 1.0);
		u = wrap01(0.5 + (phi / (2.0 * PI)) + (e.rotation / 360.0));
		v = clamp01(0.5 + (std::asin(y) / PI));
	}
	C	envLookup(const Env& e, V d)	// :456-466, sampleUV :574-598
	{
		double u, v;
		dirToUV(e, d, u, v);
		u = wrap01(u);
		v = clamp01(v);
		const double x = (u * static_cast<double>(e.W)) - 0.5, y = ((1.0 - v) * static_cast<double>(e.H)) - 0.5;
		const long long xf = static_cast<long long>(std::floor(x)), yf = static_cast<long long>(std::floor(y));
		auto wrapI = [](long long i, std::size_t n) { long long m = i % static_cast<long long>(n); if (m < 0) m += static_cast<long long>(n); return static_cast<std::size_t>(m); };
		const std::size_t x0 = wrapI(xf, e.W), x1 = wrapI(xf + 1, e.W);
		const std::size_t y0 = static_cast<std::size_t>(std::clamp<long long>(yf, 0, static_cast<long long>(e.H - 1)));
		const std::size_t y1 = static_cast<std::size_t>(std::clamp<long long>(yf + 1, 0, static_cast<long long>(e.H - 1)));
		const double tx = x - std::floor(x), ty = y - std::floor(y);
		const C top = e.px[y0 * e.W + x0] * (1.0 - tx) + e.px[y0 * e.W + x1] * tx;
		const C bot = e.px[y1 * e.W + x0] * (1.0 - tx) + e.px[y1 * e.W + x1] * tx;
		const C c = top * (1.0 - ty) + bot * ty;
		return {std::isfinite(c.r) ? std::max(0.0, c.r) : 0.0, std::isfinite(c.g) ? std::max(0.0, c.g) : 0.0,
			std::isfinite(c.b) ? std::max(0.0, c.b) : 0.0};
	}
	double	envPdf(const Env& e, V d)	// :513-539
	{
		if (e.total <= 0.0 || e.weights.empty()) return 1.0 / (4.0 * PI);
		double u, v;
		dirToUV(e, d, u, v);
		const std::size_t x = std::min(static_cast<std::size_t>(std::floor(wrap01(u) * static_cast<double>(e.W))), e.W - 1);
		const std::size_t y = std::min(static_cast<std::size_t>(std::floor((1.0 - clamp01(v)) * static_cast<double>(e.H))), e.H - 1);
		const std::size_t i = y * e.W + x;
		if (e.weights[i] <= 0.0 || e.solid[i] <= 0.0) return 0.0;
		return (e.weights[i] / e.total) / e.solid[i];
	}
	double	clampSample(double v)
	{
		if (!std::isfinite(v) || v <= 0.0) return 0.0;
		if (v >= 1.0) return std::nextafter(1.0, 0.0);
		return v;
	}

	bool	hasEnvLight(const Scene& s) { return s.hasEnv && s.env.lighting && s.env.strength > 0.0; }
	bool	hasConstBg(const Scene& s) { return !hasEnvLight(s) && s.sky == luzgpu::SKY_NONE && maxChannel(s.background) > 0.0; }

	LS	sampleInfinite(const Scene& s, V origin, const Sampler& ps)	// ColorHelper.cpp:831-903
	{
		LS out;
		if (hasEnvLight(s))
		{
			const Env& e = s.env;
			const double sel = ps.s1(D_ENV_SELECTION);
			float jx, jy;
			ps.s2(D_ENV_POINT, jx, jy);
			jx = static_cast<float>(clampSample(jx));
			jy = static_cast<float>(clampSample(jy));
			V dir;
			double pdf;
			if (e.total <= 0.0 || e.cdf.empty())
			{
				const double z = 1.0 - (2.0 * jy), rad = std::sqrt(std::max(0.0, 1.0 - z * z)), phi = 2.0 * PI * jx;
				dir = {rad * std::cos(phi), z, rad * std::sin(phi)};
				pdf = 1.0 / (4.0 * PI);
			}
			else
			{
				const double target = clampSample(sel) * e.total;
				auto it = std::lower_bound(e.cdf.begin(), e.cdf.end(), target);
				if (it == e.cdf.end()) it = e.cdf.end() - 1;
				const std::size_t index = static_cast<std::size_t>(it - e.cdf.begin());
				const std::size_t y = index / e.W, x = index % e.W;
				const double u = (static_cast<double>(x) + jx) / static_cast<double>(e.W);
				const double v = 1.0 - ((static_cast<double>(y) + jy) / static_cast<double>(e.H));
				const double phi = (wrap01(u - (e.rotation / 360.0)) - 0.5) * 2.0 * PI;	// :618-625
				const double yy = std::sin((clamp01(v) - 0.5) * PI);
				const double rad = std::sqrt(std::max(0.0, 1.0 - yy * yy));
				dir = {rad * std::cos(phi), yy, rad * std::sin(phi)};
				pdf = envPdf(e, dir);
			}
			if (!(pdf > 0.0 && std::isfinite(pdf))) return out;
			out.dir = dir;
			out.emitted = envLookup(e, dir) * e.strength;
			out.pdf = pdf;
			out.tMax = T_MAX;
			if (maxChannel(out.emitted) <= 0.0) return out;
			if (occluded(s, Ray::normalized(origin, out.dir), T_MAX, &ps)) return out;
			out.valid = true;
			return out;
		}
		if (hasConstBg(s))
		{
			float sx, sy;
			ps.s2(D_ENV_POINT, sx, sy);
			const float z = 1.0f - (2.0f * sy);	// Sampler::sphereDirection (Sampler.cpp:248-256)
			const float rr = std::sqrt(std::max(0.0f, 1.0f - z * z));
			const float phi = static_cast<float>(2.0 * PI) * sx;
			out.dir = {rr * std::cos(phi), rr * std::sin(phi), z};
			out.emitted = s.background;
			out.pdf = 1.0 / (4.0 * PI);
			out.tMax = T_MAX;
			if (maxChannel(out.emitted) <= 0.0) return out;
			if (occluded(s, Ray::normalized(origin, out.dir), T_MAX, &ps)) return out;
			out.valid = true;
		}
		return out;
	}

	C	clampRay(C c)	// ColorHelper.cpp:259-273
	{
		const double l = luminance(c);
		if (!std::isfinite(l)) return {};
		if (l > 1.0e12) c = c * (1.0e12 / l);
		return c;
	}
	C	clampBg(const Scene& s, C c)	// :290-300
	{
		const double bl = luminance(s.background);
		if (!std::isfinite(bl) || bl <= 0.0) return c;
		const double l = luminance(c), cap = bl * 4.0;
		if (!std::isfinite(l)) return {};
		if (cap > 0.0 && l > cap) c = c * (cap / l);
		return c;
	}
	double	power(double a, double b)	// :307-318
	{
		const double f = a * a, g = b * b, sum = f + g;
		if (sum <= 0.0 || !std::isfinite(sum)) return 0.0;
		return f / sum;
	}
	bool	terminated(C t) { const double m = maxChannel(t); return !std::isfinite(m) || m <= 1e-6; }	// :320-325
	bool	roulette(C& t, int b, const Sampler& ps)	// :674-693
	{
		if (b < 3) return true;
		double sp = std::max(0.05, std::min(0.95, maxChannel(t)));
		if (ps.s1(D_RR) > sp) return false;
		t = t / sp;
		return true;
	}

	C	direct(const Scene& s, const Mat& m, const Hit& h, const Scatter& sc, int b, const Sampler& ps, std::int64_t* lightMaterial = nullptr,
		V* dirOut = nullptr)	// :973-1033
	{
		if (s.lights.empty()) return {};
		const double prob = b < 2 ? 1.0 : 0.5;
		if (prob < 1.0 && ps.s1(D_LIGHT_STRATEGY) >= prob) return {};
		const LS ls = sampleLight(s, h.p, ps);
		if (!ls.valid || !leavesOpaque(m, h, ls.dir)) return {};
		const double sp = scatterPdf(m, h, sc, ls.dir);
		const C f = scatterBsdfCos(m, h, sc, ls.dir);
		if (sp <= 0.0 || !std::isfinite(sp) || maxChannel(f) <= 0.0) return {};
		if (occluded(s, Ray::normalized(h.p, ls.dir), ls.tMax - T_MIN, &ps)) return {};
		const double w = power(ls.pdf, sp);
		if (lightMaterial) *lightMaterial = ls.material;
		if (dirOut) *dirOut = ls.dir;
		return f * ls.emitted * (w / (ls.pdf * prob));
	}

	C	environment(const Scene& s, const Mat& m, const Hit& h, const Scatter& sc, int b, const Sampler& ps, V* dirOut = nullptr)	// :1035-1098
	{
		const bool cb = hasConstBg(s);
		if (cb && sc.type == PDF_BSDF) return {};
		const double prob = b < 2 ? 1.0 : 0.5;
		if (prob < 1.0 && ps.s1(D_ENV_STRATEGY) >= prob) return {};
		const LS ls = sampleInfinite(s, h.p, ps);
		if (!ls.valid || !leavesOpaque(m, h, ls.dir)) return {};
		const double sp = scatterPdf(m, h, sc, ls.dir);
		const C f = scatterBsdfCos(m, h, sc, ls.dir);
		if (sp <= 0.0 || !std::isfinite(sp) || maxChannel(f) <= 0.0) return {};
		const double w = power(ls.pdf, sp);
		C c = f * ls.emitted * (w / (ls.pdf * prob));
		if (cb) c = clampBg(s, c);
		if (dirOut) *dirOut = ls.dir;
		return c;
	}

	// ---- path-replay adjoint (detached sampling): per-vertex records resolved at the end --------
	// Probability that the metal sampler (Metal.cpp:126-170) falls back to the mirror
	// direction, for a view at cos mu and GGX alpha a, and its derivative in a. With the view
	// at polar angle tv and a half vector at polar angle th, azimuth phi from the view, the
	// fallback happens iff v.h <= cos(tv) / (2 cos th), i.e. cos(phi) <= -cot(tv) cot(2 th):
	// never for th <= pi/4 - tv/2, always for th >= pi/4 + tv/2, and for a fraction
	// 1 - acos(K)/pi of the azimuths in between. The sampler draws t = cos^2(th) with CDF
	// C(t) = a^2 t / (1 + (a^2 - 1) t), so P = C(t2) + integral over the band of the fraction
	// times the density; the band integral is Gauss-Legendre in s with th = mid + half sin(pi s / 2),
	// which smooths the square-root behaviour of the fraction at the band edges.
	struct	GaussLegendre
	{
		std::vector<double>	x, w;
		explicit GaussLegendre(int n)
		{
			for (int i = 0; i < n; i++)
			{
				double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
				for (int it = 0; it < 100; it++)
				{
					double p0 = 1.0, p1 = z;
					for (int k = 2; k <= n; k++)
					{
						const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
						p0 = p1;
						p1 = p2;
					}
					dp = n * (z * p1 - p0) / (z * z - 1.0);
					const double dz = p1 / dp;
					z -= dz;
					if (std::fabs(dz) < 1e-15) break;
				}
				x.push_back(z);
				w.push_back(2.0 / ((1.0 - z * z) * dp * dp));
			}
		}
	};
	double	mirrorFallback(double mu, double a, double* dPda, const GaussLegendre& gl)
	{
		a = std::max(1e-4, a);
		mu = std::min(1.0, std::max(0.0, mu));
		const double tv = std::acos(mu), cotv = mu / std::sqrt(std::max(1e-300, 1.0 - mu * mu)), a2 = a * a;
		const double th1 = 0.25 * PI - 0.5 * tv, th2 = 0.25 * PI + 0.5 * tv;
		const double t2 = std::cos(th2) * std::cos(th2), d2 = 1.0 + (a2 - 1.0) * t2;
		double P = a2 * t2 / d2, dP = 2.0 * a * t2 * (1.0 - t2) / (d2 * d2);
		const double mid = 0.5 * (th1 + th2), half = 0.5 * (th2 - th1);
		for (std::size_t i = 0; i < gl.x.size(); i++)
		{
			const double th = mid + half * std::sin(0.5 * PI * gl.x[i]);
			const double jac = half * 0.5 * PI * std::cos(0.5 * PI * gl.x[i]);
			const double t = std::cos(th) * std::cos(th), d = 1.0 + (a2 - 1.0) * t, s2 = std::sin(2.0 * th);
			const double K = -cotv * std::cos(2.0 * th) / s2;
			const double frac = 1.0 - std::acos(std::min(1.0, std::max(-1.0, K))) / PI;
			P += gl.w[i] * jac * frac * (a2 / (d * d)) * s2;
			dP += gl.w[i] * jac * frac * (2.0 * a * (1.0 - t - a2 * t) / (d * d * d)) * s2;
		}
		if (dPda) *dPda = dP;
		return P;
	}
	double	mirrorFallback(double mu, double a, double* dPda)
	{
		static const GaussLegendre gl(64);
		return mirrorFallback(mu, a, dPda, gl);
	}
	// d ln G1(c) / d alpha (Smith GGX, BSDF.cpp:220-236)
	double	dlogG1(double c, double a)
	{
		c = clamp01(c);
		a = std::max(1e-4, a);
		const double sq = std::sqrt(a * a + (1.0 - a * a) * c * c);
		const double d = -(a * (1.0 - c * c) / sq) / (c + sq);
		return std::isfinite(d) ? d : 0.0;
	}

	// d ln(D G1(noV) G1(noL)) / d alpha of the GGX reflection lobe (BSDF.cpp:209-236)
	double	ggxDlogAlpha(V n, V view, V l, double a)
	{
		a = std::max(1e-4, a);
		const double noV = clamp01(dot(n, view)), noL = clamp01(dot(n, l));
		const V hv = safeNormalize(view + l, n);
		const double noH = clamp01(dot(n, hv));
		const double a2 = a * a, den = noH * noH * (a2 - 1.0) + 1.0;
		const double d = 2.0 / a - 4.0 * a * noH * noH / den + dlogG1(noV, a) + dlogG1(noL, a);
		return std::isfinite(d) ? d : 0.0;
	}

	struct	Adjoint
	{
		C	weight;						// dLoss / dL of the sample
		std::vector<double>*	grads;	// 7 per material: albedo rgb, roughness, emission rgb
		const Scene*	scene;
		struct	Rec { std::uint32_t m; bool rough; C value; double dlog; };
		std::vector<Rec>	recs;
		// after a mirror fallback: the next emitter's MIS weight w depends on the metal's alpha
		// through the previous pdf; d ln w / dr = (1 - w) * pendingScale
		std::int64_t	pendingMat = -1;
		double	pendingScale = 0.0;

		static bool	roughMetal(const Mat& m) { return m.kind == luzgpu::MAT_METAL && m.roughness > 1e-4; }
		static double	dAlphaDr(const Mat& m) { const double r = clamp01(m.roughness); return r * r > 1e-4 ? 2.0 * r : 0.0; }
		void	add(std::uint32_t m, int k, double v) { if (v != 0.0 && std::isfinite(v)) (*grads)[m * 7 + k] += v; }
		double	wsum(C c) const { return weight.r * c.r + weight.g * c.g + weight.b * c.b; }
		void	emission(std::uint32_t m, C inc, C le)
		{
			if (le.r > 0.0) add(m, 4, weight.r * inc.r / le.r);
			if (le.g > 0.0) add(m, 5, weight.g * inc.g / le.g);
			if (le.b > 0.0) add(m, 6, weight.b * inc.b / le.b);
		}
		// an emitter or sky reached by a BSDF-sampled ray with MIS weight w
		void	arrival(C inc, double w)
		{
			if (pendingMat >= 0) add(static_cast<std::uint32_t>(pendingMat), 3, wsum(inc) * (1.0 - w) * pendingScale);
			pendingMat = -1;
		}
		void	specular() { pendingMat = -1; }
		void	vertex(const Mat& m, std::uint32_t mi, C acc)
		{
			pendingMat = -1;
			if (m.kind == luzgpu::MAT_LAMBERTIAN) recs.push_back(Rec{mi, false, acc, 0.0});
		}
		void	connection(const Mat& m, std::uint32_t mi, const Hit& h, const Scatter& sc, V dir, C inc, std::int64_t lightMat)
		{
			if (lightMat >= 0) emission(static_cast<std::uint32_t>(lightMat), inc, scene->mats[static_cast<std::size_t>(lightMat)].color);
			if (roughMetal(m))
			{
				const V view = safeNormalize(sc.incident * -1.0, h.n);
				add(mi, 3, (weight.r * inc.r + weight.g * inc.g + weight.b * inc.b) * ggxDlogAlpha(h.n, view, dir, m.alpha) * dAlphaDr(m));
			}
		}
		// The sampled continuation. A regular GGX sample differentiates f / p with p held: d ln f.
		// A mirror fallback is a point mass of probability P(alpha) carrying f / p at the mirror
		// direction (= F G1(v)^2, D cancels): its derivative is the score d ln P plus
		// 2 d ln G1(v), and the next emitter's MIS weight moves with p(mirror) (arrival).
		void	continuation(const Mat& m, std::uint32_t mi, const Hit& h, const Scatter& sc, V dir, C acc)
		{
			if (!roughMetal(m)) return;
			const V view = safeNormalize(sc.incident * -1.0, h.n);
			if (!sc.mirrorFallback)
			{
				recs.push_back(Rec{mi, true, acc, ggxDlogAlpha(h.n, view, dir, m.alpha) * dAlphaDr(m)});
				return;
			}
			const double mu = clamp01(dot(h.n, view));
			double dP = 0.0;
			const double P = mirrorFallback(mu, m.alpha, &dP);
			const double score = P > 0.0 ? dP / P : 0.0;
			recs.push_back(Rec{mi, true, acc, (score + 2.0 * dlogG1(mu, m.alpha)) * dAlphaDr(m)});
			const double a = std::max(1e-4, m.alpha);
			const double noH = clamp01(dot(h.n, safeNormalize(view + dir, h.n)));
			const double dlogD = 2.0 / a - 4.0 * a * noH * noH / (noH * noH * (a * a - 1.0) + 1.0);
			pendingMat = mi;
			pendingScale = 2.0 * dlogD * dAlphaDr(m);
		}
		void	finish(C L)
		{
			for (const Rec& r : recs)
			{
				const C rest = L + r.value * -1.0;
				if (r.rough) add(r.m, 3, (weight.r * rest.r + weight.g * rest.g + weight.b * rest.b) * r.dlog);
				else
				{
					const C rho = scene->mats[r.m].color;
					if (rho.r > 1e-6) add(r.m, 0, weight.r * rest.r / rho.r);
					if (rho.g > 1e-6) add(r.m, 1, weight.g * rest.g / rho.g);
					if (rho.b > 1e-6) add(r.m, 2, weight.b * rest.b / rho.b);
				}
			}
			recs.clear();
			pendingMat = -1;
		}
	};

	// henyeyGreensteinDirection (ColorHelper.cpp:115-148) for a non-zero phase direction
	V	hgDirection(V phaseDir, double gIn, float sx, float sy)
	{
		const V pd = phaseDir / std::sqrt(len2(phaseDir));
		const double g = std::max(-0.99, std::min(0.99, gIn));
		const double phi = 2.0 * PI * sx;
		double c;
		if (std::fabs(g) < 1e-3) c = 1.0 - (2.0 * sy);
		else
		{
			const double rm = (1.0 - (g * g)) / (1.0 - g + (2.0 * g * sy));
			c = (1.0 + (g * g) - (rm * rm)) / (2.0 * g);
			c = std::max(-1.0, std::min(1.0, c));
		}
		const double st = std::sqrt(std::max(0.0, 1.0 - (c * c)));
		return Onb(pd).local(st * std::cos(phi), st * std::sin(phi), c);
	}

	// relocateSubsurfaceExit (ColorHelper.cpp:376-504): after a Burley subsurface scatter the
	// vertex moves to where a probe ray, cast along -normal from above a point a sampled
	// radius away on the tangent disk, meets the same material within reach; the cosine
	// direction is drawn again about the exit's normal (refreshSubsurfaceScatterAtExit,
	// :394-426). Unchanged when any step fails.
	void	relocateSubsurface(const Scene& s, const Mat& m, Hit& h, Scatter& sc, const Sampler& ps)
	{
		if (!sc.subsurface || m.subsurfaceThin) return;
		const double d = m.subsurfaceRadius;	// subsurfaceSampleRadiusMeters, :376-392
		if (!std::isfinite(d) || d <= 0.0) return;
		const double pick = ps.s1(D_SUBSURFACE_PROFILE_COMPONENT);
		float rx, ry;
		ps.s2(D_SUBSURFACE_SPATIAL, rx, ry);
		const double u = std::max(1e-12, std::min(1.0 - 1e-12, static_cast<double>(ry)));
		const double radiusM = -(pick < 0.25 ? d : 3.0 * d) * std::log(1.0 - u);
		if (radiusM <= 0.0 || !std::isfinite(radiusM) || s.mpu <= 0.0 || !std::isfinite(s.mpu)) return;
		const double radius = radiusM / s.mpu;
		if (radius <= T_MIN) return;
		V n = h.n;
		if (len2(n) <= 1e-12) n = h.gn;
		if (len2(n) <= 1e-12) return;
		n = normalize(n);
		const double angle = 2.0 * PI * static_cast<double>(rx);
		const V candidate = h.p + Onb(n).local(std::cos(angle) * radius, std::sin(angle) * radius, 0.0);
		const double reach = std::max(radius + (4.0 * d / s.mpu), 16.0 * T_MIN);
		const Ray probe = Ray::make(candidate + n * reach, n * -1.0);
		Hit e;
		if (!closestHit(s, probe, e, &ps)) return;
		if (e.material != h.material || e.t <= T_MIN || e.t > (2.0 * reach + T_MIN)) return;
		float sx, sy;
		ps.s2(D_BSDF_DIRECTION, sx, sy);
		const float phi = static_cast<float>(2.0 * PI) * sx;
		const float rr = std::sqrt(sy);
		const float z = std::sqrt(std::max(0.0f, 1.0f - sy));
		const V dir = Onb(e.n).local(static_cast<double>(rr * std::cos(phi)), static_cast<double>(rr * std::sin(phi)),
			static_cast<double>(z));
		const double pdf = bsdfPdf(m, e, sc, dir);
		const C f = bsdfEval(m, e, sc, dir);
		if (pdf <= 0.0 || !std::isfinite(pdf) || maxChannel(f) <= 0.0) return;
		h = e;
		sc.dir = dir;
		sc.att = f / pdf;
	}

	C	trace(const Scene& s, Ray ray, Sampler& ps, Adjoint* adj = nullptr)	// calculateLightRaysColor, ColorHelper.cpp:1136-1326
	{
		C acc, thr = {1, 1, 1};
		bool prevSpec = true, prevInf = false;
		V prevOrigin;
		double prevPdf = 0.0;
		const bool sampledInf = hasEnvLight(s) || hasConstBg(s);
		for (int b = 0; b <= s.maxBounces; b++)
		{
			ps.bounce = static_cast<std::uint32_t>(b);
			Hit h;
			if (!closestHit(s, ray, h, &ps))
			{
				double w = 1.0;
				if (b > 0 && !prevSpec && prevInf)
				{
					w = power(prevPdf, hasEnvLight(s) ? envPdf(s.env, ray.d) : (hasConstBg(s) ? 1.0 / (4.0 * PI) : 0.0));
				}
				C sky;
				if (s.sky == luzgpu::SKY_LINEAR)
				{
					// _calculateSkyInterpolation (Atmospherics.cpp:44-53): skyline 0.5, no MIS weight
					const double t = 0.5 * (normalize(ray.d).y + 1.0);
					sky = C{1.0, 1.0, 1.0} * (1.0 - t) + C{0.5, 0.7, 1.0} * t;
				}
				else if (s.sky == luzgpu::SKY_ENVIRONMENT && s.hasEnv && s.env.strength > 0.0) sky = envLookup(s.env, ray.d) * s.env.strength * w;
				else sky = s.background * w;
				C c = thr * sky;
				if (b > 0 && hasConstBg(s)) c = clampBg(s, c);
				c = clampRay(c);
				if (adj) adj->arrival(c, w);
				return acc + c;
			}
			const Mat& m = s.mats[h.material];
			const C emitted = m.kind == luzgpu::MAT_EMISSIVE ? m.color : C{};
			Scatter sc;
			if (!scatter(m, h, ray, ps, sc))
			{
				const double lp = (!prevSpec && !s.lights.empty()) ? lightsPdf(s, prevOrigin, ray.d) : 0.0;
				const double w = prevSpec ? 1.0 : power(prevPdf, lp);
				const C inc = clampRay((thr * emitted) * w);
				if (adj)
				{
					adj->emission(h.material, inc, emitted);
					adj->arrival(inc, w);
				}
				return acc + inc;
			}
			relocateSubsurface(s, m, h, sc, ps);
			// effectiveScatterAttenuation / mediumTransmittance (ColorHelper.cpp:695-723): a back-face
			// hit attenuates by Beer-Lambert over the segment's length in metres
			C att = sc.att;
			if (!h.front && sc.absorbing)
			{
				const double dm = h.t * s.mpu;
				if (std::isfinite(dm) && dm > 0.0)
				{
					att = att * C{std::exp(-sc.absorption.r * dm), std::exp(-sc.absorption.g * dm), std::exp(-sc.absorption.b * dm)};
				}
			}
			if (sc.specular)
			{
				if (!leavesOpaque(m, h, sc.dir)) return acc;
				if (adj) adj->specular();
				thr = clampRay(thr * att);
				if (terminated(thr) || !roulette(thr, b, ps)) return acc;
				ray = Ray::normalized(h.p, sc.dir);
				prevSpec = true;
				prevPdf = 0.0;
				continue;
			}
			acc = acc + thr * emitted;
			if (adj) adj->vertex(m, h.material, acc);
			if (terminated(thr * att)) return acc;
			if (!s.lights.empty())
			{
				std::int64_t lm = -1;
				V ld;
				const C inc = clampRay(thr * direct(s, m, h, sc, b, ps, &lm, &ld));
				acc = acc + inc;
				if (adj && maxChannel(inc) > 0.0) adj->connection(m, h.material, h, sc, ld, inc, lm);
			}
			if (sampledInf)
			{
				V ed;
				const C inc = clampRay(thr * environment(s, m, h, sc, b, ps, &ed));
				acc = acc + inc;
				if (adj && maxChannel(inc) > 0.0) adj->connection(m, h.material, h, sc, ed, inc, -1);
			}
			V dir;
			if (sc.type == PDF_COSINE)	// Sampler::cosineHemisphere (Sampler.cpp:239-246)
			{
				float sx, sy;
				ps.s2(D_BSDF_DIRECTION, sx, sy);
				const float phi = static_cast<float>(2.0 * PI) * sx;
				const float r = std::sqrt(sy);
				const float z = std::sqrt(std::max(0.0f, 1.0f - sy));
				const V local = {static_cast<double>(r * std::cos(phi)), static_cast<double>(r * std::sin(phi)), static_cast<double>(z)};
				dir = local.x * sc.basis.u + local.y * sc.basis.v + local.z * sc.basis.w;
			}
			else if (sc.type == PDF_SPHERE || (sc.type == PDF_HG && !(len2(sc.phaseDir) > 0.0 && std::isfinite(len2(sc.phaseDir)))))
			{
				float sx, sy;	// Sampler::sphereDirection (Sampler.cpp:249-257)
				ps.s2(D_BSDF_DIRECTION, sx, sy);
				const float z = 1.0f - (2.0f * sy);
				const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
				const float phi = static_cast<float>(2.0 * PI) * sx;
				dir = {static_cast<double>(r * std::cos(phi)), static_cast<double>(r * std::sin(phi)), static_cast<double>(z)};
			}
			else if (sc.type == PDF_HG)
			{
				float sx, sy;
				ps.s2(D_BSDF_DIRECTION, sx, sy);
				dir = hgDirection(sc.phaseDir, sc.g, sx, sy);
			}
			else
			{
				dir = sc.dir;
			}
			const Ray scattered = Ray::normalized(h.p, dir);
			if (!leavesOpaque(m, h, scattered.d)) return acc;
			const double pdf = scatterPdf(m, h, sc, scattered.d);
			if (pdf <= 0.0 || !std::isfinite(pdf)) return acc;
			if (adj) adj->continuation(m, h.material, h, sc, scattered.d, acc);
			thr = clampRay(thr * att);
			if (terminated(thr) || !roulette(thr, b, ps)) return acc;
			ray = scattered;
			prevSpec = false;
			prevOrigin = h.p;
			prevPdf = pdf;
			prevInf = (s.sky == luzgpu::SKY_ENVIRONMENT && hasEnvLight(s)) || (hasConstBg(s) && sc.type != PDF_BSDF);
		}
		return acc;
	}

	Ray	cameraRay(const Scene& s, std::size_t x, std::size_t y, const Sampler& ps)	// RayHelper.cpp:85-106
	{
		float cx, cy;
		ps.s2(D_CAMERA, cx, cy);
		const double xU = double(static_cast<double>(x) + cx) * s.invW;
		const double yV = double(static_cast<double>(y) + cy) * s.invH;
		V off;
		if (s.lensRadius > 0.0)
		{
			float sx, sy;	// Sampler::unitDisk (Sampler.cpp:267-292)
			ps.s2(D_LENS, sx, sy);
			sx = (2.0f * sx) - 1.0f;
			sy = (2.0f * sy) - 1.0f;
			V rd;
			if (!(sx == 0.0f && sy == 0.0f))
			{
				float rad, th;
				const float pi = static_cast<float>(PI);
				if (std::fabs(sx) > std::fabs(sy)) { rad = sx; th = (pi / 4.0f) * (sy / sx); }
				else { rad = sy; th = (pi / 2.0f) - ((pi / 4.0f) * (sx / sy)); }
				rd = {rad * std::cos(th), rad * std::sin(th), 0.0};
			}
			rd = rd * s.lensRadius;
			off = s.camU * rd.x + s.camV * rd.y;
		}
		return Ray::make(s.camPos + off, s.llc - (s.horizontal * xU) - (s.vertical * yV) - s.camPos - off);
	}

	C	clean(C c) { return {std::isfinite(c.r) ? c.r : 0.0, std::isfinite(c.g) ? c.g : 0.0, std::isfinite(c.b) ? c.b : 0.0}; }	// Threads.cpp:122-137
}

extern "C" {

// Henyey-Greenstein phase pdf for the angle between the phase direction and `dir`
// (scatterPDFValue, ColorHelper.cpp:56-77), and the cosine of a sampled direction
// (henyeyGreensteinDirection, :115-148) for the 2-D sample (sx, sy).
double	oracle_hg_pdf(double g, double cosTheta)
{
	oracle::Scatter sc;
	sc.type = oracle::PDF_HG;
	sc.phaseDir = {0.0, 0.0, 1.0};
	sc.g = g;
	const double st = std::sqrt(std::max(0.0, 1.0 - cosTheta * cosTheta));
	oracle::Mat m;
	m.kind = luzgpu::MAT_HG;
	return oracle::scatterPdf(m, oracle::Hit{}, sc, oracle::V{st, 0.0, cosTheta});
}

double	oracle_hg_sample_cos(double g, float sx, float sy)
{
	return oracle::hgDirection(oracle::V{0.0, 0.0, 1.0}, g, sx, sy).z;
}

// Sampler values for n keys (seed, x, y, sample, stream, bounce, dim).
void	oracle_sampler(std::size_t n, const std::uint32_t* keys, double* s1, float* s2)
{
	for (std::size_t i = 0; i < n; i++)
	{
		const std::uint32_t* k = keys + 7 * i;
		oracle::Sampler ps;
		ps.begin(k[0], k[1], k[2], k[3], k[4]);
		ps.bounce = k[5];
		s1[i] = ps.s1(k[6]);
		ps.s2(k[6], s2[2 * i], s2[2 * i + 1]);
	}
}

static thread_local std::string	g_oracleError;

const char*	oracle_last_error(void)
{
	return g_oracleError.c_str();
}

// FP64 render of samples [sample_begin, sample_begin + spp) — _manageThreads/_threadRender
// (Threads.cpp:375-583) with 16-pixel blocks over `threads` workers. out: H*W*3 doubles.
int	oracle_render(const char* path, std::uint32_t seed, std::uint32_t spp, std::uint32_t sample_begin, std::uint32_t width,
	std::uint32_t height, std::uint32_t max_bounces_plus1, std::uint32_t threads, double* out, double* render_ms)
{
	try
	{
		const luzgpu::Scene in = luzgpu::readSceneFile(path);
		oracle::Scene s = oracle::build(in, width, height);
		if (max_bounces_plus1 > 0) s.maxBounces = static_cast<int>(max_bounces_plus1) - 1;
		const std::size_t W = s.width, H = s.height, total = W * H;
		std::atomic<std::size_t> next(0);
		const auto t0 = std::chrono::steady_clock::now();
		std::vector<std::thread> pool;
		for (std::uint32_t t = 0; t < std::max(1u, threads); t++)
		{
			pool.emplace_back([&]() {
				while (true)
				{
					const std::size_t start = next.fetch_add(16);
					if (start >= total) break;
					for (std::size_t i = start; i < std::min(start + 16, total); i++)
					{
						const std::size_t x = i % W, y = i / W;
						oracle::C sum;
						for (std::uint32_t k = 0; k < spp; k++)
						{
							oracle::Sampler ps;
							ps.begin(seed, x, y, sample_begin + k, 0);
							const oracle::Ray r = oracle::cameraRay(s, x, y, ps);
							sum = sum + oracle::clean(oracle::trace(s, r, ps));
						}
						const oracle::C mean = oracle::clean(sum / static_cast<double>(spp));
						out[3 * i] = mean.r;
						out[3 * i + 1] = mean.g;
						out[3 * i + 2] = mean.b;
					}
				}
			});
		}
		for (auto& t : pool) t.join();
		if (render_ms) *render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		return 0;
	}
	catch (const std::exception& e)
	{
		g_oracleError = e.what();
		return -1;
	}
}

namespace oracle
{
	// AdaptiveAccumulator (Threads.cpp:31-53)
	struct	AdaptiveAcc
	{
		C	sum, square;
		double	lumSum = 0.0, lumSquare = 0.0, maxLum = 0.0;

		void	add(const C& c)
		{
			// sampleLuminance (Threads.cpp:139-148)
			double lum = luminance(c);
			if (!std::isfinite(lum)) lum = 0.0;
			sum = sum + c;
			square = square + C{c.r * c.r, c.g * c.g, c.b * c.b};
			lumSum += lum;
			lumSquare += lum * lum;
			maxLum = std::max(maxLum, lum);
		}
	};

	// confidenceInterval95 over sampleMeanVariance (Threads.cpp:150-172)
	double	ci95(double sum, double square, unsigned count)
	{
		double mv = 0.0;
		if (count > 1)
		{
			const double n = static_cast<double>(count);
			mv = std::max(0.0, ((square - (sum * sum / n)) / (n - 1.0)) / n);
		}
		if (!std::isfinite(mv)) return std::numeric_limits<double>::infinity();
		return 1.96 * std::sqrt(mv);
	}

	// channelConverged (Threads.cpp:174-180)
	bool	channelOk(double mean, double ci, double threshold)
	{
		return ci <= threshold * std::max(std::fabs(mean), 0.02);
	}

	struct	AdaptiveRuleC
	{
		unsigned	maxSamples, minSamples, checkInterval, darkMinimum;
		double	threshold;
	};

	// adaptiveSampleConverged (Threads.cpp:194-262)
	bool	adaptiveStop(const AdaptiveRuleC& r, unsigned used, const AdaptiveAcc& a)
	{
		if (used >= r.maxSamples) return true;
		if (used < r.minSamples) return false;
		if ((used - r.minSamples) % r.checkInterval != 0) return false;
		if (used <= 1) return false;
		const double n = static_cast<double>(used);
		if (a.maxLum < 0.02 && used < r.darkMinimum) return false;
		return channelOk(a.lumSum / n, ci95(a.lumSum, a.lumSquare, used), r.threshold)
			&& channelOk(a.sum.r / n, ci95(a.sum.r, a.square.r, used), r.threshold)
			&& channelOk(a.sum.g / n, ci95(a.sum.g, a.square.g, used), r.threshold)
			&& channelOk(a.sum.b / n, ci95(a.sum.b, a.square.b, used), r.threshold);
	}
}

// FP64 render with adaptive per-pixel stopping: _threadRender's loop (Threads.cpp:551-583)
// with the rule of adaptiveSampleConverged. min_samples / check_interval / threshold: 0 (or
// <= 0) takes the scene file's value. out: H*W*3 doubles, counts: H*W samples per pixel.
// Returns 1 without rendering when the rule cannot stop a pixel (adaptiveSamplingCanStop,
// Threads.cpp:265-271, false).
int	oracle_render_adaptive(const char* path, std::uint32_t seed, std::uint32_t spp, std::uint32_t width, std::uint32_t height,
	std::uint32_t max_bounces_plus1, std::uint32_t threads, std::uint32_t min_samples, std::uint32_t check_interval,
	double threshold, double* out, std::uint32_t* counts)
{
	try
	{
		const luzgpu::Scene in = luzgpu::readSceneFile(path);
		oracle::Scene s = oracle::build(in, width, height);
		if (max_bounces_plus1 > 0) s.maxBounces = static_cast<int>(max_bounces_plus1) - 1;
		const unsigned total = spp > 0 ? spp : in.settings.samples;
		oracle::AdaptiveRuleC rule{};
		rule.maxSamples = total;
		rule.minSamples = std::min(min_samples ? min_samples : in.settings.adaptiveMinSamples, total);
		rule.checkInterval = check_interval ? check_interval : in.settings.adaptiveCheckInterval;
		rule.threshold = threshold > 0.0 ? threshold : in.settings.adaptiveThreshold;
		if ((min_samples ? min_samples : in.settings.adaptiveMinSamples) >= total) return 1;
		// darkMinimumSamples (Threads.cpp:183-192)
		rule.darkMinimum = std::min(total, std::max(rule.minSamples, std::max(256u, rule.minSamples + 3u * rule.checkInterval)));
		const std::size_t W = s.width, H = s.height, N = W * H;
		std::atomic<std::size_t> next(0);
		std::vector<std::thread> pool;
		for (std::uint32_t t = 0; t < std::max(1u, threads); t++)
		{
			pool.emplace_back([&]() {
				while (true)
				{
					const std::size_t start = next.fetch_add(16);
					if (start >= N) break;
					for (std::size_t i = start; i < std::min(start + 16, N); i++)
					{
						const std::size_t x = i % W, y = i / W;
						oracle::C sum;
						oracle::AdaptiveAcc acc;
						unsigned used = 0;
						for (unsigned k = 0; k < total; k++)
						{
							oracle::Sampler ps;
							ps.begin(seed, x, y, k, 0);
							const oracle::Ray r = oracle::cameraRay(s, x, y, ps);
							const oracle::C c = oracle::clean(oracle::trace(s, r, ps));
							sum = sum + c;
							used = k + 1;
							acc.add(c);
							if (oracle::adaptiveStop(rule, used, acc)) break;
						}
						const oracle::C mean = oracle::clean(sum / static_cast<double>(used));
						out[3 * i] = mean.r;
						out[3 * i + 1] = mean.g;
						out[3 * i + 2] = mean.b;
						counts[i] = used;
					}
				}
			});
		}
		for (auto& t : pool) t.join();
		return 0;
	}
	catch (const std::exception& e)
	{
		g_oracleError = e.what();
		return -1;
	}
}

// Path-replay adjoint of oracle_render for dLoss/dImage = dldi (W*H*3): FP64 gradients per
// material (7 each: albedo rgb, GGX roughness, emission rgb).
int	oracle_adjoint(const char* path, std::uint32_t seed, std::uint32_t spp, std::uint32_t width, std::uint32_t height,
	std::uint32_t max_bounces_plus1, std::uint32_t threads, const float* dldi, double* grads, std::uint32_t material_count)
{
	try
	{
		const luzgpu::Scene in = luzgpu::readSceneFile(path);
		oracle::Scene s = oracle::build(in, width, height);
		if (max_bounces_plus1 > 0) s.maxBounces = static_cast<int>(max_bounces_plus1) - 1;
		if (material_count != s.mats.size()) throw std::runtime_error("oracle_adjoint: material count mismatch");
		const std::size_t W = s.width, H = s.height, total = W * H;
		std::atomic<std::size_t> next(0);
		const std::uint32_t nt = std::max(1u, threads);
		std::vector<std::vector<double>> partial(nt, std::vector<double>(s.mats.size() * 7, 0.0));
		std::vector<std::thread> pool;
		for (std::uint32_t t = 0; t < nt; t++)
		{
			pool.emplace_back([&, t]() {
				oracle::Adjoint adj;
				adj.grads = &partial[t];
				adj.scene = &s;
				while (true)
				{
					const std::size_t start = next.fetch_add(16);
					if (start >= total) break;
					for (std::size_t i = start; i < std::min(start + 16, total); i++)
					{
						const std::size_t x = i % W, y = i / W;
						adj.weight = {dldi[3 * i] / static_cast<double>(spp), dldi[3 * i + 1] / static_cast<double>(spp),
							dldi[3 * i + 2] / static_cast<double>(spp)};
						for (std::uint32_t k = 0; k < spp; k++)
						{
							oracle::Sampler ps;
							ps.begin(seed, x, y, k, 0);
							const oracle::Ray r = oracle::cameraRay(s, x, y, ps);
							adj.finish(oracle::trace(s, r, ps, &adj));
						}
					}
				}
			});
		}
		for (auto& th : pool) th.join();
		for (std::size_t k = 0; k < s.mats.size() * 7; k++)
		{
			double v = 0.0;
			for (const auto& p : partial) v += p[k];
			grads[k] = v;
		}
		return 0;
	}
	catch (const std::exception& e)
	{
		g_oracleError = e.what();
		return -1;
	}
}

// Primary-ray closest hit of sample 0 / stream 0 for every pixel: ids[2*i] = top-level
// hittable, ids[2*i+1] = mesh triangle (or -1); t[i] (or -1).
int	oracle_primary(const char* path, std::uint32_t seed, std::uint32_t width, std::uint32_t height, std::int64_t* ids, double* t)
{
	try
	{
		const luzgpu::Scene in = luzgpu::readSceneFile(path);
		const oracle::Scene s = oracle::build(in, width, height);
		for (std::size_t i = 0; i < static_cast<std::size_t>(s.width) * s.height; i++)
		{
			oracle::Sampler ps;
			ps.begin(seed, i % s.width, i / s.width, 0, 0);
			const oracle::Ray r = oracle::cameraRay(s, i % s.width, i / s.width, ps);
			oracle::Hit h;
			if (oracle::closestHit(s, r, h, &ps)) { ids[2 * i] = h.top; ids[2 * i + 1] = h.tri; t[i] = h.t; }
			else { ids[2 * i] = -1; ids[2 * i + 1] = -1; t[i] = -1.0; }
		}
		return 0;
	}
	catch (const std::exception& e)
	{
		g_oracleError = e.what();
		return -1;
	}
}

// Monte-Carlo normalisation check helpers (closed forms): integral of the GGX reflection
// pdf over the sphere for a view direction at cos(theta) = mu, and the Schlick / dielectric
// Fresnel at normal incidence.
double	oracle_ggx_pdf_integral(double alpha, double mu, std::uint32_t n)
{
	const oracle::V nrm = {0, 0, 1};
	const oracle::V view = {std::sqrt(1.0 - mu * mu), 0.0, mu};
	double sum = 0.0;
	for (std::uint32_t i = 0; i < n; i++)
	{
		for (std::uint32_t j = 0; j < n; j++)
		{
			const double z = (i + 0.5) / n;				// upper hemisphere, uniform in z
			const double phi = 2.0 * oracle::PI * (j + 0.5) / n;
			const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
			sum += oracle::ggxReflPdf(nrm, view, {r * std::cos(phi), r * std::sin(phi), z}, alpha);
		}
	}
	return sum * (2.0 * oracle::PI) / (static_cast<double>(n) * n);
}

// Integral of D(h) (n.h) over the hemisphere of microfacet normals (= 1 for any alpha).
double	oracle_ggx_ndf_integral(double alpha, std::uint32_t n)
{
	double sum = 0.0;
	for (std::uint32_t i = 0; i < n; i++)
	{
		const double z = (i + 0.5) / n;	// n.h, uniform in z; the integrand is symmetric in phi
		sum += oracle::ggxD(z, alpha) * z;
	}
	return sum * (2.0 * oracle::PI) / static_cast<double>(n);
}

// d ln f / d alpha of the GGX reflection lobe used by the adjoint, for a view at cos mu_v and
// a light at cos mu_l, azimuth phi apart; *fd receives the central difference (step h) of
// ln of the oracle's own GGX reflection BSDF, the quantity it differentiates.
double	oracle_ggx_dlog_alpha(double mu_v, double mu_l, double phi, double alpha, double h, double* fd)
{
	const oracle::V n = {0, 0, 1};
	const oracle::V view = {std::sqrt(1.0 - mu_v * mu_v), 0.0, mu_v};
	const double sl = std::sqrt(1.0 - mu_l * mu_l);
	const oracle::V l = {sl * std::cos(phi), sl * std::sin(phi), mu_l};
	const oracle::C one = {1.0, 1.0, 1.0};
	const double fp = oracle::ggxReflBsdf(one, n, view, l, alpha + h).r;
	const double fm = oracle::ggxReflBsdf(one, n, view, l, alpha - h).r;
	*fd = (std::log(fp) - std::log(fm)) / (2.0 * h);
	return oracle::ggxDlogAlpha(n, view, l, alpha);
}

// Mirror-fallback probability of the metal sampler (closed form + quadrature with `nodes`
// Gauss-Legendre points, 0 = the adjoint's default) and its alpha derivative; *brute
// receives the fallback frequency of the sampler itself over an n x 20n grid of (sx, sy).
double	oracle_ggx_fallback(double mu, double alpha, std::uint32_t nodes, std::uint32_t n, double* dPda, double* brute)
{
	double P;
	if (nodes == 0) P = oracle::mirrorFallback(mu, alpha, dPda);
	else P = oracle::mirrorFallback(mu, alpha, dPda, oracle::GaussLegendre(static_cast<int>(nodes)));
	if (n > 0)
	{
		const oracle::V nrm = {0, 0, 1};
		const oracle::V view = {std::sqrt(1.0 - mu * mu), 0.0, mu};
		std::size_t hits = 0;
		for (std::uint32_t i = 0; i < n; i++)
		{
			for (std::uint32_t j = 0; j < 20 * n; j++)
			{
				const float sx = static_cast<float>((i + 0.5) / n), sy = static_cast<float>((j + 0.5) / (20.0 * n));
				const oracle::V half = oracle::ggxHalf(nrm, alpha, sx, sy);
				const oracle::V dir = oracle::reflect(view * -1.0, half);
				if (oracle::dot(view, half) <= 0.0 || oracle::dot(dir, nrm) <= 0.0) hits++;
			}
		}
		*brute = static_cast<double>(hits) / (20.0 * n * n);
	}
	return P;
}

double	oracle_dielectric_fresnel(double cos_theta, double eta_i, double eta_t)
{
	return oracle::dielectricFresnel(cos_theta, eta_i, eta_t);
}

// conductor Fresnel of one channel (BSDF.cpp:127-174)
double	oracle_conductor_fresnel(double cos_theta, double eta, double k)
{
	return oracle::conductorFresnel(oracle::C{eta, eta, eta}, oracle::C{k, k, k}, cos_theta).r;
}

}
