// Small fp32 vector type for device code (float3 arithmetic without HIP's vector
// operator overload set, so every operation and its rounding is explicit).
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace luzgpu
{
	struct	f3
	{
		float	x, y, z;
	};

	__device__ __forceinline__ f3	mk(float x, float y, float z) { return f3{x, y, z}; }
	__device__ __forceinline__ f3	operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
	__device__ __forceinline__ f3	operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
	__device__ __forceinline__ f3	operator*(f3 a, float f) { return f3{a.x * f, a.y * f, a.z * f}; }
	__device__ __forceinline__ f3	operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
	__device__ __forceinline__ f3	operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
	__device__ __forceinline__ float	dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
	__device__ __forceinline__ f3	cross(f3 a, f3 b)
	{
		return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
	}
	__device__ __forceinline__ float	len2(f3 a) { return dot(a, a); }
	__device__ __forceinline__ float	maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
	__device__ __forceinline__ f3	normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
	__device__ __forceinline__ bool	finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
	__device__ __forceinline__ f3	ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
	__device__ __forceinline__ float	luminance(f3 c)
	{
		return c.x * 0.2722287168f + c.y * 0.6740817658f + c.z * 0.0536895174f;
	}
	__device__ __forceinline__ f3	reflect(f3 v, f3 n) { return v - n * (2.0f * dot(v, n)); }
	__device__ __forceinline__ float	clamp01(float v)
	{
		return isfinite(v) ? fmaxf(0.0f, fminf(1.0f, v)) : 0.0f;
	}

	// Orthonormal basis around a normal (ONB.cpp:5-13).
	struct	Onb
	{
		f3	u, v, w;

		__device__ __forceinline__ explicit Onb(f3 n)
		{
			w = normalize(n);
			const f3 a = (fabsf(w.x) > 0.9f) ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
			v = normalize(cross(w, a));
			u = cross(w, v);
		}

		__device__ __forceinline__ f3	local(float a, float b, float c) const
		{
			return u * a + v * b + w * c;
		}
	};
}
