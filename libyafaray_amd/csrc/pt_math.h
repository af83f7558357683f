// pt_math.h -- host/device arithmetic of the path tracer: vectors and colours
// in the reference's evaluation order, its FAST_MATH / FAST_TRIG
// approximations, and its low-discrepancy samplers.
//
// The expressions keep the reference's operation order so that, compiled with
// -ffp-contract=off and IEEE division / square root, results match its FP32
// x86 arithmetic bit for bit.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define YMX_HD __host__ __device__ __forceinline__
#else
// plain C++ (the host layer's mesh smoothing uses the fast cosine)
#define YMX_HD inline
#endif

namespace ymx {

struct V3
{
	float x, y, z;
};
struct Rgb
{
	float r, g, b;
};

YMX_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
YMX_HD V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
YMX_HD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// Vec * float and float * Vec both evaluate f * component (include/geometry/vector.h:175-198)
YMX_HD V3 mul(float f, V3 a) { return V3{f * a.x, f * a.y, f * a.z}; }
YMX_HD V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// dot product left to right (include/geometry/vector.h:155-172)
YMX_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// include/geometry/vector.h:238-245
YMX_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Vec::normalize (include/geometry/vector.h:306-315)
YMX_HD V3 normalize(V3 a)
{
	const float l2 = dot(a, a);
	if(l2 != 0.f)
	{
		const float inv = 1.f / sqrtf(l2);
		a.x *= inv;
		a.y *= inv;
		a.z *= inv;
	}
	return a;
}
YMX_HD V3 load3(const float *p) { return V3{p[0], p[1], p[2]}; }
// Vec3f::reflect in place: vn = (2 v) . n, v = vn n - v (include/geometry/vector.h:357-362)
YMX_HD V3 reflect_about(V3 v, V3 n)
{
	const float vn = dot(mul(2.f, v), n);
	return V3{vn * n.x - v.x, vn * n.y - v.y, vn * n.z - v.z};
}
// Vec3f::reflectDir (include/geometry/vector.h:390-396)
YMX_HD V3 reflect_dir(V3 normal, V3 vec)
{
	const float vn = dot(vec, normal);
	if(vn < 0.f) return neg(vec);
	return sub(mul(2.f * vn, normal), vec);
}

YMX_HD Rgb rgb(float r, float g, float b) { return Rgb{r, g, b}; }
YMX_HD Rgb cadd(Rgb a, Rgb b) { return Rgb{a.r + b.r, a.g + b.g, a.b + b.b}; }
YMX_HD Rgb cmul(Rgb a, Rgb b) { return Rgb{a.r * b.r, a.g * b.g, a.b * b.b}; }
// Rgb * float evaluates f * component; Rgb / float divides (include/color/color.h:245-258)
YMX_HD Rgb cscale(Rgb a, float f) { return Rgb{f * a.r, f * a.g, f * a.b}; }
YMX_HD Rgb cdiv(Rgb a, float f) { return Rgb{a.r / f, a.g / f, a.b / f}; }
YMX_HD float cmaximum(Rgb a)
{
	const float m = (a.g < a.b) ? a.b : a.g;  // std::max(g, b)
	return (a.r < m) ? m : a.r;               // std::max(r, .)
}
YMX_HD bool cblack(Rgb a) { return a.r == 0.f && a.g == 0.f && a.b == 0.f; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float k2Pi = 6.28318530717958647692f;
constexpr float kHalfPi = 1.57079632679489661923f;
constexpr float k1Pi = 0.31830988618379067154f;

// FAST_MATH exp2 / log2 (include/math/math.h:116-145): a fifth-order
// polynomial on the fraction and the exponent written into the bits.
YMX_HD float poly_exp(float x)
{
	return x * (x * (x * (x * (x * 1.8775767e-3f + 8.9893397e-3f) + 5.5826318e-2f) + 2.4015361e-1f) + 6.9315308e-1f) +
	       9.9999994e-1f;
}
YMX_HD float fast_exp2(float x)
{
	float xl = (129.00000f < x) ? 129.00000f : x;   // std::min(x, 129)
	xl = (xl < -126.99999f) ? -126.99999f : xl;      // std::max(., -126.99999)
	const int ipart = (int)(xl - 0.5f);
	const float fpart = xl - (float)ipart;
	union
	{
		int i;
		float f;
	} e;
	e.i = (ipart + 127) << 23;
	return e.f * poly_exp(fpart);
}
YMX_HD float fast_exp(float a) { return fast_exp2(1.44269504088896340736f * a); }
YMX_HD float poly_log(float x)
{
	return x * (x * (x * (x * (x * -3.4436006e-2f + 3.1821337e-1f) + -1.2315303f) + 2.5988452f) + -3.3241990f) +
	       3.1157899f;
}
// log2 from the exponent bits and a polynomial in the mantissa (the sign bit is ignored)
YMX_HD float fast_log2(float x)
{
	union
	{
		int i;
		float f;
	} v, m;
	v.f = x;
	m.i = (v.i & 0x7FFFFF) | 0x3F800000;  // mantissa with the exponent of 1.0f
	const float e = (float)(((v.i & 0x7F800000) >> 23) - 127);
	return poly_log(m.f) * (m.f - 1.f) + e;
}
// math::pow = exp2(log2(a) * b) under FAST_MATH (include/math/math.h:176-183)
YMX_HD float fast_pow(float a, float b) { return fast_exp2(fast_log2(a) * b); }

YMX_HD float std_max(float a, float b) { return (a < b) ? b : a; }
YMX_HD float std_min(float a, float b) { return (b < a) ? b : a; }

// FAST_TRIG sine (include/math/math.h:228-255): parabola with a 0.225 correction.
YMX_HD float fast_sin(float x)
{
	if(x > k2Pi || x < -k2Pi) x -= (float)(int)(x * 0.15915494309189533577f) * k2Pi;
	if(x < -kPi) x += k2Pi;
	else if(x > kPi) x -= k2Pi;
	x = 1.27323954473516268615f * x - 0.40528473456935108578f * x * fabsf(x);
	const float result = 0.225f * (x * fabsf(x) - x) + x;
	if(result <= -1.f) return -1.f;
	if(result >= 1.f) return 1.f;
	return result;
}
YMX_HD float fast_cos(float x) { return fast_sin(x + kHalfPi); }

YMX_HD float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }

// Radical inverses (include/sampler/sample.h:100-122) scaled by 2^-32 in double.
YMX_HD float ri_vdc(uint32_t bits, uint32_t r)
{
	bits = (bits << 16) | (bits >> 16);
	bits = ((bits & 0x00ff00ffu) << 8) | ((bits & 0xff00ff00u) >> 8);
	bits = ((bits & 0x0f0f0f0fu) << 4) | ((bits & 0xf0f0f0f0u) >> 4);
	bits = ((bits & 0x33333333u) << 2) | ((bits & 0xccccccccu) >> 2);
	bits = ((bits & 0x55555555u) << 1) | ((bits & 0xaaaaaaaau) >> 1);
	return clamp01((float)((double)(bits ^ r) * 0.00000000023283064365386962890625));
}
// Sobol' second dimension (include/sampler/sample.h:110-115)
YMX_HD float ri_s(uint32_t i, uint32_t r)
{
	for(uint32_t v = 1u << 31; i; i >>= 1, v ^= v >> 1)
		if(i & 1u) r ^= v;
	return clamp01((float)((double)r * 0.00000000023283064365386962890625));
}
YMX_HD float ri_lp(uint32_t i, uint32_t r)
{
	for(uint32_t v = 1u << 31; i; i >>= 1, v |= v >> 1)
		if(i & 1u) r ^= v;
	return clamp01((float)((double)r * 0.00000000023283064365386962890625));
}
// FNV-1a of the four little-endian bytes (include/sampler/sample.h:130-146)
YMX_HD uint32_t fnv32a(uint32_t value)
{
	uint32_t h = 0x811c9dc5u;
	for(int k = 0; k < 4; ++k)
	{
		h ^= (value >> (8 * k)) & 0xffu;
		h *= 0x01000193u;
	}
	return h;
}

// Incremental Halton sequence (include/sampler/halton.h:30-81), double
// precision.  Successive next() calls accumulate exactly as the reference's
// getNext() does, which differs in the last bits from restarting at start + i.
struct Halton
{
	uint32_t base;
	double inv, value;

	YMX_HD void start(uint32_t b, uint32_t s)
	{
		base = b;
		inv = 1.0 / (double)b;
		double factor = inv;
		value = 0.0;
		while(s > 0)
		{
			value += (double)(s % base) * factor;
			s /= base;
			factor *= inv;
		}
	}
	YMX_HD float next()
	{
		const double r = 0.9999999999 - value;
		if(inv < r) value += inv;
		else
		{
			double hh = 0.0, h = inv;
			while(h >= r)
			{
				hh = h;
				h *= inv;
			}
			value += hh + h - 1.0;
		}
		return clamp01((float)value);
	}
};

// Halton(base, start).getNext()
YMX_HD float halton_next(uint32_t base, uint32_t start)
{
	Halton h;
	h.start(base, start);
	return h.next();
}

// Counter-based stream replacing the Russian-roulette generator.
YMX_HD float rr_random(uint32_t pixel, uint32_t sample, uint32_t depth, uint32_t seed)
{
	uint32_t h = pixel * 0x9E3779B1u;
	h ^= (sample + 0x7F4A7C15u) * 0x85EBCA77u;
	h ^= (depth + 0x165667B1u) * 0xC2B2AE3Du;
	h ^= seed * 0x27D4EB2Fu;
	h ^= h >> 16;
	h *= 0x85EBCA6Bu;
	h ^= h >> 13;
	h *= 0xC2B2AE35u;
	h ^= h >> 16;
	return (float)(h >> 8) * (1.0f / 16777216.0f);
}

}  // namespace ymx
