// guide_math.h -- host/device helpers of the path-guiding network: BF16
// rounding, the hash-grid lattice, and the von Mises-Fisher mixture that the
// network outputs (density and sampling).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/ymx_guide.h"

namespace ymx {
namespace guide {

constexpr int kLevels = YMX_GUIDE_LEVELS;
constexpr int kFeat = YMX_GUIDE_FEATURES;
constexpr int kLobes = YMX_GUIDE_LOBES;
constexpr int kIn = 64, kHidden = 64, kOut = 48;
constexpr int kMix = YMX_GUIDE_MIX_FLOATS;
constexpr float kKappaLogMin = -2.f, kKappaLogMax = 9.f;
constexpr uint32_t kPrimeY = 2654435761u, kPrimeZ = 805459861u;

// float -> bfloat16, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
	union
	{
		float f;
		uint32_t u;
	} v;
	v.f = f;
	if((v.u & 0x7F800000u) == 0x7F800000u && (v.u & 0x7FFFFFu)) return (uint16_t)((v.u >> 16) | 0x40u);
	const uint32_t r = v.u + 0x7FFFu + ((v.u >> 16) & 1u);
	return (uint16_t)(r >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
	union
	{
		uint32_t u;
		float f;
	} v;
	v.u = (uint32_t)h << 16;
	return v.f;
}

// Lattice entry of corner (x, y, z) at a level of resolution `res`: dense while
// (res + 1)^3 fits the table, otherwise the spatial hash of Mueller et al. 2022.
// Dense coordinates are clamped to res: a corner past it only occurs at the
// upper face of the unit cube, where its interpolation weight is 0.
__host__ __device__ __forceinline__ uint32_t grid_index(int res, uint32_t table_mask, uint32_t x, uint32_t y,
                                                        uint32_t z)
{
	const uint32_t r1 = (uint32_t)res + 1u;
	if((uint64_t)r1 * r1 * r1 <= (uint64_t)table_mask + 1u)
	{
		const uint32_t r = (uint32_t)res;
		x = x < r ? x : r;
		y = y < r ? y : r;
		z = z < r ? z : r;
		return x + y * r1 + z * r1 * r1;
	}
	return (x ^ (y * kPrimeY) ^ (z * kPrimeZ)) & table_mask;
}

// Decoded mixture of one query: per lobe mu (3), kappa, weight, normalisation
// kappa / (2 pi (1 - e^-2 kappa)).
__host__ __device__ __forceinline__ float mix_pdf(const float *m, float wx, float wy, float wz)
{
	float q = 0.f;
	for(int k = 0; k < kLobes; ++k)
	{
		const float *l = m + 6 * k;
		const float c = l[0] * wx + l[1] * wy + l[2] * wz;
		q += l[4] * l[5] * expf(l[3] * (c - 1.f));
	}
	return q;
}

// Sample the mixture: lobe by its weight (u0), then Jakob's stable vMF sample
// (u1: cosine, u2: azimuth) in a frame around mu (Duff et al. 2017 basis).
__host__ __device__ __forceinline__ void mix_sample(const float *m, float u0, float u1, float u2, float &wx, float &wy,
                                                    float &wz)
{
	int k = 0;
	float acc = 0.f;
	for(; k < kLobes - 1; ++k)
	{
		acc += m[6 * k + 4];
		if(u0 < acc) break;
	}
	const float *l = m + 6 * k;
	const float kappa = l[3];
	const float wc = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kappa)) / kappa;
	const float s = sqrtf(fmaxf(0.f, 1.f - wc * wc));
	const float phi = 6.283185307179586f * u2;
	const float mx = l[0], my = l[1], mz = l[2];
	const float sign = mz >= 0.f ? 1.f : -1.f;
	const float a = -1.f / (sign + mz);
	const float b = mx * my * a;
	const float tx = 1.f + sign * mx * mx * a, ty = sign * b, tz = -sign * mx;
	const float bx = b, by = sign + my * my * a, bz = -my;
	const float cp = cosf(phi), sp = sinf(phi);
	wx = wc * mx + s * (cp * tx + sp * bx);
	wy = wc * my + s * (cp * ty + sp * by);
	wz = wc * mz + s * (cp * tz + sp * bz);
}

}  // namespace guide
}  // namespace ymx
