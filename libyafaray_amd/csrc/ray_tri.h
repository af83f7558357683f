// ray_tri.h -- the reference's tree-bound crossing and triangle test, shared
// by the traversal kernels (accel_trace.hip), the object-light intersection of
// the path tracer (pt_kernels.hip) and the host traversal of the C++ layer
// (libyafaray_amd/host), which compiles them as plain C++.
#pragma once

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define YMX_RT_FN __host__ __device__ __forceinline__
#else
#define YMX_RT_FN inline
struct float4
{
	float x, y, z, w;
};
#endif

namespace ymx {

// Box tests of the BVH traversals only cull: slab intervals are widened by
// kSlabWiden >= 1 + 2 gamma_3 for FP32 (T. Ize, "Robust BVH Ray Traversal",
// JCGT 2013) and the accepted [t_min, t_best] window by kBoxWindow relative,
// so every triangle whose exact test would be accepted is tested.
constexpr float kSlabWiden = 1.0000004f;
constexpr float kBoxWindow = 0x1p-12f;

// include/geometry/bound.h:157-198 on the tree bound, exactly as the reference
// evaluates it.
YMX_RT_FN bool tree_cross(const float a[3], const float g[3], const float from[3],
                                           const float dir[3], float t_max, float &enter, float &leave)
{
	float p[3];
	for(int k = 0; k < 3; ++k) p[k] = from[k] - a[k];
	float lmin = -FLT_MAX, lmax = FLT_MAX;
#pragma unroll
	for(int axis = 0; axis < 3; ++axis)
	{
		if(dir[axis] != 0.f)
		{
			const float inv_dir = 1.f / dir[axis];
			float ltmin, ltmax;
			if(inv_dir > 0.f)
			{
				ltmin = -p[axis] * inv_dir;
				ltmax = ((g[axis] - a[axis]) - p[axis]) * inv_dir;
			}
			else
			{
				ltmin = ((g[axis] - a[axis]) - p[axis]) * inv_dir;
				ltmax = -p[axis] * inv_dir;
			}
			if(axis == 0)
			{
				lmin = ltmin;
				lmax = ltmax;
			}
			else
			{
				// std::max(ltmin, lmin) and std::min(ltmax, lmax) with their NaN
				// behaviour: a NaN ltmin / ltmax is kept, a NaN lmin / lmax replaced
				lmin = (ltmin < lmin) ? lmin : ltmin;
				lmax = (lmax < ltmax) ? lmax : ltmax;
			}
			if((lmax < 0.f) || (lmin > t_max)) return false;
		}
	}
	if((lmin <= lmax) && (lmax >= 0.f) && (lmin <= t_max))
	{
		enter = lmin;
		leave = lmax;
		return true;
	}
	return false;
}

// include/geometry/shape/shape_polygon.h:126-176 (N == 3) on a pre-formed
// record (v0, e1, e2).  Returns t > 0 on a hit, 0 otherwise.
YMX_RT_FN float tri_test(const float4 t0, const float4 t1, const float4 t2, const float o[3],
                                          const float d[3], float &u_out, float &v_out)
{
	const float e1x = t1.x, e1y = t1.y, e1z = t1.z;
	const float e2x = t2.x, e2y = t2.y, e2z = t2.z;
	// pvec = dir ^ edge_2
	const float px = d[1] * e2z - d[2] * e2y;
	const float py = d[2] * e2x - d[0] * e2z;
	const float pz = d[0] * e2y - d[1] * e2x;
	const float det = e1x * px + e1y * py + e1z * pz;
	if(det == 0.f) return 0.f;
	const float inv_det = 1.f / det;
	const float tx = o[0] - t0.x, ty = o[1] - t0.y, tz = o[2] - t0.z;
	const float u = (tx * px + ty * py + tz * pz) * inv_det;
	if(!(u >= 0.f && u <= 1.f)) return 0.f;
	// qvec = tvec ^ edge_1
	const float qx = ty * e1z - tz * e1y;
	const float qy = tz * e1x - tx * e1z;
	const float qz = tx * e1y - ty * e1x;
	const float v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv_det;
	if(!(v >= 0.f && (u + v) <= 1.f)) return 0.f;
	const float t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
	if(!(t > 0.f)) return 0.f;
	u_out = u;
	v_out = v;
	return t;
}

// include/geometry/shape/shape_polygon.h:126-176 (N == 4) on the record
// (v0, e1, e2) plus e3 = v3 - v0.  The second triangle (v0, v2, v3) is tried
// only when the first one's u is outside [0, 1]; det == 0 of the first, or a
// failed v / t test with u inside [0, 1], is a miss.  uv is {u + v, v} from the
// first triangle, {u, u + v} from the second.
YMX_RT_FN float quad_test(const float4 t0, const float4 t1, const float4 t2, const float4 e3, const float o[3],
                          const float d[3], float &u_out, float &v_out)
{
	const float e1x = t1.x, e1y = t1.y, e1z = t1.z;
	const float e2x = t2.x, e2y = t2.y, e2z = t2.z;
	// pvec_2 = dir ^ edge_2
	const float px = d[1] * e2z - d[2] * e2y;
	const float py = d[2] * e2x - d[0] * e2z;
	const float pz = d[0] * e2y - d[1] * e2x;
	const float det12 = e1x * px + e1y * py + e1z * pz;
	if(det12 == 0.f) return 0.f;
	const float inv12 = 1.f / det12;
	const float tx = o[0] - t0.x, ty = o[1] - t0.y, tz = o[2] - t0.z;
	const float u1 = (tx * px + ty * py + tz * pz) * inv12;
	if(u1 >= 0.f && u1 <= 1.f)
	{
		// qvec_1 = tvec ^ edge_1
		const float qx = ty * e1z - tz * e1y;
		const float qy = tz * e1x - tx * e1z;
		const float qz = tx * e1y - ty * e1x;
		const float v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv12;
		if(!(v >= 0.f && (u1 + v) <= 1.f)) return 0.f;
		const float t = (e2x * qx + e2y * qy + e2z * qz) * inv12;
		if(!(t > 0.f)) return 0.f;
		u_out = u1 + v;
		v_out = v;
		return t;
	}
	// pvec_3 = dir ^ edge_3
	const float sx = d[1] * e3.z - d[2] * e3.y;
	const float sy = d[2] * e3.x - d[0] * e3.z;
	const float sz = d[0] * e3.y - d[1] * e3.x;
	const float det23 = e2x * sx + e2y * sy + e2z * sz;
	if(det23 == 0.f) return 0.f;
	const float inv23 = 1.f / det23;
	const float u = (tx * sx + ty * sy + tz * sz) * inv23;
	if(!(u >= 0.f && u <= 1.f)) return 0.f;
	// qvec_2 = tvec ^ edge_2
	const float qx = ty * e2z - tz * e2y;
	const float qy = tz * e2x - tx * e2z;
	const float qz = tx * e2y - ty * e2x;
	const float v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv23;
	if(!(v >= 0.f && (u + v) <= 1.f)) return 0.f;
	const float t = (e3.x * qx + e3.y * qy + e3.z * qz) * inv23;
	if(!(t > 0.f)) return 0.f;
	u_out = u;
	v_out = u + v;
	return t;
}

// A primitive record: a triangle, or a quad when t2.w holds the 1-based index
// of its e3 = v3 - v0 in `quads` (0 for a triangle).
YMX_RT_FN float prim_test(const float4 t0, const float4 t1, const float4 t2, const float4 *quads, const float o[3],
                          const float d[3], float &u_out, float &v_out)
{
	uint32_t q;
	memcpy(&q, &t2.w, sizeof q);
	if(q == 0u) return tri_test(t0, t1, t2, o, d, u_out, v_out);
	return quad_test(t0, t1, t2, quads[q - 1u], o, d, u_out, v_out);
}

}  // namespace ymx
