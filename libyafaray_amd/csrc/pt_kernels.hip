// pt_kernels.hip -- wavefront kernels of the path tracer.
//
// A wavefront holds `count` AA samples of every pixel (path i traces pixel
// i % npix, sample first + i / npix).  Every call of the reference's
// PathIntegrator::integrate (src/integrator/surface/integrator_path_tracer.cc:138-314)
// is a *node*: the camera samples are the level-0 nodes; the recursive raytrace
// of a mirror, glass or glossy surface (integrator_montecarlo.cc:506-748) adds
// child nodes one level down.  Levels run in order, each in chunks of at most
// `capacity` nodes; the stages of one chunk are
//   extend        closest hit of the node rays (accel_trace.hip)
//   shade_node    emission + next-event estimation for all lights (shadow rays
//                 queued, contributions kept per slot); children written to the
//                 next level
//   shadow        any hit over the shadow queue; resolve_node sums the unoccluded
//                 slots in the reference's order
//   per sub-path: subpath_init samples the first bounce; then per segment d:
//   extend, shade_path (roulette, one-light estimation, emission, next BSDF
//   sample), shadow, resolve_path
//   finish        col += path_col / path_samples
// After the last level, combine adds each level's children into their parents
// (deepest first) in the reference's order, and film gather-splats the camera
// samples with the Gauss filter table in the order a row-major scatter would
// add them (ImageFilm::addSample).  Compiled with -ffp-contract=off and IEEE
// division: the shading keeps the reference's operation order (pt_math.h).
#include "guide_math.h"
#include "pt_internal.h"
#include "ray_tri.h"

#include <float.h>

#include <vector>

namespace ymx {
namespace {

constexpr int kBlock = 256;

enum StateBits : uint32_t
{
	kActive = 1u, kAddEmit = 2u, kCaustic = 4u, kDirect = 8u, kVertex = 16u, kPathBlock = 32u, kSampled = 64u
};

struct PtArgs
{
	const PtConst *c;
	const PtMaterial *mats;
	const PtLight *lights;
	const PtLightTri *ltris;
	const float4 *prim;
	const float4 *cnorm;   // corner normals, 3 per triangle (NULL: every triangle flat)
	const int32_t *faure;
	const double *inv_prims;
	const uint32_t *prims;
	PtState st;
	PtNodes nd;
	float4 *film;
	float *film_w;
	int32_t first_sample;  // AA sample index of the wavefront's first slice
	uint32_t npix;         // resx * resy
	uint32_t n_active;     // pixels of this wavefront's AA pass (npix, or those listed in pix_list)
	const uint32_t *pix_list;  // the pass's pixels in ascending order (NULL: every pixel)
	const uint32_t *pix_slot;  // per pixel its index in pix_list, kNoSlot if not in the pass
	float light_mult;      // the AA pass's light sample multiplier
	int32_t pass_samples;  // AA samples of the pass (single-pass sub-pixel offsets)
	uint32_t n_paths;      // n_active * count (level-0 nodes)
	int32_t count;         // AA samples in this wavefront
	uint32_t base;         // first node of the chunk: slot i is node base + i
	uint32_t cnt;          // nodes in the chunk
	uint32_t next_start;   // first node of the next level
	uint32_t recursive;    // children may be spawned from this chunk
	int32_t sub;
	int32_t depth;
	PtGuide g;             // path guiding (g.on == 0: BSDF sampling only)
};

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Camera sample `path` is sample first_sample + path / n_active of the pass's
// (path % n_active)-th pixel.
__device__ __forceinline__ uint32_t path_pixel(const PtArgs &a, uint32_t path)
{
	const uint32_t k = path % a.n_active;
	return a.pix_list ? a.pix_list[k] : k;
}

// What a kernel needs to know of the integrate() call behind slot i.
struct NodeCtx
{
	uint32_t n;          // node index
	uint32_t path;       // camera sample (node index of the level-0 ancestor)
	uint32_t pixel;      // camera_res_x * y + x
	uint32_t sample;     // AA sample index
	uint32_t offset;     // fnv32a(y * fnv32a(x))
	uint32_t level;      // ray_level of the integrate() call
	uint32_t division;   // RayDivision::division_
	int32_t rd_offset;   // RayDivision::offset_
	uint32_t tag;        // deterministic-stream tag of the call
};

__device__ __forceinline__ NodeCtx node_ctx(const PtArgs &a, uint32_t i)
{
	NodeCtx x;
	x.n = a.base + i;
	const uint4 info = a.nd.info[x.n];
	x.path = info.x;
	x.pixel = path_pixel(a, info.x);
	x.sample = (uint32_t)a.first_sample + info.x / a.n_active;
	const uint32_t px = x.pixel % (uint32_t)a.c->resx, py = x.pixel / (uint32_t)a.c->resx;
	x.offset = fnv32a(py * fnv32a(px));
	x.level = info.y & 0xFFu;
	x.division = info.y >> 8;
	x.tag = info.z;
	x.rd_offset = (int32_t)info.w;
	return x;
}

struct Surf
{
	V3 p, n, ng, u, v;
	uint32_t mat;
	uint32_t flags;
};

// Halton::lowDiscrepancySampling (src/sampler/halton.cc:419-440): Faure-scrambled
// radical inverse whose digits come from the reference's 9-decimal 1/p.
__device__ double low_discrepancy(const PtArgs &a, int dim, uint32_t n)
{
	double value = 0.0;
	if(dim >= 1 && dim < kLdsDims)
	{
		const uint32_t base = a.prims[dim];
		const double f = a.inv_prims[dim];
		const int32_t *sigma = a.faure + dim * kLdsMaxBase;
		double dn = (double)n;
		double factor = f;
		while(n > 0)
		{
			value += (double)sigma[n % base] * factor;
			dn *= f;
			n = (uint32_t)dn;
			factor *= f;
		}
	}
	return value;
}

// Vec3f::createCoordsSystem, include/geometry/vector.h:399-421
__device__ void coords_system(V3 n, V3 &u, V3 &v)
{
	if(n.x == 0.f && n.y == 0.f)
	{
		u = n.z < 0.f ? v3(-1.f, 0.f, 0.f) : v3(1.f, 0.f, 0.f);
		v = v3(0.f, 1.f, 0.f);
	}
	else
	{
		const float d = 1.f / sqrtf(n.y * n.y + n.x * n.x);
		u = v3(n.y * d, -n.x * d, 0.f);
		v = cross(n, u);
	}
}

// Accelerator::intersect(ray, camera) hit point and PrimitivePolygon::getSurfacePolygon
// for meshes without vertex normals (src/geometry/primitive/primitive_polygon.cc:37-115)
__device__ void make_surf(const PtArgs &a, V3 from, V3 dir, const ymx_hit &h, Surf &s)
{
	s.p = add(from, mul(h.t, dir));
	const float4 pi = a.prim[h.prim];
	s.ng = v3(pi.x, pi.y, pi.z);
	s.n = s.ng;
	if(a.cnorm)
	{
		// a smooth mesh or one with vertex normals (primitive_polygon.cc:45-51):
		// corner normals (the face normal for a corner without one,
		// FacePrimitive::getVertexNormal) weighted by (1 - u - v, u, v)
		const float4 *cn = a.cnorm + 3 * (size_t)h.prim;
		const float4 c0 = cn[0];
		if(c0.w >= 0.f)
		{
			const float4 c1 = cn[1], c2 = cn[2];
			const V3 n0 = c0.w > 0.f ? v3(c0.x, c0.y, c0.z) : s.ng;
			const V3 n1 = c1.w > 0.f ? v3(c1.x, c1.y, c1.z) : s.ng;
			const V3 n2 = c2.w > 0.f ? v3(c2.x, c2.y, c2.z) : s.ng;
			const float bu = 1.f - h.u - h.v, bv = h.u, bw = h.v;
			s.n = normalize(add(add(mul(bu, n0), mul(bv, n1)), mul(bw, n2)));
		}
	}
	coords_system(s.n, s.u, s.v);
	s.mat = __float_as_uint(pi.w);
	s.flags = a.mats[s.mat].flags;
}

__device__ __forceinline__ V3 face_forward(V3 ng, V3 n, V3 wo) { return dot(ng, wo) < 0.f ? neg(n) : n; }

struct BSample
{
	float s1, s2;
	uint32_t flags, sampled;
	float pdf;
};

// sample::cosHemisphere, include/sampler/sample.h:43-52
__device__ V3 cos_hemisphere(V3 n, V3 ru, V3 rv, float s1, float s2)
{
	if(s1 >= 1.0f) return n;
	const float z1 = s1;
	const float z2 = s2 * k2Pi;
	return add(mul(sqrtf(1.f - z1), add(mul(fast_cos(z2), ru), mul(fast_sin(z2), rv))), mul(sqrtf(z1), n));
}

// ---- ShinyDiffuseMaterial, diffuse component (src/material/material_shiny_diffuse.cc) ----

// eval (:297-340)
__device__ Rgb sd_eval(const PtMaterial &m, const Surf &sp, V3 wo, V3 wl, uint32_t bsdfs)
{
	const V3 n = face_forward(sp.ng, sp.n, wo);
	if(!(bsdfs & (sp.flags & kDiffuse))) return rgb(0.f, 0.f, 0.f);
	if(dot(n, wl) < 0.0) return rgb(0.f, 0.f, 0.f);
	const float m_d = 1.f * (1.f - 0.f) * m.diffuse_reflect;
	return cscale(m.color, m_d);
}

// pdf (:441-483)
__device__ float sd_pdf(const PtMaterial &m, const Surf &sp, V3 wo, V3 wi, uint32_t bsdfs)
{
	if(!(bsdfs & kDiffuse)) return 0.f;
	if(!(sp.flags & kDiffuse)) return 0.f;
	const V3 n = face_forward(sp.ng, sp.n, wo);
	float pdf = 0.f, sum = 0.f;
	int n_match = 0;
	if(bsdfs & (kDiffuse | kReflect))
	{
		const float width = m.diffuse_reflect;
		sum += width;
		pdf += fabsf(dot(wi, n)) * width;
		++n_match;
	}
	if(!n_match || sum < 0.00001) return 0.f;
	return pdf / sum;
}

// sample (:349-439)
__device__ Rgb sd_sample(const PtMaterial &m, const Surf &sp, V3 wo, V3 &wi, BSample &s, float &w)
{
	const float cos_ng_wo = dot(sp.ng, wo);
	const V3 n = face_forward(sp.ng, sp.n, wo);
	const uint32_t cflag = kDiffuse | kReflect;
	float sum = 0.f, width = 0.f;
	int n_match = 0;
	if((sp.flags & kDiffuse) && (s.flags & cflag) == cflag)
	{
		width = m.diffuse_reflect;
		sum += width;
		++n_match;
	}
	if(!n_match || sum < 0.00001)
	{
		s.sampled = kNone;
		s.pdf = 0.f;
		return rgb(1.f, 1.f, 1.f);
	}
	const float inv_sum = 1.f / sum;
	width *= inv_sum;
	const float s1 = s.s1 / width;
	Rgb scolor = rgb(0.f, 0.f, 0.f);
	wi = cos_hemisphere(n, sp.u, sp.v, s1, s.s2);
	if(cos_ng_wo * dot(sp.ng, wi) > 0) scolor = cscale(m.color, m.diffuse_reflect);
	s.pdf = fabsf(dot(wi, n)) * width;
	s.sampled = cflag;
	w = fabsf(dot(wi, sp.n)) / (s.pdf * 0.99f + 0.01f);
	w = w * 1.f + 1.f * (1.f - 1.f);  // getAlpha() == 1
	return scolor;
}

// ---- microfacet helpers (include/material/material_utils_microfacet.h) ----

__device__ __forceinline__ float pdf_divisor(float cos) { return 8.f * kPi * (cos * 0.99f + 0.04f); }
__device__ __forceinline__ float as_divisor(float cos_1, float cos_i, float cos_o)
{
	return 8.f * kPi * ((cos_1 * std_max(cos_i, cos_o)) * 0.99f + 0.04f);
}
__device__ __forceinline__ float blinn_d(float cos_h, float e) { return (e + 1.f) * fast_pow(cos_h, e); }
__device__ __forceinline__ float blinn_pdf(float costheta, float cos_w_h, float e)
{
	return blinn_d(costheta, e) / pdf_divisor(cos_w_h);
}
__device__ V3 blinn_sample(float s_1, float s_2, float exponent)
{
	const float cos_theta = fast_pow(1.f - s_2, 1.f / (exponent + 1.f));
	const float sin_theta = sqrtf(1.f - cos_theta * cos_theta);
	const float phi = s_1 * k2Pi;
	return v3(sin_theta * fast_cos(phi), sin_theta * fast_sin(phi), cos_theta);
}
__device__ __forceinline__ float schlick_fresnel(float costheta, float r)
{
	const float c_1 = (1.f - costheta);
	const float c_2 = c_1 * c_1;
	return r + ((1.f - r) * c_1 * c_2 * c_2);
}
// diffuseReflect: the scale in double (diffuse_ratio is a double constant), rounded to float
__device__ Rgb diffuse_reflect(float wi_n, float wo_n, float glossy, float diffuse, Rgb diff_base)
{
	const double diffuse_ratio = 0.387507688;
	float f_wi = 1.f - (0.5f * wi_n);
	float temp = f_wi * f_wi;
	f_wi = temp * temp * f_wi;
	float f_wo = (1.f - (0.5f * wo_n));
	temp = f_wo * f_wo;
	f_wo = temp * temp * f_wo;
	const float k = (float)(diffuse_ratio * diffuse * (1.f - glossy) * (1.f - f_wi) * (1.f - f_wo));
	return cscale(diff_base, k);
}

// ---- GlossyMaterial, isotropic (src/material/material_glossy.cc) ----

// eval (:200-250)
__device__ Rgb glossy_eval(const PtMaterial &m, const Surf &sp, V3 wo, V3 wl, uint32_t bsdfs)
{
	if(!(bsdfs & kDiffuse) || ((dot(sp.ng, wl) * dot(sp.ng, wo)) < 0.f)) return rgb(0.f, 0.f, 0.f);
	Rgb col = rgb(0.f, 0.f, 0.f);
	const bool diffuse_flag = (bsdfs & kDiffuse) != 0;
	const V3 n = face_forward(sp.ng, sp.n, wo);
	const float wi_n = fabsf(dot(wl, n));
	const float wo_n = fabsf(dot(wo, n));
	if((m.as_diffuse && diffuse_flag) || (!m.as_diffuse && (bsdfs & kGlossy)))
	{
		const V3 h = normalize(add(wo, wl));
		const float cos_wi_h = std_max(0.f, dot(wl, h));
		const float glossy =
			blinn_d(dot(h, n), m.exponent) * schlick_fresnel(cos_wi_h, m.glossy_reflect) / as_divisor(cos_wi_h, wo_n, wi_n);
		col = cscale(m.color, glossy);
	}
	if(m.diffuse_reflect > 0.f && diffuse_flag)
		col = cadd(col, cscale(m.diffuse_color, m.diffuse_reflect * (1.f - m.glossy_reflect)));
	return col;
}

// pdf (:404-440)
__device__ float glossy_pdf(const PtMaterial &m, const Surf &sp, V3 wo, V3 wi, uint32_t flags)
{
	if((dot(sp.ng, wo) * dot(sp.ng, wi)) < 0.f) return 0.f;
	const V3 n = face_forward(sp.ng, sp.n, wo);
	float pdf = 0.f;
	const float cur_p_diffuse = m.p_diffuse;
	const bool use_glossy = m.as_diffuse ? (flags & kDiffuse) != 0 : (flags & kGlossy) != 0;
	const bool use_diffuse = m.diffuse_reflect > 0.f && (flags & kDiffuse);
	if(use_diffuse)
	{
		pdf = fabsf(dot(wi, n));
		if(use_glossy)
		{
			const V3 h = normalize(add(wi, wo));
			const float cos_wo_h = dot(wo, h);
			const float cos_n_h = dot(n, h);
			pdf = pdf * cur_p_diffuse + blinn_pdf(cos_n_h, cos_wo_h, m.exponent) * (1.f - cur_p_diffuse);
		}
		return pdf;
	}
	if(use_glossy)
	{
		const V3 h = normalize(add(wi, wo));
		const float cos_wo_h = dot(wo, h);
		const float cos_n_h = dot(n, h);
		pdf = blinn_pdf(cos_n_h, cos_wo_h, m.exponent);
	}
	return pdf;
}

// sample (:252-402); early returns leave w unchanged
__device__ Rgb glossy_sample(const PtMaterial &m, const Surf &sp, V3 wo, V3 &wi, BSample &s, float &w)
{
	const float cos_ng_wo = dot(sp.ng, wo);
	const V3 n = face_forward(sp.ng, sp.n, wo);
	s.pdf = 0.f;
	float wi_n = 0.f;
	const float wo_n = fabsf(dot(wo, n));
	Rgb scolor = rgb(0.f, 0.f, 0.f);
	float s_1 = s.s1;
	const float cur_p_diffuse = m.p_diffuse;
	const bool use_glossy = m.as_diffuse ? (s.flags & kDiffuse) != 0 : (s.flags & kGlossy) != 0;
	const bool use_diffuse = m.diffuse_reflect > 0.f && (s.flags & kDiffuse);
	if(use_diffuse)
	{
		const float s_p_diffuse = use_glossy ? cur_p_diffuse : 1.f;
		if(s_1 < s_p_diffuse)
		{
			s_1 /= s_p_diffuse;
			wi = cos_hemisphere(n, sp.u, sp.v, s_1, s.s2);
			const float cos_ng_wi = dot(sp.ng, wi);
			if(cos_ng_wi * cos_ng_wo < 0.f) return scolor;
			wi_n = fabsf(dot(wi, n));
			s.pdf = wi_n;
			float glossy = 0.f;
			if(use_glossy)
			{
				const V3 h = normalize(add(wi, wo));
				const float cos_wo_h = dot(wo, h);
				const float cos_wi_h = fabsf(dot(wi, h));
				const float cos_n_h = dot(n, h);
				s.pdf = s.pdf * cur_p_diffuse + blinn_pdf(cos_n_h, cos_wo_h, m.exponent) * (1.f - cur_p_diffuse);
				glossy = blinn_d(cos_n_h, m.exponent) * schlick_fresnel(cos_wi_h, m.glossy_reflect) /
				         as_divisor(cos_wi_h, wo_n, wi_n);
			}
			s.sampled = kDiffuse | kReflect;
			if(!(s.flags & kReflect)) return rgb(0.f, 0.f, 0.f);
			scolor = cscale(m.color, glossy);
			scolor = cadd(scolor, diffuse_reflect(wi_n, wo_n, m.glossy_reflect, m.diffuse_reflect, m.diffuse_color));
			w = wi_n / (s.pdf * 0.99f + 0.01f);
			return scolor;
		}
		s_1 -= cur_p_diffuse;
		s_1 /= (1.f - cur_p_diffuse);
	}
	if(use_glossy)
	{
		const V3 hs = blinn_sample(s_1, s.s2, m.exponent);
		V3 h = add(add(mul(hs.x, sp.u), mul(hs.y, sp.v)), mul(hs.z, n));
		float cos_wo_h = dot(wo, h);
		if(cos_wo_h < 0.f)
		{
			h = reflect_about(h, n);
			cos_wo_h = dot(wo, h);
		}
		wi = reflect_dir(h, wo);
		const float cos_ng_wi = dot(sp.ng, wi);
		if(cos_ng_wo * cos_ng_wi < 0.f) return rgb(0.f, 0.f, 0.f);
		wi_n = fabsf(dot(wi, n));
		const float cos_hn = dot(h, n);
		s.pdf = blinn_pdf(cos_hn, cos_wo_h, m.exponent);
		const float glossy = blinn_d(cos_hn, m.exponent) * schlick_fresnel(cos_wo_h, m.glossy_reflect) /
		                     as_divisor(cos_wo_h, wo_n, wi_n);
		scolor = cscale(m.color, glossy);
		s.sampled = m.as_diffuse ? (kDiffuse | kReflect) : (kGlossy | kReflect);
	}
	if(use_diffuse)
	{
		const Rgb add_col = diffuse_reflect(wi_n, wo_n, m.glossy_reflect, m.diffuse_reflect, m.diffuse_color);
		s.pdf = wi_n * cur_p_diffuse + s.pdf * (1.f - cur_p_diffuse);
		scolor = cadd(scolor, add_col);
	}
	w = wi_n / (s.pdf * 0.99f + 0.01f);
	return scolor;
}

// ---- GlassMaterial without dispersion (src/material/material_glass.cc:168-378) ----

// Vec3f::refract, include/geometry/vector.h:484-504
__device__ bool refract(V3 n, V3 wi, V3 &wo, float ior)
{
	float eta;
	float cos_v_n = dot(wi, n);
	if(cos_v_n < 0.f)
	{
		eta = ior;
		n = neg(n);
		cos_v_n = -cos_v_n;
	}
	else eta = 1.f / ior;
	const float k = 1.f - eta * eta * (1.f - cos_v_n * cos_v_n);
	if(k <= 0.f) return false;
	wo = normalize(add(mul(eta, neg(wi)), mul(eta * cos_v_n - sqrtf(k), n)));
	return true;
}

// Vec3f::fresnel, include/geometry/vector.h:506-518
__device__ void fresnel(V3 i, V3 n, float ior, float &kr, float &kt)
{
	const bool negative = dot(i, n) < 0.f;
	const float c = dot(i, negative ? neg(n) : n);
	float g = ior * ior + c * c - 1.f;
	if(g <= 0.f) g = 0.f;
	else g = sqrtf(g);
	const float aux = c * (g + c);
	kr = ((0.5f * (g - c) * (g - c)) / ((g + c) * (g + c))) *
	     (1.f + ((aux - 1.f) * (aux - 1.f)) / ((aux + 1.f) * (aux + 1.f)));
	kt = (kr < 1.f) ? 1.f - kr : 0.f;
}

// the normal bent towards wo (material_glass.cc:172-174, :307-317)
__device__ V3 glass_normal(const Surf &sp, V3 wo, bool &outside)
{
	outside = dot(sp.ng, wo) > 0.f;
	const float cos_wo_n = dot(sp.n, wo);
	const bool keep = outside ? cos_wo_n >= 0.f : cos_wo_n <= 0.f;
	return keep ? sp.n : normalize(sub(sp.n, mul(1.00001f * cos_wo_n, wo)));
}

__device__ Rgb glass_sample(const PtMaterial &m, const Surf &sp, V3 wo, V3 &wi, BSample &s, float &w)
{
	if(!(s.flags & kSpecular))
	{
		s.pdf = 0.f;
		return rgb(0.f, 0.f, 0.f);
	}
	bool outside;
	const V3 n = glass_normal(sp, wo, outside);
	s.pdf = 1.f;
	V3 refdir;
	if(refract(n, wo, refdir, m.ior))
	{
		float kr, kt;
		fresnel(wo, n, m.ior, kr, kt);
		const float p_kr = (float)(0.01 + 0.99 * (double)kr), p_kt = (float)(0.01 + 0.99 * (double)kt);
		if(s.s1 < p_kt && (s.flags & m.transmit_flags))
		{
			wi = refdir;
			s.pdf = p_kt;
			s.sampled = m.transmit_flags;
			w = 1.f;
			return m.filter_color;
		}
		else if(s.flags & (kSpecular | kReflect))
		{
			wi = reflect_about(wo, n);
			s.pdf = p_kr;
			s.sampled = kSpecular | kReflect;
			w = 1.f;
			return m.mirror_color;
		}
	}
	else if(s.flags & (kSpecular | kReflect))
	{
		wi = reflect_about(wo, n);
		s.sampled = kSpecular | kReflect;
		w = 1.f;
		return rgb(1.f, 1.f, 1.f);
	}
	s.pdf = 0.f;
	return rgb(0.f, 0.f, 0.f);
}

// ---- dispatch over the material types ----

__device__ Rgb bsdf_eval(const PtArgs &a, const Surf &sp, V3 wo, V3 wl, uint32_t bsdfs)
{
	const PtMaterial &m = a.mats[sp.mat];
	if(m.type == YMX_MAT_SHINYDIFFUSE) return sd_eval(m, sp, wo, wl, bsdfs);
	if(m.type == YMX_MAT_GLOSSY) return glossy_eval(m, sp, wo, wl, bsdfs);
	return rgb(0.f, 0.f, 0.f);  // mirror, glass, light_mat
}

__device__ float bsdf_pdf(const PtArgs &a, const Surf &sp, V3 wo, V3 wi, uint32_t bsdfs)
{
	const PtMaterial &m = a.mats[sp.mat];
	if(m.type == YMX_MAT_SHINYDIFFUSE) return sd_pdf(m, sp, wo, wi, bsdfs);
	if(m.type == YMX_MAT_GLOSSY) return glossy_pdf(m, sp, wo, wi, bsdfs);
	return 0.f;  // Material::pdf default
}

__device__ Rgb bsdf_sample(const PtArgs &a, const Surf &sp, V3 wo, V3 &wi, BSample &s, float &w)
{
	const PtMaterial &m = a.mats[sp.mat];
	switch(m.type)
	{
		case YMX_MAT_SHINYDIFFUSE: return sd_sample(m, sp, wo, wi, s, w);
		case YMX_MAT_GLOSSY: return glossy_sample(m, sp, wo, wi, s, w);
		case YMX_MAT_MIRROR:  // MirrorMaterial::sample, src/material/material_mirror.cc:70-76
			wi = reflect_dir(sp.n, wo);
			s.sampled = kSpecular | kReflect;
			w = 1.f;
			return cscale(m.ref_col, 1.f / fabsf(dot(sp.n, wi)));
		case YMX_MAT_GLASS: return glass_sample(m, sp, wo, wi, s, w);
		default:  // LightMaterial::sample, src/material/material_light.cc:75-80
			s.pdf = 0.f;
			w = 0.f;
			return rgb(0.f, 0.f, 0.f);
	}
}

// Material::emit: shinydiffuse emit * color; LightMaterial::emit (src/material/material_light.cc:82-87)
__device__ Rgb mat_emit(const PtArgs &a, const Surf &sp, V3 wo)
{
	const PtMaterial &m = a.mats[sp.mat];
	if(m.type == YMX_MAT_LIGHT && !m.double_sided && !(dot(wo, sp.n) > 0.f)) return rgb(0.f, 0.f, 0.f);
	return m.emit_color;
}

// ---- path guiding (one-sample MIS of the BSDF sampler and the guide's mixture) ----

__device__ __forceinline__ bool guided_material(const PtArgs &a, const Surf &sp)
{
	return a.g.on && a.mats[sp.mat].type == YMX_MAT_SHINYDIFFUSE && (sp.flags & kDiffuse);
}

// What the shinydiffuse sampler (sd_sample) returns as colour and weight w had
// it drawn the unit direction wi.
__device__ Rgb sd_weight(const PtMaterial &m, const Surf &sp, V3 wo, V3 wi, float &w)
{
	const V3 n = face_forward(sp.ng, sp.n, wo);
	const Rgb scolor = (dot(sp.ng, wo) * dot(sp.ng, wi) > 0) ? cscale(m.color, m.diffuse_reflect) : rgb(0.f, 0.f, 0.f);
	const float pdf = fabsf(dot(wi, n));  // width 1: the diffuse component is the only one
	w = fabsf(dot(wi, sp.n)) / (pdf * 0.99f + 0.01f);
	return scolor;
}

// Uniforms of the guided sampling (selection and direction, k = 0..3): over a
// pixel's samples ps, the 4-D Kronecker sequence frac(ps g_k) with
// g_k = phi^-(k+1), phi the real root of x^5 = x + 1 (M. Roberts' R4
// sequence), in 32-bit fixed point, shifted by a random per-pixel, per-segment
// offset (Cranley-Patterson rotation, from the counter-based stream, distinct
// from the roulette's).  Each point is uniform, so the estimator stays
// unbiased, and the rotation is independent of the reference's sample
// dimensions; the sequence keeps a pixel's guided samples stratified.
__device__ __forceinline__ float guide_rand(const NodeCtx &x, uint32_t ps, uint32_t seg, uint32_t seed, uint32_t k)
{
	constexpr uint32_t kGen[4] = {3679390609u, 3152041523u, 2700274806u, 2313257605u};  // g_k * 2^32
	const float rot = rr_random(x.pixel ^ (0x632BE59Bu * (k + 1u)), 0u, seg + 256u * x.tag + 0x10000u * (k + 1u),
	                            seed ^ 0xA511E9B3u);
	const uint32_t u = ps * kGen[k & 3u] + (uint32_t)(rot * 4294967296.f);
	return (float)(u >> 8) * 0x1p-24f;
}

// Region of a position (training-weight normalisation, learned selection).
__device__ __forceinline__ uint32_t guide_cell(const PtGuide &g, const float p[3])
{
	uint32_t c[3];
	for(int k = 0; k < 3; ++k)
	{
		const float f = (p[k] - g.bmin[k]) * g.scale[k];
		c[k] = !(f > 0.f) ? 0u : (f >= (float)(kGuideCells - 1) ? (uint32_t)(kGuideCells - 1) : (uint32_t)f);
	}
	return (c[2] * kGuideCells + c[1]) * kGuideCells + c[0];
}

// The selection probability of region `cell` before the fallback: the fixed
// bsdf_prob, or the region's learned value.
__device__ __forceinline__ float region_alpha(const PtGuide &g, uint32_t cell)
{
	if(!g.learn_alpha) return g.bsdf_prob;
	return kAlphaMin + (1.f - kAlphaMin) / (1.f + __expf(-g.alpha_theta[cell]));
}

// The probability of sampling the BSDF at a guided vertex at p: 1 in a region
// fallen back to BSDF sampling (its running guided second moment above the
// BSDF's, after at least kFallbackMinRecords records), else the region's alpha.
__device__ __forceinline__ float selection_prob(const PtGuide &g, V3 p)
{
	if(!g.learn_alpha && !g.fallback) return g.bsdf_prob;
	const float pp[3] = {p.x, p.y, p.z};
	const uint32_t cell = guide_cell(g, pp);
	constexpr uint32_t kCells = (uint32_t)(kGuideCells * kGuideCells * kGuideCells);
	if(g.fallback && g.fb_run[2 * kCells + cell] >= kFallbackMinRecords && g.fb_run[cell] > g.fb_run[kCells + cell])
		return 1.f;
	return region_alpha(g, cell);
}

// The densities a guided direction was drawn with: the BSDF's cosine lobe p,
// the mixture q and the selection probability alpha (kept in the training
// record for the selection's gradient).
struct GuidePdfs
{
	float p, q, alpha;
};

// Next direction at a guided vertex: the BSDF's own sample with probability
// alpha, otherwise the mixture's; weight = F / (alpha p + (1 - alpha) q) with
// F = sw * p the plain sampler's integrand, p = cos / pi the density of its
// cosine lobe and q the mixture's.  The BSDF branch keeps the sampler's own
// direction and weight (its direction is unit only to the accuracy of the
// fast sine and cosine); both densities are evaluated at the unit direction,
// since a sharp lobe's exp(kappa (mu . w - 1)) amplifies any length error by
// kappa.  At alpha 1 (a region fallen back to BSDF sampling, or bsdf_prob 1)
// the vertex is the plain tracer's: its own low-discrepancy sample and weight
// sw.  w_plain: the sampler's w for wi (what a later vertex that leaves w
// unset reuses).  Returns false when the weight is 0.
__device__ bool guided_sample(const PtArgs &a, const Surf &sp, V3 wo, const float *mix, BSample &s, float u_sel,
                              float u0, float u1, float u2, V3 &wi, Rgb &weight, float &w_plain, GuidePdfs &pdfs)
{
	const PtMaterial &m = a.mats[sp.mat];
	const float alpha = selection_prob(a.g, sp.p);
	Rgb sw;
	if(u_sel < alpha)
	{
		// The BSDF branch draws its direction from the guide's own uniforms, not
		// from the sample's low-discrepancy dimensions: those are correlated with
		// the light samples of the next vertex, and weighting them by the MIS
		// factor p / (alpha p + (1 - alpha) q) would no longer average to the
		// plain estimate (it leaves the camera-vertex bounce about 1 % high).
		// At alpha 1 (a region fallen back, or bsdf_prob 1) the factor is 1 and
		// the vertex is the plain tracer's: it keeps the sample's own dimensions.
		if(alpha < 1.f)
		{
			s.s1 = u1;
			s.s2 = u2;
		}
		w_plain = 0.f;
		sw = cscale(sd_sample(m, sp, wo, wi, s, w_plain), w_plain);
	}
	else
	{
		float x, y, z;
		ymx::guide::mix_sample(mix, u0, u1, u2, x, y, z);
		wi = normalize(v3(x, y, z));
		s.sampled = kDiffuse | kReflect;
		s.pdf = 0.f;
		sw = cscale(sd_weight(m, sp, wo, wi, w_plain), w_plain);
	}
	const V3 wn = normalize(wi);
	const float cos_n = dot(wn, face_forward(sp.ng, sp.n, wo));
	const float p = cos_n > 0.f ? cos_n * k1Pi : 0.f;
	const float q = ymx::guide::mix_pdf(mix, wn.x, wn.y, wn.z);
	const float denom = alpha * p + (1.f - alpha) * q;
	pdfs.p = p;
	pdfs.q = q;
	pdfs.alpha = alpha;
	if(alpha >= 1.f)
	{
		weight = sw;
		return !cblack(sw);
	}
	if(!(p > 0.f) || !(denom > 0.f))
	{
		weight = rgb(0.f, 0.f, 0.f);
		return false;
	}
	weight = cscale(sw, p / denom);
	return true;
}

__device__ __forceinline__ float luminance(Rgb c) { return 0.2126f * c.r + 0.7152f * c.g + 0.0722f * c.b; }

// A pending training sample of segment r (position, normal, wo, direction) and
// the throughput before it; the path colour before it is set in base.xyz.
__device__ __forceinline__ void record_segment(const PtArgs &a, uint32_t slot, uint32_t r, const Surf &sp, V3 wo, V3 wi,
                                               float tp_lum, const GuidePdfs &pdfs)
{
	ymx_guide_sample rec;
	const V3 n = face_forward(sp.ng, sp.n, wo);
	rec.pos[0] = sp.p.x;
	rec.pos[1] = sp.p.y;
	rec.pos[2] = sp.p.z;
	rec.c = 0.f;
	rec.normal[0] = n.x;
	rec.normal[1] = n.y;
	rec.normal[2] = n.z;
	rec.pad0 = pdfs.p;      // the densities of the draw, for the selection's gradient
	rec.wo[0] = wo.x;
	rec.wo[1] = wo.y;
	rec.wo[2] = wo.z;
	rec.pad1 = pdfs.q;
	rec.w[0] = wi.x;
	rec.w[1] = wi.y;
	rec.w[2] = wi.z;
	rec.pad2 = pdfs.alpha;
	const size_t k = (size_t)r * a.cnt + slot;
	a.g.rec[k] = rec;
	a.g.base[k].w = tp_lum;
	a.g.mask[slot] |= 1u << r;
}

// ShapePolygon<float, 4>::intersect, include/geometry/shape/shape_polygon.h:126-176
__device__ float quad_intersect(const V3 q[4], V3 from, V3 dir)
{
	const V3 e1 = sub(q[1], q[0]);
	const V3 e2 = sub(q[2], q[0]);
	const V3 p2 = cross(dir, e2);
	const float det12 = dot(e1, p2);
	if(det12 != 0.f)
	{
		const float inv12 = 1.f / det12;
		const V3 tvec = sub(from, q[0]);
		float u = dot(tvec, p2) * inv12;
		if(u >= 0.f && u <= 1.f)
		{
			const V3 q1 = cross(tvec, e1);
			const float v = dot(dir, q1) * inv12;
			if(v >= 0.f && (u + v) <= 1.f)
			{
				const float t = dot(e2, q1) * inv12;
				if(t > 0.f) return t;
			}
		}
		else
		{
			const V3 e3 = sub(q[3], q[0]);
			const V3 p3 = cross(dir, e3);
			const float det23 = dot(e2, p3);
			if(det23 != 0.f)
			{
				const float inv23 = 1.f / det23;
				u = dot(tvec, p3) * inv23;
				if(u >= 0.f && u <= 1.f)
				{
					const V3 q2 = cross(tvec, e2);
					const float v = dot(dir, q2) * inv23;
					if(v >= 0.f && (u + v) <= 1.f)
					{
						const float t = dot(e3, q2) * inv23;
						if(t > 0.f) return t;
					}
				}
			}
		}
	}
	return 0.f;
}

// doLightEstimation's sample count: ceil(samples x the AA pass's light sample
// multiplier), divided by the ray division (integrator_montecarlo.cc:427-429)
__device__ __forceinline__ int light_samples(const PtLight &l, uint32_t division, float mult)
{
	int num_samples = (int)ceilf((float)l.samples * mult);
	if(division > 1) num_samples = num_samples / (int)division > 1 ? num_samples / (int)division : 1;
	return num_samples;
}

__device__ __forceinline__ uint32_t light_slots(const PtLight &l, uint32_t division, float mult)
{
	return l.type == YMX_LIGHT_POINT ? 1u : 2u * (uint32_t)light_samples(l, division, mult);
}

// Accelerator::isShadowed (include/accelerator/accelerator.h:103-111): origin
// moved by tmin, t_max = tmax - 2 tmin.  Returns the queue index, or -1 when no
// test is needed (t_max <= 0 can never be occluded).
//
// The shadow queue is segmented (kQueueSegments segments of sh_seg_cap rays):
// a wave appends to segment (its index mod kQueueSegments) with one atomic
// for the lanes appending together, so the appends spread over the
// segments' counters.  sh_seg_cap is sized for every wave of a segment
// filling all its slots, so no segment overflows.
__device__ int32_t queue_shadow(const PtArgs &a, V3 from, V3 dir, float tmin, float tmax)
{
	const float t_max = (tmax >= 0.f) ? tmax - 2 * tmin : FLT_MAX;
	if(!(t_max > 0.f)) return -1;
	const V3 f = add(from, mul(tmin, dir));
	const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * kBlock + threadIdx.x) >> 6);
	const uint32_t seg = wave % kQueueSegments;
	const unsigned long long active = __ballot(true);
	const int leader = __ffsll((long long)active) - 1;
	const int lane = (int)__lane_id();
	uint32_t base = 0;
	if(lane == leader) base = atomicAdd(a.st.sh_count + (size_t)seg * kQueueStride, (uint32_t)__popcll(active));
	base = __shfl(base, leader);
	const uint32_t off = base + (uint32_t)__popcll(active & ((1ull << lane) - 1ull));
	if(off >= a.st.sh_seg_cap) return -1;  // cannot happen with the sizing above; keeps the write in bounds
	const uint32_t q = seg * a.st.sh_seg_cap + off;
	float4 *r = reinterpret_cast<float4 *>(a.st.sh_rays + q);
	r[0] = make_float4(f.x, f.y, f.z, tmin);
	r[1] = make_float4(dir.x, dir.y, dir.z, t_max);
	return (int32_t)q;
}

// A slot holds one light sample's contribution when unoccluded.  A zero
// contribution is stored as no contribution (adding +0 changes no sum) and
// needs no shadow ray.  Slots are laid out column by column: slot j of path i
// is entry j x capacity + i, so the lanes of a wave, which shade consecutive
// paths, write and read each slot column in whole cache lines.
__device__ __forceinline__ uint32_t slot_at(const PtArgs &a, uint32_t slot0, uint32_t j)
{
	return slot0 + j * a.st.capacity;
}
__device__ __forceinline__ void set_slot_empty(const PtArgs &a, uint32_t slot)
{
	a.st.slot_contrib[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
	a.st.slot_queue[slot] = -1;
}
__device__ __forceinline__ void set_slot(const PtArgs &a, uint32_t slot, Rgb c, bool cast, V3 from, V3 dir, float tmin,
                                         float tmax)
{
	if(cblack(c))
	{
		set_slot_empty(a, slot);
		return;
	}
	a.st.slot_contrib[slot] = make_float4(c.r, c.g, c.b, 1.f);
	a.st.slot_queue[slot] = cast ? queue_shadow(a, from, dir, tmin, tmax) : -1;
}

// ObjectLight::sampleSurface (src/light/light_object_light.cc:124-139): the
// triangle from the light's area cdf by Pdf1D::dSample (std::lower_bound,
// include/sampler/sample_pdf1d.h:83-95), the first coordinate rescaled to that
// triangle's cdf interval, then ShapePolygon::sample's uniform triangle point
// (include/geometry/shape/shape_polygon.h:78-90).
__device__ void obj_sample_surface(const PtArgs &a, const PtLight &l, float s1, float s2, V3 &p, V3 &n)
{
	const PtLightTri *lt = a.ltris + l.first;
	uint32_t idx;
	if(s1 <= 0.f) idx = 0;
	else if(s1 >= 1.f) idx = l.count - 1;
	else
	{
		uint32_t lo = 0, hi = l.count;
		while(lo < hi)
		{
			const uint32_t mid = lo + (hi - lo) / 2;
			if(lt[mid].v0_cdf.w < s1) lo = mid + 1;
			else hi = mid;
		}
		idx = lo;
	}
	const float cdf = lt[idx].v0_cdf.w;
	float delta = cdf, ss1;
	if(idx > 0)
	{
		const float prev = lt[idx - 1].v0_cdf.w;
		delta -= prev;
		ss1 = (s1 - prev) / delta;
	}
	else ss1 = s1 / delta;
	const PtLightTri t = lt[idx];
	const float su1 = sqrtf(ss1);
	const float u = 1.f - su1;
	const float v = s2 * su1;
	const float w = 1.f - u - v;
	p = add(add(v3(u * t.v0_cdf.x, u * t.v0_cdf.y, u * t.v0_cdf.z), v3(v * t.v1_vis.x, v * t.v1_vis.y, v * t.v1_vis.z)),
	        v3(w * t.v2.x, w * t.v2.y, w * t.v2.z));
	n = v3(t.n.x, t.n.y, t.n.z);
}

// Light::illumSample of the area, object and portal lights (src/light/light_area.cc:99-124,
// src/light/light_object_light.cc:141-171, src/light/light_background_portal.cc:157-181;
// the portal is one-sided and has no guard on a zero area * cos).  False when
// the light gives no sample.
__device__ bool illum_sample(const PtArgs &a, const PtLight &l, V3 sp_p, float s1, float s2, V3 &ldir_out,
                             float &dist_out, float &pdf_out)
{
	V3 ldir;
	float dist_sqr, cos_angle, area_cos;
	if(l.type == YMX_LIGHT_OBJECT || l.type == YMX_LIGHT_PORTAL)
	{
		V3 p, n;
		obj_sample_surface(a, l, s1, s2, p, n);
		ldir = sub(p, sp_p);
		dist_sqr = dot(ldir, ldir);
		const float dist = sqrtf(dist_sqr);
		if(dist <= 0.f) return false;
		ldir = mul(1.f / dist, ldir);
		cos_angle = -dot(ldir, n);
		if(cos_angle <= 0.f)
		{
			if(l.double_sided) cos_angle = -cos_angle;
			else return false;
		}
		area_cos = l.area * cos_angle;
		pdf_out = dist_sqr * kPi / ((area_cos == 0.f && l.type == YMX_LIGHT_OBJECT) ? 1e-8f : area_cos);
		dist_out = dist;
	}
	else
	{
		const V3 p = add(add(l.corner, mul(s1, l.to_x)), mul(s2, l.to_y));
		ldir = sub(p, sp_p);
		dist_sqr = dot(ldir, ldir);
		const float dist = sqrtf(dist_sqr);
		if(dist <= 0.f) return false;
		ldir = mul(1.f / dist, ldir);
		cos_angle = dot(ldir, l.normal_flipped);
		if(cos_angle <= 0.f) return false;
		pdf_out = dist_sqr * kPi / (l.area * cos_angle);
		dist_out = dist;
	}
	ldir_out = ldir;
	return true;
}

// The closest visible triangle of a light's own BVH (the node and triangle
// layout of accel_internal.h, ids 0 .. count - 1) along o + t d with t in
// [t_min, FLT_MAX): the one the loop over the light's triangles in id order
// finds, since the box tests only cull (ray_tri.h) and an exact tie keeps the
// lower id.  -1 when none.  The stack holds at most the tree's depth
// (< kLightStack) entries: one deferred sibling per level of the path.
__device__ int32_t light_tree_closest(const float4 *nodes, const float4 *tris, const float o[3], const float d[3],
                                      float t_min)
{
	float idir[3];
	for(int k = 0; k < 3; ++k)
	{
		const float dk = (fabsf(d[k]) > 1e-30f) ? d[k] : copysignf(1e-30f, d[k]);
		idir[k] = 1.f / dk;
	}
	const float tlo = t_min - t_min * ymx::kBoxWindow;
	float best_t = FLT_MAX;
	int32_t best = -1;
	int32_t stack[kLightStack];
	uint32_t sp = 0;
	int32_t ref = 0;
	for(;;)
	{
		if(ref >= 0)
		{
			const float4 *np = nodes + 4 * (size_t)ref;
			const float4 f0 = np[0], f1 = np[1], f2 = np[2], f3 = np[3];
			const float thi = best_t + best_t * ymx::kBoxWindow;
			const float a0x = (f0.x - o[0]) * idir[0], b0x = (f0.y - o[0]) * idir[0];
			const float a0y = (f0.z - o[1]) * idir[1], b0y = (f0.w - o[1]) * idir[1];
			const float a0z = (f2.x - o[2]) * idir[2], b0z = (f2.y - o[2]) * idir[2];
			const float a1x = (f1.x - o[0]) * idir[0], b1x = (f1.y - o[0]) * idir[0];
			const float a1y = (f1.z - o[1]) * idir[1], b1y = (f1.w - o[1]) * idir[1];
			const float a1z = (f2.z - o[2]) * idir[2], b1z = (f2.w - o[2]) * idir[2];
			const float n0 = fmaxf(fmaxf(fminf(a0x, b0x), fminf(a0y, b0y)), fmaxf(fminf(a0z, b0z), tlo));
			const float x0 = fminf(fminf(fmaxf(a0x, b0x), fmaxf(a0y, b0y)), fmaxf(a0z, b0z)) * ymx::kSlabWiden;
			const float n1 = fmaxf(fmaxf(fminf(a1x, b1x), fminf(a1y, b1y)), fmaxf(fminf(a1z, b1z), tlo));
			const float x1 = fminf(fminf(fmaxf(a1x, b1x), fmaxf(a1y, b1y)), fmaxf(a1z, b1z)) * ymx::kSlabWiden;
			const bool h0 = n0 <= fminf(x0, thi), h1 = n1 <= fminf(x1, thi);
			const int32_t c0 = __float_as_int(f3.x), c1 = __float_as_int(f3.y);
			if(h0 && h1)
			{
				const bool swap = n1 < n0;
				if(sp < kLightStack) stack[sp++] = swap ? c0 : c1;
				ref = swap ? c1 : c0;
			}
			else ref = h0 ? c0 : (h1 ? c1 : YMX_REF_EMPTY);
			if(ref != YMX_REF_EMPTY) continue;
		}
		else if(ref != YMX_REF_EMPTY)
		{
			const uint32_t first = ymx_leaf_first(ref), count = ymx_leaf_count(ref);
			for(uint32_t k = 0; k < count; ++k)
			{
				const float4 *tp = tris + 3 * (size_t)(first + k);
				const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
				if(!(__float_as_uint(t1.w) & YMX_VIS_VISIBLE)) continue;
				float u, v;
				const float th = ymx::tri_test(t0, t1, t2, o, d, u, v);
				if(th <= 0.f || th < t_min) continue;
				const int32_t id = (int32_t)__float_as_uint(t0.w);
				if(th < best_t || (th == best_t && best >= 0 && id < best))
				{
					best_t = th;
					best = id;
				}
			}
		}
		if(sp == 0) break;
		ref = stack[--sp];
	}
	return best;
}

// ObjectLight::intersect and BackgroundPortalLight::intersect
// (src/light/light_object_light.cc:211-229, src/light/light_background_portal.cc:207-222):
// the closest hit among the light's visible triangles past max(tmin, dynamic
// bias of the light's tree bound) (include/accelerator/accelerator.h:64,122-131,
// include/accelerator/accelerator_kdtree_common.h:139), from the light's own
// BVH or, for a small mesh, by testing every triangle.  A portal is one-sided.
// The reference leaves its `t` argument, the ray's tmax_ = -1, unchanged, so
// the 1 / (t * t) factor of the inverse pdf is 1.
__device__ bool obj_intersect(const PtArgs &a, const PtLight &l, V3 from, V3 dir, float tmin, float &ipdf)
{
	const float f[3] = {from.x, from.y, from.z}, d[3] = {dir.x, dir.y, dir.z};
	float enter, leave;
	if(!ymx::tree_cross(l.bound_a, l.bound_g, f, d, FLT_MAX, enter, leave)) return false;
	const float bias = (0.1f * 0.00005f) * fabsf(leave - enter);
	const float t_min = (tmin < bias) ? bias : tmin;
	const PtLightTri *lt = a.ltris + l.first;
	int32_t hit = -1;
	if(l.tree_nodes) hit = light_tree_closest(l.tree_nodes, l.tree_tris, f, d, t_min);
	else
	{
		float best = FLT_MAX;
		for(uint32_t i = 0; i < l.count; ++i)
		{
			const PtLightTri t = lt[i];
			if(!(__float_as_uint(t.v1_vis.w) & YMX_VIS_VISIBLE)) continue;
			const float4 e1 = make_float4(t.v1_vis.x - t.v0_cdf.x, t.v1_vis.y - t.v0_cdf.y, t.v1_vis.z - t.v0_cdf.z, 0.f);
			const float4 e2 = make_float4(t.v2.x - t.v0_cdf.x, t.v2.y - t.v0_cdf.y, t.v2.z - t.v0_cdf.z, 0.f);
			float u, v;
			const float th = ymx::tri_test(t.v0_cdf, e1, e2, f, d, u, v);
			if(th <= 0.f || th < t_min || th >= best) continue;
			best = th;
			hit = (int32_t)i;
		}
	}
	if(hit < 0) return false;
	const float4 n = lt[hit].n;
	float cos_angle = dot(dir, v3(-n.x, -n.y, -n.z));
	if(cos_angle <= 0.f)
	{
		if(l.double_sided) cos_angle = fabsf(cos_angle);
		else return false;
	}
	const float t_arg = -1.f;
	const float idist_sqr = 1.f / (t_arg * t_arg);
	ipdf = idist_sqr * l.area * cos_angle * k1Pi;
	return true;
}

// MonteCarloIntegrator::doLightEstimation with its area / dirac paths
// (src/integrator/surface/integrator_montecarlo.cc:118-440), queuing one shadow
// ray per sample instead of tracing it.
__device__ void queue_light(const PtArgs &a, const Surf &sp, V3 wo, const PtLight &l, uint32_t loffs,
                            uint32_t division, uint32_t psd_sample, uint32_t psd_offset, uint32_t slot0)
{
	const PtMaterial &m = a.mats[sp.mat];
	const bool cast_shadows = l.cast_shadows && m.receive_shadows;
	const float plen = sqrtf(dot(sp.p, sp.p));
	const float pscale = (1.f < plen) ? plen : 1.f;
	const float shadow_tmin = 0.0005f * pscale;
	if(l.type == YMX_LIGHT_POINT)
	{
		// PointLight::illuminate (src/light/light_point.cc:70-81), diracLight (:118-190)
		V3 ldir = sub(l.from, sp.p);
		const float dist_sqr = ldir.x * ldir.x + ldir.y * ldir.y + ldir.z * ldir.z;
		const float dist = sqrtf(dist_sqr);
		if(dist == 0.f)
		{
			set_slot_empty(a, slot0);
			return;
		}
		const float idist_sqr = 1.f / dist_sqr;
		ldir = mul(1.f / dist, ldir);
		const Rgb lcol = cscale(l.color, idist_sqr);
		const float angle = fabsf(dot(sp.n, ldir));
		const Rgb surf_col = bsdf_eval(a, sp, wo, ldir, kAll);
		const Rgb c = cmul(cscale(cmul(surf_col, lcol), angle), rgb(1.f, 1.f, 1.f));
		set_slot(a, slot0, c, cast_shadows, sp.p, ldir, shadow_tmin, dist);
		return;
	}
	const uint32_t l_offs = loffs * 4567u;
	const int num_samples = light_samples(l, division, a.light_mult);
	const uint32_t offs = (uint32_t)num_samples * psd_sample + psd_offset + l_offs;
	// areaLightSampleLight (:192-316) with Light::illumSample
	Halton h2, h3;
	h2.start(2, offs - 1);
	h3.start(3, offs - 1);
	for(int i = 0; i < num_samples; ++i)
	{
		const float s1 = h2.next();
		const float s2 = h3.next();
		const uint32_t slot = slot_at(a, slot0, (uint32_t)i);
		V3 ldir;
		float dist, lpdf;
		if(!illum_sample(a, l, sp.p, s1, s2, ldir, dist, lpdf) || !(lpdf > 1e-6f))
		{
			set_slot_empty(a, slot);
			continue;
		}
		const Rgb surf_col = bsdf_eval(a, sp, wo, ldir, kAll);
		const float angle = fabsf(dot(sp.n, ldir));
		float w = 1.f;
		const float m_pdf = bsdf_pdf(a, sp, wo, ldir, kGlossy | kDiffuse | kDispersive | kReflect | kTransmit);
		if(m_pdf > 1e-6f)
		{
			const float l_2 = lpdf * lpdf;
			const float m_2 = m_pdf * m_pdf;
			w = l_2 / (l_2 + m_2);
		}
		const Rgb c = cdiv(cscale(cscale(cmul(surf_col, l.color), angle), w), lpdf);
		set_slot(a, slot, c, cast_shadows, sp.p, ldir, shadow_tmin, dist);
	}
	// areaLightSampleMaterial (:318-415) with Light::intersect.  The shadow ray
	// ends at the ray's tmax_, which only AreaLight::intersect sets (to the hit
	// distance); it stays -1 (no end) for an object light.
	const float btmin = 0.00005f * pscale;
	h2.start(2, offs - 1);
	h3.start(3, offs - 1);
	for(int i = 0; i < num_samples; ++i)
	{
		const uint32_t slot = slot_at(a, slot0, (uint32_t)num_samples + (uint32_t)i);
		BSample s;
		s.s1 = h2.next();
		s.s2 = h3.next();
		s.flags = kGlossy | kDiffuse | kDispersive | kReflect | kTransmit;
		s.pdf = 0.f;
		s.sampled = kNone;
		float W = 0.f;
		V3 dir = v3(0.f, 0.f, 0.f);
		const Rgb surf_col = bsdf_sample(a, sp, wo, dir, s, W);
		float t = -1.f, ipdf = 0.f;
		bool hit = false;
		if(l.type == YMX_LIGHT_OBJECT || l.type == YMX_LIGHT_PORTAL) hit = obj_intersect(a, l, sp.p, dir, btmin, ipdf);
		else
		{
			// AreaLight::intersect (src/light/light_area.cc:146-155)
			const float cos_l = dot(dir, l.normal_flipped);
			if(cos_l > 0.f)
			{
				t = quad_intersect(l.quad, sp.p, dir);
				if(t > 0.f)
				{
					hit = true;
					ipdf = 1.f / (t * t) * l.area * cos_l * k1Pi;
				}
			}
		}
		if(hit && s.pdf > 1e-6f && ipdf > 1e-6f)
		{
			const float l_pdf = 1.f / ipdf;
			const float l_2 = l_pdf * l_pdf;
			const float m_2 = s.pdf * s.pdf;
			const float w = m_2 / (l_2 + m_2);
			const Rgb c = cscale(cscale(cmul(surf_col, l.hit_color), w), W);
			set_slot(a, slot, c, cast_shadows, sp.p, dir, btmin, t);
		}
		else set_slot_empty(a, slot);
	}
}

// GlassMaterial::getTransparency (src/material/material_glass.cc:297-304) at a
// triangle a shadow ray along dir passes: kt of the Fresnel terms about the
// face normal turned towards dir, times the filter colour.
__device__ Rgb transparency(const PtArgs &a, uint32_t prim, V3 dir)
{
	const float4 pr = a.prim[prim];
	const V3 ng = v3(pr.x, pr.y, pr.z);
	const PtMaterial &m = a.mats[__float_as_uint(pr.w)];
	float kr, kt;
	fresnel(dir, face_forward(ng, ng, dir), m.ior, kr, kt);
	return cscale(m.filter_color, kt);
}

// A slot's contribution when its shadow ray is unoccluded.  With transparent
// shadows it is scaled by the product of the transparencies passed, front to
// back (the reference scales the light colour by that product before the
// other factors, and multiplies in its kd-tree's order: the result agrees to
// float rounding).
__device__ __forceinline__ bool slot_lit(const PtArgs &a, uint32_t slot, Rgb &c)
{
	const float4 v = a.st.slot_contrib[slot];
	if(v.w == 0.f) return false;
	const int32_t q = a.st.slot_queue[slot];
	if(q >= 0 && ((a.st.sh_occ[(uint32_t)q >> 5] >> ((uint32_t)q & 31u)) & 1u)) return false;
	c = rgb(v.x, v.y, v.z);
	if(q >= 0 && a.c->transparent_shadows)
	{
		const uint32_t n = a.st.sh_tcount[q];
		if(n > 0)
		{
			const float4 d = reinterpret_cast<const float4 *>(a.st.sh_rays + q)[1];
			const V3 dir = v3(d.x, d.y, d.z);
			Rgb s = rgb(1.f, 1.f, 1.f);
			for(uint32_t k = 0; k < n; ++k)
				s = cmul(s, transparency(a, a.st.sh_thits[(size_t)q * a.c->shadow_depth + k].x, dir));
			c = cmul(c, s);
		}
	}
	return true;
}

// Sum of one light's slots in the order doLightEstimation accumulates them.
__device__ Rgb resolve_light(const PtArgs &a, const PtLight &l, uint32_t division, uint32_t slot0)
{
	Rgb c;
	if(l.type == YMX_LIGHT_POINT)
	{
		Rgb col = rgb(0.f, 0.f, 0.f);
		if(slot_lit(a, slot0, c)) col = cadd(col, c);
		return col;
	}
	const int num_samples = light_samples(l, division, a.light_mult);
	const float inv = 1.f / (float)num_samples;
	Rgb col = rgb(0.f, 0.f, 0.f);
	Rgb acc = rgb(0.f, 0.f, 0.f);
	for(int i = 0; i < num_samples; ++i)
		if(slot_lit(a, slot_at(a, slot0, (uint32_t)i), c)) acc = cadd(acc, c);
	col = cadd(col, cscale(acc, inv));
	acc = rgb(0.f, 0.f, 0.f);
	for(int i = 0; i < num_samples; ++i)
		if(slot_lit(a, slot_at(a, slot0, (uint32_t)num_samples + (uint32_t)i), c)) acc = cadd(acc, c);
	col = cadd(col, cscale(acc, inv));
	return col;
}

__device__ __forceinline__ void queue_extension(const PtArgs &a, int qsel, uint32_t slot, V3 from, V3 dir, float tmin)
{
	const uint32_t q = atomicAdd(a.st.ext_count + qsel, 1u);
	float4 *r = reinterpret_cast<float4 *>(a.st.ext_rays[qsel] + q);
	r[0] = make_float4(from.x, from.y, from.z, tmin);
	r[1] = make_float4(dir.x, dir.y, dir.z, -1.f);
	a.st.ext_path[qsel][q] = slot;
}

__device__ __forceinline__ void load_ray(const PtArgs &a, uint32_t n, V3 &from, V3 &dir)
{
	const float4 *r = reinterpret_cast<const float4 *>(a.nd.ray + n);
	const float4 o = r[0], d = r[1];
	from = v3(o.x, o.y, o.z);
	dir = v3(d.x, d.y, d.z);
}

// max(1, path_samples / division) (integrator_path_tracer.cc:181)
__device__ __forceinline__ int n_subpaths(const PtConst &c, uint32_t division)
{
	const int n = c.path_samples / (int)division;
	return n > 1 ? n : 1;
}

// TiledIntegrator::renderTile sample set-up and PerspectiveCamera::shootRay
// (src/integrator/surface/integrator_tiled.cc:290-335, src/camera/camera_perspective.cc:165-180):
// the level-0 nodes.
__global__ __launch_bounds__(kBlock) void k_camera(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t n = a.n_paths;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= n) return;
	const uint32_t p = path_pixel(a, i);
	const uint32_t sample = (uint32_t)a.first_sample + i / a.n_active;
	const uint32_t x = p % (uint32_t)c.resx, y = p / (uint32_t)c.resx;
	const uint32_t offset = fnv32a(y * fnv32a(x));
	float dx = 0.5f, dy = 0.5f;
	if(c.aa_passes > 1)
	{
		// scrambled (0, 2) sequence: the total sample count is not known
		dx = ri_vdc(sample, offset);
		dy = ri_s(sample, offset);
	}
	else if(a.pass_samples > 1)
	{
		const float d_1 = 1.f / (float)a.pass_samples;
		dx = (0.5f + (float)sample) * d_1;
		dy = ri_lp(sample + offset, 0);
	}
	const float px = (float)x + dx, py = (float)y + dy;
	const V3 dir = normalize(add(add(mul(px, c.vright), mul(py, c.vup)), c.vto));
	a.st.dxy[i] = make_float2(dx, dy);
	float4 *r = reinterpret_cast<float4 *>(a.nd.ray + i);
	r[0] = make_float4(c.cam_from.x, c.cam_from.y, c.cam_from.z, 0.f);
	r[1] = make_float4(dir.x, dir.y, dir.z, -1.f);
	a.nd.info[i] = make_uint4(i, 1u << 8, 0u, 0u);
	a.nd.dec[i] = make_float2(0.f, 0.f);
}

// One child of a recursive raytrace.
__device__ void write_child(const PtArgs &a, uint32_t k, V3 from, V3 dir, Rgb wt, uint32_t path, uint32_t level,
                            uint32_t division, int32_t rd_offset, float2 dec, uint32_t tag)
{
	float4 *r = reinterpret_cast<float4 *>(a.nd.ray + k);
	r[0] = make_float4(from.x, from.y, from.z, 0.00005f);  // ray_min_dist_
	r[1] = make_float4(dir.x, dir.y, dir.z, -1.f);
	a.nd.info[k] = make_uint4(path, level | (division << 8), tag, (uint32_t)rd_offset);
	a.nd.dec[k] = dec;
	a.nd.wt[k] = make_float4(wt.r, wt.g, wt.b, 0.f);
}

// MonteCarloIntegrator::recursiveRaytrace (integrator_montecarlo.cc:698-748) for a
// node at x.level: the glossy branches (:506-589 with glossyReflectNoTransmit
// :623-645) or the perfect reflection / refraction (specularReflect/Refract,
// :647-696, Material::getSpecular).  A child whose weight is exactly zero adds
// nothing to its parent and is not traced.
__device__ void spawn_children(const PtArgs &a, const NodeCtx &x, const Surf &sp, V3 wo)
{
	const PtConst &c = *a.c;
	const int ray_level = (int)x.level + 1;
	uint2 ch = make_uint2(0u, 0u);
	if(!(ray_level <= c.raydepth && ray_level < 20 && spawns_children(sp.flags)))
	{
		a.nd.child[x.n] = ch;
		return;
	}
	const PtMaterial &m = a.mats[sp.mat];
	if(sp.flags & kGlossy)
	{
		if(!(sp.flags & kReflect))
		{
			a.nd.child[x.n] = ch;
			return;
		}
		const uint32_t rsg = x.division > 1 ? (kGlossyBranches / x.division > 1 ? kGlossyBranches / x.division : 1u)
		                                   : kGlossyBranches;
		const uint32_t ndiv = x.division * rsg;
		int32_t branch = (int32_t)ndiv * x.rd_offset;
		const uint32_t offs = rsg * x.sample + x.offset;
		Halton h2, h3;
		h2.start(2, offs);
		h3.start(3, offs);
		V3 dirs[kGlossyBranches];
		Rgb wts[kGlossyBranches];
		float2 decs[kGlossyBranches];
		int32_t offsets[kGlossyBranches];
		uint32_t tags[kGlossyBranches];
		uint32_t cnt = 0;
		for(uint32_t ns = 0; ns < rsg; ++ns)
		{
			const float dec1 = (float)low_discrepancy(a, 2 * ray_level + 1, (uint32_t)branch + x.offset);
			const float dec2 = (float)low_discrepancy(a, 2 * ray_level + 2, (uint32_t)branch + x.offset);
			const int32_t boff = branch;
			++branch;
			BSample s;
			s.s1 = h2.next();
			s.s2 = h3.next();
			s.flags = kGlossy | kReflect;
			s.pdf = 0.f;
			s.sampled = kNone;
			float w = 0.f;
			V3 wi = v3(0.f, 0.f, 0.f);
			const Rgb mcol = bsdf_sample(a, sp, wo, wi, s, w);
			const Rgb wt = cscale(mcol, w);
			if(cblack(wt)) continue;
			dirs[cnt] = wi;
			wts[cnt] = wt;
			decs[cnt] = make_float2(dec1, dec2);
			offsets[cnt] = boff;
			tags[cnt] = x.tag * 11u + 1u + ns;
			++cnt;
		}
		if(cnt)
		{
			const uint32_t first = a.next_start + atomicAdd(a.nd.spawn, cnt);
			for(uint32_t k = 0; k < cnt; ++k)
				write_child(a, first + k, sp.p, dirs[k], wts[k], x.path, (uint32_t)ray_level, ndiv, offsets[k], decs[k],
				            tags[k]);
			ch = make_uint2(first, cnt | (kChildGlossy << 8) | (rsg << 16));
		}
		a.nd.child[x.n] = ch;
		return;
	}
	// perfect specular: mirror (src/material/material_mirror.cc:78-86), glass (material_glass.cc:303-366)
	bool has_r = false, has_t = false;
	V3 dir_r = v3(0.f, 0.f, 0.f), dir_t = v3(0.f, 0.f, 0.f);
	Rgb col_r = rgb(0.f, 0.f, 0.f), col_t = rgb(0.f, 0.f, 0.f);
	if(m.type == YMX_MAT_MIRROR)
	{
		has_r = true;
		col_r = m.ref_col;
		dir_r = reflect_dir(face_forward(sp.ng, sp.n, wo), wo);
	}
	else if(m.type == YMX_MAT_GLASS)
	{
		bool outside;
		const V3 n = glass_normal(sp, wo, outside);
		V3 refdir;
		if(refract(n, wo, refdir, m.ior))
		{
			float kr, kt;
			fresnel(wo, n, m.ior, kr, kt);
			has_t = true;
			col_t = cscale(m.filter_color, kt);
			dir_t = refdir;
			if(outside || ray_level < 3)
			{
				has_r = true;
				dir_r = reflect_about(wo, n);
				col_r = cscale(m.mirror_color, kr);
			}
		}
		else
		{
			has_r = true;
			col_r = m.mirror_color;
			dir_r = reflect_about(wo, n);
		}
	}
	if(has_r && cblack(col_r)) has_r = false;
	if(has_t && cblack(col_t)) has_t = false;
	const uint32_t cnt = (has_r ? 1u : 0u) + (has_t ? 1u : 0u);
	if(cnt)
	{
		// the ray division passes down unchanged (specularReflect / specularRefract)
		const float2 dec = a.nd.dec[x.n];
		const uint32_t first = a.next_start + atomicAdd(a.nd.spawn, cnt);
		uint32_t k = first;
		if(has_r)
			write_child(a, k++, sp.p, dir_r, col_r, x.path, (uint32_t)ray_level, x.division, x.rd_offset, dec,
			            x.tag * 11u + 9u);
		if(has_t)
			write_child(a, k, sp.p, dir_t, col_t, x.path, (uint32_t)ray_level, x.division, x.rd_offset, dec,
			            x.tag * 11u + 10u);
		ch = make_uint2(first, cnt | (kChildSpecular << 8));
	}
	a.nd.child[x.n] = ch;
}

// Node vertex: emission, estimateAllDirectLight and the children of the
// recursive raytrace (integrator_path_tracer.cc:145-170, 289).
__global__ __launch_bounds__(kBlock) void k_shade_node(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.cnt) return;
	const NodeCtx x = node_ctx(a, i);
	const ymx_hit h = a.st.ext_hits[i];
	a.st.cam_hit[i] = h;
	if(h.prim == YMX_NO_HIT)
	{
		// TiledIntegrator::background (integrator_tiled.cc:693-706)
		a.nd.col[x.n] = c.has_background ? make_float4(c.background.r, c.background.g, c.background.b, 1.f)
		                                 : make_float4(0.f, 0.f, 0.f, 1.f);
		a.nd.child[x.n] = make_uint2(0u, 0u);
		a.st.flags[i] = 0u;
		return;
	}
	V3 from, dir;
	load_ray(a, x.n, from, dir);
	Surf sp;
	make_surf(a, from, dir, h, sp);
	const V3 wo = neg(dir);
	Rgb col = rgb(0.f, 0.f, 0.f);
	if(sp.flags & kEmit) col = cadd(col, mat_emit(a, sp, wo));
	a.nd.col[x.n] = make_float4(col.r, col.g, col.b, 1.f);
	uint32_t f = 0u;
	if(sp.flags & kDiffuse)
	{
		uint32_t slot = i;
		for(uint32_t l = 0; l < c.n_lights; ++l)
		{
			queue_light(a, sp, wo, a.lights[l], l, x.division, x.sample, x.offset, slot);
			slot = slot_at(a, slot, light_slots(a.lights[l], x.division, a.light_mult));
		}
		f |= kDirect;
	}
	if(sp.flags & kDiffuse) f |= kPathBlock;  // path_flags = Diffuse (no_recursive = false)
	a.st.flags[i] = f;
	if(a.recursive) spawn_children(a, x, sp, wo);
	else a.nd.child[x.n] = make_uint2(0u, 0u);
}

__global__ __launch_bounds__(kBlock) void k_resolve_node(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i < kQueueSegments)
	{
		// consumed by the launches before this one; the next enqueues come after it
		uint32_t *cnt = a.st.sh_count + (size_t)i * kQueueStride;
		atomicAdd(a.st.totals + 2, (unsigned long long)*cnt);
		*cnt = 0u;
	}
	if(i == 0)
	{
		a.st.ext_count[0] = 0u;
		a.st.ext_count[1] = 0u;
	}
	if(i >= a.cnt) return;
	a.st.path_col[i] = make_float4(0.f, 0.f, 0.f, 0.f);
	if(!(a.st.flags[i] & kDirect)) return;
	const uint32_t n = a.base + i;
	const uint32_t division = a.nd.info[n].y >> 8;
	Rgb col_all = rgb(0.f, 0.f, 0.f);
	uint32_t slot = i;
	for(uint32_t l = 0; l < c.n_lights; ++l)
	{
		col_all = cadd(col_all, resolve_light(a, a.lights[l], division, slot));
		slot = slot_at(a, slot, light_slots(a.lights[l], division, a.light_mult));
	}
	const float4 c4 = a.nd.col[n];
	const Rgb col = cadd(rgb(c4.x, c4.y, c4.z), col_all);
	a.nd.col[n] = make_float4(col.r, col.g, col.b, c4.w);
	a.st.flags[i] &= ~kDirect;
}

// First bounce of sub-path a.sub (integrator_path_tracer.cc:183-203)
__global__ __launch_bounds__(kBlock) void k_subpath_init(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.cnt) return;
	uint32_t f = a.st.flags[i];
	f &= ~(kActive | kAddEmit | kCaustic | kDirect | kVertex | kSampled);
	a.st.flags[i] = f;
	if(a.g.train) a.g.mask[i] = 0u;
	if(!(f & kPathBlock)) return;
	const NodeCtx x = node_ctx(a, i);
	if(a.sub >= n_subpaths(c, x.division)) return;
	const ymx_hit h = a.st.cam_hit[i];
	V3 from, dir;
	load_ray(a, x.n, from, dir);
	Surf sp;
	make_surf(a, from, dir, h, sp);
	const V3 wo = neg(dir);
	const uint32_t offs = (uint32_t)c.path_samples * x.sample + x.offset + (uint32_t)a.sub;
	BSample s;
	s.s1 = ri_vdc(offs, 0);
	s.s2 = (float)low_discrepancy(a, 2, offs);
	if(x.division > 1)
	{
		// math::addMod1 with the branch's decorrelation (include/math/math.h:304-309)
		const float2 dec = a.nd.dec[x.n];
		float t = s.s1 + dec.x;
		s.s1 = t > 1 ? t - 1 : t;
		t = s.s2 + dec.y;
		s.s2 = t > 1 ? t - 1 : t;
	}
	s.flags = kDiffuse | kDiffuse | kReflect | kTransmit;
	s.pdf = 0.f;
	s.sampled = kNone;
	float w = 0.f;
	V3 pdir = v3(0.f, 0.f, 0.f);
	Rgb scol;
	if(guided_material(a, sp))
	{
		const uint32_t ps = x.sample * (uint32_t)c.path_samples + (uint32_t)a.sub;
		GuidePdfs pdfs;
		guided_sample(a, sp, wo, a.g.mix_node + (size_t)i * ymx::guide::kMix, s, guide_rand(x, ps, 0, c.seed, 0),
		              guide_rand(x, ps, 0, c.seed, 1), guide_rand(x, ps, 0, c.seed, 2), guide_rand(x, ps, 0, c.seed, 3),
		              pdir, scol, w, pdfs);
		if(a.g.train)
		{
			// the path colour before this sub-path, throughput 1
			record_segment(a, i, 0, sp, wo, pdir, 1.f, pdfs);
			const float4 pc = a.st.path_col[i];
			a.g.base[i].x = pc.x;
			a.g.base[i].y = pc.y;
			a.g.base[i].z = pc.z;
		}
	}
	else
	{
		scol = bsdf_sample(a, sp, wo, pdir, s, w);
		scol = cscale(scol, w);
	}
	// w rides along in tp_next.w: a later sample call that leaves w unset reuses it
	a.st.tp_next[i] = make_float4(scol.r, scol.g, scol.b, w);
	a.st.pwo[i] = make_float4(wo.x, wo.y, wo.z, 0.f);
	a.st.flags[i] = f | kActive | (s.sampled != kNone ? kSampled : 0u);
	queue_extension(a, 0, i, sp.p, pdir, 0.00005f);
}

// One path vertex (integrator_path_tracer.cc:203-285): roulette, one-light
// estimation, emission, and the BSDF sample of the next segment.
__global__ __launch_bounds__(kBlock) void k_shade_path(PtArgs a)
{
	const PtConst &c = *a.c;
	const int d = a.depth;
	const int qsel = d & 1;
	const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
	if(q >= a.st.ext_count[qsel]) return;
	const uint32_t i = a.st.ext_path[qsel][q];
	uint32_t f = a.st.flags[i] & ~(kActive | kAddEmit | kDirect | kVertex);
	const ymx_hit h = a.st.ext_hits[q];
	if(h.prim == YMX_NO_HIT)
	{
		a.st.flags[i] = f;
		return;
	}
	const float4 *rr = reinterpret_cast<const float4 *>(a.st.ext_rays[qsel] + q);
	const float4 ro = rr[0], rd = rr[1];
	const V3 from = v3(ro.x, ro.y, ro.z), dir = v3(rd.x, rd.y, rd.z);
	Surf sp;
	make_surf(a, from, dir, h, sp);
	V3 pwo;
	const float4 tpn = a.st.tp_next[i];
	Rgb tp = rgb(tpn.x, tpn.y, tpn.z);
	uint32_t first_flags;
	const NodeCtx x = node_ctx(a, i);
	if(d == 0)
	{
		const float4 p4 = a.st.pwo[i];
		pwo = (f & kSampled) ? neg(dir) : v3(p4.x, p4.y, p4.z);
		first_flags = sp.flags;
		f = (f & ~(0x1FFu << 8)) | (first_flags << 8);
	}
	else
	{
		pwo = neg(dir);
		first_flags = (f >> 8) & 0x1FFu;
		if(d > c.rr_min_bounces)
		{
			const float random_value = rr_random(x.pixel, x.sample * (uint32_t)c.path_samples + (uint32_t)a.sub,
			                                     (uint32_t)d + 256u * x.tag, c.seed);
			// The reference divides by the throughput's maximum even above 1, which
			// scales such paths down; BSDF-sampled Lambertian paths never get there,
			// guided ones do (weights up to albedo / alpha), so with guiding on a
			// probability of 1 keeps the path unscaled.
			float probability = cmaximum(tp);
			if(a.g.on && probability > 1.f) probability = 1.f;
			if(probability <= 0.f || probability < random_value)
			{
				a.st.flags[i] = f;
				return;
			}
			if(!(a.g.on && probability == 1.f)) tp = cscale(tp, 1.f / probability);
		}
	}
	// A vertex with zero throughput adds lcol * 0 = 0 to the path colour: its
	// light estimation and emission are skipped.
	if(!cblack(tp))
	{
		a.st.tp_vertex[i] = make_float4(tp.r, tp.g, tp.b, 0.f);
		f |= kVertex;
		const bool need_direct = (d == 0) || (first_flags & kDiffuse);
		if(need_direct && c.n_lights > 0)
		{
			// estimateOneDirectLight (integrator_montecarlo.cc:108-116) with the
			// deterministic Halton base-2 light-pick index
			const uint32_t corr = (x.pixel * c.aa_total + x.sample) *
			                          ((uint32_t)c.path_samples * (uint32_t)c.bounces) +
			                      (uint32_t)a.sub * (uint32_t)c.bounces + (uint32_t)d + 1u + x.tag * 0x9E3779B9u;
			const float hv = halton_next(2, corr - 1u);
			int lnum = (int)(hv * (float)c.n_lights);
			if(lnum > (int)c.n_lights - 1) lnum = (int)c.n_lights - 1;
			queue_light(a, sp, pwo, a.lights[lnum], (uint32_t)lnum, x.division, x.sample, x.offset, i);
			a.st.light[i] = lnum;
			f |= kDirect;
		}
		const bool add_emit = (d == 0) ? (first_flags & kEmit) != 0 : ((first_flags & kEmit) && (f & kCaustic));
		if(add_emit)
		{
			const Rgb e = mat_emit(a, sp, pwo);
			a.st.vert_emit[i] = make_float4(e.r, e.g, e.b, 0.f);
			f |= kAddEmit;
		}
	}
	// next segment (loop iteration depth = d + 1)
	if(d + 1 < c.bounces)
	{
		const uint32_t offs = (uint32_t)c.path_samples * x.sample + x.offset + (uint32_t)a.sub;
		const int d_4 = 4 * (d + 1);
		BSample s;
		s.s1 = (float)low_discrepancy(a, d_4 + 3, offs);
		s.s2 = (float)low_discrepancy(a, d_4 + 4, offs);
		s.flags = kAll;
		s.pdf = 0.f;
		s.sampled = kNone;
		// The reference keeps w and the ray direction in variables of the whole
		// integrate() call: a sample that leaves them unset (shinydiffuse without a
		// diffuse component returns white) continues straight on with the last w.
		float w = tpn.w;
		V3 ndir = dir;
		Rgb scol;
		const bool guided = d < a.g.max_depth && guided_material(a, sp);
		GuidePdfs pdfs = {0.f, 0.f, 0.f};
		if(guided)
		{
			const uint32_t ps = x.sample * (uint32_t)c.path_samples + (uint32_t)a.sub;
			const uint32_t seg = (uint32_t)d + 1u;
			guided_sample(a, sp, pwo, a.g.mix_path + (size_t)q * ymx::guide::kMix, s, guide_rand(x, ps, seg, c.seed, 0),
			              guide_rand(x, ps, seg, c.seed, 1), guide_rand(x, ps, seg, c.seed, 2),
			              guide_rand(x, ps, seg, c.seed, 3), ndir, scol, w, pdfs);
		}
		else
		{
			scol = bsdf_sample(a, sp, pwo, ndir, s, w);
			scol = cscale(scol, w);
		}
		if(!cblack(scol))
		{
			if(guided && a.g.train && !cblack(tp))
				record_segment(a, i, (uint32_t)d + 1u, sp, pwo, ndir, luminance(tp), pdfs);
			const Rgb tpn2 = cmul(tp, scol);
			a.st.tp_next[i] = make_float4(tpn2.r, tpn2.g, tpn2.b, w);
			if(s.sampled & (kSpecular | kGlossy | kFilter)) f |= kCaustic;
			else f &= ~kCaustic;
			f |= kActive;
			queue_extension(a, qsel ^ 1, i, sp.p, ndir, 0.00005f);
		}
	}
	a.st.flags[i] = f;
}

__global__ __launch_bounds__(kBlock) void k_resolve_path(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i < kQueueSegments)
	{
		uint32_t *cnt = a.st.sh_count + (size_t)i * kQueueStride;
		atomicAdd(a.st.totals + 2, (unsigned long long)*cnt);
		*cnt = 0u;
	}
	if(i == 0)
	{
		a.st.totals[1] += a.st.ext_count[a.depth & 1];
		a.st.ext_count[a.depth & 1] = 0u;  // this segment's queue, consumed by shade_path
	}
	if(i >= a.cnt) return;
	const uint32_t f = a.st.flags[i];
	if(!(f & kVertex)) return;
	Rgb lcol = rgb(0.f, 0.f, 0.f);
	if(f & kDirect)
	{
		const int lnum = a.st.light[i];
		const uint32_t division = a.nd.info[a.base + i].y >> 8;
		lcol = cscale(resolve_light(a, a.lights[lnum], division, i), (float)c.n_lights);
	}
	if(f & kAddEmit)
	{
		const float4 e = a.st.vert_emit[i];
		lcol = cadd(lcol, rgb(e.x, e.y, e.z));
	}
	const float4 t4 = a.st.tp_vertex[i];
	const float4 p4 = a.st.path_col[i];
	const Rgb pc = cadd(rgb(p4.x, p4.y, p4.z), cmul(lcol, rgb(t4.x, t4.y, t4.z)));
	a.st.path_col[i] = make_float4(pc.r, pc.g, pc.b, 0.f);
	a.st.flags[i] = f & ~(kDirect | kVertex | kAddEmit);
	if(a.g.train && ((a.g.mask[i] >> (a.depth + 1)) & 1u))
	{
		// the segment sampled at this vertex starts after the vertex's own light
		float4 *b = a.g.base + (size_t)(a.depth + 1) * a.cnt + i;
		b->x = pc.r;
		b->y = pc.g;
		b->z = pc.b;
	}
}

// After a sub-path: each recorded segment's training weight is the radiance
// that reached the path through it, relative to the throughput before it
// (an estimate of f(w) / p(w) for the guide's loss).  A fraction a.g.accept
// of them is kept.
__global__ __launch_bounds__(kBlock) void k_guide_collect(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.cnt) return;
	const uint32_t mask = a.g.mask[i];
	if(!mask) return;
	a.g.mask[i] = 0u;
	const NodeCtx x = node_ctx(a, i);
	const float4 p4 = a.st.path_col[i];
	for(int r = 0; r < c.bounces; ++r)
	{
		if(!((mask >> r) & 1u)) continue;
		const size_t k = (size_t)r * a.cnt + i;
		const float4 b = a.g.base[k];
		const float cw = luminance(rgb(p4.x - b.x, p4.y - b.y, p4.z - b.z)) / b.w;
		if(!(cw > 0.f) || !(cw < 1e30f)) continue;
		const uint32_t ps = x.sample * (uint32_t)c.path_samples + (uint32_t)a.sub;
		if(guide_rand(x, ps, (uint32_t)r, c.seed, 7) >= a.g.accept) continue;
		// ring buffer of the most recent samples (train_cap is a power of two)
		const uint32_t slot = atomicAdd(a.g.train_count, 1u) & (a.g.train_cap - 1u);
		ymx_guide_sample rec = a.g.rec[k];
		rec.c = cw;
		a.g.train_buf[slot] = rec;
		float2 *cs = a.g.cell_sum + guide_cell(a.g, rec.pos);
		atomicAdd(&cs->x, cw);
		atomicAdd(&cs->y, 1.f);
	}
}

// One training mini-batch: dst[k] = ring[h(k) mod n] for a hashed index h,
// its weight divided by the mean weight of its region: the loss -c log q then
// fits the shape of every region's directional distribution with the same
// emphasis, instead of spending the network on the brightest regions.
__global__ __launch_bounds__(kBlock) void k_guide_gather(ymx_guide_sample *dst, const ymx_guide_sample *ring, uint32_t n,
                                                         uint32_t batch, uint32_t seed, PtGuide g)
{
	const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
	if(k >= batch) return;
	uint32_t h = (k + 0x9E3779B9u * seed) * 0x85EBCA6Bu;
	h ^= h >> 13;
	h *= 0xC2B2AE35u;
	h ^= h >> 16;
	h ^= seed;
	h *= 0x27D4EB2Fu;
	h ^= h >> 15;
	ymx_guide_sample s = ring[(uint32_t)(((uint64_t)h * n) >> 32)];
	const float2 cs = g.cell_sum[guide_cell(g, s.pos)];
	s.c = cs.x > 0.f ? s.c * (cs.y / cs.x) : s.c;
	dst[k] = s;
}

// Learned selection probability (ymx_pt_set_guide_selection).  A record drawn
// with densities p (BSDF), q (mixture) and alpha_old has target weight c =
// F / p_old, F the integrand and p_old = alpha_old p + (1 - alpha_old) q.
// The second moment of the one-sample estimator F / p_mix(alpha) over its own
// density is M = integral of F^2 / p_mix; its derivative in alpha, estimated
// from the records' draws, is -F^2 (p - q) / (p_mix^2 p_old) per record.  The
// region's logit theta takes the sum times d alpha / d theta; Adam (per
// region) then moves alpha towards the strategy of smaller variance.  The
// records' weights carry their region's normalisation, one factor per region,
// which Adam's scaling removes.
__global__ __launch_bounds__(kBlock) void k_alpha_grad(const ymx_guide_sample *batch, uint32_t n, PtGuide g)
{
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= n) return;
	const ymx_guide_sample r = batch[i];
	const float p = r.pad0, q = r.pad1, a_old = r.pad2;
	const float p_old = a_old * p + (1.f - a_old) * q;
	if(!(r.c > 0.f) || !(p_old > 0.f) || !(a_old > 0.f)) return;
	const uint32_t cell = guide_cell(g, r.pos);
	const float sg = 1.f / (1.f + __expf(-g.alpha_theta[cell]));
	const float alpha = kAlphaMin + (1.f - kAlphaMin) * sg;
	const float pm = alpha * p + (1.f - alpha) * q;
	const float f = r.c * p_old;
	const float d_alpha = -f * f * (p - q) / (pm * pm * p_old);
	const float d_theta = d_alpha * (1.f - kAlphaMin) * sg * (1.f - sg);
	if(isfinite(d_theta)) atomicAdd(g.alpha_grad + cell, d_theta);
}

// Per-region fallback (ymx_pt_set_guide_fallback).  A record drawn with
// densities p, q, alpha_old (p_old = alpha_old p + (1 - alpha_old) q) and
// target weight c has integrand F = c p_old.  The second moment of a
// one-sample estimator with density r, M_r = integral of F^2 / r, is
// estimated from the records' draws by the mean of F^2 / (r p_old): here for
// r = the region's mixture at its alpha (before the fallback) and r = p.
// Both means share the region's weight normalisation, so their comparison
// does not depend on it.
__global__ __launch_bounds__(kBlock) void k_fallback_sum(const ymx_guide_sample *batch, uint32_t n, PtGuide g)
{
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= n) return;
	const ymx_guide_sample r = batch[i];
	const float p = r.pad0, q = r.pad1, a_old = r.pad2;
	const float p_old = a_old * p + (1.f - a_old) * q;
	if(!(r.c > 0.f) || !(p_old > 0.f) || !(p > 0.f)) return;
	const uint32_t cell = guide_cell(g, r.pos);
	const float alpha = region_alpha(g, cell);
	const float pm = alpha * p + (1.f - alpha) * q;
	const float f = r.c * p_old;
	const float m_guided = f * f / (pm * p_old);
	const float m_bsdf = f * f / (p * p_old);
	if(!isfinite(m_guided) || !isfinite(m_bsdf)) return;
	constexpr uint32_t kCells = (uint32_t)(kGuideCells * kGuideCells * kGuideCells);
	atomicAdd(g.fb_sum + cell, m_guided);
	atomicAdd(g.fb_sum + kCells + cell, m_bsdf);
	atomicAdd(g.fb_sum + 2 * kCells + cell, 1.f);
}

// The running moments of each region with records in the mini-batch: the
// first batch's means, then kFallbackDecay * running + (1 - kFallbackDecay) *
// the batch's mean; and the count of records seen.  Clears the sums.
__global__ __launch_bounds__(kBlock) void k_fallback_update(PtGuide g)
{
	constexpr uint32_t kCells = (uint32_t)(kGuideCells * kGuideCells * kGuideCells);
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= kCells) return;
	const float cnt = g.fb_sum[2 * kCells + i];
	if(cnt == 0.f) return;
	const float mg = g.fb_sum[i] / cnt, mb = g.fb_sum[kCells + i] / cnt;
	g.fb_sum[i] = 0.f;
	g.fb_sum[kCells + i] = 0.f;
	g.fb_sum[2 * kCells + i] = 0.f;
	const float rg = g.fb_run[i], rb = g.fb_run[kCells + i];
	const bool first = rg == 0.f && rb == 0.f;
	g.fb_run[i] = first ? mg : kFallbackDecay * rg + (1.f - kFallbackDecay) * mg;
	g.fb_run[kCells + i] = first ? mb : kFallbackDecay * rb + (1.f - kFallbackDecay) * mb;
	g.fb_run[2 * kCells + i] += cnt;
}

__global__ __launch_bounds__(kBlock) void k_alpha_adam(PtGuide g, uint32_t cells, float lr, float bc1, float bc2)
{
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= cells) return;
	const float gr = g.alpha_grad[i];
	g.alpha_grad[i] = 0.f;
	if(gr == 0.f) return;  // no sample of this region in the batch
	const float m = 0.9f * g.alpha_m[i] + 0.1f * gr;
	const float v = 0.999f * g.alpha_v[i] + 0.001f * gr * gr;
	g.alpha_m[i] = m;
	g.alpha_v[i] = v;
	float t = g.alpha_theta[i] - lr * (m / bc1) / (sqrtf(v / bc2) + 1e-30f);
	// past +-8 the probability no longer moves (sigmoid within 3e-4 of its ends)
	g.alpha_theta[i] = t > 8.f ? 8.f : (t < -8.f ? -8.f : t);
}

// Queries of the guide at the node vertices that start sub-paths.
__global__ __launch_bounds__(kBlock) void k_guide_query_node(PtArgs a)
{
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.cnt) return;
	ymx_guide_sample qr;
	memset(&qr, 0, sizeof qr);
	if(a.st.flags[i] & kPathBlock)
	{
		const ymx_hit h = a.st.cam_hit[i];
		V3 from, dir;
		load_ray(a, a.base + i, from, dir);
		Surf sp;
		make_surf(a, from, dir, h, sp);
		if(guided_material(a, sp))
		{
			const V3 wo = neg(dir);
			const V3 n = face_forward(sp.ng, sp.n, wo);
			qr.pos[0] = sp.p.x;
			qr.pos[1] = sp.p.y;
			qr.pos[2] = sp.p.z;
			qr.normal[0] = n.x;
			qr.normal[1] = n.y;
			qr.normal[2] = n.z;
			qr.wo[0] = wo.x;
			qr.wo[1] = wo.y;
			qr.wo[2] = wo.z;
		}
	}
	a.g.query[i] = qr;
}

// Queries of the guide at the path vertices of segment a.depth (queue order).
__global__ __launch_bounds__(kBlock) void k_guide_query_path(PtArgs a)
{
	const int qsel = a.depth & 1;
	const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
	if(q >= a.st.ext_count[qsel]) return;
	ymx_guide_sample qr;
	memset(&qr, 0, sizeof qr);
	const ymx_hit h = a.st.ext_hits[q];
	if(h.prim != YMX_NO_HIT)
	{
		const float4 *rr = reinterpret_cast<const float4 *>(a.st.ext_rays[qsel] + q);
		const float4 ro = rr[0], rd = rr[1];
		const V3 from = v3(ro.x, ro.y, ro.z), dir = v3(rd.x, rd.y, rd.z);
		Surf sp;
		make_surf(a, from, dir, h, sp);
		if(guided_material(a, sp))
		{
			const V3 wo = neg(dir);
			const V3 n = face_forward(sp.ng, sp.n, wo);
			qr.pos[0] = sp.p.x;
			qr.pos[1] = sp.p.y;
			qr.pos[2] = sp.p.z;
			qr.normal[0] = n.x;
			qr.normal[1] = n.y;
			qr.normal[2] = n.z;
			qr.wo[0] = wo.x;
			qr.wo[1] = wo.y;
			qr.wo[2] = wo.z;
		}
	}
	a.g.query[q] = qr;
}

__global__ __launch_bounds__(kBlock) void k_finish(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.cnt) return;
	if(!(a.st.flags[i] & kPathBlock)) return;
	const uint32_t n = a.base + i;
	const int n_samples = n_subpaths(c, a.nd.info[n].y >> 8);
	const float4 p4 = a.st.path_col[i];
	const float4 c4 = a.nd.col[n];
	const Rgb col = cadd(rgb(c4.x, c4.y, c4.z), cdiv(rgb(p4.x, p4.y, p4.z), (float)n_samples));
	a.nd.col[n] = make_float4(col.r, col.g, col.b, c4.w);
}

// Adds the children's colours into the nodes of one level, in the order
// recursiveRaytrace sums them: the glossy branches as (sum of child * weight)
// / branches; the specular reflection, then the refraction.
__global__ __launch_bounds__(kBlock) void k_combine(PtArgs a)
{
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.cnt) return;
	const uint32_t n = a.base + i;
	const uint2 ch = a.nd.child[n];
	const uint32_t cnt = ch.y & 0xFFu;
	if(cnt == 0) return;
	Rgb sum = rgb(0.f, 0.f, 0.f);
	for(uint32_t k = 0; k < cnt; ++k)
	{
		const float4 cc = a.nd.col[ch.x + k];
		const float4 w = a.nd.wt[ch.x + k];
		sum = cadd(sum, cmul(rgb(cc.x, cc.y, cc.z), rgb(w.x, w.y, w.z)));
	}
	if(((ch.y >> 8) & 0xFFu) == kChildGlossy)
	{
		const uint32_t rsg = ch.y >> 16;
		sum = cadd(rgb(0.f, 0.f, 0.f), cscale(sum, 1.f / (float)rsg));
	}
	const float4 c4 = a.nd.col[n];
	const Rgb col = cadd(rgb(c4.x, c4.y, c4.z), sum);
	a.nd.col[n] = make_float4(col.r, col.g, col.b, c4.w);
}

// roundToIntWithBias, include/render/imagefilm.h:344-348
__device__ __forceinline__ int round_bias(double v) { return (int)(v + (.5 - 1.4e-11)); }

// ImageFilm::addSample (src/render/imagefilm.cc:771-822) as a gather: each
// film pixel adds the samples whose footprint covers it, visiting the source
// pixels in row-major order -- the order a row-major scatter adds them.
// Rgb::clampProportionalRgb (include/color/color.h:399-431): the largest RGB
// component clamped to max_value, the others scaled by max_value / it; alpha
// untouched.  max_value 0: no clamp.
__device__ __forceinline__ float4 clamp_proportional(float4 c, float max_value)
{
	if(max_value <= 0.f) return c;
	const float max_rgb = fmaxf(c.x, fmaxf(c.y, c.z));
	const float adj = max_value / max_rgb;
	if(max_rgb > max_value)
	{
		if(c.x >= max_rgb)
		{
			c.x = max_value;
			c.y *= adj;
			c.z *= adj;
		}
		else if(c.y >= max_rgb)
		{
			c.y = max_value;
			c.x *= adj;
			c.z *= adj;
		}
		else
		{
			c.z = max_value;
			c.x *= adj;
			c.y *= adj;
		}
	}
	return c;
}

__global__ __launch_bounds__(kBlock) void k_film(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
	if(t >= (uint32_t)(c.resx * c.resy)) return;
	const int tx = (int)(t % (uint32_t)c.resx), ty = (int)(t / (uint32_t)c.resx);
	const int radius = (int)c.filter_width + 1;
	float w = a.film_w[t];
	float4 acc = a.film[t];
	const double fw = (double)c.filter_width;
	// sample slices in AA order, then source pixels in row-major order
	for(int j = 0; j < a.count; ++j)
	for(int sy = ty - radius; sy <= ty + radius; ++sy)
	{
		if(sy < 0 || sy >= c.resy) continue;
		for(int sx = tx - radius; sx <= tx + radius; ++sx)
		{
			if(sx < 0 || sx >= c.resx) continue;
			const uint32_t src = (uint32_t)(sy * c.resx + sx);
			const uint32_t slot = a.pix_slot ? a.pix_slot[src] : src;
			if(slot == kNoSlot) continue;  // not rendered in this pass
			const uint32_t s = (uint32_t)j * a.n_active + slot;
			const float2 dd = a.st.dxy[s];
			int dx0 = round_bias((double)dd.x - fw);
			int dx1 = round_bias((double)dd.x + fw - 1.0);
			int dy0 = round_bias((double)dd.y - fw);
			int dy1 = round_bias((double)dd.y + fw - 1.0);
			if(-sx > dx0) dx0 = -sx;
			if((c.resx - 1) - sx < dx1) dx1 = (c.resx - 1) - sx;
			if(-sy > dy0) dy0 = -sy;
			if((c.resy - 1) - sy < dy1) dy1 = (c.resy - 1) - sy;
			const int ix = tx - sx, iy = ty - sy;
			if(ix < dx0 || ix > dx1 || iy < dy0 || iy > dy1) continue;
			int xi = (int)floor(fabs(((double)ix - ((double)dd.x - 0.5)) * (double)c.filter_table_scale));
			int yi = (int)floor(fabs(((double)iy - ((double)dd.y - 0.5)) * (double)c.filter_table_scale));
			xi = xi > 15 ? 15 : xi;
			yi = yi > 15 ? 15 : yi;
			const float wt = c.filter_table[yi * 16 + xi];
			const float4 col = clamp_proportional(a.nd.col[s], c.aa_clamp);  // camera sample s is node s
			w = w + wt;
			acc.x += wt * col.x;
			acc.y += wt * col.y;
			acc.z += wt * col.z;
			acc.w += wt * col.w;
		}
	}
	a.film_w[t] = w;
	a.film[t] = acc;
}

// ---- ImageFilm::nextPass (src/render/imagefilm.cc:398-513) on the device.
// d_aa holds, npix words each: flags, pair bits, the pass's pixel list and
// each pixel's slot in it; then per-block counts / offsets and the count word.
struct AaArgs
{
	const float4 *film;
	const float *film_w;
	int32_t w, h;
	float threshold;          // ImageFilm's aa_threshold_calculated_
	int32_t dark_type;        // 0 none, 1 linear, 2 curve
	float dark_factor;
	int32_t color_noise;      // AA_detect_color_noise
	int32_t var_half;         // AA_variance_edge_size / 2
	int32_t var_pixels;       // AA_variance_pixels
	uint32_t *flags, *bits, *list, *slot, *block_count, *count;
};

// Rgba::normalized (include/color/color.h:538-542) of film pixel p.
__device__ __forceinline__ float4 film_pixel(const AaArgs &g, uint32_t p)
{
	float w = g.film_w[p];
	if(w == 0.f) return make_float4(0.f, 0.f, 0.f, 0.f);
	w = 1.f / w;
	const float4 c = g.film[p];
	return make_float4(c.x * w, c.y * w, c.z * w, c.w * w);
}

// Rgba::colorDifference (include/color/color.h:434-451) with col2Bri (:57)
__device__ __forceinline__ float color_difference(float4 a, float4 b, int rgb_components)
{
	const float bri_a = 0.2126f * a.x + 0.7152f * a.y + 0.0722f * a.z;
	const float bri_b = 0.2126f * b.x + 0.7152f * b.y + 0.0722f * b.z;
	float d = fabsf(bri_b - bri_a);
	if(rgb_components)
	{
		const float dr = fabsf(b.x - a.x), dg = fabsf(b.y - a.y), db = fabsf(b.z - a.z), da = fabsf(b.w - a.w);
		if(d < dr) d = dr;
		if(d < dg) d = dg;
		if(d < db) d = db;
		if(d < da) d = da;
	}
	return d;
}

// ImageFilm::darkThresholdCurveInterpolate (src/render/imagefilm.cc:882-899)
__device__ __forceinline__ float dark_curve(float b)
{
	if(b <= 0.10f) return 0.0001f;
	else if(b <= 0.20f) return (0.0001f + (b - 0.10f) * (0.0010f - 0.0001f) / 0.10f);
	else if(b <= 0.30f) return (0.0010f + (b - 0.20f) * (0.0020f - 0.0010f) / 0.10f);
	else if(b <= 0.40f) return (0.0020f + (b - 0.30f) * (0.0035f - 0.0020f) / 0.10f);
	else if(b <= 0.50f) return (0.0035f + (b - 0.40f) * (0.0055f - 0.0035f) / 0.10f);
	else if(b <= 0.60f) return (0.0055f + (b - 0.50f) * (0.0075f - 0.0055f) / 0.10f);
	else if(b <= 0.70f) return (0.0075f + (b - 0.60f) * (0.0100f - 0.0075f) / 0.10f);
	else if(b <= 0.80f) return (0.0100f + (b - 0.70f) * (0.0150f - 0.0100f) / 0.10f);
	else if(b <= 0.90f) return (0.0150f + (b - 0.80f) * (0.0250f - 0.0150f) / 0.10f);
	else if(b <= 1.00f) return (0.0250f + (b - 0.90f) * (0.0400f - 0.0250f) / 0.10f);
	else if(b <= 1.20f) return (0.0400f + (b - 1.00f) * (0.0800f - 0.0400f) / 0.20f);
	else if(b <= 1.40f) return (0.0800f + (b - 1.20f) * (0.0950f - 0.0800f) / 0.20f);
	else if(b <= 1.80f) return (0.0950f + (b - 1.40f) * (0.1000f - 0.0950f) / 0.40f);
	return 0.1000f;
}

// Per source pixel (x, y) of the reference's loop (x < w - 1, y < h - 1):
// bit k set when its k-th forward pair -- (x+1, y), (x, y+1), (x+1, y+1),
// (x-1, y+1) -- differs by at least its threshold; bit 4 when its variance
// window counts AA_variance_pixels such differences.
__global__ __launch_bounds__(kBlock) void k_aa_pairs(AaArgs g)
{
	const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
	if(p >= (uint32_t)(g.w * g.h)) return;
	const int x = (int)(p % (uint32_t)g.w), y = (int)(p / (uint32_t)g.w);
	uint32_t b = 0u;
	if(x < g.w - 1 && y < g.h - 1)
	{
		const float4 pix = film_pixel(g, p);
		float thr = g.threshold;
		const float bri = 0.2126f * fabsf(pix.x) + 0.7152f * fabsf(pix.y) + 0.0722f * fabsf(pix.z);
		if(g.dark_type == 1 && g.dark_factor > 0.f) thr = g.threshold * ((1.f - g.dark_factor) + (bri * g.dark_factor));
		else if(g.dark_type == 2) thr = dark_curve(bri);
		const int nx[4] = {x + 1, x, x + 1, x - 1}, ny[4] = {y, y + 1, y + 1, y + 1};
		for(int k = 0; k < 4; ++k)
		{
			if(k == 3 && x == 0) continue;
			const float4 q = film_pixel(g, (uint32_t)(ny[k] * g.w + nx[k]));
			if(color_difference(pix, q, g.color_noise) >= thr) b |= 1u << k;
		}
		if(g.var_pixels > 0)
		{
			int vx = 0, vy = 0;
			for(int xd = -g.var_half; xd < g.var_half - 1; ++xd)
			{
				int xi = x + xd;
				if(xi < 0) xi = 0;
				else if(xi >= g.w - 1) xi = g.w - 2;
				const float4 c0 = film_pixel(g, (uint32_t)(y * g.w + xi));
				const float4 c1 = film_pixel(g, (uint32_t)(y * g.w + xi + 1));
				if(color_difference(c0, c1, g.color_noise) >= thr) ++vx;
			}
			for(int yd = -g.var_half; yd < g.var_half - 1; ++yd)
			{
				int yi = y + yd;
				if(yi < 0) yi = 0;
				else if(yi >= g.h - 1) yi = g.h - 2;
				const float4 c0 = film_pixel(g, (uint32_t)(yi * g.w + x));
				const float4 c1 = film_pixel(g, (uint32_t)((yi + 1) * g.w + x));
				if(color_difference(c0, c1, g.color_noise) >= thr) ++vy;
			}
			if(vx + vy >= g.var_pixels) b |= 16u;
		}
	}
	g.bits[p] = b;
}

// Pixel (px, py) is flagged when it has no weight yet, when one of the pairs
// it belongs to differs, or when it lies in the (clamped) variance window
// [sx - half, sx + half) x [sy - half, sy + half) of a source with bit 4:
// with clamping, that is sx in [px - half + 1, px + half] and likewise in y.
// Each block also counts its flagged pixels.
__global__ __launch_bounds__(kBlock) void k_aa_flags(AaArgs g)
{
	const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
	const uint32_t npix = (uint32_t)(g.w * g.h);
	uint32_t f = 0u;
	if(p < npix)
	{
		const int x = (int)(p % (uint32_t)g.w), y = (int)(p / (uint32_t)g.w);
		f = g.film_w[p] > 0.f ? 0u : 1u;
		if(g.bits[p] & 15u) f = 1u;
		if(x > 0 && (g.bits[p - 1] & 1u)) f = 1u;                               // (x-1, y) -> (x, y)
		if(y > 0 && (g.bits[p - (uint32_t)g.w] & 2u)) f = 1u;                   // (x, y-1) -> (x, y)
		if(x > 0 && y > 0 && (g.bits[p - (uint32_t)g.w - 1] & 4u)) f = 1u;      // (x-1, y-1) -> (x, y)
		if(x + 1 < g.w && y > 0 && (g.bits[p - (uint32_t)g.w + 1] & 8u)) f = 1u;  // (x+1, y-1) -> (x, y)
		if(g.var_pixels > 0 && g.var_half > 0 && !f)
		{
			const int sx0 = max(x - g.var_half + 1, 0), sx1 = min(x + g.var_half, g.w - 2);
			const int sy0 = max(y - g.var_half + 1, 0), sy1 = min(y + g.var_half, g.h - 2);
			for(int sy = sy0; sy <= sy1 && !f; ++sy)
				for(int sx = sx0; sx <= sx1; ++sx)
					if(g.bits[(uint32_t)(sy * g.w + sx)] & 16u)
					{
						f = 1u;
						break;
					}
		}
		g.flags[p] = f;
	}
	const int n = __syncthreads_count((int)f);
	if(threadIdx.x == 0) g.block_count[blockIdx.x] = (uint32_t)n;
}

// Exclusive scan of the per-block counts (one block), total into *count.
__global__ __launch_bounds__(1024) void k_aa_scan(uint32_t *block_count, uint32_t n_blocks, uint32_t *count)
{
	__shared__ uint32_t part[1024];
	const uint32_t per = (n_blocks + 1023u) / 1024u;
	const uint32_t b0 = threadIdx.x * per;
	uint32_t sum = 0u;
	for(uint32_t k = 0; k < per && b0 + k < n_blocks; ++k) sum += block_count[b0 + k];
	part[threadIdx.x] = sum;
	__syncthreads();
	for(uint32_t off = 1; off < 1024u; off <<= 1)
	{
		const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
		__syncthreads();
		part[threadIdx.x] += v;
		__syncthreads();
	}
	uint32_t run = part[threadIdx.x] - sum;  // exclusive
	for(uint32_t k = 0; k < per && b0 + k < n_blocks; ++k)
	{
		const uint32_t c = block_count[b0 + k];
		block_count[b0 + k] = run;
		run += c;
	}
	if(threadIdx.x == 1023) *count = part[1023];
}

// The flagged pixels in ascending order, and each pixel's slot.
__global__ __launch_bounds__(kBlock) void k_aa_list(AaArgs g)
{
	__shared__ uint32_t wave_base[kBlock / 64];
	const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
	const uint32_t npix = (uint32_t)(g.w * g.h);
	const bool f = p < npix && g.flags[p];
	const unsigned long long m = __ballot(f);
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	if(lane == 0) wave_base[wave] = (uint32_t)__popcll(m);
	__syncthreads();
	uint32_t base = g.block_count[blockIdx.x];
	for(uint32_t k = 0; k < wave; ++k) base += wave_base[k];
	const uint32_t pos = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
	if(p < npix)
	{
		g.slot[p] = f ? pos : kNoSlot;
		if(f) g.list[pos] = p;
	}
}

__global__ void k_count_rays(unsigned long long *total, uint32_t n)
{
	if(blockIdx.x == 0 && threadIdx.x == 0) *total += n;
}

inline unsigned blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// The queued shadow rays: any-hit, or with transparent shadows the
// transparent-shadow query (Accelerator::isShadowed / isShadowedTransparentShadow).
int shadow_pass(ymx_pt *pt, hipStream_t stream)
{
	const PtConst &hc = pt->host_const;
	if(hc.transparent_shadows)
		return ymx_launch_transparent_segmented(pt->accel, pt->st.sh_rays, pt->d_transp, hc.shadow_depth,
		                                        pt->st.sh_occ, pt->st.sh_tcount, pt->st.sh_thits, pt->st.sh_count,
		                                        pt->st.sh_seg_cap, stream);
	return ymx_launch_shadow_segmented(pt->accel, pt->st.sh_rays, pt->st.sh_occ, pt->st.sh_count, pt->st.sh_seg_cap,
	                                   pt->shadow_replacement, stream);
}

// The stages of one chunk of nodes [a.base, a.base + a.cnt).
int trace_chunk(ymx_pt *pt, PtArgs &a, hipStream_t stream)
{
	const PtConst &hc = pt->host_const;
	const uint32_t n = a.cnt;
	const unsigned g = blocks_for(n);
	k_count_rays<<<1, 64, 0, stream>>>(pt->st.totals, n);
	YMX_HIP_CHECK(hipGetLastError());
	// camera rays are coherent: the wave-ordered kernel (ray replacement costs there)
	int rc = ymx_launch_closest_ordered(pt->accel, pt->nd.ray + a.base, pt->st.ext_hits, n, stream);
	if(rc != YMX_OK) return rc;
	k_shade_node<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	if(hc.n_lights > 0)
	{
		rc = shadow_pass(pt, stream);
		if(rc != YMX_OK) return rc;
	}
	k_resolve_node<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	if(a.g.on)
	{
		k_guide_query_node<<<g, kBlock, 0, stream>>>(a);
		YMX_HIP_CHECK(hipGetLastError());
		rc = ymx_guide_infer(pt->guide, a.g.query, a.g.mix_node, n, nullptr, stream);
		if(rc != YMX_OK) return rc;
	}
	const int n_sub = hc.path_samples > 1 ? hc.path_samples : 1;
	for(int sub = 0; sub < n_sub; ++sub)
	{
		a.sub = sub;
		k_subpath_init<<<g, kBlock, 0, stream>>>(a);
		YMX_HIP_CHECK(hipGetLastError());
		for(int d = 0; d < hc.bounces; ++d)
		{
			a.depth = d;
			const int qsel = d & 1;
			rc = ymx_launch_closest_queue(pt->accel, pt->st.ext_rays[qsel], pt->st.ext_hits, pt->st.ext_count + qsel, n,
			                              pt->ext_replacement, stream);
			if(rc != YMX_OK) return rc;
			if(a.g.on && d + 1 < hc.bounces && d < a.g.max_depth)
			{
				k_guide_query_path<<<g, kBlock, 0, stream>>>(a);
				YMX_HIP_CHECK(hipGetLastError());
				rc = ymx_guide_infer(pt->guide, a.g.query, a.g.mix_path, n, pt->st.ext_count + qsel, stream);
				if(rc != YMX_OK) return rc;
			}
			k_shade_path<<<g, kBlock, 0, stream>>>(a);
			YMX_HIP_CHECK(hipGetLastError());
			if(hc.n_lights > 0)
			{
				rc = shadow_pass(pt, stream);
				if(rc != YMX_OK) return rc;
			}
			k_resolve_path<<<g, kBlock, 0, stream>>>(a);
			YMX_HIP_CHECK(hipGetLastError());
		}
		if(a.g.train)
		{
			k_guide_collect<<<g, kBlock, 0, stream>>>(a);
			YMX_HIP_CHECK(hipGetLastError());
		}
	}
	k_finish<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}

}  // namespace
}  // namespace ymx

int ymx_pt_trace_pass(ymx_pt *pt, int32_t first_sample, int32_t count, hipStream_t stream)
{
	const uint32_t npix = (uint32_t)(pt->host_const.resx * pt->host_const.resy);
	return ymx_pt_trace_subset(pt, first_sample, count, nullptr, nullptr, npix, 1.f, pt->host_const.aa_samples, stream);
}

int ymx_pt_aa_flags(ymx_pt *pt, float threshold, hipStream_t stream)
{
	using namespace ymx;
	const PtConst &hc = pt->host_const;
	const ymx_pt_settings &st = pt->settings;
	const uint32_t npix = (uint32_t)(hc.resx * hc.resy);
	const uint32_t n_blocks = blocks_for(npix);
	AaArgs g;
	g.film = pt->d_film;
	g.film_w = pt->d_film_w;
	g.w = hc.resx;
	g.h = hc.resy;
	g.threshold = threshold;
	g.dark_type = st.aa_dark_detection_type;
	g.dark_factor = st.aa_dark_threshold_factor;
	g.color_noise = st.aa_detect_color_noise ? 1 : 0;
	g.var_half = st.aa_variance_edge_size / 2;
	g.var_pixels = st.aa_variance_pixels;
	g.flags = pt->d_aa;
	g.bits = pt->d_aa + npix;
	g.list = pt->d_aa + 2 * (size_t)npix;
	g.slot = pt->d_aa + 3 * (size_t)npix;
	g.block_count = pt->d_aa + 4 * (size_t)npix;
	g.count = g.block_count + n_blocks;
	k_aa_pairs<<<n_blocks, kBlock, 0, stream>>>(g);
	YMX_HIP_CHECK(hipGetLastError());
	k_aa_flags<<<n_blocks, kBlock, 0, stream>>>(g);
	YMX_HIP_CHECK(hipGetLastError());
	k_aa_scan<<<1, 1024, 0, stream>>>(g.block_count, n_blocks, g.count);
	YMX_HIP_CHECK(hipGetLastError());
	k_aa_list<<<n_blocks, kBlock, 0, stream>>>(g);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}

int ymx_pt_trace_subset(ymx_pt *pt, int32_t first_sample, int32_t count, const uint32_t *pix_list,
                        const uint32_t *pix_slot, uint32_t n_active, float light_mult, int32_t pass_samples,
                        hipStream_t stream)
{
	using namespace ymx;
	const PtConst &hc = pt->host_const;
	const uint32_t npix = (uint32_t)(hc.resx * hc.resy);
	const uint64_t n64 = (uint64_t)n_active * (uint64_t)(count > 0 ? count : 0);
	const uint32_t n = (uint32_t)n64;
	if(count < 1 || n_active < 1 || n_active > npix || n64 > pt->st.capacity || n > pt->nd.cap)
	{
		ymx_set_error("ymx_pt_trace_pass: %d samples of %u pixels exceed the wavefront capacity", count, n_active);
		return YMX_ERR_INVALID_ARG;
	}
	PtArgs a;
	a.c = pt->d_const;
	a.mats = pt->d_materials;
	a.lights = pt->d_lights;
	a.ltris = pt->d_ltris;
	a.prim = pt->d_prim;
	a.cnorm = pt->d_cnorm;
	a.faure = pt->d_faure;
	a.inv_prims = pt->d_inv_prims;
	a.prims = pt->d_prims;
	a.st = pt->st;
	a.nd = pt->nd;
	a.film = pt->d_film;
	a.film_w = pt->d_film_w;
	a.first_sample = first_sample;
	a.npix = npix;
	a.n_active = n_active;
	a.pix_list = pix_list;
	a.pix_slot = pix_slot;
	a.light_mult = light_mult;
	a.pass_samples = pass_samples;
	a.n_paths = n;
	a.count = count;
	a.sub = 0;
	a.depth = 0;
	a.g = pt->g;
	k_camera<<<blocks_for(n), kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	// levels of integrate() calls, node ranges [start, start + size)
	std::vector<uint32_t> lvl_start{0u}, lvl_size{n};
	const uint32_t cap = pt->g.on && pt->guide_cap < pt->st.capacity ? pt->guide_cap : pt->st.capacity;
	for(int level = 0;; ++level)
	{
		const uint32_t start = lvl_start[level], size = lvl_size[level];
		const uint32_t next_start = start + size;
		const bool spawn = pt->recursive && level + 1 <= hc.raydepth;
		uint32_t spawned = 0;
		if(spawn) YMX_HIP_CHECK(hipMemsetAsync(pt->nd.spawn, 0, sizeof(uint32_t), stream));
		for(uint32_t base = start; base < next_start; base += cap)
		{
			a.base = base;
			a.cnt = (next_start - base) < cap ? (next_start - base) : cap;
			a.next_start = next_start;
			a.recursive = spawn ? 1u : 0u;
			if(spawn)
			{
				// room for every child this chunk can spawn, after those spawned so far
				const uint64_t need = (uint64_t)next_start + spawned + (uint64_t)a.cnt * pt->max_fan;
				if(need > pt->nd.cap)
				{
					const int rc = ymx_pt_grow_nodes(pt, need, next_start + spawned);
					if(rc != YMX_OK) return rc;
					a.nd = pt->nd;
				}
			}
			const int rc = trace_chunk(pt, a, stream);
			if(rc != YMX_OK) return rc;
			if(spawn)
			{
				YMX_HIP_CHECK(hipMemcpyAsync(&spawned, pt->nd.spawn, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
				YMX_HIP_CHECK(hipStreamSynchronize(stream));
			}
		}
		if(spawned == 0) break;
		lvl_start.push_back(next_start);
		lvl_size.push_back(spawned);
	}
	// children into parents, deepest level first
	for(int level = (int)lvl_start.size() - 2; level >= 0; --level)
	{
		const uint32_t end = lvl_start[level] + lvl_size[level];
		for(uint32_t base = lvl_start[level]; base < end; base += cap)
		{
			a.base = base;
			a.cnt = (end - base) < cap ? (end - base) : cap;
			k_combine<<<blocks_for(a.cnt), kBlock, 0, stream>>>(a);
			YMX_HIP_CHECK(hipGetLastError());
		}
	}
	k_film<<<blocks_for(npix), kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	if(pt->g.train) return ymx_pt_train_guide(pt, stream);
	return YMX_OK;
}

// Adam steps of the guide on mini-batches drawn from the ring buffer of recent
// samples, then the acceptance probability that lets one wavefront fill about a
// quarter of the ring (so it holds the last few wavefronts).
int ymx_pt_train_guide(ymx_pt *pt, hipStream_t stream)
{
	using namespace ymx;
	uint32_t total = 0;
	YMX_HIP_CHECK(hipMemcpyAsync(&total, pt->g.train_count, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
	YMX_HIP_CHECK(hipStreamSynchronize(stream));
	const uint32_t kept = total - pt->g_seen;  // this wavefront's samples (mod 2^32)
	pt->g_seen = total;
	pt->g_filled = pt->g_filled + kept < pt->g.train_cap ? pt->g_filled + kept : pt->g.train_cap;
	const uint32_t n = pt->g_filled;
	const uint32_t max_batch = ymx_guide_batch_limit(pt->guide);
	pt->stats.guide_samples += (uint64_t)(kept / (double)(pt->g.accept > 0.f ? pt->g.accept : 1.f));
	if(n > 0 && pt->train_steps > 0)
	{
		pt->stats.guide_steps += pt->train_steps;
		const uint32_t batch = n < max_batch ? n : max_batch;
		for(uint32_t k = 0; k < pt->train_steps; ++k)
		{
			// samples drawn uniformly (with replacement) from the whole ring: its
			// order follows the wavefront's slots, so a contiguous slice would be
			// one band of the image
			pt->g_rng = pt->g_rng * 6364136223846793005ull + 1442695040888963407ull;
			k_guide_gather<<<blocks_for(batch), kBlock, 0, stream>>>(pt->g.batch_buf, pt->g.train_buf, n, batch,
			                                                          (uint32_t)(pt->g_rng >> 32), pt->g);
			YMX_HIP_CHECK(hipGetLastError());
			const int rc = ymx_guide_train_step(pt->guide, pt->g.batch_buf, batch, nullptr, stream);
			if(rc != YMX_OK) return rc;
			if(pt->g.learn_alpha)
			{
				constexpr uint32_t cells = (uint32_t)(kGuideCells * kGuideCells * kGuideCells);
				k_alpha_grad<<<blocks_for(batch), kBlock, 0, stream>>>(pt->g.batch_buf, batch, pt->g);
				++pt->alpha_steps;
				const float bc1 = 1.f - powf(0.9f, (float)pt->alpha_steps);
				const float bc2 = 1.f - powf(0.999f, (float)pt->alpha_steps);
				k_alpha_adam<<<blocks_for(cells), kBlock, 0, stream>>>(pt->g, cells, pt->alpha_lr, bc1, bc2);
				YMX_HIP_CHECK(hipGetLastError());
			}
			if(pt->g.fallback)
			{
				// the moments at the alpha the selection step just set
				constexpr uint32_t cells = (uint32_t)(kGuideCells * kGuideCells * kGuideCells);
				k_fallback_sum<<<blocks_for(batch), kBlock, 0, stream>>>(pt->g.batch_buf, batch, pt->g);
				k_fallback_update<<<blocks_for(cells), kBlock, 0, stream>>>(pt->g);
				YMX_HIP_CHECK(hipGetLastError());
			}
		}
	}
	const double records = kept / (double)(pt->g.accept > 0.f ? pt->g.accept : 1.f);
	double acc = records > 0 ? 0.25 * pt->g.train_cap / records : 1.0;
	acc = acc > 1.0 ? 1.0 : (acc < 1e-4 ? 1e-4 : acc);
	pt->g.accept = (float)acc;
	return YMX_OK;
}
