// pt_kernels.hip -- wavefront kernels of the path tracer.
//
// One AA sample index of every pixel is one wavefront.  Stage order per pass
// (PathIntegrator::integrate, src/integrator/surface/integrator_path_tracer.cc:138-314):
//   camera        camera rays of every pixel -> extension queue 0
//   extend        closest hit (accel_trace.hip)
//   shade_camera  emission + next-event estimation for all lights: shadow
//                 rays queued, their contributions kept per slot
//   shadow        any hit over the shadow queue, then resolve_camera sums the
//                 unoccluded slot contributions in the reference's order
//   per sub-path: subpath_init samples the first bounce; then for each
//   segment d: extend, shade_path (roulette, one-light estimation, next BSDF
//   sample), shadow, resolve_path
//   finish        col += path_col / path_samples
//   film          gather-splat into the film with the Gauss filter table in
//                 the order a row-major scatter would add (ImageFilm::addSample)
// The shading arithmetic follows the reference's evaluation order; see
// pt_math.h.  Compiled with -ffp-contract=off and IEEE division.
#include "pt_internal.h"

#include <float.h>

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
	const float4 *prim;
	const int32_t *faure;
	const double *inv_prims;
	const uint32_t *prims;
	PtState st;
	float4 *film;
	float *film_w;
	int32_t first_sample;  // AA sample index of the wavefront's first slice
	uint32_t npix;         // resx * resy
	uint32_t n_paths;      // npix * count: path i traces pixel i % npix, sample first + i / npix
	int32_t count;         // AA samples in this wavefront
	int32_t sub;
	int32_t depth;
};

__device__ __forceinline__ uint32_t pix_of(const PtArgs &a, uint32_t i) { return i % a.npix; }
__device__ __forceinline__ uint32_t sample_of(const PtArgs &a, uint32_t i)
{
	return (uint32_t)a.first_sample + i / a.npix;
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
	coords_system(s.n, s.u, s.v);
	s.mat = __float_as_uint(pi.w);
	s.flags = a.mats[s.mat].flags;
}

__device__ __forceinline__ V3 face_forward(V3 ng, V3 n, V3 wo) { return dot(ng, wo) < 0.f ? neg(n) : n; }

// ShinyDiffuseMaterial::eval (src/material/material_shiny_diffuse.cc:297-340)
__device__ Rgb bsdf_eval(const PtArgs &a, const Surf &sp, V3 wo, V3 wl, uint32_t bsdfs)
{
	const V3 n = face_forward(sp.ng, sp.n, wo);
	if(!(bsdfs & (sp.flags & kDiffuse))) return rgb(0.f, 0.f, 0.f);
	if(dot(n, wl) < 0.0) return rgb(0.f, 0.f, 0.f);
	const PtMaterial &m = a.mats[sp.mat];
	const float m_d = 1.f * (1.f - 0.f) * m.diffuse_reflect;
	return cscale(m.color, m_d);
}

// ShinyDiffuseMaterial::pdf (:441-483)
__device__ float bsdf_pdf(const PtArgs &a, const Surf &sp, V3 wo, V3 wi, uint32_t bsdfs)
{
	if(!(bsdfs & kDiffuse)) return 0.f;
	if(!(sp.flags & kDiffuse)) return 0.f;
	const V3 n = face_forward(sp.ng, sp.n, wo);
	float pdf = 0.f, sum = 0.f;
	int n_match = 0;
	if(bsdfs & (kDiffuse | kReflect))
	{
		const float width = a.mats[sp.mat].diffuse_reflect;
		sum += width;
		pdf += fabsf(dot(wi, n)) * width;
		++n_match;
	}
	if(!n_match || sum < 0.00001) return 0.f;
	return pdf / sum;
}

// sample::cosHemisphere, include/sampler/sample.h:43-52
__device__ V3 cos_hemisphere(V3 n, V3 ru, V3 rv, float s1, float s2)
{
	if(s1 >= 1.0f) return n;
	const float z1 = s1;
	const float z2 = s2 * k2Pi;
	return add(mul(sqrtf(1.f - z1), add(mul(fast_cos(z2), ru), mul(fast_sin(z2), rv))), mul(sqrtf(z1), n));
}

struct BSample
{
	float s1, s2;
	uint32_t flags, sampled;
	float pdf;
};

// ShinyDiffuseMaterial::sample, diffuse component (:349-439)
__device__ Rgb bsdf_sample(const PtArgs &a, const Surf &sp, V3 wo, V3 &wi, BSample &s, float &w)
{
	const float cos_ng_wo = dot(sp.ng, wo);
	const V3 n = face_forward(sp.ng, sp.n, wo);
	const PtMaterial &m = a.mats[sp.mat];
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

__device__ __forceinline__ uint32_t light_slots(const PtLight &l)
{
	return l.type == YMX_LIGHT_AREA ? 2u * (uint32_t)l.samples : 1u;
}

// Accelerator::isShadowed (include/accelerator/accelerator.h:103-111): origin
// moved by tmin, t_max = tmax - 2 tmin.  Returns the queue index, or -1 when no
// test is needed (t_max <= 0 can never be occluded).
__device__ int32_t queue_shadow(const PtArgs &a, V3 from, V3 dir, float tmin, float tmax)
{
	const float t_max = tmax - 2 * tmin;
	if(!(t_max > 0.f)) return -1;
	const V3 f = add(from, mul(tmin, dir));
	const uint32_t q = atomicAdd(a.st.sh_count, 1u);
	float4 *r = reinterpret_cast<float4 *>(a.st.sh_rays + q);
	r[0] = make_float4(f.x, f.y, f.z, tmin);
	r[1] = make_float4(dir.x, dir.y, dir.z, t_max);
	return (int32_t)q;
}

__device__ __forceinline__ void set_slot(const PtArgs &a, uint32_t slot, Rgb c, bool valid, int32_t q)
{
	a.st.slot_contrib[slot] = make_float4(c.r, c.g, c.b, valid ? 1.f : 0.f);
	a.st.slot_queue[slot] = q;
}

// MonteCarloIntegrator::doLightEstimation with its area / dirac paths
// (src/integrator/surface/integrator_montecarlo.cc:118-440), queuing one shadow
// ray per sample instead of tracing it.
__device__ void queue_light(const PtArgs &a, const Surf &sp, V3 wo, const PtLight &l, uint32_t loffs,
                            uint32_t psd_sample, uint32_t psd_offset, uint32_t slot0)
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
			set_slot(a, slot0, rgb(0.f, 0.f, 0.f), false, -1);
			return;
		}
		const float idist_sqr = 1.f / dist_sqr;
		ldir = mul(1.f / dist, ldir);
		const Rgb lcol = cscale(l.color, idist_sqr);
		const float angle = fabsf(dot(sp.n, ldir));
		const Rgb surf_col = bsdf_eval(a, sp, wo, ldir, kAll);
		const Rgb c = cmul(cscale(cmul(surf_col, lcol), angle), rgb(1.f, 1.f, 1.f));
		const int32_t q = cast_shadows ? queue_shadow(a, sp.p, ldir, shadow_tmin, dist) : -1;
		set_slot(a, slot0, c, true, q);
		return;
	}
	const uint32_t l_offs = loffs * 4567u;
	const int num_samples = (int)ceilf((float)l.samples * 1.f);
	const uint32_t offs = (uint32_t)num_samples * psd_sample + psd_offset + l_offs;
	// areaLightSampleLight (:192-316) with AreaLight::illumSample (src/light/light_area.cc:99-124)
	Halton h2, h3;
	h2.start(2, offs - 1);
	h3.start(3, offs - 1);
	for(int i = 0; i < num_samples; ++i)
	{
		const float s1 = h2.next();
		const float s2 = h3.next();
		const uint32_t slot = slot0 + (uint32_t)i;
		const V3 p = add(add(l.corner, mul(s1, l.to_x)), mul(s2, l.to_y));
		V3 ldir = sub(p, sp.p);
		const float dist_sqr = dot(ldir, ldir);
		const float dist = sqrtf(dist_sqr);
		if(dist <= 0.f)
		{
			set_slot(a, slot, rgb(0.f, 0.f, 0.f), false, -1);
			continue;
		}
		ldir = mul(1.f / dist, ldir);
		const float cos_angle = dot(ldir, l.normal_flipped);
		if(cos_angle <= 0.f)
		{
			set_slot(a, slot, rgb(0.f, 0.f, 0.f), false, -1);
			continue;
		}
		const float lpdf = dist_sqr * kPi / (l.area * cos_angle);
		if(!(lpdf > 1e-6f))
		{
			set_slot(a, slot, rgb(0.f, 0.f, 0.f), false, -1);
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
		const int32_t q = cast_shadows ? queue_shadow(a, sp.p, ldir, shadow_tmin, dist) : -1;
		set_slot(a, slot, c, true, q);
	}
	// areaLightSampleMaterial (:318-415) with AreaLight::intersect (src/light/light_area.cc:146-155)
	const float btmin = 0.00005f * pscale;
	h2.start(2, offs - 1);
	h3.start(3, offs - 1);
	for(int i = 0; i < num_samples; ++i)
	{
		const uint32_t slot = slot0 + (uint32_t)num_samples + (uint32_t)i;
		BSample s;
		s.s1 = h2.next();
		s.s2 = h3.next();
		s.flags = kGlossy | kDiffuse | kDispersive | kReflect | kTransmit;
		s.pdf = 0.f;
		s.sampled = kNone;
		float W = 0.f;
		V3 dir = v3(0.f, 0.f, 0.f);
		const Rgb surf_col = bsdf_sample(a, sp, wo, dir, s, W);
		bool valid = false;
		Rgb c = rgb(0.f, 0.f, 0.f);
		int32_t q = -1;
		const float cos_l = dot(dir, l.normal_flipped);
		if(cos_l > 0.f)
		{
			const float t = quad_intersect(l.quad, sp.p, dir);
			if(t > 0.f && s.pdf > 1e-6f)
			{
				const float ipdf = 1.f / (t * t) * l.area * cos_l * k1Pi;
				if(ipdf > 1e-6f)
				{
					const float l_pdf = 1.f / ipdf;
					const float l_2 = l_pdf * l_pdf;
					const float m_2 = s.pdf * s.pdf;
					const float w = m_2 / (l_2 + m_2);
					c = cscale(cscale(cmul(surf_col, l.color), w), W);
					valid = true;
					q = cast_shadows ? queue_shadow(a, sp.p, dir, btmin, t) : -1;
				}
			}
		}
		set_slot(a, slot, c, valid, q);
	}
}

__device__ __forceinline__ bool slot_lit(const PtArgs &a, uint32_t slot, Rgb &c)
{
	const float4 v = a.st.slot_contrib[slot];
	if(v.w == 0.f) return false;
	const int32_t q = a.st.slot_queue[slot];
	if(q >= 0 && ((a.st.sh_occ[(uint32_t)q >> 5] >> ((uint32_t)q & 31u)) & 1u)) return false;
	c = rgb(v.x, v.y, v.z);
	return true;
}

// Sum of one light's slots in the order doLightEstimation accumulates them.
__device__ Rgb resolve_light(const PtArgs &a, const PtLight &l, uint32_t slot0)
{
	Rgb c;
	if(l.type == YMX_LIGHT_POINT)
	{
		Rgb col = rgb(0.f, 0.f, 0.f);
		if(slot_lit(a, slot0, c)) col = cadd(col, c);
		return col;
	}
	const int num_samples = (int)ceilf((float)l.samples * 1.f);
	const float inv = 1.f / (float)num_samples;
	Rgb col = rgb(0.f, 0.f, 0.f);
	Rgb acc = rgb(0.f, 0.f, 0.f);
	for(int i = 0; i < num_samples; ++i)
		if(slot_lit(a, slot0 + (uint32_t)i, c)) acc = cadd(acc, c);
	col = cadd(col, cscale(acc, inv));
	acc = rgb(0.f, 0.f, 0.f);
	for(int i = 0; i < num_samples; ++i)
		if(slot_lit(a, slot0 + (uint32_t)num_samples + (uint32_t)i, c)) acc = cadd(acc, c);
	col = cadd(col, cscale(acc, inv));
	return col;
}

__device__ __forceinline__ uint32_t path_offset(uint32_t x, uint32_t y) { return fnv32a(y * fnv32a(x)); }

__device__ __forceinline__ void queue_extension(const PtArgs &a, int qsel, uint32_t path, V3 from, V3 dir, float tmin)
{
	const uint32_t q = atomicAdd(a.st.ext_count + qsel, 1u);
	float4 *r = reinterpret_cast<float4 *>(a.st.ext_rays[qsel] + q);
	r[0] = make_float4(from.x, from.y, from.z, tmin);
	r[1] = make_float4(dir.x, dir.y, dir.z, -1.f);
	a.st.ext_path[qsel][q] = path;
}

// TiledIntegrator::renderTile sample set-up and PerspectiveCamera::shootRay
// (src/integrator/surface/integrator_tiled.cc:290-335, src/camera/camera_perspective.cc:165-180)
__global__ __launch_bounds__(kBlock) void k_camera(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t n = a.n_paths;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i == 0)
	{
		a.st.ext_count[1] = 0u;
		*a.st.sh_count = 0u;
		a.st.totals[0] += n;
	}
	if(i >= n) return;
	const uint32_t p = pix_of(a, i);
	const uint32_t sample = sample_of(a, i);
	const uint32_t x = p % (uint32_t)c.resx, y = p / (uint32_t)c.resx;
	const uint32_t offset = path_offset(x, y);
	float dx = 0.5f, dy = 0.5f;
	if(c.aa_samples > 1)
	{
		const float d_1 = 1.f / (float)c.aa_samples;
		dx = (0.5f + (float)sample) * d_1;
		dy = ri_lp(sample + offset, 0);
	}
	const float px = (float)x + dx, py = (float)y + dy;
	const V3 dir = normalize(add(add(mul(px, c.vright), mul(py, c.vup)), c.vto));
	a.st.col[i] = make_float4(0.f, 0.f, 0.f, 1.f);
	a.st.path_col[i] = make_float4(0.f, 0.f, 0.f, 0.f);
	a.st.cam_dir[i] = make_float4(dir.x, dir.y, dir.z, 0.f);
	a.st.dxy[i] = make_float2(dx, dy);
	a.st.offset[i] = offset;
	a.st.flags[i] = 0u;
	float4 *r = reinterpret_cast<float4 *>(a.st.ext_rays[0] + i);
	r[0] = make_float4(c.cam_from.x, c.cam_from.y, c.cam_from.z, 0.f);
	r[1] = make_float4(dir.x, dir.y, dir.z, -1.f);
	a.st.ext_path[0][i] = i;
}

// Camera vertex: emission and estimateAllDirectLight (integrator_path_tracer.cc:145-170)
__global__ __launch_bounds__(kBlock) void k_shade_camera(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.n_paths) return;  // the camera queue is the identity
	const ymx_hit h = a.st.ext_hits[i];
	a.st.cam_hit[i] = h;
	if(h.prim == YMX_NO_HIT)
	{
		// TiledIntegrator::background (integrator_tiled.cc:693-706)
		if(c.has_background) a.st.col[i] = make_float4(c.background.r, c.background.g, c.background.b, 1.f);
		return;
	}
	const float4 d4 = a.st.cam_dir[i];
	const V3 dir = v3(d4.x, d4.y, d4.z);
	Surf sp;
	make_surf(a, c.cam_from, dir, h, sp);
	const V3 wo = neg(dir);
	Rgb col = rgb(0.f, 0.f, 0.f);
	if(sp.flags & kEmit) col = cadd(col, a.mats[sp.mat].emit_color);
	a.st.col[i] = make_float4(col.r, col.g, col.b, 1.f);
	uint32_t f = 0u;
	if(sp.flags & kDiffuse)
	{
		const uint32_t offset = a.st.offset[i];
		const uint32_t sample = sample_of(a, i);
		uint32_t slot = i * a.st.slots;
		for(uint32_t l = 0; l < c.n_lights; ++l)
		{
			queue_light(a, sp, wo, a.lights[l], l, sample, offset, slot);
			slot += light_slots(a.lights[l]);
		}
		f |= kDirect;
	}
	if(sp.flags & kDiffuse) f |= kPathBlock;  // path_flags = Diffuse (no_recursive = false)
	a.st.flags[i] = f;
}

__global__ __launch_bounds__(kBlock) void k_resolve_camera(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i == 0)
	{
		// consumed by the launches before this one; the next enqueues come after it
		a.st.totals[2] += *a.st.sh_count;
		*a.st.sh_count = 0u;
		a.st.ext_count[0] = 0u;
		a.st.ext_count[1] = 0u;
	}
	if(i >= a.n_paths) return;
	if(!(a.st.flags[i] & kDirect)) return;
	Rgb col_all = rgb(0.f, 0.f, 0.f);
	uint32_t slot = i * a.st.slots;
	for(uint32_t l = 0; l < c.n_lights; ++l)
	{
		col_all = cadd(col_all, resolve_light(a, a.lights[l], slot));
		slot += light_slots(a.lights[l]);
	}
	const float4 c4 = a.st.col[i];
	const Rgb col = cadd(rgb(c4.x, c4.y, c4.z), col_all);
	a.st.col[i] = make_float4(col.r, col.g, col.b, c4.w);
	a.st.flags[i] &= ~kDirect;
}

// First bounce of sub-path a.sub (integrator_path_tracer.cc:183-203)
__global__ __launch_bounds__(kBlock) void k_subpath_init(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.n_paths) return;
	uint32_t f = a.st.flags[i];
	f &= ~(kActive | kAddEmit | kCaustic | kDirect | kVertex | kSampled);
	a.st.flags[i] = f;
	if(!(f & kPathBlock)) return;
	const ymx_hit h = a.st.cam_hit[i];
	const float4 d4 = a.st.cam_dir[i];
	const V3 dir = v3(d4.x, d4.y, d4.z);
	Surf sp;
	make_surf(a, c.cam_from, dir, h, sp);
	const V3 wo = neg(dir);
	const uint32_t offs = (uint32_t)c.path_samples * sample_of(a, i) + a.st.offset[i] + (uint32_t)a.sub;
	BSample s;
	s.s1 = ri_vdc(offs, 0);
	s.s2 = (float)low_discrepancy(a, 2, offs);
	s.flags = kDiffuse | kDiffuse | kReflect | kTransmit;
	s.pdf = 0.f;
	s.sampled = kNone;
	float w = 0.f;
	V3 pdir = v3(0.f, 0.f, 0.f);
	Rgb scol = bsdf_sample(a, sp, wo, pdir, s, w);
	scol = cscale(scol, w);
	a.st.tp_next[i] = make_float4(scol.r, scol.g, scol.b, 0.f);
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
	const uint32_t pixel = pix_of(a, i);  // camera_res_x * y + x
	const uint32_t sample = sample_of(a, i);
	const uint32_t offset = a.st.offset[i];
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
			const float random_value = rr_random(pixel, sample * (uint32_t)c.path_samples + (uint32_t)a.sub,
			                                     (uint32_t)d, c.seed);
			const float probability = cmaximum(tp);
			if(probability <= 0.f || probability < random_value)
			{
				a.st.flags[i] = f;
				return;
			}
			tp = cscale(tp, 1.f / probability);
		}
	}
	a.st.tp_vertex[i] = make_float4(tp.r, tp.g, tp.b, 0.f);
	a.st.vert_mat[i] = sp.mat;
	f |= kVertex;
	const bool need_direct = (d == 0) || (first_flags & kDiffuse);
	if(need_direct && c.n_lights > 0)
	{
		// estimateOneDirectLight (integrator_montecarlo.cc:108-116) with the
		// deterministic Halton base-2 light-pick index
		const uint32_t corr = (pixel * (uint32_t)c.aa_samples + sample) *
		                          ((uint32_t)c.path_samples * (uint32_t)c.bounces) +
		                      (uint32_t)a.sub * (uint32_t)c.bounces + (uint32_t)d + 1u;
		const float hv = halton_next(2, corr - 1u);
		int lnum = (int)(hv * (float)c.n_lights);
		if(lnum > (int)c.n_lights - 1) lnum = (int)c.n_lights - 1;
		queue_light(a, sp, pwo, a.lights[lnum], (uint32_t)lnum, sample, offset, i * a.st.slots);
		a.st.light[i] = lnum;
		f |= kDirect;
	}
	if(d == 0)
	{
		if(first_flags & kEmit) f |= kAddEmit;
	}
	else if((first_flags & kEmit) && (f & kCaustic)) f |= kAddEmit;
	// next segment (loop iteration depth = d + 1)
	if(d + 1 < c.bounces)
	{
		const uint32_t offs = (uint32_t)c.path_samples * sample + offset + (uint32_t)a.sub;
		const int d_4 = 4 * (d + 1);
		BSample s;
		s.s1 = (float)low_discrepancy(a, d_4 + 3, offs);
		s.s2 = (float)low_discrepancy(a, d_4 + 4, offs);
		s.flags = kAll;
		s.pdf = 0.f;
		s.sampled = kNone;
		float w = 0.f;
		V3 ndir = v3(0.f, 0.f, 0.f);
		Rgb scol = bsdf_sample(a, sp, pwo, ndir, s, w);
		scol = cscale(scol, w);
		if(!cblack(scol))
		{
			const Rgb tpn2 = cmul(tp, scol);
			a.st.tp_next[i] = make_float4(tpn2.r, tpn2.g, tpn2.b, 0.f);
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
	if(i == 0)
	{
		a.st.totals[2] += *a.st.sh_count;
		a.st.totals[1] += a.st.ext_count[a.depth & 1];
		*a.st.sh_count = 0u;
		a.st.ext_count[a.depth & 1] = 0u;  // this segment's queue, consumed by shade_path
	}
	if(i >= a.n_paths) return;
	const uint32_t f = a.st.flags[i];
	if(!(f & kVertex)) return;
	Rgb lcol = rgb(0.f, 0.f, 0.f);
	if(f & kDirect)
	{
		const int lnum = a.st.light[i];
		lcol = cscale(resolve_light(a, a.lights[lnum], i * a.st.slots), (float)c.n_lights);
	}
	if(f & kAddEmit) lcol = cadd(lcol, a.mats[a.st.vert_mat[i]].emit_color);
	const float4 t4 = a.st.tp_vertex[i];
	const float4 p4 = a.st.path_col[i];
	const Rgb pc = cadd(rgb(p4.x, p4.y, p4.z), cmul(lcol, rgb(t4.x, t4.y, t4.z)));
	a.st.path_col[i] = make_float4(pc.r, pc.g, pc.b, 0.f);
	a.st.flags[i] = f & ~(kDirect | kVertex | kAddEmit);
}

__global__ __launch_bounds__(kBlock) void k_finish(PtArgs a)
{
	const PtConst &c = *a.c;
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if(i >= a.n_paths) return;
	if(!(a.st.flags[i] & kPathBlock)) return;
	const int n_samples = c.path_samples > 1 ? c.path_samples : 1;
	const float4 p4 = a.st.path_col[i];
	const float4 c4 = a.st.col[i];
	const Rgb col = cadd(rgb(c4.x, c4.y, c4.z), cdiv(rgb(p4.x, p4.y, p4.z), (float)n_samples));
	a.st.col[i] = make_float4(col.r, col.g, col.b, c4.w);
}

// roundToIntWithBias, include/render/imagefilm.h:344-348
__device__ __forceinline__ int round_bias(double v) { return (int)(v + (.5 - 1.4e-11)); }

// ImageFilm::addSample (src/render/imagefilm.cc:771-822) as a gather: each
// film pixel adds the samples whose footprint covers it, visiting the source
// pixels in row-major order -- the order a row-major scatter adds them.
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
			const uint32_t s = (uint32_t)j * a.npix + (uint32_t)(sy * c.resx + sx);
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
			const float4 col = a.st.col[s];
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

inline unsigned blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace
}  // namespace ymx

int ymx_pt_trace_pass(ymx_pt *pt, int32_t first_sample, int32_t count, hipStream_t stream)
{
	using namespace ymx;
	const PtConst &hc = pt->host_const;
	const uint32_t npix = (uint32_t)(hc.resx * hc.resy);
	const uint32_t n = npix * (uint32_t)count;
	if(count < 1 || n > pt->st.capacity)
	{
		ymx_set_error("ymx_pt_trace_pass: %d samples exceed the wavefront capacity", count);
		return YMX_ERR_INVALID_ARG;
	}
	PtArgs a;
	a.c = pt->d_const;
	a.mats = pt->d_materials;
	a.lights = pt->d_lights;
	a.prim = pt->d_prim;
	a.faure = pt->d_faure;
	a.inv_prims = pt->d_inv_prims;
	a.prims = pt->d_prims;
	a.st = pt->st;
	a.film = pt->d_film;
	a.film_w = pt->d_film_w;
	a.first_sample = first_sample;
	a.npix = npix;
	a.n_paths = n;
	a.count = count;
	a.sub = 0;
	a.depth = 0;
	const unsigned g = blocks_for(n);
	const uint32_t slot_cap = n * pt->st.slots;
	k_camera<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	int rc = ymx_launch_closest(pt->accel, pt->st.ext_rays[0], pt->st.ext_hits, n, stream, nullptr);
	if(rc != YMX_OK) return rc;
	k_shade_camera<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	if(hc.n_lights > 0)
	{
		rc = ymx_launch_shadow_queue(pt->accel, pt->st.sh_rays, pt->st.sh_occ, pt->st.sh_count, slot_cap, stream);
		if(rc != YMX_OK) return rc;
	}
	k_resolve_camera<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
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
			                              stream);
			if(rc != YMX_OK) return rc;
			k_shade_path<<<g, kBlock, 0, stream>>>(a);
			YMX_HIP_CHECK(hipGetLastError());
			if(hc.n_lights > 0)
			{
				rc = ymx_launch_shadow_queue(pt->accel, pt->st.sh_rays, pt->st.sh_occ, pt->st.sh_count, slot_cap, stream);
				if(rc != YMX_OK) return rc;
			}
			k_resolve_path<<<g, kBlock, 0, stream>>>(a);
			YMX_HIP_CHECK(hipGetLastError());
		}
	}
	k_finish<<<g, kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	k_film<<<blocks_for(npix), kBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}
