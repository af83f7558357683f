/*
 * ora_pt.c -- CPU restatement of the reference's path tracer for the C0
 * feature set (TEST INFRASTRUCTURE ONLY; see ora_pt.h for the scope and the
 * deterministic replacements).  Each function names the reference lines it
 * follows; FP32 expressions keep the reference's evaluation order.
 */
#include "ora_pt.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* small vector / colour helpers with the reference's operation order   */

typedef struct { float x, y, z; } v3;
typedef struct { float r, g, b; } rgb;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(float f, v3 a) { return V(f * a.x, f * a.y, f * a.z); }  /* vector.h:175-187 */
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
	return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* Vec::normalize, include/geometry/vector.h:306-315 */
static inline v3 vnorm(v3 a)
{
	const float l2 = vdot(a, a);
	if(l2 != 0.f)
	{
		const float inv = 1.f / sqrtf(l2);
		a.x *= inv; a.y *= inv; a.z *= inv;
	}
	return a;
}
static inline v3 vload(const float *p) { return V(p[0], p[1], p[2]); }

static inline rgb C(float r, float g, float b) { rgb c = {r, g, b}; return c; }
static inline rgb cadd(rgb a, rgb b) { return C(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline rgb cmul(rgb a, rgb b) { return C(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline rgb cscale(rgb a, float f) { return C(f * a.r, f * a.g, f * a.b); }   /* color.h:250-253 */
static inline rgb cdiv(rgb a, float f) { return C(a.r / f, a.g / f, a.b / f); }     /* color.h:255-258 */
static inline float cmax(rgb a) { float m = a.g > a.b ? a.g : a.b; return a.r > m ? a.r : m; }  /* color.h:88 */
static inline int cblack(rgb a) { return a.r == 0.f && a.g == 0.f && a.b == 0.f; }

static const float K_PI = (float)3.1415926535897932384626433832795L;
static const float K_2PI = (float)6.283185307179586476925286766559L;
static const float K_1_PI = (float)0.31830988618379067153776752674503L;

/* ------------------------------------------------------------------ */
/* samplers                                                             */

/* include/sampler/sample.h:100-108 */
float ora_ri_vdc(uint32_t bits, uint32_t r)
{
	bits = (bits << 16) | (bits >> 16);
	bits = ((bits & 0x00ff00ffu) << 8) | ((bits & 0xff00ff00u) >> 8);
	bits = ((bits & 0x0f0f0f0fu) << 4) | ((bits & 0xf0f0f0f0u) >> 4);
	bits = ((bits & 0x33333333u) << 2) | ((bits & 0xccccccccu) >> 2);
	bits = ((bits & 0x55555555u) << 1) | ((bits & 0xaaaaaaaau) >> 1);
	const float v = (float)((double)(bits ^ r) * 0.00000000023283064365386962890625);
	return v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
}

/* include/sampler/sample.h:110-115 */
float ora_ri_s(uint32_t i, uint32_t r)
{
	for(uint32_t v = 1u << 31; i; i >>= 1, v ^= v >> 1)
		if(i & 1u) r ^= v;
	const float f = (float)((double)r * 0.00000000023283064365386962890625);
	return f < 0.f ? 0.f : (f > 1.f ? 1.f : f);
}

/* include/sampler/sample.h:117-122 */
float ora_ri_lp(uint32_t i, uint32_t r)
{
	for(uint32_t v = 1u << 31; i; i >>= 1, v |= v >> 1)
		if(i & 1u) r ^= v;
	const float f = (float)((double)r * 0.00000000023283064365386962890625);
	return f < 0.f ? 0.f : (f > 1.f ? 1.f : f);
}

/* include/sampler/sample.h:130-146 (FNV-1a over the 4 little-endian bytes) */
uint32_t ora_fnv32a(uint32_t value)
{
	uint32_t hash = 0x811c9dc5u;
	for(int k = 0; k < 4; ++k)
	{
		hash ^= (value >> (8 * k)) & 0xffu;
		hash *= 0x01000193u;
	}
	return hash;
}

/* Faure's scrambling permutations (H. Faure, "Good permutations for extreme
 * discrepancy", J. Number Theory 42, 1992): sigma_2 = (0 1); for b = 2c,
 * sigma_b(i) = 2 sigma_c(i), sigma_b(c + i) = 2 sigma_c(i) + 1; for b = 2c + 1,
 * sigma_b(c) = c and the other entries are sigma_{2c} with values >= c shifted up
 * by one.  src/sampler/halton.cc tabulates the prime bases up to 229. */
#define FAURE_MAX_BASE 230
static int g_faure[FAURE_MAX_BASE][FAURE_MAX_BASE];
static pthread_once_t g_faure_once = PTHREAD_ONCE_INIT;

static void faure_init(void)
{
	g_faure[1][0] = 0;
	g_faure[2][0] = 0;
	g_faure[2][1] = 1;
	for(int b = 3; b < FAURE_MAX_BASE; ++b)
	{
		if(b % 2 == 0)
		{
			const int c = b / 2;
			for(int i = 0; i < c; ++i)
			{
				g_faure[b][i] = 2 * g_faure[c][i];
				g_faure[b][c + i] = 2 * g_faure[c][i] + 1;
			}
		}
		else
		{
			const int c = (b - 1) / 2;
			for(int i = 0; i < b; ++i)
			{
				if(i == c) { g_faure[b][i] = c; continue; }
				const int s = g_faure[b - 1][i < c ? i : i - 1];
				g_faure[b][i] = s >= c ? s + 1 : s;
			}
		}
	}
}

int ora_faure_perm(int base, int digit)
{
	pthread_once(&g_faure_once, faure_init);
	return g_faure[base][digit];
}

/* faure::prims and faure::inv_prims, src/sampler/halton.cc:401-416: the inverse
 * bases are 1/p rounded to nine decimals (not exact reciprocals). */
static const int k_prims[50] = {1, 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73,
                                79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131, 137, 139, 149, 151, 157, 163, 167,
                                173, 179, 181, 191, 193, 197, 199, 211, 223, 227};

static double inv_prim9(int p)
{
	/* nearest double to round(1e9 / p) / 1e9, as a 9-decimal literal parses */
	return floor(1e9 / (double)p + 0.5) / 1e9;
}

/* Halton::lowDiscrepancySampling, src/sampler/halton.cc:419-440 (dim < 50). */
double ora_low_discrepancy(int dim, uint32_t n)
{
	double value = 0.0;
	if(dim < 50 && dim >= 1)
	{
		pthread_once(&g_faure_once, faure_init);
		const uint32_t base = (uint32_t)k_prims[dim];
		const int table = dim <= 2 ? 3 : (int)base;  /* all_tables[0..2] are base_3 */
		const double f = inv_prim9(k_prims[dim]);
		double dn = (double)n;
		double factor = f;
		while(n > 0)
		{
			value += (double)g_faure[table][n % base] * factor;
			dn *= f;
			n = (uint32_t)dn;
			factor *= f;
		}
	}
	return value;
}

/* Halton(base, start).getNext(), include/sampler/halton.h:54-80 */
typedef struct { uint32_t base; double inv_base; double value; } halton;

static void halton_start(halton *h, int base, uint32_t start)
{
	h->base = (uint32_t)base;
	h->inv_base = 1.0 / (double)base;
	double factor = h->inv_base;
	h->value = 0.0;
	while(start > 0)
	{
		h->value += (double)(start % h->base) * factor;
		start /= h->base;
		factor *= h->inv_base;
	}
}

static float halton_next(halton *h)
{
	const double r = 0.9999999999 - h->value;
	if(h->inv_base < r) h->value += h->inv_base;
	else
	{
		double hh = 0.0, hv = h->inv_base;
		while(hv >= r)
		{
			hh = hv;
			hv *= h->inv_base;
		}
		h->value += hh + hv - 1.0;
	}
	const float v = (float)h->value;
	return v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
}

float ora_halton_next(int base, uint32_t start)
{
	halton h;
	halton_start(&h, base, start);
	return halton_next(&h);
}

/* Counter-based replacement of the Russian-roulette generator. */
float ora_rr_random(uint32_t pixel, uint32_t sample, uint32_t depth, uint32_t seed)
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

/* math::filter::gauss with FAST_MATH exp (include/math/filter.h:32-37) and the
 * table of src/render/imagefilm.cc:231-264. */
void ora_filter_table(float pixelwidth, float table[256], float *table_scale, float *filter_width)
{
	float fw = pixelwidth * 0.5f;
	fw *= 2.f;                                   /* Gauss */
	if(fw < 0.501f) fw = 0.501f;
	if(fw > 0.5f * 8.f) fw = 0.5f * 8.f;         /* max_filter_size_ = 8 */
	const float scale = 1.f / 16.f;
	int k = 0;
	for(int y = 0; y < 16; ++y)
	{
		for(int x = 0; x < 16; ++x)
		{
			const float dx = ((float)x + .5f) * scale, dy = ((float)y + .5f) * scale;
			const float r_2 = dx * dx + dy * dy;
			const float e = ora_fast_exp2((float)1.4426950408889634073599246810019L * ((float)-6 * r_2));
			const float v = e - 0.00247875f;
			table[k++] = v > 0.f ? v : 0.f;
		}
	}
	*table_scale = 0.9999f * 16.f / fw;
	*filter_width = fw;
}

/* ------------------------------------------------------------------ */
/* scene access                                                          */

typedef struct
{
	const ora_kdtree *kd;
	const float *verts;
	const uint32_t *idx;
	const uint32_t *mat_id;
	const ora_material *mats;
	const uint8_t *vis;
	const ora_light *lights;
	uint32_t n_lights;
	struct obj_light *objs;   /* [n_lights], set up for object and portal lights */
	uint8_t *transparent;     /* per triangle, with transparent shadows: its material isTransparent() */
	const ora_pt_settings *st;
	/* camera (src/camera/camera.cc:89-110, camera_perspective.cc:97-111) */
	v3 cam_from, vright, vup, vto;
	uint32_t total_samples;   /* the most AA samples a pixel can receive (all passes) */
	float light_mult;         /* the AA pass's light sample multiplier */
	int pass_samples;         /* AA samples of the pass being rendered */
} pt_ctx;

enum
{
	F_NONE = 0, F_SPECULAR = 1, F_GLOSSY = 2, F_DIFFUSE = 4, F_DISPERSIVE = 8, F_REFLECT = 16, F_TRANSMIT = 32,
	F_FILTER = 64, F_EMIT = 128, F_VOLUMETRIC = 256, F_ALL = 127
}; /* include/material/bsdf.h:25-67 (All excludes Emit and Volumetric) */

/* RayDivision, include/render/render_data.h:25-31 */
typedef struct { int division, offset; float dec1, dec2; } ray_div;

typedef struct
{
	v3 p, n, ng;
	v3 u, v;     /* uvn_ */
	uint32_t prim;
	const ora_material *mat;
	uint32_t flags;
} surf;

static uint32_t mat_flags(const ora_material *m)
{
	uint32_t f = 0;
	switch(m->type)
	{
		case ORA_MAT_GLOSSY:  /* GlossyMaterial::GlossyMaterial, material_glossy.cc:130-140 */
			if(m->diffuse_reflect > 0.f) f = F_DIFFUSE | F_REFLECT;
			f |= m->as_diffuse ? (F_DIFFUSE | F_REFLECT) : (F_GLOSSY | F_REFLECT);
			break;
		case ORA_MAT_MIRROR: f = F_SPECULAR; break;  /* material_mirror.cc:64-68 */
		case ORA_MAT_GLASS:  /* AllSpecular (+ Filter), material_glass.cc:157-161 */
			f = F_SPECULAR | F_REFLECT | F_TRANSMIT;
			if(m->fake_shadows) f |= F_FILTER;
			break;
		case ORA_MAT_LIGHT: f = F_EMIT; break;  /* material_light.cc:69-73 */
		default:
			if(m->diffuse_reflect > 0.00001f) f |= F_DIFFUSE | F_REFLECT;  /* material_shiny_diffuse.cc:175-183 */
			if(m->emit > 0.f) f |= F_EMIT;                                 /* :134 */
	}
	return f;
}

/* Vec3f::createCoordsSystem, include/geometry/vector.h:399-421 */
static void coords_system(v3 n, v3 *u, v3 *v)
{
	if(n.x == 0.f && n.y == 0.f)
	{
		*u = n.z < 0.f ? V(-1.f, 0.f, 0.f) : V(1.f, 0.f, 0.f);
		*v = V(0.f, 1.f, 0.f);
	}
	else
	{
		const float d = 1.f / sqrtf(n.y * n.y + n.x * n.x);
		*u = V(n.y * d, -n.x * d, 0.f);
		*v = vcross(n, *u);
	}
}

/* Accelerator::intersect(ray, camera) + PrimitivePolygon::getSurfacePolygon
 * (include/accelerator/accelerator.h:89-101, src/geometry/primitive/primitive_polygon.cc:37-115) */
static int intersect_surface(const pt_ctx *c, v3 from, v3 dir, float tmin, float tmax, surf *sp)
{
	ora_ray r = {{from.x, from.y, from.z}, tmin, {dir.x, dir.y, dir.z}, tmax};
	const float t_max = (tmax >= 0.f) ? tmax : FLT_MAX;
	ora_hit h;
	if(c->st->trace_brute)
	{
		/* every triangle in id order: exact ties in t keep the lowest id */
		ora_closest_brute(c->verts, c->idx, c->vis, ora_kd_size(c->kd), &r, 1, &h, 1);
		if(h.prim == ORA_NO_HIT) return 0;
	}
	else if(!ora_kd_intersect1(c->kd, &r, t_max, 0, &h) || h.prim == ORA_NO_HIT) return 0;
	sp->p = vadd(from, vmul(h.t, dir));
	const uint32_t *ix = c->idx + 3 * (size_t)h.prim;
	const v3 p0 = vload(c->verts + 3 * (size_t)ix[0]);
	const v3 p1 = vload(c->verts + 3 * (size_t)ix[1]);
	const v3 p2 = vload(c->verts + 3 * (size_t)ix[2]);
	sp->ng = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));   /* shape_polygon.h:58-62 */
	sp->n = sp->ng;
	const float *cn = c->st->corner_normals ? c->st->corner_normals + 12 * (size_t)h.prim : NULL;
	if(cn && cn[3] >= 0.f)
	{
		/* a smooth mesh or one with vertex normals (primitive_polygon.cc:45-51):
		 * the corners' normals (the face normal where a corner has none,
		 * FacePrimitive::getVertexNormal) weighted by the barycentric
		 * coordinates (1 - u - v, u, v) of the hit, then normalised */
		v3 vn[3];
		for(int k = 0; k < 3; ++k) vn[k] = cn[4 * k + 3] > 0.f ? vload(cn + 4 * k) : sp->ng;
		const float bu = 1.f - h.u - h.v, bv = h.u, bw = h.v;
		sp->n = vnorm(vadd(vadd(vmul(bu, vn[0]), vmul(bv, vn[1])), vmul(bw, vn[2])));
	}
	coords_system(sp->n, &sp->u, &sp->v);
	sp->prim = h.prim;
	sp->mat = &c->mats[c->mat_id[h.prim]];
	sp->flags = mat_flags(sp->mat);
	return 1;
}

/* Accelerator::isShadowed, include/accelerator/accelerator.h:103-111 */
static int is_shadowed(const pt_ctx *c, v3 from, v3 dir, float tmin, float tmax)
{
	const v3 f = vadd(from, vmul(tmin, dir));
	const float t_max = (tmax >= 0.f) ? tmax - 2 * tmin : FLT_MAX;
	ora_ray r = {{f.x, f.y, f.z}, tmin, {dir.x, dir.y, dir.z}, tmax};
	ora_hit h;
	return ora_kd_intersect1(c->kd, &r, t_max, 1, &h);
}

/* SurfacePoint::normalFaceForward, include/geometry/surface.h:202-205 */
static inline v3 face_forward(v3 ng, v3 n, v3 wo) { return vdot(ng, wo) < 0.f ? vneg(n) : n; }

/* ShinyDiffuseMaterial::eval (diffuse only), material_shiny_diffuse.cc:297-340 */
static rgb sd_eval(const surf *sp, v3 wo, v3 wl, uint32_t bsdfs)
{
	const float cos_ng_wo = vdot(sp->ng, wo);
	const float cos_ng_wl = vdot(sp->ng, wl);
	const v3 n = face_forward(sp->ng, sp->n, wo);
	if(!(bsdfs & (sp->flags & F_DIFFUSE))) return C(0.f, 0.f, 0.f);
	(void)cos_ng_wl;
	(void)cos_ng_wo;
	/* kr = 1 (no Fresnel), m_t = 1, no translucency */
	if(vdot(n, wl) < 0.0) return C(0.f, 0.f, 0.f);
	const float m_d = 1.f * (1.f - 0.f) * sp->mat->diffuse_reflect;
	return cscale(C(sp->mat->color[0], sp->mat->color[1], sp->mat->color[2]), m_d);
}

/* ShinyDiffuseMaterial::pdf, material_shiny_diffuse.cc:441-483 */
static float sd_pdf(const surf *sp, v3 wo, v3 wi, uint32_t bsdfs)
{
	if(!(bsdfs & F_DIFFUSE)) return 0.f;
	const v3 n = face_forward(sp->ng, sp->n, wo);
	if(!(sp->flags & F_DIFFUSE)) return 0.f;
	float pdf = 0.f, sum = 0.f;
	int n_match = 0;
	if(bsdfs & (F_DIFFUSE | F_REFLECT))
	{
		const float width = sp->mat->diffuse_reflect;  /* accumulate(): accum[3] */
		sum += width;
		pdf += fabsf(vdot(wi, n)) * width;
		++n_match;
	}
	if(!n_match || sum < 0.00001) return 0.f;
	return pdf / sum;
}

/* sample::cosHemisphere, include/sampler/sample.h:43-52 */
static v3 cos_hemisphere(v3 n, v3 ru, v3 rv, float s1, float s2)
{
	if(s1 >= 1.0f) return n;
	const float z1 = s1;
	const float z2 = s2 * K_2PI;
	return vadd(vmul(sqrtf(1.f - z1), vadd(vmul(ora_fast_cos(z2), ru), vmul(ora_fast_sin(z2), rv))),
	            vmul(sqrtf(z1), n));
}

typedef struct { float s1, s2; uint32_t flags; uint32_t sampled; float pdf; } bsample;

/* ShinyDiffuseMaterial::sample (diffuse component), material_shiny_diffuse.cc:349-439 */
static rgb sd_sample(const surf *sp, v3 wo, v3 *wi, bsample *s, float *w)
{
	const float cos_ng_wo = vdot(sp->ng, wo);
	const v3 n = face_forward(sp->ng, sp->n, wo);
	const uint32_t cflag = F_DIFFUSE | F_REFLECT;
	float sum = 0.f, width = 0.f, val = 0.f;
	int n_match = 0;
	if((sp->flags & F_DIFFUSE) && (s->flags & cflag) == cflag)
	{
		width = sp->mat->diffuse_reflect;
		sum += width;
		val = sum;
		++n_match;
	}
	if(!n_match || sum < 0.00001)
	{
		s->sampled = F_NONE;
		s->pdf = 0.f;
		return C(1.f, 1.f, 1.f);
	}
	const float inv_sum = 1.f / sum;
	val *= inv_sum;
	width *= inv_sum;
	(void)val;
	const float s1 = s->s1 / width;
	rgb scolor = C(0.f, 0.f, 0.f);
	*wi = cos_hemisphere(n, sp->u, sp->v, s1, s->s2);
	if(cos_ng_wo * vdot(sp->ng, *wi) > 0)
		scolor = cscale(C(sp->mat->color[0], sp->mat->color[1], sp->mat->color[2]), sp->mat->diffuse_reflect);
	s->pdf = fabsf(vdot(*wi, n)) * width;
	s->sampled = cflag;
	*w = fabsf(vdot(*wi, sp->n)) / (s->pdf * 0.99f + 0.01f);
	/* getAlpha() == 1 for a non-transparent shinydiffuse: w * 1 + 1 * 0 */
	*w = *w * 1.f + 1.f * (1.f - 1.f);
	return scolor;
}

/* ---- microfacet helpers, include/material/material_utils_microfacet.h ---- */

static inline float std_max(float a, float b) { return (a < b) ? b : a; }  /* std::max */
static inline float std_min(float a, float b) { return (b < a) ? b : a; }  /* std::min */

/* pdfDivisor :26-29 and asDivisor :31-34 */
static inline float pdf_divisor(float cos) { return 8.f * K_PI * (cos * 0.99f + 0.04f); }
static inline float as_divisor(float cos_1, float cos_i, float cos_o)
{
	return 8.f * K_PI * ((cos_1 * std_max(cos_i, cos_o)) * 0.99f + 0.04f);
}
/* blinnD :83-86, blinnPdf :88-91, blinnSample :93-101 */
static inline float blinn_d(float cos_h, float e) { return (e + 1.f) * ora_fast_pow(cos_h, e); }
static inline float blinn_pdf(float costheta, float cos_w_h, float e) { return blinn_d(costheta, e) / pdf_divisor(cos_w_h); }
static v3 blinn_sample(float s_1, float s_2, float exponent)
{
	const float cos_theta = ora_fast_pow(1.f - s_2, 1.f / (exponent + 1.f));
	const float sin_theta = sqrtf(1.f - cos_theta * cos_theta);
	const float phi = s_1 * K_2PI;
	return V(sin_theta * ora_fast_cos(phi), sin_theta * ora_fast_sin(phi), cos_theta);
}
/* schlickFresnel :181-186 */
static inline float schlick_fresnel(float costheta, float r)
{
	const float c_1 = (1.f - costheta);
	const float c_2 = c_1 * c_1;
	return r + ((1.f - r) * c_1 * c_2 * c_2);
}
/* diffuseReflect :188-198: the scale is evaluated in double (diffuse_ratio is a
 * double constant) and rounded to float before multiplying the colour. */
static rgb diffuse_reflect(float wi_n, float wo_n, float glossy, float diffuse, rgb diff_base)
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

/* Vec3f::reflect (in place, include/geometry/vector.h:357-362) and
 * Vec3f::reflectDir (:390-396) */
static inline v3 reflect_about(v3 v, v3 normal)
{
	const float vn = vdot(vmul(2.f, v), normal);
	return V(vn * normal.x - v.x, vn * normal.y - v.y, vn * normal.z - v.z);
}
static inline v3 reflect_dir(v3 normal, v3 vec)
{
	const float vn = vdot(vec, normal);
	if(vn < 0.f) return vneg(vec);
	return vsub(vmul(2.f * vn, normal), vec);
}

/* ---- GlossyMaterial (isotropic Blinn, Lambertian diffuse), material_glossy.cc ---- */

static inline rgb mat_rgb(const float c[3]) { return C(c[0], c[1], c[2]); }

/* initBsdf :142-150: p_diffuse = min(0.6, 1 - g / (g + (1 - g) d)) */
static float glossy_p_diffuse(const ora_material *m)
{
	const float m_glossy = m->glossy_reflect, m_diffuse = m->diffuse_reflect;
	return std_min(0.6f, 1.f - (m_glossy / (m_glossy + (1.f - m_glossy) * m_diffuse)));
}

/* GlossyMaterial::eval :200-250 */
static rgb glossy_eval(const surf *sp, v3 wo, v3 wl, uint32_t bsdfs)
{
	const ora_material *m = sp->mat;
	if(!(bsdfs & F_DIFFUSE) || ((vdot(sp->ng, wl) * vdot(sp->ng, wo)) < 0.f)) return C(0.f, 0.f, 0.f);
	rgb col = C(0.f, 0.f, 0.f);
	const int diffuse_flag = (bsdfs & F_DIFFUSE) != 0;
	const v3 n = face_forward(sp->ng, sp->n, wo);
	const float wi_n = fabsf(vdot(wl, n));
	const float wo_n = fabsf(vdot(wo, n));
	if((m->as_diffuse && diffuse_flag) || (!m->as_diffuse && (bsdfs & F_GLOSSY)))
	{
		const v3 h = vnorm(vadd(wo, wl));
		const float cos_wi_h = std_max(0.f, vdot(wl, h));
		const float glossy = blinn_d(vdot(h, n), m->exponent) * schlick_fresnel(cos_wi_h, m->glossy_reflect) /
		                     as_divisor(cos_wi_h, wo_n, wi_n);
		col = cscale(mat_rgb(m->color), glossy);
	}
	if(m->diffuse_reflect > 0.f && diffuse_flag)
	{
		const rgb add_col = cscale(mat_rgb(m->diffuse_color), m->diffuse_reflect * (1.f - m->glossy_reflect));
		col = cadd(col, add_col);
	}
	return col;
}

/* GlossyMaterial::pdf :404-440 */
static float glossy_pdf(const surf *sp, v3 wo, v3 wi, uint32_t flags)
{
	const ora_material *m = sp->mat;
	if((vdot(sp->ng, wo) * vdot(sp->ng, wi)) < 0.f) return 0.f;
	const v3 n = face_forward(sp->ng, sp->n, wo);
	float pdf = 0.f;
	const float cur_p_diffuse = glossy_p_diffuse(m);
	const int use_glossy = m->as_diffuse ? (flags & F_DIFFUSE) != 0 : (flags & F_GLOSSY) != 0;
	const int use_diffuse = m->diffuse_reflect > 0.f && (flags & F_DIFFUSE);
	if(use_diffuse)
	{
		pdf = fabsf(vdot(wi, n));
		if(use_glossy)
		{
			const v3 h = vnorm(vadd(wi, wo));
			const float cos_wo_h = vdot(wo, h);
			const float cos_n_h = vdot(n, h);
			pdf = pdf * cur_p_diffuse + blinn_pdf(cos_n_h, cos_wo_h, m->exponent) * (1.f - cur_p_diffuse);
		}
		return pdf;
	}
	if(use_glossy)
	{
		const v3 h = vnorm(vadd(wi, wo));
		const float cos_wo_h = vdot(wo, h);
		const float cos_n_h = vdot(n, h);
		pdf = blinn_pdf(cos_n_h, cos_wo_h, m->exponent);
	}
	return pdf;
}

/* GlossyMaterial::sample :252-402.  Early returns leave w as the caller had it. */
static rgb glossy_sample(const surf *sp, v3 wo, v3 *wi, bsample *s, float *w)
{
	const ora_material *m = sp->mat;
	const float cos_ng_wo = vdot(sp->ng, wo);
	const v3 n = face_forward(sp->ng, sp->n, wo);
	s->pdf = 0.f;
	float wi_n = 0.f;
	const float wo_n = fabsf(vdot(wo, n));
	rgb scolor = C(0.f, 0.f, 0.f);
	float s_1 = s->s1;
	const float cur_p_diffuse = glossy_p_diffuse(m);
	const int use_glossy = m->as_diffuse ? (s->flags & F_DIFFUSE) != 0 : (s->flags & F_GLOSSY) != 0;
	const int use_diffuse = m->diffuse_reflect > 0.f && (s->flags & F_DIFFUSE);
	const rgb gcolor = mat_rgb(m->color), dcolor = mat_rgb(m->diffuse_color);
	if(use_diffuse)
	{
		const float s_p_diffuse = use_glossy ? cur_p_diffuse : 1.f;
		if(s_1 < s_p_diffuse)
		{
			s_1 /= s_p_diffuse;
			*wi = cos_hemisphere(n, sp->u, sp->v, s_1, s->s2);
			const float cos_ng_wi = vdot(sp->ng, *wi);
			if(cos_ng_wi * cos_ng_wo < 0.f) return scolor;
			wi_n = fabsf(vdot(*wi, n));
			s->pdf = wi_n;
			float glossy = 0.f;
			if(use_glossy)
			{
				const v3 h = vnorm(vadd(*wi, wo));
				const float cos_wo_h = vdot(wo, h);
				const float cos_wi_h = fabsf(vdot(*wi, h));
				const float cos_n_h = vdot(n, h);
				s->pdf = s->pdf * cur_p_diffuse + blinn_pdf(cos_n_h, cos_wo_h, m->exponent) * (1.f - cur_p_diffuse);
				glossy = blinn_d(cos_n_h, m->exponent) * schlick_fresnel(cos_wi_h, m->glossy_reflect) /
				         as_divisor(cos_wi_h, wo_n, wi_n);
			}
			s->sampled = F_DIFFUSE | F_REFLECT;
			if(!(s->flags & F_REFLECT)) return C(0.f, 0.f, 0.f);
			scolor = cscale(gcolor, glossy);
			scolor = cadd(scolor, diffuse_reflect(wi_n, wo_n, m->glossy_reflect, m->diffuse_reflect, dcolor));
			*w = wi_n / (s->pdf * 0.99f + 0.01f);
			return scolor;
		}
		s_1 -= cur_p_diffuse;
		s_1 /= (1.f - cur_p_diffuse);
	}
	if(use_glossy)
	{
		const v3 hs = blinn_sample(s_1, s->s2, m->exponent);
		v3 h = vadd(vadd(vmul(hs.x, sp->u), vmul(hs.y, sp->v)), vmul(hs.z, n));
		float cos_wo_h = vdot(wo, h);
		if(cos_wo_h < 0.f)
		{
			h = reflect_about(h, n);
			cos_wo_h = vdot(wo, h);
		}
		*wi = reflect_dir(h, wo);
		const float cos_ng_wi = vdot(sp->ng, *wi);
		if(cos_ng_wo * cos_ng_wi < 0.f) return C(0.f, 0.f, 0.f);
		wi_n = fabsf(vdot(*wi, n));
		const float cos_hn = vdot(h, n);
		s->pdf = blinn_pdf(cos_hn, cos_wo_h, m->exponent);
		const float glossy = blinn_d(cos_hn, m->exponent) * schlick_fresnel(cos_wo_h, m->glossy_reflect) /
		                     as_divisor(cos_wo_h, wo_n, wi_n);
		scolor = cscale(gcolor, glossy);
		s->sampled = m->as_diffuse ? (F_DIFFUSE | F_REFLECT) : (F_GLOSSY | F_REFLECT);
	}
	if(use_diffuse)
	{
		const rgb add_col = diffuse_reflect(wi_n, wo_n, m->glossy_reflect, m->diffuse_reflect, dcolor);
		s->pdf = wi_n * cur_p_diffuse + s->pdf * (1.f - cur_p_diffuse);
		scolor = cadd(scolor, add_col);
	}
	*w = wi_n / (s->pdf * 0.99f + 0.01f);
	return scolor;
}

/* ---- MirrorMaterial, material_mirror.cc:64-87, include/material/material_mirror.h:64 ---- */

static rgb mirror_col(const ora_material *m) { return cscale(mat_rgb(m->color), std_min(m->reflect, 1.f)); }

static rgb mirror_sample(const surf *sp, v3 wo, v3 *wi, bsample *s, float *w)
{
	*wi = reflect_dir(sp->n, wo);
	s->sampled = F_SPECULAR | F_REFLECT;
	*w = 1.f;
	return cscale(mirror_col(sp->mat), 1.f / fabsf(vdot(sp->n, *wi)));
}

/* ---- GlassMaterial without dispersion, material_glass.cc:168-378 ---- */

/* Vec3f::refract, include/geometry/vector.h:484-504 */
static int refract(v3 n, v3 wi, v3 *wo, float ior)
{
	float eta;
	float cos_v_n = vdot(wi, n);
	if(cos_v_n < 0.f)
	{
		eta = ior;
		n = vneg(n);
		cos_v_n = -cos_v_n;
	}
	else eta = 1.f / ior;
	const float k = 1.f - eta * eta * (1.f - cos_v_n * cos_v_n);
	if(k <= 0.f) return 0;
	*wo = vnorm(vadd(vmul(eta, vneg(wi)), vmul(eta * cos_v_n - sqrtf(k), n)));
	return 1;
}

/* Vec3f::fresnel, include/geometry/vector.h:506-518 */
static void fresnel(v3 i, v3 n, float ior, float *kr_out, float *kt_out)
{
	const int negative = vdot(i, n) < 0.f;
	const float c = vdot(i, negative ? vneg(n) : n);
	float g = ior * ior + c * c - 1.f;
	if(g <= 0.f) g = 0.f;
	else g = sqrtf(g);
	const float aux = c * (g + c);
	const float kr = ((0.5f * (g - c) * (g - c)) / ((g + c) * (g + c))) *
	                 (1.f + ((aux - 1.f) * (aux - 1.f)) / ((aux + 1.f) * (aux + 1.f)));
	*kr_out = kr;
	*kt_out = (kr < 1.f) ? 1.f - kr : 0.f;
}

/* the normal bent towards wo (material_glass.cc:172-174, :307-317) */
static v3 glass_normal(const surf *sp, v3 wo, int *outside)
{
	*outside = vdot(sp->ng, wo) > 0.f;
	const float cos_wo_n = vdot(sp->n, wo);
	const int keep = *outside ? cos_wo_n >= 0.f : cos_wo_n <= 0.f;
	return keep ? sp->n : vnorm(vsub(sp->n, vmul(1.00001f * cos_wo_n, wo)));
}

/* filter_color_ = transmit_filter * filter_color + (1 - transmit_filter), material_glass.h:92 */
static rgb glass_filter(const ora_material *m)
{
	const float tf = m->transmit_filter;
	return C(tf * m->filter_color[0] + (1.f - tf), tf * m->filter_color[1] + (1.f - tf),
	         tf * m->filter_color[2] + (1.f - tf));
}

static rgb glass_sample(const surf *sp, v3 wo, v3 *wi, bsample *s, float *w)
{
	const ora_material *m = sp->mat;
	if(!(s->flags & F_SPECULAR))  /* no dispersion: the chromatic branch never applies */
	{
		s->pdf = 0.f;
		return C(0.f, 0.f, 0.f);
	}
	int outside;
	const v3 n = glass_normal(sp, wo, &outside);
	s->pdf = 1.f;
	const uint32_t transmit_flags = m->fake_shadows ? (F_FILTER | F_TRANSMIT) : (F_SPECULAR | F_TRANSMIT);
	v3 refdir;
	if(refract(n, wo, &refdir, m->ior))
	{
		float kr, kt;
		fresnel(wo, n, m->ior, &kr, &kt);
		const float p_kr = (float)(0.01 + 0.99 * kr), p_kt = (float)(0.01 + 0.99 * kt);
		if(s->s1 < p_kt && (s->flags & transmit_flags))
		{
			*wi = refdir;
			s->pdf = p_kt;
			s->sampled = transmit_flags;
			*w = 1.f;
			return glass_filter(m);
		}
		else if(s->flags & (F_SPECULAR | F_REFLECT))
		{
			*wi = reflect_about(wo, n);
			s->pdf = p_kr;
			s->sampled = F_SPECULAR | F_REFLECT;
			*w = 1.f;
			return mat_rgb(m->mirror_color);
		}
	}
	else if(s->flags & (F_SPECULAR | F_REFLECT))  /* total inner reflection */
	{
		*wi = reflect_about(wo, n);
		s->sampled = F_SPECULAR | F_REFLECT;
		*w = 1.f;
		return C(1.f, 1.f, 1.f);
	}
	s->pdf = 0.f;
	return C(0.f, 0.f, 0.f);
}

/* Material::getSpecular: perfect reflection / refraction directions and colours */
typedef struct { int has_r, has_t; v3 dir_r, dir_t; rgb col_r, col_t; } specular;

static specular get_specular(int ray_level, const surf *sp, v3 wo)
{
	specular sp_out;
	memset(&sp_out, 0, sizeof sp_out);
	const ora_material *m = sp->mat;
	if(m->type == ORA_MAT_MIRROR)  /* MirrorMaterial::getSpecular, material_mirror.cc:78-86 */
	{
		sp_out.has_r = 1;
		sp_out.col_r = mirror_col(m);
		sp_out.dir_r = reflect_dir(face_forward(sp->ng, sp->n, wo), wo);
		return sp_out;
	}
	if(m->type != ORA_MAT_GLASS) return sp_out;
	/* GlassMaterial::getSpecular, material_glass.cc:303-366 */
	int outside;
	const v3 n = glass_normal(sp, wo, &outside);
	v3 refdir;
	if(refract(n, wo, &refdir, m->ior))
	{
		float kr, kt;
		fresnel(wo, n, m->ior, &kr, &kt);
		sp_out.has_t = 1;
		sp_out.col_t = cscale(glass_filter(m), kt);
		sp_out.dir_t = refdir;
		if(outside || ray_level < 3)
		{
			sp_out.has_r = 1;
			sp_out.dir_r = reflect_about(wo, n);
			sp_out.col_r = cscale(mat_rgb(m->mirror_color), kr);
		}
	}
	else
	{
		sp_out.has_r = 1;
		sp_out.col_r = mat_rgb(m->mirror_color);
		sp_out.dir_r = reflect_about(wo, n);
	}
	return sp_out;
}

/* ---- dispatch over the material types ---- */

static rgb bsdf_eval(const surf *sp, v3 wo, v3 wl, uint32_t bsdfs)
{
	switch(sp->mat->type)
	{
		case ORA_MAT_SHINYDIFFUSE: return sd_eval(sp, wo, wl, bsdfs);
		case ORA_MAT_GLOSSY: return glossy_eval(sp, wo, wl, bsdfs);
		default: return C(0.f, 0.f, 0.f);  /* mirror, glass, light: eval is 0 */
	}
}

static float bsdf_pdf(const surf *sp, v3 wo, v3 wi, uint32_t bsdfs)
{
	switch(sp->mat->type)
	{
		case ORA_MAT_SHINYDIFFUSE: return sd_pdf(sp, wo, wi, bsdfs);
		case ORA_MAT_GLOSSY: return glossy_pdf(sp, wo, wi, bsdfs);
		default: return 0.f;  /* Material::pdf default */
	}
}

static rgb bsdf_sample(const surf *sp, v3 wo, v3 *wi, bsample *s, float *w)
{
	switch(sp->mat->type)
	{
		case ORA_MAT_SHINYDIFFUSE: return sd_sample(sp, wo, wi, s, w);
		case ORA_MAT_GLOSSY: return glossy_sample(sp, wo, wi, s, w);
		case ORA_MAT_MIRROR: return mirror_sample(sp, wo, wi, s, w);
		case ORA_MAT_GLASS: return glass_sample(sp, wo, wi, s, w);
		default:  /* LightMaterial::sample, material_light.cc:75-80 */
			s->pdf = 0.f;
			*w = 0.f;
			return C(0.f, 0.f, 0.f);
	}
}

/* Material::emit: shinydiffuse emit_color_ = emit * color (material_shiny_diffuse.cc:342-347);
 * LightMaterial::emit, material_light.cc:82-87 (light_col_ = color * power) */
static rgb mat_emit(const surf *sp, v3 wo)
{
	const ora_material *m = sp->mat;
	if(m->type == ORA_MAT_LIGHT)
	{
		const rgb lc = cscale(mat_rgb(m->color), m->power);
		if(m->double_sided) return lc;
		return vdot(wo, sp->n) > 0.f ? lc : C(0.f, 0.f, 0.f);
	}
	if(m->type == ORA_MAT_SHINYDIFFUSE) return cscale(mat_rgb(m->color), m->emit);
	return C(0.f, 0.f, 0.f);
}

/* GlassMaterial::getTransparency (src/material/material_glass.cc:297-304) at
 * a triangle met by a shadow ray along dir: kt of the Fresnel terms about the
 * face normal turned towards dir, times the filter colour. */
static rgb transparency(const pt_ctx *c, uint32_t prim, v3 dir)
{
	const uint32_t *ix = c->idx + 3 * (size_t)prim;
	const v3 p0 = vload(c->verts + 3 * (size_t)ix[0]);
	const v3 ng = vnorm(vcross(vsub(vload(c->verts + 3 * (size_t)ix[1]), p0), vsub(vload(c->verts + 3 * (size_t)ix[2]), p0)));
	const ora_material *m = &c->mats[c->mat_id[prim]];
	float kr, kt;
	fresnel(dir, face_forward(ng, ng, dir), m->ior, &kr, &kt);
	return cscale(glass_filter(m), kt);
}

/* Accelerator::isShadowed, or with the integrator's "transpShad"
 * isShadowedTransparentShadow (include/accelerator/accelerator.h:103-119):
 * returns 1 when shadowed; otherwise *scol is the product, in the order the
 * kd-tree meets them, of the transparencies of the transparent shadow casters
 * on the way (IntersectData::color_, starting at 1). */
static int shadow_ray(const pt_ctx *c, v3 from, v3 dir, float tmin, float tmax, rgb *scol)
{
	*scol = C(1.f, 1.f, 1.f);
	if(!c->st->transparent_shadows) return is_shadowed(c, from, dir, tmin, tmax);
	const v3 f = vadd(from, vmul(tmin, dir));
	const float t_max = (tmax >= 0.f) ? tmax - 2 * tmin : FLT_MAX;
	ora_ray r = {{f.x, f.y, f.z}, tmin, {dir.x, dir.y, dir.z}, tmax};
	uint32_t prims[ORA_MAX_SHADOW_DEPTH];
	int cnt = 0;
	if(ora_kd_transparent1(c->kd, &r, t_max, c->transparent, c->st->shadow_depth, prims, &cnt)) return 1;
	rgb col = C(1.f, 1.f, 1.f);
	for(int k = 0; k < cnt; ++k) col = cmul(col, transparency(c, prims[k], dir));
	*scol = col;
	return 0;
}

/* ------------------------------------------------------------------ */
/* lights                                                                */

static rgb light_color(const pt_ctx *c, const ora_light *l)
{
	/* BackgroundPortalLight: the background's colour (ConstantBackground::eval,
	 * src/background/background_constant.cc:63-66) times the portal's power
	 * (src/light/light_background_portal.cc:173) */
	if(l->type == ORA_LIGHT_PORTAL)
		return cscale(C(c->st->background[0], c->st->background[1], c->st->background[2]), l->power);
	const rgb col = cscale(C(l->color[0], l->color[1], l->color[2]), l->power);
	/* light_area.h:78, light_object_light.h:82, light_point.h:64 */
	return l->type == ORA_LIGHT_POINT ? col : cscale(col, K_PI);
}

static void clamp_proportional(float col[4], float max_value);

/* ObjectLight's double_sided param; a portal is one-sided */
static int two_sided(const ora_light *l) { return l->type == ORA_LIGHT_OBJECT && l->double_sided; }

static v3 area_normal_flipped(const ora_light *l)
{
	/* (point2 - corner) ^ (point1 - corner), light_area.h:76 */
	return vcross(vsub(vload(l->point2), vload(l->corner)), vsub(vload(l->point1), vload(l->corner)));
}

/* area_ = normal_flipped_.normalized().length(), light_area.h:79 */
static float area_area(const ora_light *l)
{
	const v3 nf = vnorm(area_normal_flipped(l));
	return sqrtf(vdot(nf, nf));
}

/* ShapePolygon<float, 4>::intersect, include/geometry/shape/shape_polygon.h:126-176 */
static float quad_intersect(const v3 q[4], v3 from, v3 dir)
{
	const v3 e1 = vsub(q[1], q[0]);
	const v3 e2 = vsub(q[2], q[0]);
	const v3 p2 = vcross(dir, e2);
	const float det12 = vdot(e1, p2);
	if(det12 != 0.f)
	{
		const float inv12 = 1.f / det12;
		const v3 tvec = vsub(from, q[0]);
		float u = vdot(tvec, p2) * inv12;
		if(u >= 0.f && u <= 1.f)
		{
			const v3 q1 = vcross(tvec, e1);
			const float v = vdot(dir, q1) * inv12;
			if(v >= 0.f && (u + v) <= 1.f)
			{
				const float t = vdot(e2, q1) * inv12;
				if(t > 0.f) return t;
			}
		}
		else
		{
			const v3 e3 = vsub(q[3], q[0]);
			const v3 p3 = vcross(dir, e3);
			const float det23 = vdot(e2, p3);
			if(det23 != 0.f)
			{
				const float inv23 = 1.f / det23;
				u = vdot(tvec, p3) * inv23;
				if(u >= 0.f && u <= 1.f)
				{
					const v3 q2 = vcross(tvec, e2);
					const float v = vdot(dir, q2) * inv23;
					if(v >= 0.f && (u + v) <= 1.f)
					{
						const float t = vdot(e3, q2) * inv23;
						if(t > 0.f) return t;
					}
				}
			}
		}
	}
	return 0.f;
}

/* AreaLight::intersect, src/light/light_area.cc:146-155.  Returns ipdf > 0 on hit. */
static int area_intersect(const ora_light *l, v3 from, v3 dir, float *t, float *ipdf)
{
	const v3 nf = area_normal_flipped(l);
	const float cos_angle = vdot(dir, nf);
	if(cos_angle <= 0.f) return 0;
	const v3 c0 = vload(l->corner), p1 = vload(l->point1), p2 = vload(l->point2);
	const v3 q[4] = {c0, p1, vsub(vadd(p1, p2), c0), p2};
	const float th = quad_intersect(q, from, dir);
	if(th <= 0.f) return 0;
	*t = th;
	*ipdf = 1.f / (th * th) * area_area(l) * cos_angle * K_1_PI;
	return 1;
}

/* ObjectLight::initIs (src/light/light_object_light.cc:87-104): the
 * triangle areas, their Pdf1D (include/sampler/sample_pdf1d.h:60-81: a
 * double running sum of area / n, each entry rounded to float and divided by
 * the float total), area_ = the double sum of the float areas, and the tree
 * bound of the light's own kd-tree. */
struct obj_light
{
	float *cdf;
	float area;
	float a[3], g[3];
};

static v3 tri_vertex(const pt_ctx *c, uint32_t prim, int k)
{
	return vload(c->verts + 3 * (size_t)c->idx[3 * (size_t)prim + k]);
}

static int obj_light_init(const pt_ctx *c, const ora_light *l, struct obj_light *o)
{
	const uint32_t n = l->n_tris;
	o->cdf = (float *)malloc(sizeof(float) * (n ? n : 1));
	float *areas = (float *)malloc(sizeof(float) * (n ? n : 1));
	if(!o->cdf || !areas) { free(areas); return 0; }
	double total = 0.0;
	for(uint32_t i = 0; i < n; ++i)
	{
		const uint32_t p = l->first_tri + i;
		const v3 v0 = tri_vertex(c, p, 0);
		const v3 cr = vcross(vsub(tri_vertex(c, p, 1), v0), vsub(tri_vertex(c, p, 2), v0));
		areas[i] = 0.5f * sqrtf(vdot(cr, cr));   /* ShapePolygon::surfaceArea, shape_polygon.h:64-76 */
		total += (double)areas[i];
	}
	const double delta = 1.0 / (double)n;
	double acc = 0.0;
	for(uint32_t i = 0; i < n; ++i)
	{
		acc += (double)areas[i] * delta;
		o->cdf[i] = (float)acc;
	}
	const float integral = (float)acc;
	for(uint32_t i = 0; i < n; ++i) o->cdf[i] /= integral;
	o->area = (float)total;
	ora_tree_bound(c->verts, c->idx + 3 * (size_t)l->first_tri, n, o->a, o->g);
	free(areas);
	return 1;
}

/* ObjectLight::sampleSurface (src/light/light_object_light.cc:124-139) with
 * Pdf1D::dSample (sample_pdf1d.h:83-95, std::lower_bound) and the triangle
 * sample of ShapePolygon::sample (shape_polygon.h:78-90). */
static void obj_sample_surface(const pt_ctx *c, const ora_light *l, const struct obj_light *o, float s1, float s2,
                               v3 *p, v3 *n)
{
	const uint32_t cnt = l->n_tris;
	uint32_t idx;
	if(s1 <= 0.f) idx = 0;
	else if(s1 >= 1.f) idx = cnt - 1;
	else
	{
		uint32_t lo = 0, hi = cnt;   /* first entry not less than s1 */
		while(lo < hi)
		{
			const uint32_t mid = lo + (hi - lo) / 2;
			if(o->cdf[mid] < s1) lo = mid + 1;
			else hi = mid;
		}
		idx = lo;
	}
	float delta = o->cdf[idx];
	float ss1;
	if(idx > 0)
	{
		delta -= o->cdf[idx - 1];
		ss1 = (s1 - o->cdf[idx - 1]) / delta;
	}
	else ss1 = s1 / delta;
	const uint32_t prim = l->first_tri + idx;
	const v3 v0 = tri_vertex(c, prim, 0), v1 = tri_vertex(c, prim, 1), v2 = tri_vertex(c, prim, 2);
	const float su1 = sqrtf(ss1);
	const float u = 1.f - su1;
	const float v = s2 * su1;
	*p = vadd(vadd(vmul(u, v0), vmul(v, v1)), vmul(1.f - u - v, v2));
	*n = vnorm(vcross(vsub(v1, v0), vsub(v2, v0)));   /* face_normal_geometric_ */
}

/* Light::illumSample of the area, object and portal lights (src/light/light_area.cc:99-124,
 * src/light/light_object_light.cc:141-171, src/light/light_background_portal.cc:157-181):
 * the direction and distance to a point on the light and the sample's pdf; 0
 * when the light gives no sample.  The portal is one-sided and divides by
 * area * cos without the object light's 1e-8 floor. */
static int illum_sample(const pt_ctx *c, const ora_light *l, const struct obj_light *o, v3 sp_p, float s1, float s2,
                        v3 *ldir_out, float *dist_out, float *pdf_out)
{
	if(o)
	{
		v3 p, n;
		obj_sample_surface(c, l, o, s1, s2, &p, &n);
		v3 ldir = vsub(p, sp_p);
		const float dist_sqr = vdot(ldir, ldir);
		const float dist = sqrtf(dist_sqr);
		if(dist <= 0.f) return 0;
		ldir = vmul(1.f / dist, ldir);
		float cos_angle = -vdot(ldir, n);
		if(cos_angle <= 0.f)
		{
			if(two_sided(l)) cos_angle = -cos_angle;
			else return 0;
		}
		const float area_mul_cosangle = o->area * cos_angle;
		if(l->type == ORA_LIGHT_PORTAL) *pdf_out = dist_sqr * K_PI / area_mul_cosangle;
		else *pdf_out = dist_sqr * K_PI / ((area_mul_cosangle == 0.f) ? 1e-8f : area_mul_cosangle);
		*ldir_out = ldir;
		*dist_out = dist;
		return 1;
	}
	const v3 c0 = vload(l->corner);
	const v3 to_x = vsub(vload(l->point1), c0), to_y = vsub(vload(l->point2), c0);
	const v3 p = vadd(vadd(c0, vmul(s1, to_x)), vmul(s2, to_y));
	v3 ldir = vsub(p, sp_p);
	const float dist_sqr = vdot(ldir, ldir);
	const float dist = sqrtf(dist_sqr);
	if(dist <= 0.f) return 0;
	ldir = vmul(1.f / dist, ldir);
	const float cos_angle = vdot(ldir, area_normal_flipped(l));
	if(cos_angle <= 0.f) return 0;
	*pdf_out = dist_sqr * K_PI / (area_area(l) * cos_angle);
	*ldir_out = ldir;
	*dist_out = dist;
	return 1;
}

/* ObjectLight::intersect and BackgroundPortalLight::intersect
 * (src/light/light_object_light.cc:211-229, src/light/light_background_portal.cc:207-222):
 * the closest hit of the light's own kd-tree (t_min = max(ray tmin, dynamic
 * bias), visible triangles only, include/accelerator/accelerator.h:122-131),
 * found here by testing every triangle; the lowest index wins an exact tie.
 * The reference never writes the hit distance back into its `t` argument,
 * which areaLightSampleMaterial passes as the ray's tmax_ = -1, so 1 / (t * t)
 * is 1: the returned inverse pdf ignores the distance. */
static int obj_intersect(const pt_ctx *c, const ora_light *l, const struct obj_light *o, v3 from, v3 dir, float tmin,
                         float *ipdf)
{
	const float f[3] = {from.x, from.y, from.z}, d[3] = {dir.x, dir.y, dir.z};
	float enter, leave;
	if(!ora_bound_cross(o->a, o->g, f, d, FLT_MAX, &enter, &leave)) return 0;
	const float bias = ora_dynamic_bias(enter, leave);
	const float t_min = (tmin < bias) ? bias : tmin;
	float best = FLT_MAX;
	uint32_t hit = 0xFFFFFFFFu;
	for(uint32_t i = 0; i < l->n_tris; ++i)
	{
		const uint32_t p = l->first_tri + i;
		if(c->vis && !(c->vis[p] & ORA_VIS_VISIBLE)) continue;
		const v3 v0 = tri_vertex(c, p, 0), v1 = tri_vertex(c, p, 1), v2 = tri_vertex(c, p, 2);
		const float a0[3] = {v0.x, v0.y, v0.z}, a1[3] = {v1.x, v1.y, v1.z}, a2[3] = {v2.x, v2.y, v2.z};
		float t, u, v;
		if(!ora_triangle_intersect(a0, a1, a2, f, d, &t, &u, &v)) continue;
		if(t <= 0.f || t < t_min || t >= best) continue;
		best = t;
		hit = p;
	}
	if(hit == 0xFFFFFFFFu) return 0;
	const v3 v0 = tri_vertex(c, hit, 0);
	const v3 n = vnorm(vcross(vsub(tri_vertex(c, hit, 1), v0), vsub(tri_vertex(c, hit, 2), v0)));
	float cos_angle = vdot(dir, vneg(n));
	if(cos_angle <= 0.f)
	{
		if(two_sided(l)) cos_angle = fabsf(cos_angle);
		else return 0;
	}
	const float t_arg = -1.f;
	const float idist_sqr = 1.f / (t_arg * t_arg);
	*ipdf = idist_sqr * o->area * cos_angle * K_1_PI;
	return 1;
}

/* MonteCarloIntegrator::doLightEstimation and the area / dirac paths,
 * src/integrator/surface/integrator_montecarlo.cc:118-440. */
static rgb light_estimation(const pt_ctx *c, const ora_light *l, const surf *sp, v3 wo, uint32_t loffs,
                            const ray_div *rd, uint32_t psd_sample, uint32_t psd_offset)
{
	rgb col = C(0.f, 0.f, 0.f);
	const int cast_shadows = l->cast_shadows && sp->mat->receive_shadows;
	const float plen = sqrtf(vdot(sp->p, sp->p));
	const float shadow_tmin = 0.0005f * (1.f > plen ? 1.f : plen);   /* shadow_bias_auto_ */
	if(l->type == ORA_LIGHT_POINT)
	{
		/* PointLight::illuminate, src/light/light_point.cc:70-81; diracLight :118-190 */
		v3 ldir = vsub(vload(l->from), sp->p);
		const float dist_sqr = ldir.x * ldir.x + ldir.y * ldir.y + ldir.z * ldir.z;
		const float dist = sqrtf(dist_sqr);
		if(dist == 0.f) return col;
		const float idist_sqr = 1.f / dist_sqr;
		ldir = vmul(1.f / dist, ldir);
		const rgb lcol = cscale(light_color(c, l), idist_sqr);
		int shadowed = 0;
		rgb scol = C(1.f, 1.f, 1.f);
		if(cast_shadows) shadowed = shadow_ray(c, sp->p, ldir, shadow_tmin, dist, &scol);
		const float angle = fabsf(vdot(sp->n, ldir));
		if(!shadowed)
		{
			const rgb surf_col = bsdf_eval(sp, wo, ldir, F_ALL);
			const rgb lc = (c->st->transparent_shadows && cast_shadows) ? cmul(lcol, scol) : lcol;
			col = cadd(col, cmul(cscale(cmul(surf_col, lc), angle), C(1.f, 1.f, 1.f)));
		}
		return col;
	}
	const uint32_t l_offs = loffs * 4567u;
	int num_samples = (int)ceilf((float)l->samples * c->light_mult);
	if(rd->division > 1) num_samples = num_samples / rd->division > 1 ? num_samples / rd->division : 1;
	const float inv_num_samples = 1.f / (float)num_samples;
	const uint32_t offs = (uint32_t)num_samples * psd_sample + psd_offset + l_offs;
	const rgb lcolor = light_color(c, l);
	/* the colour Light::intersect returns: the portal clamps it
	 * (src/light/light_background_portal.cc:219-220) */
	float hc[4] = {lcolor.r, lcolor.g, lcolor.b, 0.f};
	if(l->type == ORA_LIGHT_PORTAL) clamp_proportional(hc, l->ibl_clamp_sampling);
	const rgb hit_color = C(hc[0], hc[1], hc[2]);
	const struct obj_light *o =
	    (l->type == ORA_LIGHT_OBJECT || l->type == ORA_LIGHT_PORTAL) ? &c->objs[l - c->lights] : NULL;
	halton h2, h3;
	/* areaLightSampleLight, :192-316 */
	halton_start(&h2, 2, offs - 1);
	halton_start(&h3, 3, offs - 1);
	rgb acc = C(0.f, 0.f, 0.f);
	for(int i = 0; i < num_samples; ++i)
	{
		const float s1 = halton_next(&h2);
		const float s2 = halton_next(&h3);
		v3 ldir;
		float dist, lpdf;
		if(!illum_sample(c, l, o, sp->p, s1, s2, &ldir, &dist, &lpdf)) continue;
		int shadowed = 0;
		rgb scol = C(1.f, 1.f, 1.f);
		if(cast_shadows) shadowed = shadow_ray(c, sp->p, ldir, shadow_tmin, dist, &scol);
		if(!shadowed && lpdf > 1e-6f)
		{
			const rgb lc = (c->st->transparent_shadows && cast_shadows) ? cmul(lcolor, scol) : lcolor;
			const rgb surf_col = bsdf_eval(sp, wo, ldir, F_ALL);
			const float angle = fabsf(vdot(sp->n, ldir));
			float w = 1.f;
			const float m_pdf = bsdf_pdf(sp, wo, ldir, F_GLOSSY | F_DIFFUSE | F_DISPERSIVE | F_REFLECT | F_TRANSMIT);
			if(m_pdf > 1e-6f)
			{
				const float l_2 = lpdf * lpdf;
				const float m_2 = m_pdf * m_pdf;
				w = l_2 / (l_2 + m_2);
			}
			acc = cadd(acc, cdiv(cscale(cscale(cmul(surf_col, lc), angle), w), lpdf));
		}
	}
	col = cadd(col, cscale(acc, inv_num_samples));
	/* areaLightSampleMaterial, :318-415 */
	halton_start(&h2, 2, offs - 1);
	halton_start(&h3, 3, offs - 1);
	acc = C(0.f, 0.f, 0.f);
	const float btmin = 0.00005f * (1.f > plen ? 1.f : plen);   /* ray_min_dist_auto_ */
	for(int i = 0; i < num_samples; ++i)
	{
		bsample s;
		s.s1 = halton_next(&h2);
		s.s2 = halton_next(&h3);
		s.flags = F_GLOSSY | F_DIFFUSE | F_DISPERSIVE | F_REFLECT | F_TRANSMIT;
		s.pdf = 0.f;
		float W = 0.f;
		v3 dir = V(0.f, 0.f, 0.f);
		const rgb surf_col = bsdf_sample(sp, wo, &dir, &s, &W);
		/* Light::intersect; the shadow ray ends at the ray's tmax_, which only
		 * AreaLight::intersect sets (to the hit distance; -1 otherwise: no end) */
		float t = -1.f, ipdf = 0.f;
		const int light_hit = o ? obj_intersect(c, l, o, sp->p, dir, btmin, &ipdf)
		                        : area_intersect(l, sp->p, dir, &t, &ipdf);
		if(light_hit && s.pdf > 1e-6f)
		{
			int shadowed = 0;
			rgb scol = C(1.f, 1.f, 1.f);
			if(cast_shadows) shadowed = shadow_ray(c, sp->p, dir, btmin, t, &scol);
			if(!shadowed && ipdf > 1e-6f)
			{
				const rgb lc = (c->st->transparent_shadows && cast_shadows) ? cmul(hit_color, scol) : hit_color;
				const float l_pdf = 1.f / ipdf;
				const float l_2 = l_pdf * l_pdf;
				const float m_2 = s.pdf * s.pdf;
				const float w = m_2 / (l_2 + m_2);
				acc = cadd(acc, cscale(cscale(cmul(surf_col, lc), w), W));
			}
		}
	}
	col = cadd(col, cscale(acc, inv_num_samples));
	return col;
}

/* per camera sample: the PixelSamplingData fields the integrator reads */
typedef struct { uint32_t pixel, sample, offset; } psd_t;

/* estimateAllDirectLight, integrator_montecarlo.cc:92-106 */
static rgb direct_all(const pt_ctx *c, const surf *sp, v3 wo, const ray_div *rd, const psd_t *psd)
{
	rgb col = C(0.f, 0.f, 0.f);
	for(uint32_t l = 0; l < c->n_lights; ++l)
		col = cadd(col, light_estimation(c, &c->lights[l], sp, wo, l, rd, psd->sample, psd->offset));
	return col;
}

/* estimateOneDirectLight, integrator_montecarlo.cc:108-116, with the
 * deterministic light-pick index of ora_pt.h. */
static rgb direct_one(const pt_ctx *c, const surf *sp, v3 wo, const ray_div *rd, const psd_t *psd, uint32_t sub,
                      uint32_t depth, uint32_t tag)
{
	const int num_lights = (int)c->n_lights;
	if(num_lights == 0) return C(0.f, 0.f, 0.f);
	const uint32_t bounces = (uint32_t)c->st->bounces;
	const uint32_t corr = (psd->pixel * c->total_samples + psd->sample) * ((uint32_t)c->st->path_samples * bounces) +
	                      sub * bounces + depth + 1u + tag * 0x9E3779B9u;
	const float h = ora_halton_next(2, corr - 1u);
	int lnum = (int)(h * (float)num_lights);
	if(lnum > num_lights - 1) lnum = num_lights - 1;
	return cscale(light_estimation(c, &c->lights[lnum], sp, wo, (uint32_t)lnum, rd, psd->sample, psd->offset),
	              (float)num_lights);
}

/* math::addMod1, include/math/math.h:304-309 */
static inline float add_mod1(float a, float b)
{
	const float s = a + b;
	return s > 1 ? s - 1 : s;
}

static void integrate(const pt_ctx *c, v3 from, v3 dir, float tmin, int ray_level, ray_div rd, const psd_t *psd,
                      uint32_t tag, rgb *out);

/* MonteCarloIntegrator::glossy with glossyReflectNoTransmit (the glossy
 * material has no transmission), integrator_montecarlo.cc:506-589,623-645 */
static rgb glossy_branches(const pt_ctx *c, int ray_level, const surf *sp, v3 wo, const ray_div *rd,
                           const psd_t *psd, uint32_t tag)
{
	if(!(sp->flags & F_REFLECT)) return C(0.f, 0.f, 0.f);
	const int ray_samples_glossy = rd->division > 1 ? (8 / rd->division > 1 ? 8 / rd->division : 1) : 8;
	ray_div nrd = *rd;
	nrd.division *= ray_samples_glossy;
	int branch = nrd.division * rd->offset;
	uint32_t offs = (uint32_t)ray_samples_glossy * psd->sample + psd->offset;
	const float inverse_ray_samples_glossy = 1.f / (float)ray_samples_glossy;
	rgb gcol = C(0.f, 0.f, 0.f);
	halton h2, h3;
	halton_start(&h2, 2, offs);
	halton_start(&h3, 3, offs);
	for(int ns = 0; ns < ray_samples_glossy; ++ns)
	{
		nrd.dec1 = (float)ora_low_discrepancy(2 * ray_level + 1, (uint32_t)branch + psd->offset);
		nrd.dec2 = (float)ora_low_discrepancy(2 * ray_level + 2, (uint32_t)branch + psd->offset);
		nrd.offset = branch;
		++offs;
		++branch;
		const float s_1 = halton_next(&h2);
		const float s_2 = halton_next(&h3);
		/* glossyReflectNoTransmit */
		float w = 0.f;
		bsample s = {s_1, s_2, F_GLOSSY | F_REFLECT, F_NONE, 0.f};
		v3 wi = V(0.f, 0.f, 0.f);
		const rgb mcol = bsdf_sample(sp, wo, &wi, &s, &w);
		rgb integ;
		integrate(c, sp->p, wi, 0.00005f, ray_level, nrd, psd, tag * 11u + 1u + (uint32_t)ns, &integ);
		integ = cmul(integ, cscale(mcol, w));
		gcol = cadd(gcol, integ);
	}
	return cscale(gcol, inverse_ray_samples_glossy);
}

/* MonteCarloIntegrator::recursiveRaytrace with specularReflect / specularRefract,
 * integrator_montecarlo.cc:647-748 (no dispersion, no volumes, additional depth 0) */
static rgb recursive_raytrace(const pt_ctx *c, int ray_level, uint32_t bsdfs, const surf *sp, v3 wo,
                              const ray_div *rd, const psd_t *psd, uint32_t tag)
{
	rgb col = C(0.f, 0.f, 0.f);
	if(ray_level <= c->st->raydepth && ray_level < 20 && (bsdfs & (F_GLOSSY | F_SPECULAR | F_FILTER)))
	{
		if(bsdfs & F_GLOSSY) col = cadd(col, glossy_branches(c, ray_level, sp, wo, rd, psd, tag));
		if(bsdfs & (F_SPECULAR | F_FILTER))
		{
			const specular spec = get_specular(ray_level, sp, wo);
			if(spec.has_r)
			{
				rgb integ;
				integrate(c, sp->p, spec.dir_r, 0.00005f, ray_level, *rd, psd, tag * 11u + 9u, &integ);
				col = cadd(col, cmul(integ, spec.col_r));
			}
			if(spec.has_t)
			{
				rgb integ;
				integrate(c, sp->p, spec.dir_t, 0.00005f, ray_level, *rd, psd, tag * 11u + 10u, &integ);
				col = cadd(col, cmul(integ, spec.col_t));
			}
		}
	}
	return col;
}

/* PathIntegrator::integrate, src/integrator/surface/integrator_path_tracer.cc:138-314.
 * The alpha it returns is always 1 here (no transparent background). */
static void integrate(const pt_ctx *c, v3 from, v3 dir, float tmin, int ray_level, ray_div rd, const psd_t *psd,
                      uint32_t tag, rgb *out)
{
	const ora_pt_settings *st = c->st;
	rgb col = C(0.f, 0.f, 0.f);
	float w = 0.f;
	surf sp;
	if(!intersect_surface(c, from, dir, tmin, -1.f, &sp))
	{
		/* TiledIntegrator::background, integrator_tiled.cc:693-706 */
		if(st->has_background) col = C(st->background[0], st->background[1], st->background[2]);
		*out = col;
		return;
	}
	const uint32_t mat_bsdfs = sp.flags;
	const v3 wo = vneg(dir);
	if(mat_bsdfs & F_EMIT) col = cadd(col, mat_emit(&sp, wo));
	if(mat_bsdfs & F_DIFFUSE) col = cadd(col, direct_all(c, &sp, wo, &rd, psd));
	uint32_t path_flags = F_DIFFUSE;
	if(mat_bsdfs & path_flags)
	{
		rgb path_col = C(0.f, 0.f, 0.f);
		path_flags |= F_DIFFUSE | F_REFLECT | F_TRANSMIT;
		const int n_samples = st->path_samples / rd.division > 1 ? st->path_samples / rd.division : 1;
		for(int i = 0; i < n_samples; ++i)
		{
			const uint32_t offs = (uint32_t)st->path_samples * psd->sample + psd->offset + (uint32_t)i;
			rgb throughput;
			v3 pwo = wo;
			bsample s;
			float s_1 = ora_ri_vdc(offs, 0);
			float s_2 = (float)ora_low_discrepancy(2, offs);
			if(rd.division > 1)
			{
				s_1 = add_mod1(s_1, rd.dec1);
				s_2 = add_mod1(s_2, rd.dec2);
			}
			s.s1 = s_1;
			s.s2 = s_2;
			s.flags = path_flags;
			s.pdf = 0.f;
			s.sampled = F_NONE;
			v3 pdir = V(0.f, 0.f, 0.f);
			rgb scol = bsdf_sample(&sp, pwo, &pdir, &s, &w);
			scol = cscale(scol, w);
			throughput = scol;
			surf hit;
			if(!intersect_surface(c, sp.p, pdir, 0.00005f, -1.f, &hit)) continue;
			if(s.sampled != F_NONE) pwo = vneg(pdir);
			rgb lcol = direct_one(c, &hit, pwo, &rd, psd, (uint32_t)i, 0u, tag);
			const uint32_t mat_bsd_fs = hit.flags;
			if(mat_bsd_fs & F_EMIT) lcol = cadd(lcol, mat_emit(&hit, pwo));
			path_col = cadd(path_col, cmul(lcol, throughput));
			int caustic = 0;
			for(int depth = 1; depth < st->bounces; ++depth)
			{
				const int d_4 = 4 * depth;
				s.s1 = (float)ora_low_discrepancy(d_4 + 3, offs);
				s.s2 = (float)ora_low_discrepancy(d_4 + 4, offs);
				s.flags = F_ALL;
				scol = bsdf_sample(&hit, pwo, &pdir, &s, &w);
				scol = cscale(scol, w);
				if(cblack(scol)) break;
				throughput = cmul(throughput, scol);
				caustic = (s.sampled & (F_SPECULAR | F_GLOSSY | F_FILTER)) != 0;
				surf nh;
				if(!intersect_surface(c, hit.p, pdir, 0.00005f, -1.f, &nh)) break;
				hit = nh;
				pwo = vneg(pdir);
				if(mat_bsd_fs & F_DIFFUSE)
					lcol = direct_one(c, &hit, pwo, &rd, psd, (uint32_t)i, (uint32_t)depth, tag);
				else lcol = C(0.f, 0.f, 0.f);
				if(depth > st->rr_min_bounces)
				{
					const float random_value =
						ora_rr_random(psd->pixel, psd->sample * (uint32_t)st->path_samples + (uint32_t)i,
						              (uint32_t)depth + 256u * tag, st->seed);
					const float probability = cmax(throughput);
					if(probability <= 0.f || probability < random_value) break;
					throughput = cscale(throughput, 1.f / probability);
				}
				if((mat_bsd_fs & F_EMIT) && caustic) lcol = cadd(lcol, mat_emit(&hit, pwo));
				path_col = cadd(path_col, cmul(lcol, throughput));
			}
		}
		col = cadd(col, cdiv(path_col, (float)n_samples));
	}
	col = cadd(col, recursive_raytrace(c, ray_level + 1, mat_bsdfs, &sp, wo, &rd, psd, tag));
	*out = col;
}

static void free_ctx(pt_ctx *c)
{
	free(c->transparent);
	c->transparent = NULL;
	if(!c->objs) return;
	for(uint32_t l = 0; l < c->n_lights; ++l) free(c->objs[l].cdf);
	free(c->objs);
	c->objs = NULL;
}

/* Returns 0 when an object light's tables cannot be allocated. */
static int setup_ctx(pt_ctx *c, const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx,
                     const uint32_t *mat_id, const uint8_t *vis, const ora_material *mats, const ora_light *lights,
                     uint32_t n_lights, const ora_pt_settings *st)
{
	c->kd = kd; c->verts = verts; c->idx = tri_idx; c->mat_id = mat_id; c->vis = vis; c->mats = mats;
	c->lights = lights; c->n_lights = n_lights; c->st = st;
	c->objs = NULL;
	c->transparent = NULL;
	if(st->transparent_shadows && kd)
	{
		/* isTransparent(): glass with fake_shadows (include/material/material_glass.h:82) */
		const uint32_t n = ora_kd_num_prims(kd);
		if(st->shadow_depth < 0 || st->shadow_depth > ORA_MAX_SHADOW_DEPTH ||
		   !(c->transparent = (uint8_t *)calloc(n ? n : 1, 1)))
			return 0;
		for(uint32_t p = 0; p < n; ++p)
		{
			const ora_material *m = &mats[mat_id[p]];
			c->transparent[p] = (m->type == ORA_MAT_GLASS && m->fake_shadows) ? 1 : 0;
		}
	}
	for(uint32_t l = 0; l < n_lights; ++l)
	{
		if(lights[l].type != ORA_LIGHT_OBJECT && lights[l].type != ORA_LIGHT_PORTAL) continue;
		/* a portal sends the scene's background through: none, no light */
		if(lights[l].type == ORA_LIGHT_PORTAL && !st->has_background)
		{
			free_ctx(c);
			return 0;
		}
		if(!c->objs && !(c->objs = (struct obj_light *)calloc(n_lights, sizeof(struct obj_light))))
		{
			free_ctx(c);
			return 0;
		}
		if(lights[l].n_tris == 0 || !obj_light_init(c, &lights[l], &c->objs[l]))
		{
			free_ctx(c);
			return 0;
		}
	}
	c->total_samples = (uint32_t)ora_aa_max_samples(st);
	c->light_mult = 1.f;
	c->pass_samples = st->aa_samples;
	/* Camera::Camera, src/camera/camera.cc:89-104 */
	const v3 from = vload(st->from);
	v3 cy = vsub(vload(st->up), from);
	v3 cz = vsub(vload(st->to), from);
	v3 cx = vcross(cz, cy);
	cy = vcross(cz, cx);
	cx = vnorm(cx);
	cy = vnorm(cy);
	cz = vnorm(cz);
	const float aspect = st->aspect_factor * (float)st->resy / (float)st->resx;
	/* PerspectiveCamera::setAxis, camera_perspective.cc:97-111 */
	c->vright = cx;
	c->vup = vmul(aspect, cy);
	c->vto = vsub(vmul(st->focal, cz), vmul(0.5f, vadd(c->vup, c->vright)));
	c->vup = V(c->vup.x / (float)st->resy, c->vup.y / (float)st->resy, c->vup.z / (float)st->resy);
	c->vright = V(c->vright.x / (float)st->resx, c->vright.y / (float)st->resx, c->vright.z / (float)st->resx);
	c->cam_from = from;
	return 1;
}

/* TiledIntegrator::renderTile per-sample set-up, integrator_tiled.cc:290-335:
 * sample is the pass offset plus the sample's index in the pass. */
static void pixel_sample(const pt_ctx *c, int x, int y, int sample, float col[4], float *dx_out, float *dy_out)
{
	const ora_pt_settings *st = c->st;
	const uint32_t pixel = (uint32_t)(st->resx * y + x);
	const uint32_t offset = ora_fnv32a((uint32_t)y * ora_fnv32a((uint32_t)x));
	const uint32_t psd_sample = (uint32_t)sample;
	float dx = 0.5f, dy = 0.5f;
	if(st->aa_passes > 1)
	{
		/* scrambled (0, 2) sequence, as the total sample count is not known */
		dx = ora_ri_vdc(psd_sample, offset);
		dy = ora_ri_s(psd_sample, offset);
	}
	else if(c->pass_samples > 1)
	{
		const float d_1 = 1.f / (float)c->pass_samples;
		dx = (0.5f + (float)sample) * d_1;
		dy = ora_ri_lp((uint32_t)sample + offset, 0);
	}
	/* PerspectiveCamera::shootRay, camera_perspective.cc:165-180 (pinhole) */
	const float px = (float)x + dx, py = (float)y + dy;
	v3 dir = vadd(vadd(vmul(px, c->vright), vmul(py, c->vup)), c->vto);
	dir = vnorm(dir);
	const psd_t psd = {pixel, psd_sample, offset};
	const ray_div rd = {1, 0, 0.f, 0.f};
	rgb out;
	integrate(c, c->cam_from, dir, 0.f, 0, rd, &psd, 0u, &out);
	col[0] = out.r;
	col[1] = out.g;
	col[2] = out.b;
	col[3] = 1.f;
	*dx_out = dx;
	*dy_out = dy;
}

void ora_pt_camera_dir(const ora_pt_settings *st, float px, float py, float dir[3])
{
	pt_ctx c;
	memset(&c, 0, sizeof c);
	setup_ctx(&c, NULL, NULL, NULL, NULL, NULL, NULL, NULL, 0, st);
	const v3 d = vnorm(vadd(vadd(vmul(px, c.vright), vmul(py, c.vup)), c.vto));
	dir[0] = d.x;
	dir[1] = d.y;
	dir[2] = d.z;
}

void ora_pt_pixel_sample(const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx, const uint32_t *mat_id,
                         const uint8_t *vis, const ora_material *mats, const ora_light *lights, uint32_t n_lights,
                         const ora_pt_settings *st, int x, int y, int sample, float col_alpha[4], float dxdy[2])
{
	pt_ctx c;
	if(!setup_ctx(&c, kd, verts, tri_idx, mat_id, vis, mats, lights, n_lights, st)) return;
	pixel_sample(&c, x, y, sample, col_alpha, &dxdy[0], &dxdy[1]);
	free_ctx(&c);
}

/* ------------------------------------------------------------------ */
/* film                                                                  */

typedef struct
{
	const pt_ctx *c;
	int sample;
	const uint8_t *flags;  /* pixels of an adaptive pass (NULL: all) */
	float *cols;   /* resx * resy * 4 */
	float *dxy;    /* resx * resy * 2 */
	int tid, nthreads;
} pass_job;

static void *pass_worker(void *arg)
{
	pass_job *j = (pass_job *)arg;
	const ora_pt_settings *st = j->c->st;
	/* pixels dealt out one by one, so a sparse flag set (a window) still
	 * spreads over every thread; the film is splatted afterwards in pixel order */
	const size_t npix = (size_t)st->resx * st->resy;
	for(size_t k = (size_t)j->tid; k < npix; k += (size_t)j->nthreads)
	{
		if(j->flags && !j->flags[k]) continue;
		const int x = (int)(k % (size_t)st->resx), y = (int)(k / (size_t)st->resx);
		pixel_sample(j->c, x, y, j->sample, j->cols + 4 * k, &j->dxy[2 * k], &j->dxy[2 * k + 1]);
	}
	return NULL;
}

/* Rgb::clampProportionalRgb, include/color/color.h:399-431 (alpha untouched) */
static void clamp_proportional(float col[4], float max_value)
{
	if(max_value <= 0.f) return;
	const float max_rgb = fmaxf(col[0], fmaxf(col[1], col[2]));
	const float adj = max_value / max_rgb;
	if(max_rgb > max_value)
	{
		if(col[0] >= max_rgb)
		{
			col[0] = max_value;
			col[1] *= adj;
			col[2] *= adj;
		}
		else if(col[1] >= max_rgb)
		{
			col[1] = max_value;
			col[0] *= adj;
			col[2] *= adj;
		}
		else
		{
			col[2] = max_value;
			col[0] *= adj;
			col[1] *= adj;
		}
	}
}

/* roundToIntWithBias, include/render/imagefilm.h:344-348 */
static inline int round_bias(double v) { return (int)(v + (.5 - 1.4e-11)); }

/* ImageFilm::addSample, src/render/imagefilm.cc:771-822 */
static void add_sample(int resx, int resy, const float table[256], float table_scale, float filter_width, int x,
                       int y, float dx, float dy, const float col_in[4], float clamp, float *film_rgba, float *film_w)
{
	float col[4] = {col_in[0], col_in[1], col_in[2], col_in[3]};
	clamp_proportional(col, clamp);
	int dx0 = round_bias((double)dx - filter_width);
	int dx1 = round_bias((double)dx + filter_width - 1.0);
	int dy0 = round_bias((double)dy - filter_width);
	int dy1 = round_bias((double)dy + filter_width - 1.0);
	if(-x > dx0) dx0 = -x;
	if((resx - 1) - x < dx1) dx1 = (resx - 1) - x;
	if(-y > dy0) dy0 = -y;
	if((resy - 1) - y < dy1) dy1 = (resy - 1) - y;
	int xi[9], yi[9];
	const double x_offs = dx - 0.5;
	for(int i = dx0, n = 0; i <= dx1; ++i, ++n) xi[n] = (int)floor(fabs(((double)i - x_offs) * table_scale));
	const double y_offs = dy - 0.5;
	for(int i = dy0, n = 0; i <= dy1; ++i, ++n) yi[n] = (int)floor(fabs(((double)i - y_offs) * table_scale));
	for(int j = y + dy0; j <= y + dy1; ++j)
		for(int i = x + dx0; i <= x + dx1; ++i)
		{
			const float wt = table[yi[j - (y + dy0)] * 16 + xi[i - (x + dx0)]];
			const size_t p = (size_t)j * resx + i;
			film_w[p] = film_w[p] + wt;
			for(int k = 0; k < 4; ++k) film_rgba[4 * p + k] += wt * col[k];
		}
}

/* renderPass of n samples from sample `offset` over the flagged pixels (all
 * when flags is NULL), splatted sample by sample, pixels in row-major order. */
static void render_pass(pt_ctx *c, int offset, int n, const uint8_t *flags, const float table[256], float table_scale,
                        float filter_width, float *cols, float *dxy, float *film_rgba, float *film_w, int threads)
{
	const ora_pt_settings *st = c->st;
	pass_job jobs[256];
	pthread_t th[256];
	c->pass_samples = n;
	for(int s = 0; s < n; ++s)
	{
		for(int t = 0; t < threads; ++t)
		{
			jobs[t].c = c; jobs[t].sample = offset + s; jobs[t].flags = flags; jobs[t].cols = cols; jobs[t].dxy = dxy;
			jobs[t].tid = t; jobs[t].nthreads = threads;
		}
		if(threads == 1) pass_worker(&jobs[0]);
		else
		{
			for(int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, pass_worker, &jobs[t]);
			for(int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
		}
		for(int y = 0; y < st->resy; ++y)
			for(int x = 0; x < st->resx; ++x)
			{
				const size_t k = (size_t)y * st->resx + x;
				if(flags && !flags[k]) continue;
				add_sample(st->resx, st->resy, table, table_scale, filter_width, x, y, dxy[2 * k], dxy[2 * k + 1],
				           cols + 4 * k, st->aa_clamp_samples, film_rgba, film_w);
			}
	}
}

int ora_aa_max_samples(const ora_pt_settings *st)
{
	/* integrator_tiled.cc:258-262: int += ceilf(int * pow(float, int)) */
	int m = st->aa_samples;
	for(int i = 1; i < st->aa_passes; ++i)
		m = (int)((float)m + ceilf((float)((double)st->aa_inc_samples * pow((double)st->aa_sample_multiplier_factor, i))));
	return m;
}

/* Rgba::col2Bri / abscol2Bri / colorDifference, include/color/color.h:57-58,434-451 */
static inline float col2bri(const float c[4]) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }
static inline float abscol2bri(const float c[4])
{
	return 0.2126f * fabsf(c[0]) + 0.7152f * fabsf(c[1]) + 0.0722f * fabsf(c[2]);
}
static float color_difference(const float a[4], const float b[4], int rgb_components)
{
	float d = fabsf(col2bri(b) - col2bri(a));
	if(rgb_components)
		for(int k = 0; k < 4; ++k)
		{
			const float dk = fabsf(b[k] - a[k]);
			if(d < dk) d = dk;
		}
	return d;
}

/* Rgba::normalized (include/color/color.h:538-542) of a film pixel: the sums
 * times 1 / weight, zero for a zero weight. */
static void film_pixel(const float *film_rgba, const float *film_w, int resx, int x, int y, float out[4])
{
	const size_t p = (size_t)y * resx + x;
	float w = film_w[p];
	if(w != 0.f)
	{
		w = 1.f / w;
		for(int k = 0; k < 4; ++k) out[k] = film_rgba[4 * p + k] * w;
	}
	else out[0] = out[1] = out[2] = out[3] = 0.f;
}

/* ImageFilm::darkThresholdCurveInterpolate, src/render/imagefilm.cc:882-899 */
static float dark_curve(float b)
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

int ora_aa_next_pass(const ora_pt_settings *st, float threshold, const float *film_rgba, const float *film_w,
                     uint8_t *flags)
{
	const int w = st->resx, h = st->resy;
	if(threshold <= 0.f)
	{
		memset(flags, 1, (size_t)w * h);
		return w * h;
	}
	for(size_t p = 0; p < (size_t)w * h; ++p) flags[p] = film_w[p] > 0.f ? 0 : 1;
	const int half = st->aa_variance_edge_size / 2;
	float thr = threshold;
	for(int y = 0; y < h - 1; ++y)
		for(int x = 0; x < w - 1; ++x)
		{
			float pix[4], q[4];
			film_pixel(film_rgba, film_w, w, x, y, pix);
			const float bri = abscol2bri(pix);
			if(st->aa_dark_detection_type == 1 && st->aa_dark_threshold_factor > 0.f)
				thr = threshold * ((1.f - st->aa_dark_threshold_factor) + (bri * st->aa_dark_threshold_factor));
			else if(st->aa_dark_detection_type == 2) thr = dark_curve(bri);
			const int nb[4][2] = {{x + 1, y}, {x, y + 1}, {x + 1, y + 1}, {x - 1, y + 1}};
			for(int k = 0; k < 4; ++k)
			{
				if(k == 3 && x == 0) continue;
				film_pixel(film_rgba, film_w, w, nb[k][0], nb[k][1], q);
				if(color_difference(pix, q, st->aa_detect_color_noise) >= thr)
				{
					flags[(size_t)y * w + x] = 1;
					flags[(size_t)nb[k][1] * w + nb[k][0]] = 1;
				}
			}
			if(st->aa_variance_pixels > 0)
			{
				int var_x = 0, var_y = 0;
				for(int xd = -half; xd < half - 1; ++xd)
				{
					int xi = x + xd;
					if(xi < 0) xi = 0;
					else if(xi >= w - 1) xi = w - 2;
					film_pixel(film_rgba, film_w, w, xi, y, pix);
					film_pixel(film_rgba, film_w, w, xi + 1, y, q);
					if(color_difference(pix, q, st->aa_detect_color_noise) >= thr) ++var_x;
				}
				for(int yd = -half; yd < half - 1; ++yd)
				{
					int yi = y + yd;
					if(yi < 0) yi = 0;
					else if(yi >= h - 1) yi = h - 2;
					film_pixel(film_rgba, film_w, w, x, yi, pix);
					film_pixel(film_rgba, film_w, w, x, yi + 1, q);
					if(color_difference(pix, q, st->aa_detect_color_noise) >= thr) ++var_y;
				}
				if(var_x + var_y >= st->aa_variance_pixels)
					for(int xd = -half; xd < half; ++xd)
						for(int yd = -half; yd < half; ++yd)
						{
							int xi = x + xd, yi = y + yd;
							xi = xi < 0 ? 0 : (xi >= w ? w - 1 : xi);
							yi = yi < 0 ? 0 : (yi >= h ? h - 1 : yi);
							flags[(size_t)yi * w + xi] = 1;
						}
			}
		}
	int n = 0;
	for(size_t p = 0; p < (size_t)w * h; ++p) n += flags[p] ? 1 : 0;
	return n;
}

int ora_pt_render(const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx, const uint32_t *mat_id,
                   const uint8_t *vis, const ora_material *mats, const ora_light *lights, uint32_t n_lights,
                   const ora_pt_settings *st, float *film_rgba, float *film_w, int threads)
{
	pt_ctx c;
	if(!setup_ctx(&c, kd, verts, tri_idx, mat_id, vis, mats, lights, n_lights, st)) return 0;
	float table[256], table_scale, filter_width;
	ora_filter_table(st->aa_pixelwidth, table, &table_scale, &filter_width);
	const size_t npix = (size_t)st->resx * st->resy;
	float *cols = (float *)malloc(npix * 4 * sizeof(float));
	float *dxy = (float *)malloc(npix * 2 * sizeof(float));
	uint8_t *flags = (uint8_t *)malloc(npix);
	if(!cols || !dxy || !flags) { free(cols); free(dxy); free(flags); free_ctx(&c); return 0; }
	if(threads < 1) threads = 1;
	if(threads > 256) threads = 256;
	/* TiledIntegrator::render, integrator_tiled.cc:160-214 (not resumed) */
	c.light_mult = 1.f;
	render_pass(&c, 0, st->aa_samples, NULL, table, table_scale, filter_width, cols, dxy, film_rgba, film_w, threads);
	const int floor_pixels = (int)floorf(st->aa_resampled_floor * (float)npix / 100.f);
	float sample_mult = 1.f, light_mult = 1.f, threshold = st->aa_threshold;
	int acum = st->aa_samples, resampled = 0, threshold_changed = 1;
	for(int i = 1; i < st->aa_passes; ++i)
	{
		sample_mult *= st->aa_sample_multiplier_factor;
		light_mult *= st->aa_light_sample_multiplier_factor;
		if(resampled <= 0 && !threshold_changed) resampled = 0;  /* nextPass skipped: nothing rendered */
		else
		{
			resampled = ora_aa_next_pass(st, threshold, film_rgba, film_w, flags);
			threshold_changed = 0;
		}
		const int n = (int)ceilf((float)st->aa_inc_samples * sample_mult);
		if(resampled > 0)
		{
			c.light_mult = light_mult;
			render_pass(&c, acum, n, flags, table, table_scale, filter_width, cols, dxy, film_rgba, film_w, threads);
		}
		acum += n;
		if(resampled < floor_pixels)
		{
			const float ratio = fminf(8.f, (float)floor_pixels / (float)resampled);
			threshold = threshold * (1.f - 0.1f * ratio);
			if(st->aa_threshold > 0.f) threshold_changed = 1;
		}
	}
	free(cols);
	free(dxy);
	free(flags);
	free_ctx(&c);
	return 1;
}

int ora_pt_render_pixels(const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx, const uint32_t *mat_id,
                          const uint8_t *vis, const ora_material *mats, const ora_light *lights, uint32_t n_lights,
                          const ora_pt_settings *st, const uint8_t *flags, float *film_rgba, float *film_w,
                          int threads)
{
	pt_ctx c;
	if(!setup_ctx(&c, kd, verts, tri_idx, mat_id, vis, mats, lights, n_lights, st)) return 0;
	float table[256], table_scale, filter_width;
	ora_filter_table(st->aa_pixelwidth, table, &table_scale, &filter_width);
	const size_t npix = (size_t)st->resx * st->resy;
	float *cols = (float *)malloc(npix * 4 * sizeof(float));
	float *dxy = (float *)malloc(npix * 2 * sizeof(float));
	if(!cols || !dxy) { free(cols); free(dxy); free_ctx(&c); return 0; }
	if(threads < 1) threads = 1;
	if(threads > 256) threads = 256;
	c.light_mult = 1.f;
	render_pass(&c, 0, st->aa_samples, flags, table, table_scale, filter_width, cols, dxy, film_rgba, film_w, threads);
	free(cols);
	free(dxy);
	free_ctx(&c);
	return 1;
}

/* ------------------------------------------------------------------ */
/* single-function access for the oracle's own pinning tests            */

static surf flat_surf(const ora_material *m, const float ng[3])
{
	surf sp;
	memset(&sp, 0, sizeof sp);
	sp.ng = vload(ng);
	sp.n = sp.ng;
	coords_system(sp.n, &sp.u, &sp.v);
	sp.mat = m;
	sp.flags = mat_flags(m);
	return sp;
}

uint32_t ora_mat_flags(const ora_material *m) { return mat_flags(m); }

void ora_bsdf_eval(const ora_material *m, const float ng[3], const float wo[3], const float wl[3], uint32_t bsdfs,
                   float out[3])
{
	const surf sp = flat_surf(m, ng);
	const rgb c = bsdf_eval(&sp, vload(wo), vload(wl), bsdfs);
	out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

float ora_bsdf_pdf(const ora_material *m, const float ng[3], const float wo[3], const float wi[3], uint32_t bsdfs)
{
	const surf sp = flat_surf(m, ng);
	return bsdf_pdf(&sp, vload(wo), vload(wi), bsdfs);
}

/* out: wi[3], colour[3], w, pdf, sampled flags (as float) */
void ora_bsdf_sample(const ora_material *m, const float ng[3], const float wo[3], float s1, float s2, uint32_t flags,
                     float out[9])
{
	const surf sp = flat_surf(m, ng);
	bsample s = {s1, s2, flags, F_NONE, 0.f};
	v3 wi = V(0.f, 0.f, 0.f);
	float w = 0.f;
	const rgb c = bsdf_sample(&sp, vload(wo), &wi, &s, &w);
	out[0] = wi.x; out[1] = wi.y; out[2] = wi.z;
	out[3] = c.r; out[4] = c.g; out[5] = c.b;
	out[6] = w; out[7] = s.pdf; out[8] = (float)s.sampled;
}

/* out: has_r, dir_r[3], col_r[3], has_t, dir_t[3], col_t[3] */
void ora_specular(const ora_material *m, const float ng[3], const float wo[3], int ray_level, float out[14])
{
	const surf sp = flat_surf(m, ng);
	const specular r = get_specular(ray_level, &sp, vload(wo));
	out[0] = (float)r.has_r;
	out[1] = r.dir_r.x; out[2] = r.dir_r.y; out[3] = r.dir_r.z;
	out[4] = r.col_r.r; out[5] = r.col_r.g; out[6] = r.col_r.b;
	out[7] = (float)r.has_t;
	out[8] = r.dir_t.x; out[9] = r.dir_t.y; out[10] = r.dir_t.z;
	out[11] = r.col_t.r; out[12] = r.col_t.g; out[13] = r.col_t.b;
}

void ora_fresnel(const float i[3], const float n[3], float ior, float out[2])
{
	fresnel(vload(i), vload(n), ior, &out[0], &out[1]);
}

int ora_refract(const float n[3], const float wi[3], float ior, float wo[3])
{
	v3 r = V(0.f, 0.f, 0.f);
	const int ok = refract(vload(n), vload(wi), &r, ior);
	wo[0] = r.x; wo[1] = r.y; wo[2] = r.z;
	return ok;
}

void ora_mat_emit(const ora_material *m, const float ng[3], const float wo[3], float out[3])
{
	const surf sp = flat_surf(m, ng);
	const rgb c = mat_emit(&sp, vload(wo));
	out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

/* ObjectLight (and BackgroundPortalLight) pieces on their own: set-up (area_
 * and the area cdf), illumSample from a point, and intersect along a direction. */
static int light_ctx(pt_ctx *c, const float *verts, const uint32_t *tri_idx, const uint8_t *vis, const ora_light *l,
                     struct obj_light *o)
{
	memset(c, 0, sizeof *c);
	c->verts = verts; c->idx = tri_idx; c->vis = vis;
	return (l->type == ORA_LIGHT_OBJECT || l->type == ORA_LIGHT_PORTAL) && l->n_tris > 0 && obj_light_init(c, l, o);
}

float ora_object_light_setup(const float *verts, const uint32_t *tri_idx, const ora_light *l, float *cdf)
{
	pt_ctx c;
	struct obj_light o;
	if(!light_ctx(&c, verts, tri_idx, NULL, l, &o)) return -1.f;
	memcpy(cdf, o.cdf, sizeof(float) * l->n_tris);
	free(o.cdf);
	return o.area;
}

int ora_object_light_sample(const float *verts, const uint32_t *tri_idx, const ora_light *l, const float p[3],
                            float s1, float s2, float out[5])
{
	pt_ctx c;
	struct obj_light o;
	if(!light_ctx(&c, verts, tri_idx, NULL, l, &o)) return -1;
	v3 ldir = V(0.f, 0.f, 0.f);
	float dist = 0.f, pdf = 0.f;
	const int hit = illum_sample(&c, l, &o, vload(p), s1, s2, &ldir, &dist, &pdf);
	out[0] = ldir.x; out[1] = ldir.y; out[2] = ldir.z; out[3] = dist; out[4] = pdf;
	free(o.cdf);
	return hit;
}

int ora_object_light_intersect(const float *verts, const uint32_t *tri_idx, const uint8_t *vis, const ora_light *l,
                               const float from[3], const float dir[3], float tmin, float *ipdf)
{
	pt_ctx c;
	struct obj_light o;
	if(!light_ctx(&c, verts, tri_idx, vis, l, &o)) return -1;
	*ipdf = 0.f;
	const int hit = obj_intersect(&c, l, &o, vload(from), vload(dir), tmin, ipdf);
	free(o.cdf);
	return hit;
}
