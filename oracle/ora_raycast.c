/*
 * ora_raycast.c -- CPU restatement of the reference's triangle test, slab test,
 * tree bound, dynamic ray bias and the brute-force closest/any-hit queries.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg as the checker; never by the shipped library.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* include/geometry/vector.h:155-172 -- dot product, left to right. */
static inline float dot3(const float a[3], const float b[3])
{
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* include/geometry/vector.h:238-245 -- cross product. */
static inline void cross3(const float a[3], const float b[3], float r[3])
{
	r[0] = a[1] * b[2] - a[2] * b[1];
	r[1] = a[2] * b[0] - a[0] * b[2];
	r[2] = a[0] * b[1] - a[1] * b[0];
}

static inline void sub3(const float a[3], const float b[3], float r[3])
{
	r[0] = a[0] - b[0];
	r[1] = a[1] - b[1];
	r[2] = a[2] - b[2];
}

/* include/geometry/shape/shape_polygon.h:126-176, N == 3 (Moller-Trumbore,
 * double sided, no epsilon on det).  The returned pair is (t, {u, v}). */
int ora_triangle_intersect(const float v0[3], const float v1[3], const float v2[3],
                           const float from[3], const float dir[3],
                           float *t_out, float *u_out, float *v_out)
{
	float edge_1[3], edge_2[3], pvec[3], tvec[3], qvec[3];
	sub3(v1, v0, edge_1);
	sub3(v2, v0, edge_2);
	cross3(dir, edge_2, pvec);
	const float det = dot3(edge_1, pvec);
	if(det != 0.f)
	{
		const float inv_det = 1.f / det;
		sub3(from, v0, tvec);
		const float u = dot3(tvec, pvec) * inv_det;
		if(u >= 0.f && u <= 1.f)
		{
			cross3(tvec, edge_1, qvec);
			const float v = dot3(dir, qvec) * inv_det;
			if(v >= 0.f && (u + v) <= 1.f)
			{
				const float t = dot3(edge_2, qvec) * inv_det;
				if(t > 0.f)
				{
					*t_out = t;
					*u_out = u;
					*v_out = v;
					return 1;
				}
			}
		}
	}
	*t_out = 0.f;
	*u_out = 0.f;
	*v_out = 0.f;
	return 0;
}

/* include/geometry/shape/shape_polygon.h:126-176, N == 4.  The first
 * triangle (v0, v1, v2) decides alone unless its u leaves [0, 1]: det == 0, or
 * u inside [0, 1] with v or t failing, is a miss.  Only for u outside [0, 1]
 * (or NaN) is the second triangle (v0, v2, v3) tried.  The uv returned is
 * {u + v, v} from the first triangle and {u, u + v} from the second. */
int ora_quad_intersect(const float v0[3], const float v1[3], const float v2[3], const float v3[3],
                       const float from[3], const float dir[3], float *t_out, float *u_out, float *v_out)
{
	float edge_1[3], edge_2[3], pvec_2[3], tvec[3], qvec[3];
	sub3(v1, v0, edge_1);
	sub3(v2, v0, edge_2);
	cross3(dir, edge_2, pvec_2);
	const float det_1_2 = dot3(edge_1, pvec_2);
	*t_out = 0.f;
	*u_out = 0.f;
	*v_out = 0.f;
	if(det_1_2 == 0.f) return 0;
	const float inv_det_1_2 = 1.f / det_1_2;
	sub3(from, v0, tvec);
	float u = dot3(tvec, pvec_2) * inv_det_1_2;
	if(u >= 0.f && u <= 1.f)
	{
		cross3(tvec, edge_1, qvec);
		const float v = dot3(dir, qvec) * inv_det_1_2;
		if(v >= 0.f && (u + v) <= 1.f)
		{
			const float t = dot3(edge_2, qvec) * inv_det_1_2;
			if(t > 0.f)
			{
				*t_out = t;
				*u_out = u + v;
				*v_out = v;
				return 1;
			}
		}
		return 0;
	}
	float edge_3[3], pvec_3[3];
	sub3(v3, v0, edge_3);
	cross3(dir, edge_3, pvec_3);
	const float det_2_3 = dot3(edge_2, pvec_3);
	if(det_2_3 == 0.f) return 0;
	const float inv_det_2_3 = 1.f / det_2_3;
	u = dot3(tvec, pvec_3) * inv_det_2_3;
	if(!(u >= 0.f && u <= 1.f)) return 0;
	cross3(tvec, edge_2, qvec);
	const float v = dot3(dir, qvec) * inv_det_2_3;
	if(!(v >= 0.f && (u + v) <= 1.f)) return 0;
	const float t = dot3(edge_3, qvec) * inv_det_2_3;
	if(!(t > 0.f)) return 0;
	*t_out = t;
	*u_out = u;
	*v_out = u + v;
	return 1;
}

int ora_face_nverts(const uint32_t *idx, uint32_t stride, uint32_t p)
{
	return (stride == 4 && idx[4 * (size_t)p + 3] != ORA_NO_VERTEX) ? 4 : 3;
}

int ora_face_intersect(const float *verts, const uint32_t *idx, uint32_t stride, uint32_t p, const float from[3],
                       const float dir[3], float *t, float *u, float *v)
{
	const uint32_t *ix = idx + (size_t)stride * p;
	if(ora_face_nverts(idx, stride, p) == 4)
		return ora_quad_intersect(verts + 3 * (size_t)ix[0], verts + 3 * (size_t)ix[1], verts + 3 * (size_t)ix[2],
		                          verts + 3 * (size_t)ix[3], from, dir, t, u, v);
	return ora_triangle_intersect(verts + 3 * (size_t)ix[0], verts + 3 * (size_t)ix[1], verts + 3 * (size_t)ix[2],
	                              from, dir, t, u, v);
}

/* include/geometry/bound.h:157-198 (Smits' method).  Axes whose direction
 * component is exactly zero are skipped, so no rejection happens for them even
 * when the origin lies outside that slab. */
int ora_bound_cross(const float a[3], const float g[3], const float from[3], const float dir[3],
                    float t_max, float *enter, float *leave)
{
	float p[3];
	sub3(from, a, p);
	float lmin = -FLT_MAX;
	float lmax = FLT_MAX;
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
				/* std::max(ltmin, lmin) is (ltmin < lmin) ? lmin : ltmin, so a NaN
				 * ltmin (0 * inf for an origin on a face and a denormal direction
				 * component) is kept, and a NaN lmin is replaced by ltmin. */
				lmin = (ltmin < lmin) ? lmin : ltmin;
				/* std::min(ltmax, lmax) returns ltmax unless lmax < ltmax. */
				lmax = (lmax < ltmax) ? lmax : ltmax;
			}
			if((lmax < 0.f) || (lmin > t_max)) return 0;
		}
	}
	if((lmin <= lmax) && (lmax >= 0.f) && (lmin <= t_max))
	{
		*enter = lmin;
		*leave = lmax;
		return 1;
	}
	return 0;
}

/* src/accelerator/accelerator_kdtree_original.cc:89-103 with the per-prim
 * bound of include/geometry/primitive/primitive_face.h:139-155: union of the
 * triangle vertex bounds, then each axis grown by 0.001 * extent, where the
 * extent is a float difference widened to double. */
void ora_tree_bound_faces(const float *verts, const uint32_t *idx, uint32_t stride, uint32_t n_tris,
                          float a_out[3], float g_out[3])
{
	float a[3] = {0.f, 0.f, 0.f}, g[3] = {0.f, 0.f, 0.f};
	for(uint32_t i = 0; i < n_tris; ++i)
	{
		const int nv = ora_face_nverts(idx, stride, i);
		for(int k = 0; k < nv; ++k)
		{
			const float *p = verts + 3 * (size_t)idx[(size_t)stride * i + k];
			for(int ax = 0; ax < 3; ++ax)
			{
				if(i == 0 && k == 0) { a[ax] = p[ax]; g[ax] = p[ax]; }
				else
				{
					if(p[ax] < a[ax]) a[ax] = p[ax];
					if(p[ax] > g[ax]) g[ax] = p[ax];
				}
			}
		}
	}
	for(int ax = 0; ax < 3; ++ax)
	{
		const double offset = (double)(g[ax] - a[ax]) * 0.001;
		a[ax] -= (float)offset;
		g[ax] += (float)offset;
	}
	memcpy(a_out, a, sizeof a);
	memcpy(g_out, g, sizeof g);
}

void ora_tree_bound(const float *verts, const uint32_t *tri_idx, uint32_t n_tris, float a_out[3], float g_out[3])
{
	ora_tree_bound_faces(verts, tri_idx, 3, n_tris, a_out, g_out);
}

/* include/accelerator/accelerator.h:64, with min_raydist_ = 5e-5f (:85). */
float ora_dynamic_bias(float enter, float leave)
{
	const float k = 0.1f * 0.00005f;
	return k * fabsf(leave - enter);
}

/* ------------------------------------------------------------------ */
/* brute force over all primitives                                     */

typedef struct
{
	const float *verts;
	const uint32_t *idx;
	uint32_t stride;
	const uint8_t *vis;
	uint32_t n_tris;
	const ora_ray *rays;
	uint32_t n_rays;
	ora_hit *hits;
	uint8_t *occluded;
	float a[3], g[3];
	int shadow;
	int tid, nthreads;
} brute_job;

static void brute_one_closest(const brute_job *j, uint32_t r)
{
	const ora_ray *ray = &j->rays[r];
	ora_hit h = {ORA_NO_HIT, 0.f, 0.f, 0.f};
	/* include/accelerator/accelerator.h:91 */
	const float t_max = (ray->tmax >= 0.f) ? ray->tmax : FLT_MAX;
	float enter, leave;
	if(ora_bound_cross(j->a, j->g, ray->org, ray->dir, t_max, &enter, &leave))
	{
		/* include/accelerator/accelerator_kdtree_common.h:139 */
		const float bias = ora_dynamic_bias(enter, leave);
		const float t_min = (ray->tmin < bias) ? bias : ray->tmin;
		float best = t_max;
		for(uint32_t p = 0; p < j->n_tris; ++p)
		{
			if(!(j->vis[p] & ORA_VIS_VISIBLE)) continue;
			float t, u, v;
			if(!ora_face_intersect(j->verts, j->idx, j->stride, p, ray->org, ray->dir, &t, &u, &v)) continue;
			/* include/accelerator/accelerator.h:125 */
			if(t <= 0.f || t < t_min || t >= best) continue;
			best = t;
			h.prim = p; h.t = t; h.u = u; h.v = v;
		}
	}
	j->hits[r] = h;
}

static void brute_one_shadow(const brute_job *j, uint32_t r)
{
	const ora_ray *ray = &j->rays[r];
	uint8_t occ = 0;
	const float t_max = (ray->tmax >= 0.f) ? ray->tmax : FLT_MAX;
	float enter, leave;
	if(ora_bound_cross(j->a, j->g, ray->org, ray->dir, t_max, &enter, &leave))
	{
		const float t_min = ora_dynamic_bias(enter, leave);
		for(uint32_t p = 0; p < j->n_tris && !occ; ++p)
		{
			if(!(j->vis[p] & ORA_VIS_SHADOWS)) continue;
			float t, u, v;
			if(!ora_face_intersect(j->verts, j->idx, j->stride, p, ray->org, ray->dir, &t, &u, &v)) continue;
			/* include/accelerator/accelerator.h:137 */
			if(t <= 0.f || t < t_min || t >= t_max) continue;
			occ = 1;
		}
	}
	j->occluded[r] = occ;
}

static void *brute_worker(void *arg)
{
	const brute_job *j = (const brute_job *)arg;
	for(uint32_t r = (uint32_t)j->tid; r < j->n_rays; r += (uint32_t)j->nthreads)
	{
		if(j->shadow) brute_one_shadow(j, r);
		else brute_one_closest(j, r);
	}
	return NULL;
}

static void brute_run(brute_job *proto, int threads, int bound_given)
{
	if(threads < 1) threads = 1;
	if(threads > 256) threads = 256;
	if(!bound_given) ora_tree_bound_faces(proto->verts, proto->idx, proto->stride, proto->n_tris, proto->a, proto->g);
	brute_job jobs[256];
	pthread_t th[256];
	for(int i = 0; i < threads; ++i)
	{
		jobs[i] = *proto;
		jobs[i].tid = i;
		jobs[i].nthreads = threads;
	}
	if(threads == 1) { brute_worker(&jobs[0]); return; }
	for(int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, brute_worker, &jobs[i]);
	for(int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

void ora_closest_brute_faces(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                             uint32_t n_tris, const ora_ray *rays, uint32_t n_rays, ora_hit *hits, int threads)
{
	brute_job j;
	memset(&j, 0, sizeof j);
	j.verts = verts; j.idx = idx; j.stride = stride; j.vis = vis; j.n_tris = n_tris;
	j.rays = rays; j.n_rays = n_rays; j.hits = hits; j.shadow = 0;
	brute_run(&j, threads, 0);
}

void ora_shadow_brute_faces(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                            uint32_t n_tris, const ora_ray *rays, uint32_t n_rays, uint8_t *occluded, int threads)
{
	brute_job j;
	memset(&j, 0, sizeof j);
	j.verts = verts; j.idx = idx; j.stride = stride; j.vis = vis; j.n_tris = n_tris;
	j.rays = rays; j.n_rays = n_rays; j.occluded = occluded; j.shadow = 1;
	brute_run(&j, threads, 0);
}

void ora_closest_brute_in_bound(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                                uint32_t n_faces, const float a[3], const float g[3], const ora_ray *rays,
                                uint32_t n_rays, ora_hit *hits, int threads)
{
	brute_job j;
	memset(&j, 0, sizeof j);
	j.verts = verts; j.idx = idx; j.stride = stride; j.vis = vis; j.n_tris = n_faces;
	j.rays = rays; j.n_rays = n_rays; j.hits = hits; j.shadow = 0;
	for(int k = 0; k < 3; ++k) { j.a[k] = a[k]; j.g[k] = g[k]; }
	brute_run(&j, threads, 1);
}

void ora_closest_brute(const float *verts, const uint32_t *tri_idx, const uint8_t *vis,
                       uint32_t n_tris, const ora_ray *rays, uint32_t n_rays,
                       ora_hit *hits, int threads)
{
	ora_closest_brute_faces(verts, tri_idx, 3, vis, n_tris, rays, n_rays, hits, threads);
}

void ora_shadow_brute(const float *verts, const uint32_t *tri_idx, const uint8_t *vis,
                      uint32_t n_tris, const ora_ray *rays, uint32_t n_rays,
                      uint8_t *occluded, int threads)
{
	ora_shadow_brute_faces(verts, tri_idx, 3, vis, n_tris, rays, n_rays, occluded, threads);
}

void ora_retest_faces(const float *verts, const uint32_t *idx, uint32_t stride, const ora_ray *rays,
                      const uint32_t *prim, uint32_t n_rays, ora_hit *out)
{
	for(uint32_t i = 0; i < n_rays; ++i)
	{
		ora_hit h = {ORA_NO_HIT, 0.f, 0.f, 0.f};
		float t, u, v;
		if(prim[i] != ORA_NO_HIT && ora_face_intersect(verts, idx, stride, prim[i], rays[i].org, rays[i].dir, &t, &u, &v))
		{
			h.prim = prim[i]; h.t = t; h.u = u; h.v = v;
		}
		out[i] = h;
	}
}

void ora_retest(const float *verts, const uint32_t *tri_idx, const ora_ray *rays, const uint32_t *prim,
                uint32_t n_rays, ora_hit *out)
{
	ora_retest_faces(verts, tri_idx, 3, rays, prim, n_rays, out);
}

/* ------------------------------------------------------------------ */
/* FAST_MATH / FAST_TRIG (include/math/math.h:104-265)                 */

typedef union { int i; float f; } bits32;

static inline float polyexp(float x)
{
	return x * (x * (x * (x * (x * 1.8775767e-3f + 8.9893397e-3f) + 5.5826318e-2f) + 2.4015361e-1f) + 6.9315308e-1f) + 9.9999994e-1f;
}

float ora_fast_exp2(float x)
{
	const float f_hi = 129.00000f;
	const float f_lo = -126.99999f;
	float xl = (x < f_hi) ? x : f_hi;           /* std::min(x, f_hi) */
	xl = (xl < f_lo) ? f_lo : xl;               /* std::max(.., f_lo) */
	const int ipart = (int)(xl - 0.5f);
	const float fpart = xl - (float)ipart;
	bits32 e; e.i = (ipart + 127) << 23;
	return e.f * polyexp(fpart);
}

static inline float polylog(float x)
{
	return x * (x * (x * (x * (x * -3.4436006e-2f + 3.1821337e-1f) + -1.2315303f) + 2.5988452f) + -3.3241990f) + 3.1157899f;
}

float ora_fast_log2(float x)
{
	bits32 i; i.f = x;
	bits32 one; one.f = 1.f;
	bits32 m; m.i = (i.i & 0x7FFFFF) | one.i;
	const float e = (float)(((i.i & 0x7F800000) >> 23) - 127);
	return polylog(m.f) * (m.f - 1.f) + e;
}

float ora_fast_pow(float a, float b) { return ora_fast_exp2(ora_fast_log2(a) * b); }

/* math::ln2<> is the float rounding of ln 2 (include/math/math.h:39-60). */
float ora_fast_log(float a) { return ora_fast_log2(a) * (float)0.69314718055994530941723212145818L; }

float ora_fast_sin(float x)
{
	const float two_pi = (float)6.2831853071795864769252867665590L;
	const float pi = (float)3.1415926535897932384626433832795L;
	const float div_1_by_2pi = (float)0.15915494309189533576888376337251L;
	const float div_4_by_pi = (float)1.2732395447351626861510701069801L;
	const float div_4_by_sq_pi = (float)0.40528473456935108577551785283891L;
	if(x > two_pi || x < -two_pi) x -= (float)(int)(x * div_1_by_2pi) * two_pi;
	if(x < -pi) x += two_pi;
	else if(x > pi) x -= two_pi;
	x = div_4_by_pi * x - div_4_by_sq_pi * x * fabsf(x);
	const float result = 0.225f * (x * fabsf(x) - x) + x;
	if(result <= -1.f) return -1.f;
	else if(result >= 1.f) return 1.f;
	return result;
}

float ora_fast_cos(float x)
{
	const float half_pi = (float)1.5707963267948966192313216916398L;
	return ora_fast_sin(x + half_pi);
}
