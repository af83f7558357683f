/*
 * oracle.h -- CPU restatement of libYafaRay's ray-cast path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product.  Only tests/, the smoke()
 * entry point and the cpu_baseline leg of bench.py may load it.  The shipped
 * path (libyafaray_amd) never links or calls it.
 *
 * Every function names the reference file:line whose behaviour it restates.
 * Arithmetic is FP32 in the reference's operation order, compiled without FMA
 * contraction (-ffp-contract=off), so results are bit-identical to the
 * reference's x86-64 SSE build of the same expressions.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   ora_triangle_intersect  -- closed-form hits (axis-aligned, exactly
 *                              representable), double-precision re-derivation,
 *                              barycentric reconstruction, mutation cases.
 *   ora_quad_intersect      -- closed-form hits on both triangles with their
 *                              uv mappings, the second triangle tried only for
 *                              u outside [0, 1] (a first-triangle v or t miss
 *                              is final), det == 0, non-planar quads.
 *   ora_bound_cross         -- closed-form slab intervals, the zero-direction
 *                              quirk, early-out cases.
 *   ora_tree_bound          -- closed form of the 0.1% enlargement.
 *   ora_closest_brute       -- analytic heightfield hits (vertical rays),
 *                              double-precision brute force, kd-tree port.
 *   ora_kd_*                -- equals ora_*_brute on every non-tie ray.
 */
#ifndef YMX_ORACLE_H
#define YMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ray record shared with the product ABI: origin, tmin, direction, tmax
 * (tmax < 0 means "unbounded", include/geometry/ray.h:45-47). */
typedef struct { float org[3]; float tmin; float dir[3]; float tmax; } ora_ray;

/* Hit record: prim id (0xFFFFFFFF = miss), t, u, v. */
typedef struct { uint32_t prim; float t; float u; float v; } ora_hit;

#define ORA_NO_HIT 0xFFFFFFFFu
#define ORA_NO_VERTEX 0xFFFFFFFFu  /* 4th index of a triangle in a face list */
#define ORA_VIS_VISIBLE 1u   /* include/common/visibility.h:28-33 */
#define ORA_VIS_SHADOWS 2u

/* include/geometry/shape/shape_polygon.h:126-176 (N == 3). Returns 1 on hit. */
int ora_triangle_intersect(const float v0[3], const float v1[3], const float v2[3],
                           const float from[3], const float dir[3],
                           float *t, float *u, float *v);

/* include/geometry/shape/shape_polygon.h:126-176 (N == 4): the second
 * triangle (v0, v2, v3) only when the first triangle's u is outside [0, 1];
 * uv {u + v, v} or {u, u + v}.  Returns 1 on hit. */
int ora_quad_intersect(const float v0[3], const float v1[3], const float v2[3], const float v3[3],
                       const float from[3], const float dir[3], float *t, float *u, float *v);

/* Face lists: `stride` 3 (triangles) or 4 (idx[4 p + 3] == ORA_NO_VERTEX for
 * a triangle, else a quad v0 v1 v2 v3).  A face is one primitive, as a
 * PrimitivePolygon<float, 3 or 4> is (src/geometry/object/object_mesh.cc:96-112). */
int ora_face_nverts(const uint32_t *idx, uint32_t stride, uint32_t p);
int ora_face_intersect(const float *verts, const uint32_t *idx, uint32_t stride, uint32_t p, const float from[3],
                       const float dir[3], float *t, float *u, float *v);

/* include/geometry/bound.h:157-198. Returns 1 when crossed. */
int ora_bound_cross(const float a[3], const float g[3], const float from[3], const float dir[3],
                    float t_max, float *enter, float *leave);

/* src/accelerator/accelerator_kdtree_original.cc:89-103 */
void ora_tree_bound(const float *verts, const uint32_t *tri_idx, uint32_t n_tris,
                    float a_out[3], float g_out[3]);

void ora_tree_bound_faces(const float *verts, const uint32_t *idx, uint32_t stride, uint32_t n_faces,
                          float a_out[3], float g_out[3]);

/* include/accelerator/accelerator.h:64 */
float ora_dynamic_bias(float enter, float leave);

/* Brute-force closest hit: Accelerator::intersect(ray, camera) semantics
 * (include/accelerator/accelerator.h:89-101, 122-133) over every primitive in
 * id order; an exact tie in t keeps the lowest id. */
void ora_closest_brute(const float *verts, const uint32_t *tri_idx, const uint8_t *vis,
                       uint32_t n_tris, const ora_ray *rays, uint32_t n_rays,
                       ora_hit *hits, int threads);

/* Brute-force any hit: Accelerator::intersectShadow(ray, t_max) semantics with
 * t_max = ray.tmax (< 0: FLT_MAX), include/accelerator/accelerator_kdtree_common.h:139,
 * include/accelerator/accelerator.h:135-145.  occluded[i] = 0/1. */
void ora_shadow_brute(const float *verts, const uint32_t *tri_idx, const uint8_t *vis,
                      uint32_t n_tris, const ora_ray *rays, uint32_t n_rays,
                      uint8_t *occluded, int threads);

/* The same over a face list (triangles and quads). */
void ora_closest_brute_faces(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                             uint32_t n_faces, const ora_ray *rays, uint32_t n_rays, ora_hit *hits, int threads);
void ora_shadow_brute_faces(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                            uint32_t n_faces, const ora_ray *rays, uint32_t n_rays, uint8_t *occluded, int threads);
void ora_retest_faces(const float *verts, const uint32_t *idx, uint32_t stride, const ora_ray *rays,
                      const uint32_t *prim, uint32_t n_rays, ora_hit *out);

/* Closest hit by brute force over a subset of a scene's faces, with the
 * whole scene's tree bound a, g (ora_tree_bound of all faces): the bound the
 * queries clip to and take their dynamic bias from.  When the subset holds
 * every face a ray segment can reach, this is the whole scene's answer (with
 * ids into the subset). */
void ora_closest_brute_in_bound(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                                uint32_t n_faces, const float a[3], const float g[3], const ora_ray *rays,
                                uint32_t n_rays, ora_hit *hits, int threads);

/* Re-run the triangle test for given (ray, primitive) pairs: out[i] is the
 * ora_hit of ray i against primitive prim[i] (prim[i] == ORA_NO_HIT: skipped,
 * out[i] = {ORA_NO_HIT, 0, 0, 0}). */
void ora_retest(const float *verts, const uint32_t *tri_idx, const ora_ray *rays, const uint32_t *prim,
                uint32_t n_rays, ora_hit *out);

/* ---- kd-tree port (src/accelerator/accelerator_kdtree_original.cc) ---- */
typedef struct ora_kdtree ora_kdtree;

/* Build with the reference's SAH (pigeonhole bins for > 128 prims, sorted
 * edges below), max depth 7 + 1.66 ln N, leaf size 1, cost ratio 0.8, empty
 * bonus 0.33.  Polygon clipping of small nodes (PRIMITIVE_CLIPPING) is not
 * restated; nodes keep whole-primitive bounds, which changes tree shape but
 * not query results.  Returns NULL on allocation failure. */
ora_kdtree *ora_kd_build(const float *verts, const uint32_t *tri_idx, const uint8_t *vis,
                         uint32_t n_tris);
ora_kdtree *ora_kd_build_faces(const float *verts, const uint32_t *idx, uint32_t stride, const uint8_t *vis,
                               uint32_t n_faces);
void ora_kd_free(ora_kdtree *kd);
void ora_kd_bound(const ora_kdtree *kd, float a[3], float g[3]);
uint32_t ora_kd_num_nodes(const ora_kdtree *kd);
uint32_t ora_kd_num_prims(const ora_kdtree *kd);

/* One ray through the kd-tree port: shadow = 0 closest (t_max rule of
 * Accelerator::intersect applied by the caller), 1 any hit.  Returns 1 on hit. */
int ora_kd_intersect1(const ora_kdtree *kd, const ora_ray *ray, float t_max, int shadow, ora_hit *hit);
/* Triangles the tree was built over. */
uint32_t ora_kd_size(const ora_kdtree *kd);

/* accelerator_kdtree_common.h:107-255, TransparentShadow, with
 * Accelerator::primitiveIntersectionTransparentShadow (accelerator.h:147-168):
 * t_min = max(ray tmin, dynamic bias); returns 1 when occluded (an opaque
 * shadow caster, or a transparent one beyond max_depth of them), else 0 with
 * the transparent[] shadow casters met, in traversal order, in prims[0 .. *count). */
int ora_kd_transparent1(const ora_kdtree *kd, const ora_ray *ray, float t_max, const uint8_t *transparent,
                        int max_depth, uint32_t *prims, int *count);

/* include/accelerator/accelerator_kdtree_common.h:107-255, Nearest. */
void ora_kd_closest(const ora_kdtree *kd, const ora_ray *rays, uint32_t n_rays,
                    ora_hit *hits, int threads);
/* Same, Shadow: occluded[i] = 0/1 and, if prim != NULL, the occluder id. */
void ora_kd_shadow(const ora_kdtree *kd, const ora_ray *rays, uint32_t n_rays,
                   uint8_t *occluded, uint32_t *prim, int threads);

/* include/math/math.h:116-145 (FAST_MATH log2/exp2/pow) and :184-190 (log). */
float ora_fast_log2(float x);
float ora_fast_exp2(float x);
float ora_fast_pow(float a, float b);
float ora_fast_log(float a);
/* include/math/math.h:228-265 (FAST_TRIG sin/cos). */
float ora_fast_sin(float x);
float ora_fast_cos(float x);

#ifdef __cplusplus
}
#endif
#endif
