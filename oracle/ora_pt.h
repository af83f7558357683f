/*
 * ora_pt.h -- CPU restatement of libYafaRay's `pathtracing` SurfaceIntegrator
 * (TEST INFRASTRUCTURE ONLY: the checker of the GPU path tracer, and the CPU
 * baseline of bench.py; never linked into the shipped library).
 *
 * Restated: TiledIntegrator::renderTile's sampling
 * (src/integrator/surface/integrator_tiled.cc:253-395), PathIntegrator::integrate
 * (src/integrator/surface/integrator_path_tracer.cc:138-314), the Monte-Carlo
 * direct lighting (src/integrator/surface/integrator_montecarlo.cc:92-440),
 * ShinyDiffuseMaterial (src/material/material_shiny_diffuse.cc:141-485),
 * GlossyMaterial without anisotropy or Oren-Nayar (src/material/material_glossy.cc:130-430,
 * include/material/material_utils_microfacet.h), MirrorMaterial
 * (src/material/material_mirror.cc:64-87), GlassMaterial without dispersion or
 * absorption (src/material/material_glass.cc:157-378), LightMaterial
 * (src/material/material_light.cc:69-92), the recursive raytrace with its
 * glossy ray division (src/integrator/surface/integrator_montecarlo.cc:506-748),
 * AreaLight / PointLight / ObjectLight (src/light/light_area.cc:99-163,
 * src/light/light_point.cc:70-97, src/light/light_object_light.cc:87-237 with
 * Pdf1D, include/sampler/sample_pdf1d.h:60-92), PerspectiveCamera::shootRay
 * (src/camera/camera_perspective.cc:97-111,165-180), the samplers
 * (include/sampler/sample.h:43-148, include/sampler/halton.h:30-81,
 * src/sampler/halton.cc:419-440) and ImageFilm::addSample with the Gauss
 * filter table (src/render/imagefilm.cc:231-264,771-822).
 *
 * Deterministic replacements for the reference's two non-reproducible
 * streams (the same in the GPU path tracer):
 *   - Russian roulette draws ymx_rr_random(pixel, sample, depth, seed) instead
 *     of the tile-seeded multiply-with-carry generator seeded by rand()
 *     (integrator_tiled.cc:256);
 *   - the one-light pick of estimateOneDirectLight uses the Halton base-2
 *     index c = (pixel * total_samples + sample) * (path_samples * bounces)
 *     + sub * bounces + k + 1 (k: the path vertex) instead of the per-thread
 *     correlative counter (integrator_montecarlo.cc:110-114).
 * Both also take the integrate() call's node tag: 0 for the camera ray, and
 * tag * 11 + j for its recursive children (j = 1..8 the glossy branches,
 * 9 the specular reflection, 10 the refraction).  The roulette hashes
 * depth + 256 * tag; the light pick adds tag * 0x9E3779B9 to c (mod 2^32).
 * Film samples are splatted sample pass by sample pass, pixels in row-major
 * order (the reference's order depends on its thread schedule).
 */
#ifndef YMX_ORA_PT_H
#define YMX_ORA_PT_H

#include <stdint.h>

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_MAT_SHINYDIFFUSE 0u
#define ORA_MAT_GLOSSY 1u
#define ORA_MAT_MIRROR 2u
#define ORA_MAT_GLASS 3u
#define ORA_MAT_LIGHT 4u

/* One record for every material type; each type reads its own fields (the
 * reference's param names and defaults in the comments). */
typedef struct
{
	uint32_t type;            /* ORA_MAT_* */
	float color[3];           /* shinydiffuse "color" (1), glossy "color" (1), mirror "color" (1), light "color" (1) */
	float diffuse_reflect;    /* shinydiffuse (1), glossy (0) "diffuse_reflect" */
	float emit;               /* shinydiffuse "emit" (0) */
	uint32_t receive_shadows; /* "receive_shadows" (true) */
	float diffuse_color[3];   /* glossy "diffuse_color" (1) */
	float glossy_reflect;     /* glossy "glossy_reflect" (1) */
	float exponent;           /* glossy "exponent" (50) */
	uint32_t as_diffuse;      /* glossy "as_diffuse" (true) */
	float reflect;            /* mirror "reflect" (1) */
	float ior;                /* glass "IOR" (1.4) */
	float filter_color[3];    /* glass "filter_color" (1) */
	float transmit_filter;    /* glass "transmit_filter" (0) */
	float mirror_color[3];    /* glass "mirror_color" (1) */
	uint32_t fake_shadows;    /* glass "fake_shadows" (false) */
	float power;              /* light "power" (1) */
	uint32_t double_sided;    /* light "double_sided" (false) */
} ora_material;

#define ORA_LIGHT_AREA 0u
#define ORA_LIGHT_POINT 1u
#define ORA_LIGHT_OBJECT 2u
#define ORA_LIGHT_PORTAL 3u   /* "bgPortalLight" */

typedef struct
{
	uint32_t type;
	float corner[3], point1[3], point2[3]; /* area light */
	float from[3];                          /* point light */
	float color[3];
	float power;
	int32_t samples;                        /* area / object light samples (default 4), portal (16) */
	uint32_t cast_shadows;
	/* object and portal lights: the object is the triangle range [first_tri, first_tri + n_tris) */
	uint32_t first_tri, n_tris;
	uint32_t double_sided;                  /* object light "double_sided" (false) */
	float ibl_clamp_sampling;               /* portal "ibl_clamp_sampling" (0: off) */
} ora_light;

typedef struct
{
	float from[3], to[3], up[3];
	float focal;
	float aspect_factor;
	int32_t resx, resy;
	int32_t path_samples;
	int32_t bounces;
	int32_t rr_min_bounces;
	int32_t aa_samples;      /* AA_minsamples (one pass) */
	float aa_pixelwidth;     /* Gauss filter, default 1.5 */
	int32_t has_background;
	float background[3];
	uint32_t seed;
	int32_t raydepth;        /* MonteCarloIntegrator "raydepth" (5) */
	int32_t transparent_shadows;  /* "transpShad" (false) */
	int32_t shadow_depth;    /* "shadowDepth" (4), at most ORA_MAX_SHADOW_DEPTH */
	/* multi-pass adaptive anti-aliasing, ImageFilm's AA params
	 * (include/render/imagefilm.h:212-226; 0 for aa_passes means 1) */
	int32_t aa_passes;                        /* "AA_passes" (1) */
	int32_t aa_inc_samples;                   /* "AA_inc_samples" (1) */
	float aa_threshold;                       /* "AA_threshold" (0.05) */
	float aa_resampled_floor;                 /* "AA_resampled_floor" (0, % of the pixels) */
	float aa_sample_multiplier_factor;        /* "AA_sample_multiplier_factor" (1) */
	float aa_light_sample_multiplier_factor;  /* "AA_light_sample_multiplier_factor" (1) */
	int32_t aa_detect_color_noise;            /* "AA_detect_color_noise" (false) */
	int32_t aa_dark_detection_type;           /* "AA_dark_detection_type": 0 none, 1 linear, 2 curve */
	float aa_dark_threshold_factor;           /* "AA_dark_threshold_factor" (0) */
	int32_t aa_variance_edge_size;            /* "AA_variance_edge_size" (10) */
	int32_t aa_variance_pixels;               /* "AA_variance_pixels" (0) */
	float aa_clamp_samples;                   /* "AA_clamp_samples" (0: no clamp) */
	/* vertex normals at the triangles' corners, 12 floats per triangle (NULL:
	 * every triangle flat): corner k is (nx, ny, nz, w) at 4 k; w < 0 in
	 * corner 0 marks a flat triangle (n = ng), w == 0 a corner without a
	 * normal (the face normal stands in), w > 0 a corner normal */
	const float *corner_normals;
	/* closest hits by brute force over every triangle (exact ties in t to the
	 * lowest id, as the GPU accelerator and ora_closest_brute resolve them)
	 * instead of the kd-tree port, whose leaf order decides ties and whose
	 * unclipped leaves can miss a ray grazing an edge; for scenes small
	 * enough to test every triangle per ray */
	int32_t trace_brute;
} ora_pt_settings;

#define ORA_MAX_SHADOW_DEPTH 64

/* TiledIntegrator::render (src/integrator/surface/integrator_tiled.cc:98-216):
 * pass 0 renders aa_samples samples of every pixel; each further pass of
 * aa_passes renders ceil(aa_inc_samples * multiplier) more samples of the
 * pixels ImageFilm::nextPass flags as noisy, with the light sample multiplier
 * of that pass.  Film sums: film_rgba[4 * (y * resx + x)]
 * (colour * weight, alpha * weight) and film_w[y * resx + x] (weights).
 * Buffers are accumulated into (zero them first).  vis: the visibility mask of
 * every triangle (the one kd was built with; object lights read it), NULL: all
 * ORA_VIS_VISIBLE | ORA_VIS_SHADOWS.  Returns 0 (film untouched) for a scene
 * it cannot set up (a portal light without a background, an empty object
 * light, too deep transparent shadows, no memory), else 1. */
int ora_pt_render(const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx, const uint32_t *mat_id,
                   const uint8_t *vis, const ora_material *mats, const ora_light *lights, uint32_t n_lights,
                   const ora_pt_settings *st, float *film_rgba, float *film_w, int threads);

/* The first pass of ora_pt_render (aa_samples samples, light multiplier 1)
 * over the pixels with flags[y * resx + x] set only: each of their samples is
 * splatted into the full-size film as in ora_pt_render, so a pixel whose
 * neighbours within the filter width are all flagged gets the sums a full
 * render gives it.  For checks on windows of frames too large to render on
 * the CPU.  Returns 0 or 1 as ora_pt_render. */
int ora_pt_render_pixels(const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx, const uint32_t *mat_id,
                          const uint8_t *vis, const ora_material *mats, const ora_light *lights, uint32_t n_lights,
                          const ora_pt_settings *st, const uint8_t *flags, float *film_rgba, float *film_w,
                          int threads);

/* Same, but only the sample values of one pixel sample (no film), for
 * sample-level checks: col_alpha[4], and the sub-pixel offsets dx, dy. */
void ora_pt_pixel_sample(const ora_kdtree *kd, const float *verts, const uint32_t *tri_idx, const uint32_t *mat_id,
                         const uint8_t *vis, const ora_material *mats, const ora_light *lights, uint32_t n_lights,
                         const ora_pt_settings *st, int x, int y, int sample, float col_alpha[4], float dxdy[2]);

/* ImageFilm::nextPass's resampling flags (src/render/imagefilm.cc:398-513)
 * for the film sums given and the current AA threshold: flags[y * resx + x]
 * set for the pixels the next pass renders; returns their count.  A
 * threshold <= 0 flags every pixel (ImageFilm::doMoreSamples). */
int ora_aa_next_pass(const ora_pt_settings *st, float threshold, const float *film_rgba, const float *film_w,
                     uint8_t *flags);

/* The most AA samples a pixel can receive over all passes (integrator_tiled.cc:258-262). */
int ora_aa_max_samples(const ora_pt_settings *st);

/* PerspectiveCamera::shootRay direction for film coordinates (px, py). */
void ora_pt_camera_dir(const ora_pt_settings *st, float px, float py, float dir[3]);

/* Samplers (include/sampler/sample.h:100-148, halton.h, halton.cc:419-440). */
float ora_ri_vdc(uint32_t bits, uint32_t r);
float ora_ri_s(uint32_t i, uint32_t r);
float ora_ri_lp(uint32_t i, uint32_t r);
uint32_t ora_fnv32a(uint32_t value);
double ora_low_discrepancy(int dim, uint32_t n);
float ora_halton_next(int base, uint32_t start);  /* Halton(base, start).getNext() */
int ora_faure_perm(int base, int digit);
float ora_rr_random(uint32_t pixel, uint32_t sample, uint32_t depth, uint32_t seed);
/* Gauss filter table of ImageFilm (16 x 16 floats) and its scale. */
void ora_filter_table(float pixelwidth, float table[256], float *table_scale, float *filter_width);

/* Single material functions on a flat surface with normal ng (n = ng), for the
 * oracle's own tests: eval, pdf, sample (out: wi[3], colour[3], w, pdf,
 * sampled flags), getSpecular (out: has_r, dir_r[3], col_r[3], has_t,
 * dir_t[3], col_t[3]), emit; Fresnel (out: kr, kt) and refraction. */
uint32_t ora_mat_flags(const ora_material *m);
void ora_bsdf_eval(const ora_material *m, const float ng[3], const float wo[3], const float wl[3], uint32_t bsdfs,
                   float out[3]);
float ora_bsdf_pdf(const ora_material *m, const float ng[3], const float wo[3], const float wi[3], uint32_t bsdfs);
void ora_bsdf_sample(const ora_material *m, const float ng[3], const float wo[3], float s1, float s2, uint32_t flags,
                     float out[9]);
void ora_specular(const ora_material *m, const float ng[3], const float wo[3], int ray_level, float out[14]);
void ora_fresnel(const float i[3], const float n[3], float ior, float out[2]);
int ora_refract(const float n[3], const float wi[3], float ior, float wo[3]);
void ora_mat_emit(const ora_material *m, const float ng[3], const float wo[3], float out[3]);

/* ObjectLight (src/light/light_object_light.cc:87-237) on its own, for the
 * oracle's pinning tests.  setup: returns area_ and writes the area cdf
 * (n_tris floats).  sample: illumSample from p; out = (ldir, dist, pdf),
 * returns 1 on a sample.  intersect: Light::intersect along dir; returns 1 on
 * a hit with its inverse pdf.  -1: not an object light. */
float ora_object_light_setup(const float *verts, const uint32_t *tri_idx, const ora_light *l, float *cdf);
int ora_object_light_sample(const float *verts, const uint32_t *tri_idx, const ora_light *l, const float p[3],
                            float s1, float s2, float out[5]);
int ora_object_light_intersect(const float *verts, const uint32_t *tri_idx, const uint8_t *vis, const ora_light *l,
                               const float from[3], const float dir[3], float tmin, float *ipdf);

#ifdef __cplusplus
}
#endif
#endif
