/*
 * ymx_pt.h -- C ABI of the MI355X wavefront path tracer (libymx.so).
 *
 * Replaces libYafaRay's `pathtracing` SurfaceIntegrator render loop:
 * SurfaceIntegrator::render -> TiledIntegrator::render/renderTile ->
 * PathIntegrator::integrate -> ImageFilm::addSample
 * (src/integrator/surface/integrator_surface.cc:227-242,
 *  src/integrator/surface/integrator_tiled.cc:98-395,
 *  src/integrator/surface/integrator_path_tracer.cc:138-314,
 *  src/render/imagefilm.cc:771-822), with the scene's accelerator built by
 * ymx_accel_create (include/ymx_accel.h).
 *
 * Several AA sample indices of all pixels are traced as one wavefront: camera
 * rays, closest-hit extension, shading with next-event-estimation shadow rays
 * queued and traced in bulk, the recursive specular / glossy rays level by
 * level, then a deterministic film gather.  Conventions
 * as in ymx_accel.h: int status, ymx_last_error(), hipStream_t as void *.
 * A handle's calls must not overlap in time (it owns one set of wavefront
 * buffers and the film).
 *
 * Materials: shinydiffuse (diffuse colour, diffuse_reflect, emit), glossy
 * (isotropic Blinn / Ashikhmin-Shirley with optional Lambert diffuse), mirror,
 * glass (no dispersion or absorption) and light_mat; specular and glossy
 * (as_diffuse = 0) surfaces recurse as the reference's recursiveRaytrace does,
 * up to `raydepth`.  Lights: area (corner / point1 / point2, colour, power,
 * samples), point (from, colour, power), object (a mesh of the scene: a
 * triangle range, colour, power, samples, double_sided) and background portal
 * (a mesh of the scene through which the constant background lights the
 * scene: a triangle range, power, samples, ibl_clamp_sampling; one-sided, the
 * background's colour times power).  Object and portal lights of more than 16
 * triangles test BSDF-sampled rays against a BVH of their own, as the
 * reference's lights do with their private kd-trees.  Param meanings and defaults follow
 * the reference's ParamMaps (include/material/material_shiny_diffuse.h:85-96,
 * include/material/material_glossy.h:68-80, include/material/material_mirror.h:57-58,
 * include/material/material_glass.h:68-75, include/material/material_light.h:59-61,
 * include/integrator/surface/integrator_montecarlo.h:47,
 * include/light/light_area.h:50-57, include/light/light_point.h:52-54,
 * include/light/light_object_light.h:53-62, include/light/light_background_portal.h:58-61,
 * include/integrator/surface/integrator_path_tracer.h:55-58,
 * include/render/imagefilm.h:196-225, include/camera/camera.h:82-89).
 */
#ifndef YMX_PT_H
#define YMX_PT_H

#include <stdint.h>

#include "ymx_accel.h"
#include "ymx_guide.h"

#ifdef __cplusplus
extern "C" {
#endif

#define YMX_MAT_SHINYDIFFUSE 0u /* "shinydiffusemat" */
#define YMX_MAT_GLOSSY 1u       /* "glossy" */
#define YMX_MAT_MIRROR 2u       /* "mirror" */
#define YMX_MAT_GLASS 3u        /* "glass" */
#define YMX_MAT_LIGHT 4u        /* "light_mat" */
#define YMX_LIGHT_AREA 0u
#define YMX_LIGHT_POINT 1u
#define YMX_LIGHT_OBJECT 2u     /* "objectlight" */
#define YMX_LIGHT_PORTAL 3u     /* "bgPortalLight" */

/* One record for every material type; each type reads its own fields, named
 * after the reference's params (defaults in brackets). */
typedef struct ymx_material
{
	uint32_t type;            /* YMX_MAT_* */
	float color[3];           /* "color" of shinydiffuse, glossy, mirror, light_mat [1] */
	float diffuse_reflect;    /* "diffuse_reflect": shinydiffuse [1], glossy [0] */
	float emit;               /* shinydiffuse "emit" [0] */
	uint32_t receive_shadows; /* "receive_shadows" [1] */
	float diffuse_color[3];   /* glossy "diffuse_color" [1] */
	float glossy_reflect;     /* glossy "glossy_reflect" [1] */
	float exponent;           /* glossy "exponent" (Blinn) [50] */
	uint32_t as_diffuse;      /* glossy "as_diffuse" [1] */
	float reflect;            /* mirror "reflect" [1] */
	float ior;                /* glass "IOR" [1.4] */
	float filter_color[3];    /* glass "filter_color" [1] */
	float transmit_filter;    /* glass "transmit_filter" [0] */
	float mirror_color[3];    /* glass "mirror_color" [1] */
	uint32_t fake_shadows;    /* glass "fake_shadows" [0] */
	float power;              /* light_mat "power" [1] */
	uint32_t double_sided;    /* light_mat "double_sided" [0] */
} ymx_material;

typedef struct ymx_light
{
	uint32_t type;                          /* YMX_LIGHT_* */
	float corner[3], point1[3], point2[3];  /* area light quad */
	float from[3];                          /* point light position */
	float color[3];
	float power;
	int32_t samples;                        /* area / object light samples, default 4; portal 16 */
	uint32_t cast_shadows;                  /* default 1 */
	/* object and portal lights: the object ("object_name") is the triangle range
	 * [first_tri, first_tri + n_tris) of the path tracer's triangles */
	uint32_t first_tri, n_tris;
	uint32_t double_sided;                  /* object light "double_sided", default 0 */
	/* portal light "ibl_clamp_sampling": the colour its intersect returns is
	 * clamped proportionally to this maximum, 0 off [0] */
	float ibl_clamp_sampling;
} ymx_light;

typedef struct ymx_pt_settings
{
	float from[3], to[3], up[3];  /* perspective camera (pinhole) */
	float focal;                  /* "focal", default 1 */
	float aspect_factor;          /* "aspect_ratio_factor", default 1 */
	int32_t resx, resy;
	int32_t path_samples;         /* "path_samples", default 32 */
	int32_t bounces;              /* "bounces", default 3 */
	int32_t rr_min_bounces;       /* "russian_roulette_min_bounces", default 0 */
	int32_t aa_samples;           /* "AA_minsamples" (single AA pass) */
	float aa_pixelwidth;          /* "AA_pixelwidth", Gauss filter, default 1.5 */
	int32_t has_background;       /* constant background seen by camera rays */
	float background[3];
	uint32_t seed;                /* seed of the Russian-roulette stream */
	int32_t raydepth;             /* "raydepth" of the recursive raytrace, default 5 (0..19) */
	int32_t transparent_shadows;  /* "transpShad": shadow rays pass transparent materials, default 0 */
	int32_t shadow_depth;         /* "shadowDepth": most transparent casters a shadow ray passes, default 4 (0..64) */
	/* Multi-pass adaptive anti-aliasing (TiledIntegrator::render, ymx_pt_render_passes),
	 * ImageFilm's AA params (include/render/imagefilm.h:212-226); aa_passes 0 means 1 */
	int32_t aa_passes;                        /* "AA_passes" [1] */
	int32_t aa_inc_samples;                   /* "AA_inc_samples": samples of each further pass [1] */
	float aa_threshold;                       /* "AA_threshold" [0.05] */
	float aa_resampled_floor;                 /* "AA_resampled_floor": % of the pixels [0] */
	float aa_sample_multiplier_factor;        /* "AA_sample_multiplier_factor" [1] */
	float aa_light_sample_multiplier_factor;  /* "AA_light_sample_multiplier_factor" [1] */
	int32_t aa_detect_color_noise;            /* "AA_detect_color_noise" [0] */
	int32_t aa_dark_detection_type;           /* "AA_dark_detection_type": 0 none, 1 linear, 2 curve [0] */
	float aa_dark_threshold_factor;           /* "AA_dark_threshold_factor" [0] */
	int32_t aa_variance_edge_size;            /* "AA_variance_edge_size" [10] */
	int32_t aa_variance_pixels;               /* "AA_variance_pixels" [0] */
	float aa_clamp_samples;                   /* "AA_clamp_samples": proportional clamp of each sample, 0 off [0] */
	/* Device memory for the wavefront state in MiB: a wavefront holds as many AA
	 * samples of every pixel as fit (at least one); 0 = a quarter of the free
	 * device memory at creation, at most 64 GiB.  No counterpart in the reference
	 * (its tiles hold one sample at a time). */
	int32_t wavefront_mb;
} ymx_pt_settings;

typedef struct ymx_pt ymx_pt;

/* Create a path tracer over an accelerator built from the same triangles.
 * verts/tri_idx: host arrays as given to ymx_accel_create (for face normals);
 * mat_id: host, one material index per triangle; vis: host, the visibility
 * masks given to ymx_accel_create (NULL: all visible and casting shadows), read by object
 * lights, whose own intersection test sees their YMX_VIS_VISIBLE triangles only. */
int ymx_pt_create(const ymx_accel *accel, const float *verts, uint32_t n_verts, const uint32_t *tri_idx,
                  const uint32_t *mat_id, const uint8_t *vis, uint32_t n_tris, const ymx_material *materials, uint32_t n_materials,
                  const ymx_light *lights, uint32_t n_lights, const ymx_pt_settings *settings, ymx_pt **out);

int ymx_pt_destroy(ymx_pt *pt);

/* Shading normals of smooth meshes and meshes with vertex normals
 * (PrimitivePolygon::getSurfacePolygon, src/geometry/primitive/primitive_polygon.cc:45-51;
 * the normals come from yafaray_addNormal or yafaray_smoothObjectMesh,
 * src/geometry/object/object_mesh.cc:150-264).  corner_normals: host, 12
 * floats per triangle, corner k at 4 k as (nx, ny, nz, w): w > 0 a corner
 * normal, w == 0 a corner without one (the face normal stands in, as
 * FacePrimitive::getVertexNormal does), and w < 0 in corner 0 a flat triangle
 * (shading normal = face normal).  A hit's shading normal is the corners'
 * normals weighted by its barycentric coordinates (1 - u - v, u, v), then
 * normalised.  NULL makes every triangle flat again.  Not with transparent
 * shadows through smooth fake-shadow glass (YMX_ERR_UNSUPPORTED). */
int ymx_pt_set_normals(ymx_pt *pt, const float *corner_normals);

/* Zero the film (colour and weight sums). */
int ymx_pt_clear_film(ymx_pt *pt, void *stream);

/* Trace AA samples [first_sample, first_sample + n_samples) of every pixel and
 * splat them into the film, asynchronously on `stream`. */
int ymx_pt_render(ymx_pt *pt, int32_t first_sample, int32_t n_samples, void *stream);

/* The whole multi-pass render of TiledIntegrator::render
 * (src/integrator/surface/integrator_tiled.cc:160-214): pass 0 renders
 * AA_minsamples samples of every pixel; each further pass of aa_passes renders
 * ceil(AA_inc_samples x sample multiplier) more samples of the pixels that
 * ImageFilm::nextPass (src/render/imagefilm.cc:365-520) flags from the film so
 * far, with that pass's light sample multiplier, and lowers the threshold when
 * fewer pixels than AA_resampled_floor were flagged.  The flags are computed
 * on the device; one host synchronisation per pass reads their count.  With
 * aa_passes > 1 every pass takes its sub-pixel offsets from the scrambled
 * (0, 2) sequence (riVdC, riS), as the reference does.  Renders into the film
 * as it is (clear it first for a new image).  Returns the pixels resampled in
 * each pass in resampled[] (aa_passes entries; may be NULL). */
int ymx_pt_render_passes(ymx_pt *pt, int32_t *resampled, void *stream);

/* Copy the film sums to host memory (synchronous): rgba[4 * (y * resx + x)]
 * holds sum(weight * colour, weight * alpha) and weight[y * resx + x] the
 * filter weights; the pixel value is rgba / weight (ImageFilm::flush). */
int ymx_pt_read_film(ymx_pt *pt, float *rgba, float *weight);

/* Save the film as ImageFilm::imageFilmSave writes it
 * (src/render/imagefilm.cc:1099-1176): the NUL-terminated tag
 * "YAF_FILMv4_0_0", ten int32 (computer node, base sampling offset, sampling
 * offset, width, height, x0, x1, y0, y1, image layers = 1), the filter
 * weights row by row, then the combined layer's weighted RGBA sums row by
 * row, all in host byte order.  The film covers the whole image (x0 = y0 =
 * 0).  Synchronous. */
int ymx_pt_save_film(ymx_pt *pt, const char *path, int32_t computer_node, int32_t base_sampling_offset,
                     int32_t sampling_offset);

/* Load a film written in that format (ImageFilm::imageFilmLoad,
 * src/render/imagefilm.cc:910-1018): the tag and the size, borders and layer
 * count must match this handle's film (YMX_ERR_INVALID_ARG otherwise, the
 * film unchanged).  add = 0 replaces the film; add != 0 adds the loaded sums
 * to it, as imageFilmLoadAllInFolder merges the films of several render
 * nodes (:1072-1094).  The three offsets read are returned where the pointers
 * are not NULL.  Synchronous. */
int ymx_pt_load_film(ymx_pt *pt, const char *path, int32_t add, int32_t *computer_node,
                     int32_t *base_sampling_offset, int32_t *sampling_offset);

/* Path guiding.  No counterpart in the reference, whose path tracer samples
 * the BSDF only (src/integrator/surface/integrator_path_tracer.cc:190-240).
 * With a guide (include/ymx_guide.h) set, the next direction at shinydiffuse
 * path vertices is drawn by one-sample MIS: the BSDF with probability
 * bsdf_prob (clamped to [0.2, 1]), the guide's mixture q otherwise, with the
 * weight F / (bsdf_prob p + (1 - bsdf_prob) q), where p is the BSDF sampler's
 * density and F / p its plain weight -- the image converges to the same
 * expectation as without a guide.  train != 0 records (position, normal, wo,
 * direction, radiance weight) samples during renders and runs train_steps Adam
 * steps on them after each wavefront.  guide NULL switches guiding off.  The
 * guide must outlive its use by this handle. */
int ymx_pt_set_guide(ymx_pt *pt, ymx_guide *guide, float bsdf_prob, int32_t train, uint32_t train_steps);

/* Limit guiding to the sub-path's start vertex and its next max_depth path
 * vertices (deeper vertices sample the BSDF alone and cost no guide query);
 * max_depth < 0: every vertex (the default).  No counterpart in the
 * reference.  Applies to the guide set now and to later ones. */
int ymx_pt_set_guide_depth(ymx_pt *pt, int32_t max_depth);

/* Learned selection probability (no counterpart in the reference).  The
 * probability of sampling the BSDF at a guided vertex is kept per region of
 * the guide's bounds (16 x 16 x 16 cells) in a table of the guide, and learned
 * while the guide trains: Adam (learning rate lr) on the gradient of the
 * second moment of the one-sample MIS estimator, estimated from the training
 * samples, between 0.2 and 1.  The estimator stays unbiased for any value.
 * mode 1: start every region at the bsdf_prob of ymx_pt_set_guide and learn;
 * mode 2: use (and while training, go on learning) the guide's table as
 * another tracer left it; mode 0: back to the fixed bsdf_prob.  Call after
 * ymx_pt_set_guide; later ymx_pt_set_guide calls on this tracer keep the mode. */
int ymx_pt_set_guide_selection(ymx_pt *pt, int32_t mode, float lr);

/* The selection probability of every region in use (1 in a region fallen
 * back to BSDF sampling): n >= 4096 floats, cell (x, y, z) at (z 16 + y) 16 + x. */
int ymx_pt_get_guide_selection(const ymx_pt *pt, float *alpha, uint32_t n);

/* Per-region fallback to BSDF sampling (no counterpart in the reference).
 * While the guide trains, each region of the selection table keeps running
 * estimates (weight 0.9 on the past, 0.1 on each training step) of the second
 * moment of its guided one-sample estimator, at the region's selection
 * probability, and of BSDF sampling alone, both from the training samples'
 * draws; a region whose guided moment is the larger, once it has seen 256
 * training records, samples the BSDF alone (selection probability 1) until
 * the estimates cross back.  The estimator
 * stays unbiased either way.  mode 1: start every region with no estimate
 * (guided); mode 2: use (and while training, go on updating) the estimates
 * another tracer left in the guide's table; mode 0: off.  Call after
 * ymx_pt_set_guide; later ymx_pt_set_guide calls on this tracer keep it. */
int ymx_pt_set_guide_fallback(ymx_pt *pt, int32_t mode);

/* The fallback's state, n >= 12288 floats: the running second moments of the
 * guided estimator for the 4096 regions, then BSDF sampling's (0: no estimate
 * yet), then the training records each region has seen. */
int ymx_pt_get_guide_fallback(const ymx_pt *pt, float *moments, uint32_t n);

/* Counters of the last ymx_pt_render: rays cast, and its duration in ms on
 * the stream (events around the whole render). */
typedef struct ymx_pt_stats
{
	uint32_t struct_size;
	uint64_t camera_rays;         /* camera rays, plus the rays of the recursive raytrace */
	uint64_t extension_rays;
	uint64_t shadow_rays;
	float render_ms;
	uint64_t guide_samples;       /* training samples the paths produced (before subsampling) */
	uint64_t guide_steps;         /* guide training steps taken */
	uint32_t light_trees;         /* object and portal lights tested through their own BVH */
} ymx_pt_stats;

int ymx_pt_get_stats(const ymx_pt *pt, ymx_pt_stats *stats);

/* Launch configuration of the wavefront, chosen per scene by ymx_pt_autotune
 * or set directly.  No counterpart in the reference (its tiles run on CPU
 * threads).  Every configuration gives the same film bit for bit. */
#define YMX_TRACE_WAVE_ORDERED 0    /* each wave traces 64 consecutive queued rays */
#define YMX_TRACE_REPLACEMENT 1     /* a lane whose ray is done takes the next one */
#define YMX_PT_TUNING_TRIALS 8
typedef struct ymx_pt_tuning
{
	uint32_t struct_size;
	int32_t samples_per_wavefront;  /* AA samples of every pixel per wavefront, 1 .. the allocated number */
	int32_t shadow_kernel;          /* YMX_TRACE_*, for the shadow queues (wave-ordered only with transparent shadows) */
	int32_t extension_kernel;       /* YMX_TRACE_*, for the extension queues */
	/* ymx_pt_autotune's measurements: per trial configuration, ms per AA sample */
	int32_t n_trials;
	int32_t trial_samples_per_wavefront[YMX_PT_TUNING_TRIALS];
	int32_t trial_shadow_kernel[YMX_PT_TUNING_TRIALS];
	int32_t trial_extension_kernel[YMX_PT_TUNING_TRIALS];
	float trial_ms_per_sample[YMX_PT_TUNING_TRIALS];
} ymx_pt_tuning;

/* Times short renders of the scene under each candidate configuration (both
 * kernels for each queue kind at the allocated wavefront size, then the best
 * of those at half and a quarter of that size), keeps the fastest, and
 * returns it with the measurements.  trial_samples: AA samples per trial
 * (0: the allocated samples per wavefront).  Synchronous; the film is left as
 * it was, and guide training is paused during the trials. */
int ymx_pt_autotune(ymx_pt *pt, int32_t trial_samples, ymx_pt_tuning *result, void *stream);

/* The configuration in use (struct_size set by the caller), or a new one. */
int ymx_pt_get_tuning(const ymx_pt *pt, ymx_pt_tuning *tuning);
int ymx_pt_set_tuning(ymx_pt *pt, const ymx_pt_tuning *tuning);

#ifdef __cplusplus
}
#endif
#endif
