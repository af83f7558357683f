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
 * All pixels of one AA sample index are traced as one wavefront: camera
 * rays, closest-hit extension, shading with next-event-estimation shadow rays
 * queued and traced in bulk, then a deterministic film gather.  Conventions
 * as in ymx_accel.h: int status, ymx_last_error(), hipStream_t as void *.
 * A handle's calls must not overlap in time (it owns one set of wavefront
 * buffers and the film).
 *
 * Materials: shinydiffuse (diffuse colour, diffuse_reflect, emit).  Lights:
 * area (corner / point1 / point2, colour, power, samples) and point (from,
 * colour, power).  Param meanings and defaults follow the reference's
 * ParamMaps (include/material/material_shiny_diffuse.h:85-96,
 * include/light/light_area.h:50-57, include/light/light_point.h:52-54,
 * include/integrator/surface/integrator_path_tracer.h:55-58,
 * include/render/imagefilm.h:196-225, include/camera/camera.h:82-89).
 */
#ifndef YMX_PT_H
#define YMX_PT_H

#include <stdint.h>

#include "ymx_accel.h"

#ifdef __cplusplus
extern "C" {
#endif

#define YMX_MAT_SHINYDIFFUSE 0u
#define YMX_LIGHT_AREA 0u
#define YMX_LIGHT_POINT 1u

typedef struct ymx_material
{
	uint32_t type;            /* YMX_MAT_* */
	float color[3];           /* "color" */
	float diffuse_reflect;    /* "diffuse_reflect", default 1 */
	float emit;               /* "emit", default 0 */
	uint32_t receive_shadows; /* "receive_shadows", default 1 */
} ymx_material;

typedef struct ymx_light
{
	uint32_t type;                          /* YMX_LIGHT_* */
	float corner[3], point1[3], point2[3];  /* area light quad */
	float from[3];                          /* point light position */
	float color[3];
	float power;
	int32_t samples;                        /* area light samples, default 4 */
	uint32_t cast_shadows;                  /* default 1 */
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
} ymx_pt_settings;

typedef struct ymx_pt ymx_pt;

/* Create a path tracer over an accelerator built from the same triangles.
 * verts/tri_idx: host arrays as given to ymx_accel_create (for face normals);
 * mat_id: host, one material index per triangle. */
int ymx_pt_create(const ymx_accel *accel, const float *verts, uint32_t n_verts, const uint32_t *tri_idx,
                  const uint32_t *mat_id, uint32_t n_tris, const ymx_material *materials, uint32_t n_materials,
                  const ymx_light *lights, uint32_t n_lights, const ymx_pt_settings *settings, ymx_pt **out);

int ymx_pt_destroy(ymx_pt *pt);

/* Zero the film (colour and weight sums). */
int ymx_pt_clear_film(ymx_pt *pt, void *stream);

/* Trace AA samples [first_sample, first_sample + n_samples) of every pixel and
 * splat them into the film, asynchronously on `stream`. */
int ymx_pt_render(ymx_pt *pt, int32_t first_sample, int32_t n_samples, void *stream);

/* Copy the film sums to host memory (synchronous): rgba[4 * (y * resx + x)]
 * holds sum(weight * colour, weight * alpha) and weight[y * resx + x] the
 * filter weights; the pixel value is rgba / weight (ImageFilm::flush). */
int ymx_pt_read_film(ymx_pt *pt, float *rgba, float *weight);

/* Counters of the last ymx_pt_render: rays cast (camera + extension +
 * shadow), and kernel time in ms of the extension and shadow launches. */
typedef struct ymx_pt_stats
{
	uint32_t struct_size;
	uint64_t camera_rays;
	uint64_t extension_rays;
	uint64_t shadow_rays;
	float render_ms;
} ymx_pt_stats;

int ymx_pt_get_stats(const ymx_pt *pt, ymx_pt_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
