// pt_internal.h -- device data of the wavefront path tracer.
#pragma once

#include "../../include/ymx_pt.h"
#include "accel_internal.h"
#include "pt_math.h"

namespace ymx {

enum BsdfBits : uint32_t
{
	kNone = 0, kSpecular = 1, kGlossy = 2, kDiffuse = 4, kDispersive = 8, kReflect = 16, kTransmit = 32,
	kFilter = 64, kEmit = 128, kVolumetric = 256, kAll = 511
};  // include/material/bsdf.h:25-67

struct PtMaterial
{
	Rgb color;
	float diffuse_reflect;
	Rgb emit_color;       // emit * color
	uint32_t flags;       // BsdfBits
	uint32_t receive_shadows;
};

struct PtLight
{
	uint32_t type;
	uint32_t cast_shadows;
	int32_t samples;
	float area;           // normal_flipped_.normalized().length()
	V3 corner, to_x, to_y, normal_flipped;
	V3 quad[4];           // corner, point1, point1 + point2 - corner, point2
	V3 from;              // point light
	Rgb color;            // colour * power (* pi for area lights)
};

constexpr int kMaxLights = 512;
constexpr int kLdsDims = 50;
constexpr int kLdsMaxBase = 228;

struct PtConst
{
	// camera (src/camera/camera.cc:89-110, src/camera/camera_perspective.cc:97-111)
	V3 cam_from, vright, vup, vto;
	int32_t resx, resy;
	int32_t path_samples, bounces, rr_min_bounces, aa_samples;
	uint32_t n_lights;
	uint32_t slots_camera;   // shadow slots of a camera vertex (all lights)
	uint32_t slots_path;     // shadow slots of a path vertex (one light)
	uint32_t has_background;
	Rgb background;
	uint32_t seed;
	// film
	float filter_width, filter_table_scale;
	float filter_table[256];
};

// Per-path state of one wavefront (one AA sample index for every pixel).
struct PtState
{
	float4 *col;        // rgb, alpha
	float4 *path_col;   // rgb
	float4 *tp_next;    // throughput including the next segment's BSDF weight
	float4 *tp_vertex;  // throughput at the current vertex (after roulette)
	float4 *pwo;        // outgoing direction at the current vertex
	float4 *cam_dir;    // camera ray direction
	ymx_hit *cam_hit;   // camera hit
	float2 *dxy;        // sub-pixel offsets
	uint32_t *offset;   // fnv32a(y * fnv32a(x))
	uint32_t *flags;    // bit 0 active, bit 1 add emission, bit 2 caustic, bit 3 direct light queued,
	                    // bits 8..16 first path vertex BSDF flags, bit 20 sampled != None
	int32_t *light;     // light picked at the current path vertex
	uint32_t *vert_mat; // material of the current path vertex
	// extension queues (double-buffered: segment d reads queue d % 2)
	ymx_ray *ext_rays[2];
	uint32_t *ext_path[2]; // path index of each queued extension ray
	uint32_t *ext_count;   // [2]
	ymx_hit *ext_hits;
	// shadow slots and queue
	float4 *slot_contrib;  // rgb, w = 1 if the slot contributes when unoccluded
	int32_t *slot_queue;   // queue index, -1: contributes without a shadow test
	ymx_ray *sh_rays;
	uint32_t *sh_occ;      // bit per queued shadow ray
	uint32_t *sh_count;
	uint32_t capacity;     // paths per wavefront
	uint32_t slots;        // shadow slots per path
	unsigned long long *totals;  // rays cast: [0] camera, [1] extension, [2] shadow
};

}  // namespace ymx

struct ymx_pt
{
	const ymx_accel *accel;
	int device;
	uint32_t n_tris, n_materials;
	float4 *d_prim;                  // (face normal xyz, material id bits) per primitive
	ymx::PtMaterial *d_materials;
	ymx::PtLight *d_lights;
	ymx::PtConst *d_const;
	int32_t *d_faure;                // [kLdsDims][kLdsMaxBase] scrambling permutations
	double *d_inv_prims;             // [kLdsDims]
	uint32_t *d_prims;               // [kLdsDims]
	float4 *d_film;                  // resx * resy (colour sums, alpha sum)
	float *d_film_w;                 // resx * resy
	ymx::PtState st;
	uint32_t batch;                  // AA samples per wavefront
	ymx::PtConst host_const;
	ymx_pt_stats stats;
	hipEvent_t ev0, ev1;
};

// Kernel launches of one wavefront: AA samples [first_sample, first_sample + count)
// of every pixel (pt_kernels.hip); count * resx * resy <= st.capacity.
int ymx_pt_trace_pass(ymx_pt *pt, int32_t first_sample, int32_t count, hipStream_t stream);
