// pt_internal.h -- device data of the wavefront path tracer.
#pragma once

#include "../../include/ymx_guide.h"
#include "../../include/ymx_pt.h"
#include "accel_internal.h"
#include "pt_math.h"

#include <vector>

namespace ymx {

enum BsdfBits : uint32_t
{
	kNone = 0, kSpecular = 1, kGlossy = 2, kDiffuse = 4, kDispersive = 8, kReflect = 16, kTransmit = 32,
	kFilter = 64, kEmit = 128, kVolumetric = 256, kAll = 127
};  // include/material/bsdf.h:25-67 (All = the seven scattering bits, without Emit and Volumetric)

// Material record on the device: the params of ymx_material plus the values
// the reference derives from them at construction.
struct PtMaterial
{
	uint32_t type;            // YMX_MAT_*
	uint32_t flags;           // BsdfBits of the material (bsdf_flags_)
	uint32_t receive_shadows;
	uint32_t as_diffuse;      // glossy
	Rgb color;                // shinydiffuse diffuse colour, glossy colour
	float diffuse_reflect;
	Rgb emit_color;           // shinydiffuse emit * color; light_mat color * power
	float glossy_reflect;
	Rgb diffuse_color;        // glossy
	float exponent;           // glossy
	Rgb ref_col;              // mirror: color * min(reflect, 1)
	float ior;                // glass
	Rgb filter_color;         // glass: transmit_filter * filter_color + (1 - transmit_filter)
	float p_diffuse;          // glossy: min(0.6, 1 - g / (g + (1 - g) d))
	Rgb mirror_color;         // glass
	uint32_t transmit_flags;  // glass: Specular|Transmit, or Filter|Transmit with fake shadows
	uint32_t double_sided;    // light_mat
	uint32_t pad[3];
};

// True when surfaces of this material spawn recursive rays (recursiveRaytrace):
// glossy without as_diffuse, mirror, glass.
YMX_HD bool spawns_children(uint32_t flags) { return (flags & (kGlossy | kSpecular | kFilter)) != 0; }

// raydepth accepted: recursive rays stop below level 20 whatever raydepth says
// (src/integrator/surface/integrator_montecarlo.cc:713), so 19 is the deepest
// that changes anything; the sampler dimensions 2 level + 2 stay below kLdsDims
constexpr int kMaxRayDepth = 19;
constexpr int kGuideCells = 16;  // regions per axis over which training weights are normalised
constexpr uint32_t kGlossyBranches = 8;  // MonteCarloIntegrator::initial_ray_samples_glossy_

struct PtLight
{
	uint32_t type;
	uint32_t cast_shadows;
	int32_t samples;
	float area;           // area light: normal_flipped_.normalized().length(); object / portal light: its surface area
	V3 corner, to_x, to_y, normal_flipped;
	V3 quad[4];           // corner, point1, point1 + point2 - corner, point2
	V3 from;              // point light
	Rgb color;            // colour * power (* pi for area and object lights); portal: background * power
	Rgb hit_color;        // the colour Light::intersect returns: color, clamped for a portal
	// object and portal lights: the triangles are ltris[first, first + count),
	// with the tree bound of the light's own kd-tree
	// (include/accelerator/accelerator_kdtree_original.cc) and, for a mesh of
	// more than kLightTreeMin triangles, the light's own BVH over them
	// (accel_internal.h layout, ids 0 .. count - 1; NULL: test every triangle)
	uint32_t first, count;
	uint32_t double_sided;
	float bound_a[3], bound_g[3];
	const float4 *tree_nodes;
	const float4 *tree_tris;
};

// Object and portal lights with more triangles than this get their own BVH;
// its traversal stack holds kLightStack entries, so a tree deeper than that
// keeps the triangle loop.
constexpr uint32_t kLightTreeMin = 16;
constexpr uint32_t kLightStack = 64;

// One triangle of an object light (64 B): v0 and its cumulative area
// fraction (the light's Pdf1D cdf), v1 and its visibility mask, v2, and the
// face normal.
struct PtLightTri
{
	float4 v0_cdf, v1_vis, v2, n;
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
	int32_t aa_passes;       // "AA_passes"
	uint32_t aa_total;       // the most AA samples a pixel can receive over all passes
	float aa_clamp;          // "AA_clamp_samples"
	int32_t raydepth;
	uint32_t transparent_shadows;  // "transpShad"
	uint32_t shadow_depth;         // "shadowDepth"
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

// The integrate() calls of one wavefront ("nodes"): first the camera samples
// (node i = path i), then their recursive children level by level.  A node's
// children are the contiguous range child.x .. child.x + count.
struct PtNodes
{
	ymx_ray *ray;       // origin, tmin, direction, tmax (-1)
	uint4 *info;        // x: path (j * npix + pixel), y: level | division << 8, z: tag, w: ray-division offset
	float2 *dec;        // ray-division decorrelation 1, 2
	float4 *col;        // colour (local part, then with the children's added), w: alpha
	float4 *wt;         // weight of this node's colour in its parent's sum
	uint2 *child;       // x: first child, y: count | kind << 16 (kind 1 glossy branches, 2 specular)
	uint32_t *spawn;    // children allocated so far in the level being shaded
	uint32_t cap;       // nodes allocated
};

enum ChildKind : uint32_t { kChildGlossy = 1u, kChildSpecular = 2u };

// Per-slot state of one chunk of nodes (slot i = node base + i).
struct PtState
{
	float4 *path_col;   // rgb
	float4 *tp_next;    // throughput including the next segment's BSDF weight
	float4 *tp_vertex;  // throughput at the current vertex (after roulette)
	float4 *pwo;        // outgoing direction at the current vertex
	ymx_hit *cam_hit;   // hit of the node's ray
	float2 *dxy;        // sub-pixel offsets of the camera samples (by path)
	uint32_t *flags;    // StateBits of pt_kernels.hip; bits 8..16 first path vertex BSDF flags
	int32_t *light;     // light picked at the current path vertex
	float4 *vert_emit;  // emission of the current path vertex, when it is added
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
	uint32_t *sh_count;    // segmented queue: kQueueSegments counters, kQueueStride words apart
	uint32_t sh_seg_cap;   // rays per segment (a multiple of 64)
	uint32_t *sh_tcount;   // transparent shadows: casters kept per queued shadow ray
	uint2 *sh_thits;       // ... and their (primitive, t bits), shadow_depth per ray
	uint32_t capacity;     // slots (nodes per chunk); also the camera paths per wavefront
	uint32_t slots;        // shadow slots per path
	unsigned long long *totals;  // rays cast: [0] camera + recursive, [1] extension, [2] shadow
};

// Path guiding inside the wavefront (see ymx_pt_set_guide): mixtures of the
// node vertices (by slot) and of the current path vertices (by queue entry),
// and the training samples the paths produce.
struct PtGuide
{
	uint32_t on;               // guided sampling at shinydiffuse vertices
	uint32_t train;            // record training samples
	float bsdf_prob;           // one-sample MIS: probability of sampling the BSDF
	int32_t max_depth;         // path vertices past the sub-path start that are guided (segments d < max_depth)
	float accept;              // probability that a finished record is kept for training
	ymx_guide_sample *query;   // [capacity] queries of one inference batch
	float *mix_node;           // [capacity][48]
	float *mix_path;           // [capacity][48]
	ymx_guide_sample *rec;     // [bounces][capacity] pending records of the current sub-path
	float4 *base;              // [bounces][capacity] path colour before the record's segment, w: lum(throughput)
	uint32_t *mask;            // [capacity] bit r: record r is pending
	ymx_guide_sample *train_buf;  // collected training samples
	ymx_guide_sample *batch_buf;  // one mini-batch gathered from the ring
	float2 *cell_sum;             // [kGuideCells^3] running (sum of weights, count) of kept samples per region
	float bmin[3], scale[3];      // region of a position: floor((p - bmin) * scale) per axis
	uint32_t *train_count;
	uint32_t train_cap;
	uint32_t *query_count;     // queries of the current path-vertex batch
	// learned selection probability (ymx_pt_set_guide_selection): per region,
	// alpha = kAlphaMin + (1 - kAlphaMin) sigmoid(theta); the arrays are the
	// guide's (ymx_guide_selection_table)
	uint32_t learn_alpha;
	float *alpha_theta;        // [kGuideCells^3]
	float *alpha_m, *alpha_v;  // Adam moments
	float *alpha_grad;         // gradient sums of one mini-batch
	// per-region fallback to BSDF sampling (ymx_pt_set_guide_fallback): running
	// second moments of the guided estimator (at the region's selection
	// probability) and of BSDF sampling alone, estimated from the training
	// records; a region whose guided moment is the larger samples the BSDF only
	uint32_t fallback;
	float *fb_sum;             // [3][kGuideCells^3] one mini-batch's sums: guided, BSDF, records
	float *fb_run;             // [3][kGuideCells^3] running moments (guided, BSDF) and records seen
};
constexpr float kAlphaMin = 0.2f;  // the BSDF is sampled with probability >= 0.2 at a guided vertex
constexpr float kFallbackDecay = 0.9f;  // weight of the running moments against one mini-batch's
constexpr float kFallbackMinRecords = 256.f;  // records a region needs before it may fall back
// The guide's selection table: theta, the two Adam moments and the gradient
// sums of the learned selection, then the fallback's sums and running moments.
constexpr size_t kSelectionTableFloats = (size_t)10 * kGuideCells * kGuideCells * kGuideCells;

}  // namespace ymx

struct ymx_pt
{
	const ymx_accel *accel;
	int device;
	uint32_t n_tris, n_materials;
	float4 *d_prim;                  // (face normal xyz, material id bits) per primitive
	float4 *d_cnorm;                 // corner normals, 3 per triangle (NULL: all flat; ymx_pt_set_normals)
	ymx::PtMaterial *d_materials;
	ymx::PtLight *d_lights;
	ymx::PtLightTri *d_ltris;        // triangles of the object and portal lights
	std::vector<ymx_accel *> light_trees;  // their own BVHs (PtLight::tree_nodes), destroyed with the tracer
	uint32_t *d_transp;              // bit per primitive: its material is transparent (transparent shadows)
	ymx::PtConst *d_const;
	int32_t *d_faure;                // [kLdsDims][kLdsMaxBase] scrambling permutations
	double *d_inv_prims;             // [kLdsDims]
	uint32_t *d_prims;               // [kLdsDims]
	float4 *d_film;                  // resx * resy (colour sums, alpha sum)
	float *d_film_w;                 // resx * resy
	ymx::PtState st;
	ymx::PtNodes nd;
	bool recursive;                  // some material spawns recursive rays and raydepth >= 1
	uint32_t max_fan;                // most children one node can spawn (8 glossy branches, 2 specular)
	ymx_guide *guide;                // path guiding (NULL: BSDF sampling only)
	ymx::PtGuide g;
	uint32_t guide_cap;              // slots the guiding buffers hold
	uint32_t train_steps;            // guide training steps after each wavefront
	int32_t guide_depth;             // ymx_pt_set_guide_depth (-1: every vertex)
	uint32_t g_seen, g_filled;       // training samples written so far (counter value), valid in the ring
	uint64_t g_rng;                  // mini-batch offsets
	bool alpha_learned;              // ymx_pt_set_guide_selection: the guide's learned selection is used
	bool fallback_on;                // ymx_pt_set_guide_fallback: regions may fall back to BSDF sampling
	float alpha_lr;
	uint32_t alpha_steps;
	uint32_t batch;                  // AA samples per wavefront (<= batch_max; ymx_pt_autotune may lower it)
	uint32_t batch_max;              // AA samples per wavefront the state is allocated for
	bool shadow_replacement;         // shadow queues traced with ray replacement (k_dyn), else wave-ordered
	bool ext_replacement;            // the same for the extension queues
	ymx_pt_settings settings;        // as given at create (the AA pass params)
	float max_light_mult;            // largest light sample multiplier of any AA pass
	uint32_t *d_aa;                  // AA passes: flags, pair bits, variance hits, pixel list, slots (npix each), count
	ymx::PtConst host_const;
	ymx_pt_stats stats;
	hipEvent_t ev0, ev1;
};

// Kernel launches of one wavefront: AA samples [first_sample, first_sample + count)
// of every pixel (pt_kernels.hip); count * resx * resy <= st.capacity.
int ymx_pt_trace_pass(ymx_pt *pt, int32_t first_sample, int32_t count, hipStream_t stream);

// One AA pass's wavefront over a subset of the pixels: pix_list (device, n_active
// pixel indices) and pix_slot (device, per pixel its position in the list or
// 0xFFFFFFFF), NULL for every pixel; light_mult the pass's light sample
// multiplier; pass_samples its sample count (the sub-pixel offsets of a
// single-pass render depend on it).
int ymx_pt_trace_subset(ymx_pt *pt, int32_t first_sample, int32_t count, const uint32_t *pix_list,
                        const uint32_t *pix_slot, uint32_t n_active, float light_mult, int32_t pass_samples,
                        hipStream_t stream);

// ImageFilm::nextPass on the device: flags the pixels to resample, lists them
// and leaves their count in d_aa's count word (pt_kernels.hip).
int ymx_pt_aa_flags(ymx_pt *pt, float threshold, hipStream_t stream);

// Grows the node pool to hold at least `need` nodes, keeping the first `keep`
// (pt_api.hip).  Synchronous.
int ymx_pt_grow_nodes(ymx_pt *pt, uint64_t need, uint32_t keep);

// Guide training on the samples of the last wavefront (pt_kernels.hip).
int ymx_pt_train_guide(ymx_pt *pt, hipStream_t stream);

// Most samples one training step of the guide takes, and its bounding box (guide.hip).
uint32_t ymx_guide_batch_limit(const ymx_guide *g);
void ymx_guide_bounds(const ymx_guide *g, float bmin[3], float bmax[3]);
// The guide's table of the path tracer's learned selection probability
// (`floats` floats, allocated on first use and freed with the guide; NULL
// when the allocation fails).  Tracers sharing the guide share the table.
float *ymx_guide_selection_table(ymx_guide *g, size_t floats);
