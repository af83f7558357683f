// accel_internal.h -- device data layout shared by the LBVH build and the
// traversal kernels of libymx.so.
//
// HBM layout (all 16-byte aligned, read with float4 loads):
//   node i (64 B, 4 x float4), a binary BVH node holding BOTH children's boxes
//   so one node fetch decides the next step for a ray:
//     f0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//     f1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//     f2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//     f3 = (ref c0, ref c1, -, -) as int bits
//   child ref >= 0 : internal node index
//   child ref <  0 : leaf = 0x80000000 | first_tri << 4 | (count - 1)
//   child ref == YMX_REF_EMPTY : no child (its box is empty)
//   triangle j (48 B, 3 x float4), in leaf order:
//     t0 = (v0.xyz, prim id bits)   t1 = (e1 = v1 - v0, visibility bits)
//     t2 = (e2 = v2 - v0, quad)
//   quad = 0 for a triangle; for a quad primitive (v0, v1, v2, v3) the
//   1-based index of its e3 = v3 - v0 in the quad array (float4 each)
//   e1/e2 are the same FP32 differences the reference forms in
//   ShapePolygon::intersect (include/geometry/shape/shape_polygon.h:130-131).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <deque>
#include <mutex>

#include "../../include/ymx_accel.h"

#define YMX_REF_EMPTY ((int32_t)0xFFFFFFFF)
#define YMX_LEAF_BIT 0x80000000u
#define YMX_LEAF_MAX_FIRST ((1u << 27) - 2u)
#define YMX_LDS_STACK 16

static __host__ __device__ inline int32_t ymx_leaf_ref(uint32_t first, uint32_t count)
{
	return (int32_t)(YMX_LEAF_BIT | (first << 4) | (count - 1u));
}
static __host__ __device__ inline uint32_t ymx_leaf_first(int32_t ref) { return ((uint32_t)ref & 0x7FFFFFFFu) >> 4; }
static __host__ __device__ inline uint32_t ymx_leaf_count(int32_t ref) { return ((uint32_t)ref & 15u) + 1u; }

// Per-stream work area of the persistent query kernels: a ray counter and a
// finished-wave counter (both left at zero by the last wave of each launch,
// so a launch needs no reset before it) and the stack spill area, one column
// of spill_depth entries per persistent lane.  Launches on one stream are
// ordered, so they share one area; a launch on another stream takes an area
// whose last launch has completed (its event) or a new one, so queries on one
// handle may overlap across streams.
struct ymx_work_area
{
	hipStream_t stream;
	hipEvent_t done;    // recorded after each launch on the area
	uint32_t *d_work;   // [0] ray counter, [1] finished waves
	int32_t *d_spill;
	void *d_sort;       // ray reordering: change count, buckets, order, per-tile bucket counts, bucket bases
	size_t sort_bytes;
};

struct ymx_accel
{
	int device;
	uint32_t n_tris;
	uint32_t n_nodes;
	uint32_t max_depth;
	float bmin[3], bmax[3];  // tree bound with the 0.1 % enlargement
	float build_ms;
	float4 *d_nodes;
	float4 *d_tris;
	float4 *d_quads;         // e3 of each quad primitive (NULL without quads)
	uint32_t n_quads;
	int32_t ray_reorder;          // ymx_accel_params.ray_reorder: 0 auto, 1 always, -1 never
	int32_t reorder_cell_bits;    // its sort granularity: 3 or 4 bits per axis
	uint32_t treelet_passes;      // SAH treelet restructuring passes run at build
	uint32_t treelets_rewritten;  // treelets whose topology they changed
	// Traversal stack: YMX_LDS_STACK entries in LDS plus spill_depth in HBM,
	// max_depth + 1 in all.  A lane at an internal node of depth D holds at
	// most D deferred siblings (one per level of its root path), and internal
	// nodes lie at depth <= max_depth - 1, so the stack never holds more than
	// max_depth entries and cannot overflow.
	uint32_t spill_depth;
	uint32_t grid_threads_closest;
	uint32_t grid_threads_shadow;
	uint32_t grid_threads_max;
	// query work areas; work_mutex is held while a launch picks its area and
	// enqueues its kernels and event
	mutable std::mutex work_mutex;
	mutable std::deque<ymx_work_area> work_areas;
};

// The work area for a launch on `stream` (work_mutex held by the caller): the
// stream's own, else one whose last launch has completed, else a new one.
// Returns NULL and sets the error when a new area cannot be allocated.
ymx_work_area *ymx_work_area_for(const ymx_accel *acc, hipStream_t stream);
void ymx_free_work_areas(ymx_accel *acc);

// A segmented ray queue: its producers spread their appends over
// kQueueSegments counters, a page (kQueueStride words) apart, so that no one
// address takes every append; segment s holds counts[s * kQueueStride] rays
// at [s * seg_cap, s * seg_cap + count).  The trace kernels claim their work
// segment by segment, so they read such a queue as it is.
constexpr uint32_t kQueueSegments = 64;
constexpr uint32_t kQueueStride = 1024;

struct ymx_trace_args
{
	const float4 *nodes;
	const float4 *tris;
	const float4 *quads;
	const ymx_ray *rays;
	const uint32_t *order;    // optional: the k-th ray traced is rays[order[k]] (ray reordering)
	const uint32_t *order_changes;  // with order: the batch was reordered when this count is >= the threshold
	uint32_t n_rays;          // ray count, or the capacity when n_dev or seg_counts is set
	const uint32_t *n_dev;    // optional device-side ray count (queues)
	const uint32_t *seg_counts;  // optional: a segmented queue's counters (n_rays = kQueueSegments x seg_cap)
	uint32_t seg_cap;            // its rays per segment, a multiple of 64
	float bmin[3], bmax[3];
	uint32_t *work;
	int32_t *spill;
	uint32_t spill_depth;
	uint32_t grid_threads;
	ymx_hit *hits;
	uint32_t *occluded;
	uint32_t *occ_prim;
	unsigned long long *stats;
	// transparent shadows: bit p of transp set when primitive p's material is
	// transparent; per ray up to max_depth (prim, t bits) kept, sorted by t
	const uint32_t *transp;
	uint32_t max_depth;
	uint32_t *t_count;
	uint2 *t_hits;
};

// Build entry (accel_build.hip) and launch entries (accel_trace.hip).
// idx: `stride` vertex indices per primitive (3, or 4 with quad_slot);
// quad_slot (NULL: no quads): 0 for a triangle, else the 1-based slot of the
// quad's e3 in acc->d_quads (acc->n_quads slots).
int ymx_build_device(ymx_accel *acc, const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t stride,
                     const uint32_t *quad_slot, const uint8_t *vis, uint32_t n_tris, int max_leaf, int builder,
                     int treelet_passes);
int ymx_launch_closest(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t n,
                       hipStream_t stream, unsigned long long *stats);
int ymx_launch_shadow(const ymx_accel *acc, const ymx_ray *rays, uint32_t *occ, uint32_t *prim,
                      uint32_t n, hipStream_t stream, unsigned long long *stats);
int ymx_trace_setup(ymx_accel *acc);
// Transparent-shadow any-hit of n rays (or *n_dev when given) (accel_trace.hip).
int ymx_launch_transparent(const ymx_accel *acc, const ymx_ray *rays, const uint32_t *transp, uint32_t max_depth,
                           uint32_t *occ, uint32_t *counts, uint2 *hits, uint32_t n, const uint32_t *n_dev,
                           hipStream_t stream);
// Closest hit of n coherent rays with the wave-ordered kernel (no ray replacement).
int ymx_launch_closest_ordered(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t n,
                               hipStream_t stream);
// Queue variants: the ray count is read on the device from *n_dev (<= capacity);
// replacement: the ray-replacement kernel (k_dyn) instead of the wave-ordered one.
int ymx_launch_closest_queue(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, const uint32_t *n_dev,
                             uint32_t capacity, bool replacement, hipStream_t stream);
int ymx_launch_shadow_queue(const ymx_accel *acc, const ymx_ray *rays, uint32_t *occ, const uint32_t *n_dev,
                            uint32_t capacity, hipStream_t stream);
// Segmented-queue variants (see kQueueSegments): occlusion bits and caster
// lists by queue position.
int ymx_launch_shadow_segmented(const ymx_accel *acc, const ymx_ray *rays, uint32_t *occ, const uint32_t *seg_counts,
                                uint32_t seg_cap, bool replacement, hipStream_t stream);
int ymx_launch_transparent_segmented(const ymx_accel *acc, const ymx_ray *rays, const uint32_t *transp,
                                     uint32_t max_depth, uint32_t *occ, uint32_t *counts, uint2 *hits,
                                     const uint32_t *seg_counts, uint32_t seg_cap, hipStream_t stream);
// Allocation-free closest hit over the fixed per-device workspace of
// accel_trace.hip; ymx_noalloc_setup (run by ymx_trace_setup with the handle's
// device current) looks up that device's workspace address once.
int ymx_noalloc_setup(int device);
int ymx_launch_closest_noalloc(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t n,
                               hipStream_t stream);

void ymx_set_error(const char *fmt, ...);

#define YMX_HIP_CHECK(expr)                                                                  \
	do                                                                                       \
	{                                                                                        \
		hipError_t ymx_e_ = (expr);                                                          \
		if(ymx_e_ != hipSuccess)                                                             \
		{                                                                                    \
			ymx_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(ymx_e_), __FILE__, \
			              __LINE__);                                                         \
			return ymx_e_ == hipErrorOutOfMemory ? YMX_ERR_OUT_OF_MEMORY : YMX_ERR_HIP;        \
		}                                                                                    \
	} while(0)
