/*
 * ymx_accel.h -- C ABI of the MI355X ray-cast accelerator (libymx.so).
 *
 * This is the drop-in boundary for libYafaRay's Accelerator plugin
 * (include/accelerator/accelerator.h:37-87 in the reference).  A scene selects
 * it through the accelerator "type" param ("mi355x-lbvh"); the host-side
 * Accelerator subclass in libyafaray_amd/host wraps these entry points.
 *
 * Conventions
 *   - Every entry point returns an int status (YMX_OK == 0, negative on error);
 *     nothing throws across the ABI.  ymx_last_error() describes the last
 *     failure of the calling thread.
 *   - Ray and hit buffers passed to the query calls are DEVICE pointers owned
 *     by the caller.  `stream` is a hipStream_t (NULL: the default stream);
 *     queries are asynchronous on it.
 *   - A handle is immutable after ymx_accel_create, and its queries may be
 *     issued concurrently, from several host threads and on several streams,
 *     as the reference's const queries are.  Each stream's queries use a work
 *     area of their own (a ray counter and a traversal-stack spill area,
 *     allocated on the stream's first query and kept until
 *     ymx_accel_destroy; a stream whose queries have all completed hands its
 *     area to the next new stream).
 *   - The traversal stack holds max_depth + 1 entries (ymx_accel_info) and a
 *     traversal never needs more than max_depth, so it cannot overflow.
 *   - FP32 throughout.  Closest-hit primitive ids are bit-exact with the
 *     reference's kd-tree except for exact ties in t, which this library
 *     resolves to the lowest primitive id.
 */
#ifndef YMX_ACCEL_H
#define YMX_ACCEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YMX_OK 0
#define YMX_ERR_INVALID_ARG (-1)
#define YMX_ERR_HIP (-2)
#define YMX_ERR_OUT_OF_MEMORY (-3)
#define YMX_ERR_UNSUPPORTED (-4)

#define YMX_NO_HIT 0xFFFFFFFFu
#define YMX_NO_VERTEX 0xFFFFFFFFu  /* 4th index of a triangle in a face list */

/* Primitive visibility bits, as include/common/visibility.h:28-33 (Visible,
 * CastsShadows).  The mask passed per primitive is the AND of the primitive's
 * and its material's visibility (include/accelerator/accelerator.h:126-128). */
#define YMX_VIS_VISIBLE 1u
#define YMX_VIS_CASTS_SHADOWS 2u

/* Ray record (32 B): include/geometry/ray.h:36-51 without time/differentials.
 * tmax < 0 means unbounded, as Ray::tmax_ = -1 (include/geometry/ray.h:47). */
typedef struct ymx_ray
{
	float org[3];
	float tmin;
	float dir[3];
	float tmax;
} ymx_ray;

/* Hit record (16 B): the fields of IntersectData
 * (include/accelerator/intersect_data.h:30-39) a caller needs, with the
 * Primitive pointer replaced by the primitive's index in the list given to
 * ymx_accel_create (YMX_NO_HIT on a miss; t, u, v are then 0). */
typedef struct ymx_hit
{
	uint32_t prim;
	float t;
	float u;
	float v;
} ymx_hit;

typedef struct ymx_accel ymx_accel;

typedef struct ymx_accel_params
{
	uint32_t struct_size;   /* sizeof(ymx_accel_params) */
	int32_t max_leaf_size;  /* at most this many triangles per leaf (1..8), 0 = default (4) */
	int32_t builder;        /* YMX_BUILDER_LBVH (0, default) or YMX_BUILDER_PLOC */
	int32_t treelet_passes; /* SAH treelet restructuring passes over the built tree
	                           (Karras & Aila 2013): 1..8, -1 = none,
	                           0 = default (YMX_TREELET_PASSES_DEFAULT) */
	int32_t ray_reorder;    /* closest- and any-hit batches trace their rays grouped by
	                           direction octant and origin cell (a device counting
	                           sort of the ray indices; results stay in the caller's
	                           order): 0 = closest-hit batches of at least 2^18 rays
	                           that are not already coherent (decided on the device,
	                           no host wait), 1 = always (any hit too), -1 = never */
	int32_t reorder_cell_bits; /* sort granularity of ray reordering: origin cells per
	                              axis of the tree bound, 2^bits (3: 4,096 buckets with
	                              the octant, 4: 32,768); 0 = default
	                              (YMX_REORDER_CELL_BITS_DEFAULT) */
	int32_t reserved[2];
} ymx_accel_params;

/* Tree builders.  LBVH (Karras 2012) splits at Morton-code prefixes; PLOC
 * (Meister & Bittner 2018) clusters Morton-ordered boxes by smallest union
 * area.  Both give the same query results; they trade speed per query: on
 * the 2.19M-triangle heightfield (4-triangle leaves) LBVH traces incoherent
 * closest-hit rays 11 % faster and PLOC any-hit rays 4 % faster, while
 * path-traced rooms render 4-12 % faster on PLOC trees. */
#define YMX_BUILDER_LBVH 0
#define YMX_BUILDER_PLOC 1
/* Two treelet passes: on the 2.19M-triangle heightfield (LBVH, 4-triangle
 * leaves) 45.8 -> 42.8 nodes and 20.2 -> 18.0 triangles per incoherent ray,
 * +9 % incoherent and +5 % coherent closest hit, +8 % any hit, for 155 ms
 * more build time (one pass: +7 %, three: +9.5 %). */
#define YMX_TREELET_PASSES_DEFAULT 2
#define YMX_REORDER_CELL_BITS_DEFAULT 3

typedef struct ymx_accel_info
{
	uint32_t struct_size;
	uint32_t n_tris;
	uint32_t n_nodes;        /* 64-byte nodes */
	uint32_t max_depth;      /* longest root-to-leaf path in nodes */
	uint64_t node_bytes;
	uint64_t tri_bytes;
	float build_ms;          /* device time of the GPU build kernels */
	float bound_min[3];      /* tree bound (getBound()) */
	float bound_max[3];
	uint32_t n_quads;        /* quad primitives (ymx_accel_create_faces) */
	uint32_t treelet_passes; /* SAH treelet restructuring passes run */
	uint32_t treelets_rewritten;  /* treelets whose topology those passes changed */
} ymx_accel_info;

/* Build the accelerator over n_tris triangles on the current HIP device.
 * verts: host, 3 floats per vertex; tri_idx: host, 3 vertex indices per
 * triangle (all < n_verts, checked); vis: host, one visibility mask per
 * triangle, or NULL for "normal" (both bits).
 * Replaces Accelerator::factory -> AcceleratorKdTree::AcceleratorKdTree
 * (src/accelerator/accelerator.cc:44-55,
 *  src/accelerator/accelerator_kdtree_original.cc:65-130). */
int ymx_accel_create(const float *verts, uint32_t n_verts, const uint32_t *tri_idx,
                     const uint8_t *vis, uint32_t n_tris, const ymx_accel_params *params,
                     ymx_accel **out);

/* The same over a face list of triangles and quads, as the reference's meshes
 * hold them (yafaray_addTriangle / yafaray_addQuad,
 * include/public_api/yafaray_c_api.h:219-222, one PrimitivePolygon<float, 3 or
 * 4> per face, src/geometry/object/object_mesh.cc:96-112).  face_idx: host, 4
 * vertex indices per face; face_idx[4 i + 3] == YMX_NO_VERTEX makes face i a
 * triangle, otherwise it is the quad (v0, v1, v2, v3).  A quad is ONE
 * primitive: its id is its face index, and a hit on it is the reference's quad
 * test (include/geometry/shape/shape_polygon.h:126-176): the second triangle
 * (v0, v2, v3) is tried only when the first triangle's u is outside [0, 1],
 * and uv comes back as {u + v, v} from the first triangle and {u, u + v} from
 * the second.  Every query below accepts handles of either kind. */
int ymx_accel_create_faces(const float *verts, uint32_t n_verts, const uint32_t *face_idx,
                           const uint8_t *vis, uint32_t n_faces, const ymx_accel_params *params,
                           ymx_accel **out);

/* Release the handle and its device memory (blocks until its work is done). */
int ymx_accel_destroy(ymx_accel *accel);

/* Accelerator::getBound() (include/accelerator/accelerator.h:53): the union of
 * the triangle bounds grown by 0.001 x extent per axis
 * (src/accelerator/accelerator_kdtree_original.cc:89-103). */
int ymx_accel_bound(const ymx_accel *accel, float bound_min[3], float bound_max[3]);

int ymx_accel_get_info(const ymx_accel *accel, ymx_accel_info *info);

/* Closest hit for n rays: Accelerator::intersect(ray, t_max) with
 * t_max = ray.tmax >= 0 ? ray.tmax : FLT_MAX
 * (include/accelerator/accelerator.h:50,89-101; traversal semantics of
 * include/accelerator/accelerator_kdtree_common.h:107-255, Nearest).
 * t_min = max(ray.tmin, dynamic bias of the tree-bound crossing). */
int ymx_accel_intersect(const ymx_accel *accel, const ymx_ray *d_rays, ymx_hit *d_hits,
                        uint32_t n_rays, void *stream);

/* Closest hit as ymx_accel_intersect (same kernels, same results), allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The traversal's ray counters and stack
 * spill area live in a fixed 64 MiB device workspace inside libymx.so, which
 * every GPU the library runs on holds from the library's load until the
 * process ends (its address is looked up by ymx_accel_create on the handle's
 * device).  A tree too deep for the workspace (spill columns of more than
 * 64 MiB for one 128-lane block, max_depth above about 131,000) is refused
 * with YMX_ERR_UNSUPPORTED.  Every call on a GPU shares that GPU's workspace:
 * calls that overlap on one GPU (on different streams or from different
 * threads) race, and the caller must order them; calls on one stream are
 * ordered by the stream. */
int ymx_accel_intersect_noalloc(const ymx_accel *accel, const ymx_ray *d_rays, ymx_hit *d_hits,
                                uint32_t n_rays, void *stream);

/* Any hit for n rays: Accelerator::intersectShadow(ray, t_max) with
 * t_max = ray.tmax >= 0 ? ray.tmax : FLT_MAX; ray.tmin is ignored, the
 * dynamic bias alone is t_min (include/accelerator/accelerator.h:51,
 * include/accelerator/accelerator_kdtree_common.h:139).
 * d_occluded: bit i % 32 of word i / 32 is ray i's occlusion; (n + 31) / 32
 * words are written.  d_prim (may be NULL): occluder id or YMX_NO_HIT; any
 * occluder may be reported, as the reference returns the first it meets. */
int ymx_accel_intersect_shadow(const ymx_accel *accel, const ymx_ray *d_rays,
                               uint32_t *d_occluded, uint32_t *d_prim, uint32_t n_rays,
                               void *stream);

/* Transparent-shadow query for n rays: Accelerator::intersectTransparentShadow
 * (include/accelerator/accelerator.h:113-119,147-168,
 * include/accelerator/accelerator_kdtree_common.h:107-255, TransparentShadow),
 * called by isShadowedTransparentShadow when the integrator's "transpShad" is
 * set.  t_min = max(ray.tmin, dynamic bias), t_max as in the any-hit query.
 * d_transparent: device bitmask, bit p % 32 of word p / 32 set when primitive
 * p's material is transparent (Material::isTransparent).  A ray is occluded
 * by an opaque shadow caster, or by more than max_depth transparent ones.
 * d_occluded as in ymx_accel_intersect_shadow; d_counts[i]: the transparent
 * casters kept for an unoccluded ray (0 when occluded); d_hits[i * max_depth +
 * k] for k < d_counts[i]: (primitive id, t as float bits), sorted by t (the
 * caller multiplies their Material::getTransparency, as the reference's
 * IntersectData::color_ does in the order its kd-tree meets them). */
int ymx_accel_intersect_transparent(const ymx_accel *accel, const ymx_ray *d_rays, const uint32_t *d_transparent,
                                    uint32_t max_depth, uint32_t *d_occluded, uint32_t *d_counts, uint32_t *d_hits,
                                    uint32_t n_rays, void *stream);

/* Copy the built tree to host memory (no counterpart in the reference: it
 * serves the per-ray queries of the C++ host layer, which traverse the same
 * tree on the CPU; libyafaray_amd/host).  nodes: n_nodes x 64 bytes, node i =
 * 4 x float4: (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y), (c1.lo.x, c1.hi.x, c1.lo.y,
 * c1.hi.y), (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z), (ref0, ref1, -, -) as int32
 * bits; the root is node 0; a ref >= 0 is a node, 0xFFFFFFFF no child, and
 * any other negative ref a leaf 0x80000000 | first << 4 | (count - 1) over
 * triangle records [first, first + count).  tris: n_tris x 48 bytes (one
 * record when n_tris is 0), 3 x float4: (v0, primitive id bits), (v1 - v0,
 * visibility bits), (v2 - v0, quad index bits: 0 for a triangle).  The byte counts must equal those of
 * ymx_accel_get_info. */
int ymx_accel_export(const ymx_accel *accel, void *nodes, uint64_t node_bytes, void *tris, uint64_t tri_bytes);

/* The quad records a handle from ymx_accel_create_faces adds to ymx_accel_export's
 * (no counterpart in the reference; for the same host walk): a primitive
 * record whose third float4 has a non-zero w (as uint32 bits) q is a quad, and
 * quads[q - 1] (16 bytes: e3 = v3 - v0, 0) completes it.  quad_bytes must be
 * 16 x n_quads of ymx_accel_get_info. */
int ymx_accel_export_quads(const ymx_accel *accel, void *quads, uint64_t quad_bytes);

/* Counting replay of the query kernels: total 64-byte nodes fetched and
 * triangles tested over the batch.  query: 0 closest, 1 shadow.  Synchronous;
 * the counts are written to host memory. */
int ymx_accel_traversal_stats(const ymx_accel *accel, const ymx_ray *d_rays, uint32_t n_rays,
                              int query, uint64_t *node_fetches, uint64_t *tri_tests, void *stream);

/* Message of the calling thread's last failure ("" if none). */
const char *ymx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
