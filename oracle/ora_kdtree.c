/*
 * ora_kdtree.c -- C port of the reference's SAH kd-tree
 * (src/accelerator/accelerator_kdtree_original.cc) and of its stack-based
 * traversal (include/accelerator/accelerator_kdtree_common.h:107-255).
 *
 * TEST INFRASTRUCTURE ONLY.  It is the second oracle for large ray batches
 * (its ties resolve in kd leaf order, like the reference's) and the CPU
 * baseline that bench.py times on the host cores.  Never linked into the
 * shipped library.
 *
 * Not restated: polygon clipping of nodes with <= 32 prims
 * (PRIMITIVE_CLIPPING, accelerator_kdtree_original.cc:451-497).  Those nodes
 * keep the whole-primitive bounds, as the reference does for primitives without
 * clipping support (:486-491).  Tree shape differs slightly; query results do
 * not, except for the order in which exactly tied hits are met.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define KD_MAX_STACK 64  /* kd_max_stack_, accelerator_kdtree_original.h:98 */
#define PIGEON_THRESH 128 /* pigeonhole_sort_thresh_, :97 */
#define AXIS_NONE (-1)

typedef struct { float a[3], g[3]; } kbound;

/* Node (accelerator_kdtree_original.h:106-125): 2 flag bits (axis or 3 = leaf),
 * 30 bits of right-child index or primitive count. */
typedef struct
{
	union { float split; uint32_t one_prim; uint32_t prim_off; } u;
	uint32_t flags;
} knode;

struct ora_kdtree
{
	const float *verts;
	const uint32_t *idx;
	uint32_t stride;  /* 3: triangles; 4: faces with idx[3] == ORA_NO_VERTEX for triangles */
	const uint8_t *vis;
	uint32_t n_tris;
	kbound tree_bound;
	kbound *prim_bounds;
	knode *nodes;
	uint32_t n_nodes, cap_nodes;
	uint32_t *leaf_prims;
	size_t n_leaf_prims, cap_leaf_prims;
	int max_depth;
	uint32_t max_leaf_size;
	float cost_ratio;
	float e_bonus;
	int failed;
};

typedef struct { int axis; int edge_offset; float cost; float t; int num_edges; } split_cost;

/* kdtree::BoundEdge, accelerator_kdtree_common.h:34-48 (end: 0 Left, 1 Both, 2 Right). */
typedef struct { float pos; uint32_t index; int end; } bound_edge;

static int edge_cmp(const void *pa, const void *pb)
{
	const bound_edge *a = (const bound_edge *)pa, *b = (const bound_edge *)pb;
	if(a->pos == b->pos)
	{
		/* operator<: at equal position the larger end sorts first */
		if(a->end > b->end) return -1;
		if(a->end < b->end) return 1;
		return (a->index < b->index) ? -1 : (a->index > b->index);
	}
	return (a->pos < b->pos) ? -1 : 1;
}

typedef struct { int n, c_left, c_right, c_bleft, c_both; float t; } tree_bin;

/* pigeonMinCost, accelerator_kdtree_original.cc:154-292 */
static split_cost pigeon_min_cost(float e_bonus, float cost_ratio, uint32_t n_idx, const kbound *bounds,
                                  const kbound *nb, const uint32_t *prim_idx)
{
	enum { MAX_BIN = 1024, NUM_BINS = MAX_BIN + 1 };
	static __thread tree_bin bins[NUM_BINS];
	memset(bins, 0, sizeof bins);
	const float len[3] = {nb->g[0] - nb->a[0], nb->g[1] - nb->a[1], nb->g[2] - nb->a[2]};
	const float inv_len[3] = {1.f / len[0], 1.f / len[1], 1.f / len[2]};
	split_cost split = {AXIS_NONE, -1, FLT_MAX, 0.f, 0};
	const float inv_total_sa = 1.f / (len[0] * len[1] + len[0] * len[2] + len[1] * len[2]);
	for(int axis = 0; axis < 3; ++axis)
	{
		const float s = (float)MAX_BIN * inv_len[axis];
		const float mn = nb->a[axis];
		for(uint32_t pn = 0; pn < n_idx; ++pn)
		{
			const kbound *bb = &bounds[prim_idx[pn]];
			const float t_low = bb->a[axis];
			const float t_up = bb->g[axis];
			int b_left = (int)((t_low - mn) * s);
			int b_right = (int)((t_up - mn) * s);
			if(b_left < 0) b_left = 0; else if(b_left > MAX_BIN) b_left = MAX_BIN;
			if(b_right < 0) b_right = 0; else if(b_right > MAX_BIN) b_right = MAX_BIN;
			if(t_low == t_up)
			{
				tree_bin *bl = &bins[b_left];
				if(bl->n == 0 || (t_low >= bl->t && bl->n != 0))
				{
					bl->t = t_low;
					bl->c_both++;
				}
				else
				{
					bl->c_left++;
					bl->c_right++;
				}
				bl->n += 2;
			}
			else
			{
				tree_bin *bl = &bins[b_left];
				if(bl->n == 0 || (t_low > bl->t && bl->n != 0))
				{
					bl->t = t_low;
					bl->c_left += bl->c_both + bl->c_bleft;
					bl->c_right += bl->c_both;
					bl->c_both = bl->c_bleft = 0;
					bl->c_bleft++;
				}
				else if(t_low == bl->t) bl->c_bleft++;
				else bl->c_left++;
				bl->n++;
				tree_bin *br = &bins[b_right];
				br->c_right++;
				if(br->n == 0 || t_up > br->t)
				{
					br->t = t_up;
					br->c_left += br->c_both + br->c_bleft;
					br->c_right += br->c_both;
					br->c_both = br->c_bleft = 0;
				}
				br->n++;
			}
		}
		const int ax1 = (axis + 1) % 3, ax2 = (axis + 2) % 3;
		const float cap_area = len[ax1] * len[ax2];
		const float cap_perim = len[ax1] + len[ax2];
		uint32_t num_left = 0, num_right = n_idx;
		for(int b = 0; b < NUM_BINS; ++b)
		{
			const tree_bin *bin = &bins[b];
			if(bin->n == 0) continue;
			num_left += (uint32_t)bin->c_left;
			num_right -= (uint32_t)bin->c_right;
			const float edget = bin->t;
			if(edget > nb->a[axis] && edget < nb->g[axis])
			{
				const float l_below = edget - nb->a[axis];
				const float l_above = nb->g[axis] - edget;
				const float below_sa = cap_area + l_below * cap_perim;
				const float above_sa = cap_area + l_above * cap_perim;
				const float raw_costs = (below_sa * (float)num_left + above_sa * (float)num_right);
				float eb;
				if(num_right == 0) eb = (0.1f + l_above * inv_len[axis]) * e_bonus * raw_costs;
				else if(num_left == 0) eb = (0.1f + l_below * inv_len[axis]) * e_bonus * raw_costs;
				else eb = 0.f;
				const float cost = cost_ratio + inv_total_sa * (raw_costs - eb);
				if(cost < split.cost)
				{
					split.t = edget;
					split.cost = cost;
					split.axis = axis;
				}
			}
			num_left += (uint32_t)(bin->c_both + bin->c_bleft);
			num_right -= (uint32_t)bin->c_both;
		}
		memset(bins, 0, sizeof bins);
	}
	return split;
}

/* minimalCost, accelerator_kdtree_original.cc:298-427 (non-clipping branch). */
static split_cost minimal_cost(float e_bonus, float cost_ratio, uint32_t n_idx, const kbound *nb,
                               const uint32_t *prim_idx, const kbound *all_bounds, bound_edge *edges[3])
{
	const float len[3] = {nb->g[0] - nb->a[0], nb->g[1] - nb->a[1], nb->g[2] - nb->a[2]};
	const float inv_len[3] = {1.f / len[0], 1.f / len[1], 1.f / len[2]};
	split_cost split = {AXIS_NONE, -1, FLT_MAX, 0.f, 0};
	const float inv_total_sa = 1.f / (len[0] * len[1] + len[0] * len[2] + len[1] * len[2]);
	for(int axis = 0; axis < 3; ++axis)
	{
		int num_edges = 0;
		bound_edge *e = edges[axis];
		for(uint32_t i = 0; i < n_idx; ++i)
		{
			const uint32_t pn = prim_idx[i];
			const kbound *bb = &all_bounds[pn];
			if(bb->a[axis] == bb->g[axis])
			{
				e[num_edges].pos = bb->a[axis]; e[num_edges].index = pn; e[num_edges].end = 1;
				++num_edges;
			}
			else
			{
				e[num_edges].pos = bb->a[axis]; e[num_edges].index = pn; e[num_edges].end = 0;
				e[num_edges + 1].pos = bb->g[axis]; e[num_edges + 1].index = pn; e[num_edges + 1].end = 2;
				num_edges += 2;
			}
		}
		qsort(e, (size_t)num_edges, sizeof(bound_edge), edge_cmp);
		const int ax1 = (axis + 1) % 3, ax2 = (axis + 2) % 3;
		const float cap_area = len[ax1] * len[ax2];
		const float cap_perim = len[ax1] + len[ax2];
		uint32_t num_left = 0, num_right = n_idx;
		if(n_idx > 5)
		{
			float edget = e[0].pos;
			float l_below = edget - nb->a[axis];
			float l_above = nb->g[axis] - edget;
			if(l_below > l_above * (float)n_idx && l_above > 0.f)
			{
				const float raw_costs = (cap_area + l_above * cap_perim) * (float)n_idx;
				const float cost = cost_ratio + inv_total_sa * (raw_costs - e_bonus);
				if(cost < split.cost)
				{
					split.cost = cost; split.axis = axis; split.edge_offset = 0; split.num_edges = num_edges;
				}
				continue;
			}
			edget = e[num_edges - 1].pos;
			l_below = edget - nb->a[axis];
			l_above = nb->g[axis] - edget;
			if(l_above > l_below * (float)n_idx && l_below > 0.f)
			{
				const float raw_costs = (cap_area + l_below * cap_perim) * (float)n_idx;
				const float cost = cost_ratio + inv_total_sa * (raw_costs - e_bonus);
				if(cost < split.cost)
				{
					split.cost = cost; split.axis = axis; split.edge_offset = num_edges - 1; split.num_edges = num_edges;
				}
				continue;
			}
		}
		for(int ei = 0; ei < num_edges; ++ei)
		{
			if(e[ei].end == 2) --num_right;
			const float edget = e[ei].pos;
			if(edget > nb->a[axis] && edget < nb->g[axis])
			{
				const float l_below = edget - nb->a[axis];
				const float l_above = nb->g[axis] - edget;
				const float below_sa = cap_area + l_below * cap_perim;
				const float above_sa = cap_area + l_above * cap_perim;
				const float raw_costs = (below_sa * (float)num_left + above_sa * (float)num_right);
				float eb;
				if(num_right == 0) eb = (0.1f + l_above * inv_len[axis]) * e_bonus * raw_costs;
				else if(num_left == 0) eb = (0.1f + l_below * inv_len[axis]) * e_bonus * raw_costs;
				else eb = 0.f;
				const float cost = cost_ratio + inv_total_sa * (raw_costs - eb);
				if(cost < split.cost)
				{
					split.cost = cost; split.axis = axis; split.edge_offset = ei; split.num_edges = num_edges;
				}
			}
			if(e[ei].end != 2)
			{
				++num_left;
				if(e[ei].end == 1) --num_right;
			}
		}
	}
	return split;
}

static uint32_t new_node(ora_kdtree *kd)
{
	if(kd->n_nodes == kd->cap_nodes)
	{
		uint32_t nc = kd->cap_nodes ? kd->cap_nodes * 2 : 256;
		knode *nn = (knode *)realloc(kd->nodes, (size_t)nc * sizeof(knode));
		if(!nn) { kd->failed = 1; return 0; }
		kd->nodes = nn;
		kd->cap_nodes = nc;
	}
	memset(&kd->nodes[kd->n_nodes], 0, sizeof(knode));
	return kd->n_nodes++;
}

/* Node::createLeaf, accelerator_kdtree_original.h:145-162 */
static void create_leaf(ora_kdtree *kd, const uint32_t *prims, uint32_t np)
{
	const uint32_t id = new_node(kd);
	if(kd->failed) return;
	knode *n = &kd->nodes[id];
	n->flags = (np << 2) | 3u;
	if(np > 1)
	{
		if(kd->n_leaf_prims + np > kd->cap_leaf_prims)
		{
			size_t nc = kd->cap_leaf_prims ? kd->cap_leaf_prims * 2 : 1024;
			while(nc < kd->n_leaf_prims + np) nc *= 2;
			uint32_t *nl = (uint32_t *)realloc(kd->leaf_prims, nc * sizeof(uint32_t));
			if(!nl) { kd->failed = 1; return; }
			kd->leaf_prims = nl;
			kd->cap_leaf_prims = nc;
		}
		n->u.prim_off = (uint32_t)kd->n_leaf_prims;
		memcpy(kd->leaf_prims + kd->n_leaf_prims, prims, np * sizeof(uint32_t));
		kd->n_leaf_prims += np;
	}
	else if(np == 1) n->u.one_prim = prims[0];
}

/* buildTree, accelerator_kdtree_original.cc:437-645 (without clipping). */
static void build_tree(ora_kdtree *kd, uint32_t n_prims, const kbound *node_bound, uint32_t *prim_nums,
                       int depth, int bad_refines, bound_edge *edges[3])
{
	if(kd->failed) return;
	if(n_prims <= kd->max_leaf_size || depth >= kd->max_depth)
	{
		create_leaf(kd, prim_nums, n_prims);
		return;
	}
	const float base_bonus = kd->e_bonus;
	kd->e_bonus = (float)((double)kd->e_bonus * (1.1 - (double)((float)depth / (float)kd->max_depth)));
	split_cost split;
	if(n_prims > PIGEON_THRESH)
		split = pigeon_min_cost(kd->e_bonus, kd->cost_ratio, n_prims, kd->prim_bounds, node_bound, prim_nums);
	else
		split = minimal_cost(kd->e_bonus, kd->cost_ratio, n_prims, node_bound, prim_nums, kd->prim_bounds, edges);
	kd->e_bonus = base_bonus;
	if(split.cost > (float)n_prims) ++bad_refines;
	if((split.cost > 1.6f * (float)n_prims && n_prims < 16) || split.axis == AXIS_NONE || bad_refines == 2)
	{
		create_leaf(kd, prim_nums, n_prims);
		return;
	}
	uint32_t *left = (uint32_t *)malloc((size_t)n_prims * sizeof(uint32_t));
	uint32_t *right = (uint32_t *)malloc((size_t)n_prims * sizeof(uint32_t));
	if(!left || !right) { free(left); free(right); kd->failed = 1; return; }
	uint32_t n0 = 0, n1 = 0;
	float split_pos;
	const int ax = split.axis;
	if(n_prims > PIGEON_THRESH)
	{
		for(uint32_t i = 0; i < n_prims; ++i)
		{
			const uint32_t pn = prim_nums[i];
			if(kd->prim_bounds[pn].a[ax] >= split.t) right[n1++] = pn;
			else
			{
				left[n0++] = pn;
				if(kd->prim_bounds[pn].g[ax] > split.t) right[n1++] = pn;
			}
		}
		split_pos = split.t;
	}
	else
	{
		const bound_edge *e = edges[ax];
		for(int i = 0; i < split.edge_offset; ++i)
			if(e[i].end != 2) left[n0++] = e[i].index;
		if(e[split.edge_offset].end == 1) right[n1++] = e[split.edge_offset].index;
		for(int i = split.edge_offset + 1; i < split.num_edges; ++i)
			if(e[i].end != 0) right[n1++] = e[i].index;
		split_pos = e[split.edge_offset].pos;
	}
	const uint32_t cur = new_node(kd);
	if(kd->failed) { free(left); free(right); return; }
	/* Node::createInterior, accelerator_kdtree_original.h:164-169 */
	kd->nodes[cur].u.split = split_pos;
	kd->nodes[cur].flags = (kd->nodes[cur].flags & ~3u) | (uint32_t)ax;
	kbound bl = *node_bound, br = *node_bound;
	bl.g[ax] = split_pos;
	br.a[ax] = split_pos;
	build_tree(kd, n0, &bl, left, depth + 1, bad_refines, edges);
	free(left);
	kd->nodes[cur].flags = (kd->nodes[cur].flags & 3u) | (kd->n_nodes << 2);
	build_tree(kd, n1, &br, right, depth + 1, bad_refines, edges);
	free(right);
}

ora_kdtree *ora_kd_build_faces(const float *verts, const uint32_t *tri_idx, uint32_t stride, const uint8_t *vis,
                               uint32_t n_tris)
{
	ora_kdtree *kd = (ora_kdtree *)calloc(1, sizeof(ora_kdtree));
	if(!kd) return NULL;
	kd->verts = verts; kd->idx = tri_idx; kd->stride = stride; kd->vis = vis; kd->n_tris = n_tris;
	/* accelerator_kdtree_original.cc:75-86 */
	kd->cost_ratio = 0.8f;
	kd->e_bonus = 0.33f;
	kd->max_leaf_size = 1;
	kd->max_depth = 0;
	if(n_tris > 0) kd->max_depth = (int)(7.0f + 1.66f * ora_fast_log((float)n_tris));
	const double log_leaves = 1.442695 * log((double)n_tris);
	if(kd->max_depth > KD_MAX_STACK) kd->max_depth = KD_MAX_STACK;
	if(log_leaves > 16.0) kd->cost_ratio = (float)((double)kd->cost_ratio + 0.25f * (log_leaves - 16.0));
	kd->prim_bounds = (kbound *)malloc((size_t)(n_tris ? n_tris : 1) * sizeof(kbound));
	uint32_t *prims = (uint32_t *)malloc((size_t)(n_tris ? n_tris : 1) * sizeof(uint32_t));
	bound_edge *edges[3];
	for(int i = 0; i < 3; ++i) edges[i] = (bound_edge *)malloc(514 * sizeof(bound_edge));
	if(!kd->prim_bounds || !prims || !edges[0] || !edges[1] || !edges[2])
	{
		free(prims); for(int i = 0; i < 3; ++i) free(edges[i]);
		ora_kd_free(kd);
		return NULL;
	}
	for(uint32_t i = 0; i < n_tris; ++i)
	{
		kbound *b = &kd->prim_bounds[i];
		const int nv = ora_face_nverts(tri_idx, stride, i);
		for(int k = 0; k < nv; ++k)
		{
			const float *p = verts + 3 * (size_t)tri_idx[(size_t)stride * i + k];
			for(int ax = 0; ax < 3; ++ax)
			{
				if(k == 0) { b->a[ax] = p[ax]; b->g[ax] = p[ax]; }
				else
				{
					if(p[ax] < b->a[ax]) b->a[ax] = p[ax];
					if(p[ax] > b->g[ax]) b->g[ax] = p[ax];
				}
			}
		}
		prims[i] = i;
	}
	ora_tree_bound_faces(verts, tri_idx, stride, n_tris, kd->tree_bound.a, kd->tree_bound.g);
	build_tree(kd, n_tris, &kd->tree_bound, prims, 0, 0, edges);
	free(prims);
	for(int i = 0; i < 3; ++i) free(edges[i]);
	if(kd->failed) { ora_kd_free(kd); return NULL; }
	return kd;
}

ora_kdtree *ora_kd_build(const float *verts, const uint32_t *tri_idx, const uint8_t *vis, uint32_t n_tris)
{
	return ora_kd_build_faces(verts, tri_idx, 3, vis, n_tris);
}

void ora_kd_free(ora_kdtree *kd)
{
	if(!kd) return;
	free(kd->prim_bounds);
	free(kd->nodes);
	free(kd->leaf_prims);
	free(kd);
}

void ora_kd_bound(const ora_kdtree *kd, float a[3], float g[3])
{
	memcpy(a, kd->tree_bound.a, 3 * sizeof(float));
	memcpy(g, kd->tree_bound.g, 3 * sizeof(float));
}

uint32_t ora_kd_num_nodes(const ora_kdtree *kd) { return kd->n_nodes; }
uint32_t ora_kd_num_prims(const ora_kdtree *kd) { return kd->n_tris; }

/* ------------------------------------------------------------------ */
/* traversal: kdtree::intersect<Node, Stack, T>                         */

typedef struct { int32_t node; float t; float point[3]; int prev; } kstack;

enum { TEST_NEAREST = 0, TEST_SHADOW = 1, TEST_TRANSPARENT = 2 };

/* TransparentShadow state: the per-primitive transparency flags, the depth
 * limit and the transparent primitives met so far (the reference's `filtered`
 * set and depth counter). */
typedef struct { const uint8_t *transparent; int max_depth; uint32_t *prims; int count; } transp_state;

static inline int prim_test(const ora_kdtree *kd, uint32_t p, const float from[3], const float dir[3],
                            float *t, float *u, float *v)
{
	return ora_face_intersect(kd->verts, kd->idx, kd->stride, p, from, dir, t, u, v);
}

/* accelerator_kdtree_common.h:107-255.  Returns 1 for a hit (closest: *h set;
 * shadow and transparent shadow: *h->prim is the occluder). */
static int kd_intersect(const ora_kdtree *kd, const ora_ray *ray, float t_max, int test, ora_hit *h,
                        transp_state *ts)
{
	h->prim = ORA_NO_HIT; h->t = 0.f; h->u = 0.f; h->v = 0.f;
	float enter, leave;
	if(!ora_bound_cross(kd->tree_bound.a, kd->tree_bound.g, ray->org, ray->dir, t_max, &enter, &leave)) return 0;
	float inv_dir[3];
	for(int ax = 0; ax < 3; ++ax) inv_dir[ax] = (ray->dir[ax] == 0.f) ? FLT_MAX : 1.f / ray->dir[ax];
	kstack stack[KD_MAX_STACK];
	const knode *nodes = kd->nodes;
	int32_t cur = 0;
	int entry_id = 0;
	stack[entry_id].t = enter;
	if(enter >= 0.f)
		for(int ax = 0; ax < 3; ++ax) stack[entry_id].point[ax] = ray->org[ax] + ray->dir[ax] * enter;
	else
		for(int ax = 0; ax < 3; ++ax) stack[entry_id].point[ax] = ray->org[ax];
	int exit_id = 1;
	stack[exit_id].t = leave;
	for(int ax = 0; ax < 3; ++ax) stack[exit_id].point[ax] = ray->org[ax] + ray->dir[ax] * leave;
	stack[exit_id].node = -1;
	float cur_tmax = t_max; /* intersect_data.t_max_ */
	int hit = 0;
	const float bias = ora_dynamic_bias(enter, leave);
	const float t_min = (test == TEST_SHADOW) ? bias : ((ray->tmin < bias) ? bias : ray->tmin);

	while(cur >= 0 && stack[entry_id].t <= t_max)
	{
		while((nodes[cur].flags & 3u) != 3u)
		{
			const int axis = (int)(nodes[cur].flags & 3u);
			const float split_val = nodes[cur].u.split;
			int32_t far_child;
			if(stack[entry_id].point[axis] <= split_val)
			{
				if(stack[exit_id].point[axis] <= split_val) { ++cur; continue; }
				far_child = (int32_t)(nodes[cur].flags >> 2);
				++cur;
			}
			else
			{
				if(stack[exit_id].point[axis] > split_val) { cur = (int32_t)(nodes[cur].flags >> 2); continue; }
				far_child = cur + 1;
				cur = (int32_t)(nodes[cur].flags >> 2);
			}
			const float t = (split_val - ray->org[axis]) * inv_dir[axis];
			const int exit_prev = exit_id;
			++exit_id;
			if(exit_id == entry_id) ++exit_id;
			const int next_axis = (axis + 1) % 3, prev_axis = (axis + 2) % 3;
			stack[exit_id].prev = exit_prev;
			stack[exit_id].t = t;
			stack[exit_id].node = far_child;
			stack[exit_id].point[axis] = split_val;
			stack[exit_id].point[next_axis] = ray->org[next_axis] + t * ray->dir[next_axis];
			stack[exit_id].point[prev_axis] = ray->org[prev_axis] + t * ray->dir[prev_axis];
		}
		const uint32_t np = nodes[cur].flags >> 2;
		const uint32_t *plist = (np == 1) ? &nodes[cur].u.one_prim : kd->leaf_prims + nodes[cur].u.prim_off;
		for(uint32_t i = 0; i < np; ++i)
		{
			const uint32_t p = plist[i];
			float t, u, v;
			if(test == TEST_NEAREST)
			{
				/* Accelerator::primitiveIntersection, accelerator.h:122-133 */
				if(!prim_test(kd, p, ray->org, ray->dir, &t, &u, &v)) continue;
				if(t <= 0.f || t < t_min || t >= cur_tmax) continue;
				if(!(kd->vis[p] & ORA_VIS_VISIBLE)) continue;
				cur_tmax = t;
				h->prim = p; h->t = t; h->u = u; h->v = v;
				hit = 1;
			}
			else if(test == TEST_SHADOW)
			{
				/* Accelerator::primitiveIntersectionShadow, accelerator.h:135-145 */
				if(!prim_test(kd, p, ray->org, ray->dir, &t, &u, &v)) continue;
				if(t <= 0.f || t < t_min || t >= t_max) continue;
				if(!(kd->vis[p] & ORA_VIS_SHADOWS)) continue;
				h->prim = p; h->t = t; h->u = u; h->v = v;
				return 1;
			}
			else
			{
				/* Accelerator::primitiveIntersectionTransparentShadow, accelerator.h:147-168:
				 * an opaque caster occludes; a transparent one met for the first time
				 * is kept, and occludes once max_depth of them are kept */
				if(!prim_test(kd, p, ray->org, ray->dir, &t, &u, &v)) continue;
				if(t <= 0.f || t < t_min || t >= t_max) continue;
				if(!(kd->vis[p] & ORA_VIS_SHADOWS)) continue;
				if(!ts->transparent[p])
				{
					h->prim = p; h->t = t; h->u = u; h->v = v;
					return 1;
				}
				int seen = 0;
				for(int k = 0; k < ts->count; ++k) seen |= (ts->prims[k] == p);
				if(seen) continue;
				if(ts->count >= ts->max_depth)
				{
					h->prim = p; h->t = t; h->u = u; h->v = v;
					return 1;
				}
				ts->prims[ts->count++] = p;
			}
		}
		if(test == TEST_NEAREST && hit && cur_tmax <= stack[exit_id].t) return 1;
		entry_id = exit_id;
		cur = stack[exit_id].node;
		exit_id = stack[entry_id].prev;
	}
	if(test == TEST_NEAREST) return hit;
	h->prim = ORA_NO_HIT;
	return 0;
}

int ora_kd_intersect1(const ora_kdtree *kd, const ora_ray *ray, float t_max, int shadow, ora_hit *hit)
{
	return kd_intersect(kd, ray, t_max, shadow ? TEST_SHADOW : TEST_NEAREST, hit, NULL);
}

int ora_kd_transparent1(const ora_kdtree *kd, const ora_ray *ray, float t_max, const uint8_t *transparent,
                        int max_depth, uint32_t *prims, int *count)
{
	transp_state ts = {transparent, max_depth, prims, 0};
	ora_hit h;
	const int occ = kd_intersect(kd, ray, t_max, TEST_TRANSPARENT, &h, &ts);
	*count = ts.count;
	return occ;
}

typedef struct
{
	const ora_kdtree *kd;
	const ora_ray *rays;
	uint32_t n_rays;
	ora_hit *hits;
	uint8_t *occ;
	uint32_t *occ_prim;
	int test, tid, nthreads;
} kd_job;

static void *kd_worker(void *arg)
{
	const kd_job *j = (const kd_job *)arg;
	for(uint32_t r = (uint32_t)j->tid; r < j->n_rays; r += (uint32_t)j->nthreads)
	{
		const ora_ray *ray = &j->rays[r];
		const float t_max = (ray->tmax >= 0.f) ? ray->tmax : FLT_MAX;
		ora_hit h;
		const int hit = kd_intersect(j->kd, ray, t_max, j->test, &h, NULL);
		if(j->test == TEST_NEAREST) j->hits[r] = h;
		else
		{
			j->occ[r] = (uint8_t)hit;
			if(j->occ_prim) j->occ_prim[r] = hit ? h.prim : ORA_NO_HIT;
		}
	}
	return NULL;
}

static void kd_run(kd_job *proto, int threads)
{
	if(threads < 1) threads = 1;
	if(threads > 256) threads = 256;
	kd_job jobs[256];
	pthread_t th[256];
	for(int i = 0; i < threads; ++i) { jobs[i] = *proto; jobs[i].tid = i; jobs[i].nthreads = threads; }
	if(threads == 1) { kd_worker(&jobs[0]); return; }
	for(int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, kd_worker, &jobs[i]);
	for(int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

void ora_kd_closest(const ora_kdtree *kd, const ora_ray *rays, uint32_t n_rays, ora_hit *hits, int threads)
{
	kd_job j;
	memset(&j, 0, sizeof j);
	j.kd = kd; j.rays = rays; j.n_rays = n_rays; j.hits = hits; j.test = TEST_NEAREST;
	kd_run(&j, threads);
}

void ora_kd_shadow(const ora_kdtree *kd, const ora_ray *rays, uint32_t n_rays, uint8_t *occluded,
                   uint32_t *prim, int threads)
{
	kd_job j;
	memset(&j, 0, sizeof j);
	j.kd = kd; j.rays = rays; j.n_rays = n_rays; j.occ = occluded; j.occ_prim = prim; j.test = TEST_SHADOW;
	kd_run(&j, threads);
}

uint32_t ora_kd_size(const ora_kdtree *kd) { return kd ? kd->n_tris : 0u; }
