// treelet.h -- SAH treelet restructuring of a binary BVH (T. Karras and
// T. Aila, "Fast Parallel Construction of High-Quality Bounding Volume
// Hierarchies", HPG 2013), one treelet at a time.
//
// A treelet rooted at internal node R is R plus up to five of its internal
// descendants, grown by repeatedly opening the treelet leaf of largest
// surface area; its (up to seven) treelet leaves are whole subtrees or
// triangles, and they stay as they are.  Over the 2^n subsets of the n
// treelet leaves a dynamic program finds the topology of least SAH cost
//     c(S) = C_node * A(S) + min over partitions {P, S \ P} of c(P) + c(S \ P)
// (c of one treelet leaf: its subtree's cost, C_tri * A for a triangle), and
// when that beats the current topology the treelet's internal nodes are
// rewritten with it, reusing the same node slots (R stays R).
//
// Node layout: that of accel_internal.h (both children's boxes in the
// parent, refs in the fourth float4).  This header is shared by the build
// kernel (one thread per treelet, DP tables in LDS) and the CPU check of
// the restructuring, so it holds no device-only code.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define YMX_TL_FN __host__ __device__ inline
#else
#include "ray_tri.h"  // float4 for plain C++
#define YMX_TL_FN inline
#endif

namespace ymx {

constexpr int kTreeletLeaves = 7;
constexpr int kTreeletSubsets = 1 << kTreeletLeaves;
// SAH weights: a node visit against one triangle test (Karras and Aila's values)
constexpr float kSahNode = 1.2f;
constexpr float kSahTri = 1.0f;

struct TreeletBox
{
	float lo[3], hi[3];
};

YMX_TL_FN float tl_half_area(const TreeletBox &b)
{
	const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
	return dx * dy + dy * dz + dz * dx;
}

YMX_TL_FN void tl_read_slot(const float4 *nodes, int32_t node, int slot, TreeletBox &b)
{
	const float *f = reinterpret_cast<const float *>(nodes + 4 * (size_t)node);
	b.lo[0] = f[4 * slot + 0];
	b.hi[0] = f[4 * slot + 1];
	b.lo[1] = f[4 * slot + 2];
	b.hi[1] = f[4 * slot + 3];
	b.lo[2] = f[8 + 2 * slot];
	b.hi[2] = f[9 + 2 * slot];
}

YMX_TL_FN void tl_refs(const float4 *nodes, int32_t node, int32_t &c0, int32_t &c1)
{
	const float *f = reinterpret_cast<const float *>(nodes + 4 * (size_t)node);
	const int32_t *r = reinterpret_cast<const int32_t *>(f + 12);
	c0 = r[0];
	c1 = r[1];
}

YMX_TL_FN void tl_write_node(float4 *nodes, int32_t node, const TreeletBox &b0, const TreeletBox &b1,
                                              int32_t c0, int32_t c1)
{
	float *f = reinterpret_cast<float *>(nodes + 4 * (size_t)node);
	f[0] = b0.lo[0];
	f[1] = b0.hi[0];
	f[2] = b0.lo[1];
	f[3] = b0.hi[1];
	f[4] = b1.lo[0];
	f[5] = b1.hi[0];
	f[6] = b1.lo[1];
	f[7] = b1.hi[1];
	f[8] = b0.lo[2];
	f[9] = b0.hi[2];
	f[10] = b1.lo[2];
	f[11] = b1.hi[2];
	int32_t *r = reinterpret_cast<int32_t *>(f + 12);
	r[0] = c0;
	r[1] = c1;
	r[2] = 0;
	r[3] = 0;
}

YMX_TL_FN void tl_union(TreeletBox &a, const TreeletBox &b)
{
	for(int k = 0; k < 3; ++k)
	{
		a.lo[k] = a.lo[k] < b.lo[k] ? a.lo[k] : b.lo[k];
		a.hi[k] = a.hi[k] > b.hi[k] ? a.hi[k] : b.hi[k];
	}
}

// The tree as the restructuring sees it.  Leaves are one-triangle refs
// (0x80000000 | first << 4); parent_leaf is indexed by the triangle.
struct TreeletTree
{
	float4 *nodes;
	int32_t *parent_int;   // internal node -> parent (root: any)
	int32_t *parent_leaf;  // triangle -> parent
	float *cost;           // SAH cost of each internal node's subtree
	uint32_t *count;       // triangles below each internal node
};

// Sets cost and count of internal node r from its children (final subtrees).
YMX_TL_FN void tl_node_cost(const TreeletTree &t, int32_t r)
{
	TreeletBox b0, b1;
	tl_read_slot(t.nodes, r, 0, b0);
	tl_read_slot(t.nodes, r, 1, b1);
	int32_t c0, c1;
	tl_refs(t.nodes, r, c0, c1);
	TreeletBox u = b0;
	tl_union(u, b1);
	const float cc0 = c0 >= 0 ? t.cost[c0] : kSahTri * tl_half_area(b0);
	const float cc1 = c1 >= 0 ? t.cost[c1] : kSahTri * tl_half_area(b1);
	t.cost[r] = kSahNode * tl_half_area(u) + cc0 + cc1;
	t.count[r] = (c0 >= 0 ? t.count[c0] : 1u) + (c1 >= 0 ? t.count[c1] : 1u);
}

// Restructures the treelet rooted at r (both of r's subtrees final, their
// cost and count set), then sets cost[r] and count[r].  dp_cost / dp_part:
// kTreeletSubsets entries each, element s at [s * stride].  Returns 1 when the
// treelet was rewritten.
YMX_TL_FN int tl_restructure(const TreeletTree &t, int32_t r, float *dp_cost, uint8_t *dp_part,
                                              int stride)
{
	int32_t leaf[kTreeletLeaves];
	TreeletBox box[kTreeletLeaves];
	float area[kTreeletLeaves];
	int32_t inner[kTreeletLeaves - 1];
	int n = 2, n_inner = 1;
	inner[0] = r;
	tl_refs(t.nodes, r, leaf[0], leaf[1]);
	tl_read_slot(t.nodes, r, 0, box[0]);
	tl_read_slot(t.nodes, r, 1, box[1]);
	area[0] = tl_half_area(box[0]);
	area[1] = tl_half_area(box[1]);
	TreeletBox root_box = box[0];
	tl_union(root_box, box[1]);
	float old_cost = kSahNode * tl_half_area(root_box);
	// grow: open the internal treelet leaf of largest area
	while(n < kTreeletLeaves)
	{
		int k = -1;
		float best = -1.f;
		for(int i = 0; i < n; ++i)
			if(leaf[i] >= 0 && area[i] > best)
			{
				best = area[i];
				k = i;
			}
		if(k < 0) break;
		const int32_t x = leaf[k];
		inner[n_inner++] = x;
		old_cost += kSahNode * area[k];
		int32_t c0, c1;
		tl_refs(t.nodes, x, c0, c1);
		leaf[k] = c0;
		tl_read_slot(t.nodes, x, 0, box[k]);
		area[k] = tl_half_area(box[k]);
		leaf[n] = c1;
		tl_read_slot(t.nodes, x, 1, box[n]);
		area[n] = tl_half_area(box[n]);
		++n;
	}
	float leaf_cost[kTreeletLeaves];
	uint32_t leaf_count[kTreeletLeaves];
	for(int i = 0; i < n; ++i)
	{
		leaf_cost[i] = leaf[i] >= 0 ? t.cost[leaf[i]] : kSahTri * area[i];
		leaf_count[i] = leaf[i] >= 0 ? t.count[leaf[i]] : 1u;
		old_cost += leaf_cost[i];
	}
	int rewritten = 0;
	if(n >= 3)
	{
		const uint32_t full = (1u << n) - 1u;
		for(uint32_t s = 1; s <= full; ++s)
		{
			if((s & (s - 1u)) == 0u)
			{
				int k = 0;
				while(!((s >> k) & 1u)) ++k;
				dp_cost[s * stride] = leaf_cost[k];
				continue;
			}
			TreeletBox u;
			bool first = true;
			for(int k = 0; k < n; ++k)
				if((s >> k) & 1u)
				{
					if(first) u = box[k];
					else tl_union(u, box[k]);
					first = false;
				}
			// partitions {P, S \ P} with P holding the lowest member of S
			const uint32_t low = s & (0u - s), rest = s ^ low;
			float best = 3.4e38f;
			uint32_t best_p = low;
			for(uint32_t q = (rest - 1u) & rest;; q = (q - 1u) & rest)
			{
				const uint32_t p = low | q;
				const float v = dp_cost[p * stride] + dp_cost[(s ^ p) * stride];
				if(v < best)
				{
					best = v;
					best_p = p;
				}
				if(q == 0u) break;
			}
			dp_cost[s * stride] = kSahNode * tl_half_area(u) + best;
			dp_part[s * stride] = (uint8_t)best_p;
		}
		if(dp_cost[full * stride] < old_cost * (1.f - 1e-6f))
		{
			// rewrite top-down; internal nodes of the new topology take the
			// treelet's node slots in order (r first)
			uint32_t stack_set[kTreeletLeaves];
			int32_t stack_node[kTreeletLeaves];
			int sp = 0, next = 1;
			stack_set[sp] = full;
			stack_node[sp] = r;
			++sp;
			while(sp > 0)
			{
				--sp;
				const uint32_t s = stack_set[sp];
				const int32_t nd = stack_node[sp];
				const uint32_t p = dp_part[s * stride];
				const uint32_t sets[2] = {p, s ^ p};
				TreeletBox cb[2];
				int32_t cr[2];
				for(int c = 0; c < 2; ++c)
				{
					const uint32_t cs = sets[c];
					bool first = true;
					uint32_t cnt = 0;
					for(int k = 0; k < n; ++k)
						if((cs >> k) & 1u)
						{
							if(first) cb[c] = box[k];
							else tl_union(cb[c], box[k]);
							first = false;
							cnt += leaf_count[k];
						}
					if((cs & (cs - 1u)) == 0u)
					{
						int k = 0;
						while(!((cs >> k) & 1u)) ++k;
						cr[c] = leaf[k];
						if(leaf[k] >= 0) t.parent_int[leaf[k]] = nd;
						else t.parent_leaf[((uint32_t)leaf[k] & 0x7FFFFFFFu) >> 4] = nd;
					}
					else
					{
						const int32_t m = inner[next++];
						cr[c] = m;
						t.parent_int[m] = nd;
						t.cost[m] = dp_cost[cs * stride];
						t.count[m] = cnt;
						stack_set[sp] = cs;
						stack_node[sp] = m;
						++sp;
					}
				}
				tl_write_node(t.nodes, nd, cb[0], cb[1], cr[0], cr[1]);
			}
			rewritten = 1;
		}
	}
	tl_node_cost(t, r);
	return rewritten;
}

}  // namespace ymx
