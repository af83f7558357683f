// accel_build.hip -- GPU construction of the linear BVH used by the
// "mi355x-lbvh" accelerator, and the tree bound the reference's kd-tree
// exposes through getBound().
//
// Pipeline (all on the device, one stream):
//   prep     one thread per triangle: exact FP32 bounds, centroid, the 48-byte
//            triangle record (v0, v1 - v0, v2 - v0); wave-reduced scene and
//            centroid bounds folded in with ordered-int atomics.
//   morton   63-bit Morton code of the centroid (21 bits per axis).
//   sort     radix sort of (code, triangle) pairs.
//   reorder  triangle records into leaf order.
//   topology, by the builder chosen at create (ymx_accel_params.builder):
//     LBVH (default): one thread per internal node finds its key range and
//            split (T. Karras, "Maximizing Parallelism in the Construction of
//            BVHs, Octrees, and k-d Trees", HPG 2012), writing child refs and
//            parent links; then one thread per leaf climbs to the root and
//            the second arrival at a node unions both child boxes (refit).
//     PLOC:  agglomerative clustering in Morton order (see k_ploc_nn),
//            writing nodes, boxes and parent links as pairs merge.
//   depth    longest root-to-leaf chain, sizing the traversal stack.
//
// The reference builds a SAH kd-tree on the CPU instead
// (src/accelerator/accelerator_kdtree_original.cc:65-645); only its query
// results and its tree bound (:89-103) are part of the contract.
#include "accel_internal.h"
#include "treelet.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <float.h>
#include <math.h>
#include <string.h>

#include <utility>
#include <vector>

namespace {

constexpr int kBuildBlock = 256;

__device__ __forceinline__ int ordered_from_float(float f)
{
	const int i = __float_as_int(f);
	return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__host__ inline float float_from_ordered(int i)
{
	const int b = i >= 0 ? i : i ^ 0x7FFFFFFF;
	float f;
	memcpy(&f, &b, sizeof f);
	return f;
}

// bounds buffer layout (ordered ints): [0..2] scene min, [3..5] scene max,
// [6..8] centroid min, [9..11] centroid max
__global__ void k_init_bounds(int *b)
{
	const int i = threadIdx.x;
	if(i < 12) b[i] = ((i / 3) % 2 == 0) ? 0x7FFFFFFF : (int)0x80000000;
}

__global__ __launch_bounds__(kBuildBlock) void k_prep(const float *__restrict__ verts,
                                                      const uint32_t *__restrict__ idx, uint32_t stride,
                                                      const uint32_t *__restrict__ quad_slot,
                                                      const uint8_t *__restrict__ vis, uint32_t n,
                                                      float4 *__restrict__ tris_unsorted,
                                                      float4 *__restrict__ quads,
                                                      float *__restrict__ prim_box,
                                                      float *__restrict__ centroid, int *bounds)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	if(i < n)
	{
		const uint32_t a = idx[stride * (size_t)i], b = idx[stride * (size_t)i + 1], c = idx[stride * (size_t)i + 2];
		const uint32_t slot = quad_slot ? quad_slot[i] : 0u;
		float v[4][3];
		for(int k = 0; k < 3; ++k)
		{
			v[0][k] = verts[3 * (size_t)a + k];
			v[1][k] = verts[3 * (size_t)b + k];
			v[2][k] = verts[3 * (size_t)c + k];
		}
		for(int k = 0; k < 3; ++k)
		{
			lo[k] = fminf(v[0][k], fminf(v[1][k], v[2][k]));
			hi[k] = fmaxf(v[0][k], fmaxf(v[1][k], v[2][k]));
		}
		if(slot)
		{
			// FacePrimitive::getBound over all four vertices
			const uint32_t d = idx[stride * (size_t)i + 3];
			for(int k = 0; k < 3; ++k)
			{
				v[3][k] = verts[3 * (size_t)d + k];
				lo[k] = fminf(lo[k], v[3][k]);
				hi[k] = fmaxf(hi[k], v[3][k]);
			}
			quads[slot - 1u] = make_float4(v[3][0] - v[0][0], v[3][1] - v[0][1], v[3][2] - v[0][2], 0.f);
		}
		for(int k = 0; k < 3; ++k)
		{
			clo[k] = chi[k] = 0.5f * (lo[k] + hi[k]);
			prim_box[6 * (size_t)i + k] = lo[k];
			prim_box[6 * (size_t)i + 3 + k] = hi[k];
			centroid[3 * (size_t)i + k] = clo[k];
		}
		const uint32_t vmask = vis ? (uint32_t)vis[i] : 3u;
		tris_unsorted[3 * (size_t)i + 0] = make_float4(v[0][0], v[0][1], v[0][2], __uint_as_float(i));
		tris_unsorted[3 * (size_t)i + 1] = make_float4(v[1][0] - v[0][0], v[1][1] - v[0][1],
		                                               v[1][2] - v[0][2], __uint_as_float(vmask));
		tris_unsorted[3 * (size_t)i + 2] = make_float4(v[2][0] - v[0][0], v[2][1] - v[0][1],
		                                               v[2][2] - v[0][2], __uint_as_float(slot));
	}
	// wave64 reduction, then one atomic per wave
	for(int off = 32; off > 0; off >>= 1)
	{
		for(int k = 0; k < 3; ++k)
		{
			lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
			hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
			clo[k] = fminf(clo[k], __shfl_xor(clo[k], off));
			chi[k] = fmaxf(chi[k], __shfl_xor(chi[k], off));
		}
	}
	if((threadIdx.x & 63) == 0)
	{
		for(int k = 0; k < 3; ++k)
		{
			atomicMin(&bounds[k], ordered_from_float(lo[k]));
			atomicMax(&bounds[3 + k], ordered_from_float(hi[k]));
			atomicMin(&bounds[6 + k], ordered_from_float(clo[k]));
			atomicMax(&bounds[9 + k], ordered_from_float(chi[k]));
		}
	}
}

__device__ __forceinline__ uint64_t spread_bits_21(uint32_t x)
{
	uint64_t v = x & 0x1FFFFFu;
	v = (v | (v << 32)) & 0x1F00000000FFFFull;
	v = (v | (v << 16)) & 0x1F0000FF0000FFull;
	v = (v | (v << 8)) & 0x100F00F00F00F00Full;
	v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
	v = (v | (v << 2)) & 0x1249249249249249ull;
	return v;
}

__global__ __launch_bounds__(kBuildBlock) void k_morton(const float *__restrict__ centroid, uint32_t n,
                                                        float cx, float cy, float cz, float sx, float sy,
                                                        float sz, uint64_t *__restrict__ keys,
                                                        uint32_t *__restrict__ vals)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const float scale = 2097151.f;
	const float fx = fminf(fmaxf((centroid[3 * (size_t)i + 0] - cx) * sx, 0.f), 1.f) * scale;
	const float fy = fminf(fmaxf((centroid[3 * (size_t)i + 1] - cy) * sy, 0.f), 1.f) * scale;
	const float fz = fminf(fmaxf((centroid[3 * (size_t)i + 2] - cz) * sz, 0.f), 1.f) * scale;
	const uint64_t code = (spread_bits_21((uint32_t)fx) << 2) | (spread_bits_21((uint32_t)fy) << 1) |
	                      spread_bits_21((uint32_t)fz);
	keys[i] = code;
	vals[i] = i;
}

__device__ __forceinline__ int karras_delta(const uint64_t *__restrict__ keys, int n, int i, int j)
{
	if(j < 0 || j >= n) return -1;
	const uint64_t a = keys[i], b = keys[j];
	if(a == b) return 64 + __clz((uint32_t)(i ^ j));
	return __clzll((long long)(a ^ b));
}

__global__ __launch_bounds__(kBuildBlock) void k_karras(const uint64_t *__restrict__ keys, int n,
                                                        float4 *__restrict__ nodes,
                                                        int32_t *__restrict__ parent_internal,
                                                        int32_t *__restrict__ parent_leaf, uint2 *__restrict__ range)
{
	const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if(i >= n - 1) return;
	const int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = karras_delta(keys, n, i, i - d);
	int lmax = 2;
	while(karras_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
	int l = 0;
	for(int t = lmax >> 1; t >= 1; t >>= 1)
		if(karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
	const int j = i + l * d;
	const int dnode = karras_delta(keys, n, i, j);
	int s = 0;
	int t = l;
	do
	{
		t = (t + 1) >> 1;
		if(karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
	} while(t > 1);
	const int gamma = i + s * d + (d < 0 ? -1 : 0);
	const int lo = i < j ? i : j;
	const int hi = i < j ? j : i;
	range[i] = make_uint2((uint32_t)lo, (uint32_t)hi);
	int32_t c0, c1;
	if(lo == gamma)
	{
		c0 = ymx_leaf_ref((uint32_t)gamma, 1u);
		parent_leaf[gamma] = i;
	}
	else
	{
		c0 = gamma;
		parent_internal[gamma] = i;
	}
	if(hi == gamma + 1)
	{
		c1 = ymx_leaf_ref((uint32_t)(gamma + 1), 1u);
		parent_leaf[gamma + 1] = i;
	}
	else
	{
		c1 = gamma + 1;
		parent_internal[gamma + 1] = i;
	}
	nodes[4 * (size_t)i + 3] = make_float4(__int_as_float(c0), __int_as_float(c1), 0.f, 0.f);
}

__global__ __launch_bounds__(kBuildBlock) void k_reorder(const float4 *__restrict__ src,
                                                         const uint32_t *__restrict__ order, uint32_t n,
                                                         float4 *__restrict__ dst)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= n) return;
	const uint32_t s = order[j];
	dst[3 * (size_t)j + 0] = src[3 * (size_t)s + 0];
	dst[3 * (size_t)j + 1] = src[3 * (size_t)s + 1];
	dst[3 * (size_t)j + 2] = src[3 * (size_t)s + 2];
}

__device__ __forceinline__ float load_coherent(const float *p)
{
	return __uint_as_float(__hip_atomic_load((const uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void store_slot(float4 *nodes, int node, int slot, const float lo[3], const float hi[3])
{
	float *f = (float *)(nodes + 4 * (size_t)node);
	// slot 0: f0 = (lo.x, hi.x, lo.y, hi.y), f2.xy = (lo.z, hi.z); slot 1: f1, f2.zw
	float *xy = f + 4 * slot;
	xy[0] = lo[0];
	xy[1] = hi[0];
	xy[2] = lo[1];
	xy[3] = hi[1];
	f[8 + 2 * slot] = lo[2];
	f[9 + 2 * slot] = hi[2];
}

__global__ __launch_bounds__(kBuildBlock) void k_refit(float4 *nodes, const float *__restrict__ prim_box,
                                                       const uint32_t *__restrict__ order, uint32_t n,
                                                       const int32_t *__restrict__ parent_internal,
                                                       const int32_t *__restrict__ parent_leaf,
                                                       uint32_t *flags, float pad)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= n) return;
	const uint32_t p = order[j];
	float lo[3], hi[3];
	for(int k = 0; k < 3; ++k)
	{
		lo[k] = prim_box[6 * (size_t)p + k] - pad;
		hi[k] = prim_box[6 * (size_t)p + 3 + k] + pad;
	}
	int32_t child_ref = ymx_leaf_ref(j, 1u);
	int32_t node = parent_leaf[j];
	while(node >= 0)
	{
		const int32_t c0 = __float_as_int(nodes[4 * (size_t)node + 3].x);
		const int slot = (c0 == child_ref) ? 0 : 1;
		store_slot(nodes, node, slot, lo, hi);
		__threadfence();
		const uint32_t old = atomicAdd(&flags[node], 1u);
		if(old == 0) return;  // sibling not done yet; it will continue upwards
		__threadfence();
		const float *f = (const float *)(nodes + 4 * (size_t)node);
		const int other = 1 - slot;
		const float olo[3] = {load_coherent(f + 4 * other + 0), load_coherent(f + 4 * other + 2),
		                      load_coherent(f + 8 + 2 * other)};
		const float ohi[3] = {load_coherent(f + 4 * other + 1), load_coherent(f + 4 * other + 3),
		                      load_coherent(f + 9 + 2 * other)};
		for(int k = 0; k < 3; ++k)
		{
			lo[k] = fminf(lo[k], olo[k]);
			hi[k] = fmaxf(hi[k], ohi[k]);
		}
		child_ref = node;
		node = (node == 0) ? -1 : parent_internal[node];
	}
}

// Leaves of up to max_leaf triangles: an LBVH node covers the contiguous leaf
// range [lo, hi], so a child whose range is that small becomes one leaf over
// those triangles (its box, set by refit, is already their union).  The
// nodes below it are no longer reached.
__global__ __launch_bounds__(kBuildBlock) void k_collapse_leaves(float4 *__restrict__ nodes,
                                                                 const uint2 *__restrict__ range, uint32_t n,
                                                                 uint32_t max_leaf)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i + 1 >= n) return;
	float4 f3 = nodes[4 * (size_t)i + 3];
	int32_t c[2] = {__float_as_int(f3.x), __float_as_int(f3.y)};
	for(int k = 0; k < 2; ++k)
	{
		if(c[k] < 0) continue;
		const uint2 r = range[c[k]];
		const uint32_t size = r.y - r.x + 1u;
		if(size <= max_leaf) c[k] = ymx_leaf_ref(r.x, size);
	}
	f3.x = __int_as_float(c[0]);
	f3.y = __int_as_float(c[1]);
	nodes[4 * (size_t)i + 3] = f3;
}

__global__ __launch_bounds__(kBuildBlock) void k_depth(const int32_t *__restrict__ parent_internal,
                                                       const int32_t *__restrict__ parent_leaf, uint32_t n,
                                                       uint32_t *max_depth)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t d = 0;
	if(j < n)
	{
		int32_t node = parent_leaf[j];
		while(node >= 0 && d <= n)  // at most n - 1 internal nodes on a path
		{
			++d;
			node = (node == 0) ? -1 : parent_internal[node];
		}
	}
	for(int off = 32; off > 0; off >>= 1) d = max(d, (uint32_t)__shfl_xor((int)d, off));
	if((threadIdx.x & 63) == 0) atomicMax(max_depth, d);
}

// ---- SAH treelet restructuring (treelet.h), bottom-up over the whole tree:
// one thread per leaf climbs towards the root; the second arrival at a node
// finds both of its subtrees final, restructures the treelet rooted there
// (when it holds at least min_tris triangles) and goes on up.  The treelet's
// DP tables live in LDS, subset-major so that the lanes of a wave touch
// adjacent words.
constexpr int kTreeletBlock = 64;

__global__ __launch_bounds__(kTreeletBlock) void k_treelets(float4 *nodes, int32_t *parent_int, int32_t *parent_leaf,
                                                            uint32_t n, uint32_t *flags, float *cost, uint32_t *count,
                                                            uint32_t min_tris, uint32_t *rewritten)
{
	// rewritten[0]: treelets rewritten; rewritten[1]: set on an inconsistent tree
	__shared__ float dp_cost[ymx::kTreeletSubsets * kTreeletBlock];
	__shared__ uint8_t dp_part[ymx::kTreeletSubsets * kTreeletBlock];
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= n) return;
	const ymx::TreeletTree t{nodes, parent_int, parent_leaf, cost, count};
	uint32_t done = 0;
	int32_t node = parent_leaf[j];
	while(node >= 0)
	{
		// the first arrival leaves; the second finds both subtrees final (the
		// fences order the other thread's restructuring writes before its
		// arrival and this thread's reads after it)
		__threadfence();
		const uint32_t arrivals = atomicAdd(&flags[node], 1u);
		if(arrivals == 0u) break;
		if(arrivals > 1u)
		{
			// a node reached three times: the parent links are not a tree
			atomicOr(rewritten + 1, 1u);
			break;
		}
		__threadfence();
		uint32_t c = 0;
		int32_t c0, c1;
		ymx::tl_refs(nodes, node, c0, c1);
		c = (c0 >= 0 ? count[c0] : 1u) + (c1 >= 0 ? count[c1] : 1u);
		if(c >= min_tris) done += (uint32_t)ymx::tl_restructure(t, node, dp_cost + threadIdx.x, dp_part + threadIdx.x,
		                                                        kTreeletBlock);
		else ymx::tl_node_cost(t, node);
		node = (node == 0) ? -1 : parent_int[node];
	}
	if(done) atomicAdd(rewritten, done);
}

// ---- PLOC: parallel locally-ordered clustering (D. Meister, J. Bittner,
// "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy
// Construction", IEEE TVCG 2018).  Clusters stay in Morton order; each
// iteration every cluster finds the neighbour within kPlocRadius positions
// whose union box has the smallest surface area, mutual nearest neighbours
// merge into a new internal node, and the survivors are compacted.  Node
// indices are handed out downwards from n - 2, so the last merge, the root,
// is node 0.
constexpr int kPlocRadius = 16;

struct Cluster
{
	float4 lo, hi;  // xyz: box; lo.w: node ref bits
};

__device__ __forceinline__ float union_area(const Cluster &a, const Cluster &b)
{
	const float dx = fmaxf(a.hi.x, b.hi.x) - fminf(a.lo.x, b.lo.x);
	const float dy = fmaxf(a.hi.y, b.hi.y) - fminf(a.lo.y, b.lo.y);
	const float dz = fmaxf(a.hi.z, b.hi.z) - fminf(a.lo.z, b.lo.z);
	return dx * dy + dy * dz + dz * dx;
}

__global__ __launch_bounds__(kBuildBlock) void k_ploc_init(const float *__restrict__ prim_box,
                                                           const uint32_t *__restrict__ order, uint32_t n, float pad,
                                                           Cluster *__restrict__ c)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= n) return;
	const uint32_t p = order[j];
	const float *b = prim_box + 6 * (size_t)p;
	Cluster k;
	k.lo = make_float4(b[0] - pad, b[1] - pad, b[2] - pad, __int_as_float(ymx_leaf_ref(j, 1u)));
	k.hi = make_float4(b[3] + pad, b[4] + pad, b[5] + pad, 0.f);
	c[j] = k;
}

// Nearest neighbour of each cluster (smallest union area; ties to the lower index).
__global__ __launch_bounds__(kBuildBlock) void k_ploc_nn(const Cluster *__restrict__ c, uint32_t count,
                                                         int32_t *__restrict__ nn)
{
	__shared__ Cluster tile[kBuildBlock + 2 * kPlocRadius];
	const int base = (int)(blockIdx.x * blockDim.x) - kPlocRadius;
	for(int t = threadIdx.x; t < kBuildBlock + 2 * kPlocRadius; t += kBuildBlock)
	{
		const int g = base + t;
		if(g >= 0 && g < (int)count) tile[t] = c[g];
	}
	__syncthreads();
	const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if(i >= (int)count) return;
	const Cluster me = tile[i - base];
	float best = INFINITY;
	int best_j = -1;
	const int j0 = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
	const int j1 = i + kPlocRadius >= (int)count ? (int)count - 1 : i + kPlocRadius;
	for(int j = j0; j <= j1; ++j)
	{
		if(j == i) continue;
		const float a = union_area(me, tile[j - base]);
		if(a < best)  // ascending j: ties keep the lower index
		{
			best = a;
			best_j = j;
		}
	}
	nn[i] = best_j;
}

// keep[i]: cluster i survives (merged or not); merge[i]: i starts a merge.
__global__ __launch_bounds__(kBuildBlock) void k_ploc_flags(const int32_t *__restrict__ nn, uint32_t count,
                                                            uint32_t *__restrict__ keep, uint32_t *__restrict__ merge)
{
	const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if(i >= (int)count) return;
	const int j = nn[i];
	const bool mutual = j >= 0 && nn[j] == i;
	keep[i] = (mutual && j < i) ? 0u : 1u;
	merge[i] = (mutual && i < j) ? 1u : 0u;
}

// Every internal node must hold two child refs in range (d_nodes starts as
// 0xFF bytes, so a node no merge wrote has YMX_REF_EMPTY refs).
__global__ __launch_bounds__(kBuildBlock) void k_check_tree(const float4 *__restrict__ nodes, uint32_t n_tris,
                                                            uint32_t *__restrict__ bad)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i + 1 >= n_tris) return;
	const float4 f3 = nodes[4 * (size_t)i + 3];
	const int32_t r[2] = {__float_as_int(f3.x), __float_as_int(f3.y)};
	for(int k = 0; k < 2; ++k)
	{
		const bool ok = r[k] >= 0 ? (uint32_t)r[k] < n_tris - 1u && (uint32_t)r[k] != i
		                          : r[k] != YMX_REF_EMPTY && ymx_leaf_first(r[k]) < n_tris && ymx_leaf_count(r[k]) == 1u;
		if(!ok) atomicOr(bad, 1u);
	}
}

// ---- node relayout: depth-first order in which the internal children of a
// node take adjacent slots (a node's children first, then the whole subtree
// of its first internal child, then that of its second).  Root stays node 0.
// Nodes no ref reaches (below a collapsed leaf) are dropped.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t load_u32_coherent(const uint32_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBuildBlock) void k_lay_parents(const float4 *__restrict__ nodes, uint32_t n_int,
                                                             int32_t *__restrict__ par, uint32_t *__restrict__ nint)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n_int) return;
	const float4 f3 = nodes[4 * (size_t)i + 3];
	const int32_t r[2] = {__float_as_int(f3.x), __float_as_int(f3.y)};
	uint32_t k = 0;
	for(int s = 0; s < 2; ++s)
		if(r[s] >= 0 && (uint32_t)r[s] < n_int)
		{
			par[r[s]] = (int32_t)i;
			++k;
		}
	nint[i] = k;
}

__device__ __forceinline__ void lay_sums(const float4 *nodes, uint32_t n, uint32_t *size, uint32_t *tcount)
{
	const float4 f3 = nodes[4 * (size_t)n + 3];
	const int32_t r[2] = {__float_as_int(f3.x), __float_as_int(f3.y)};
	uint32_t s = 1u, t = 0u;
	for(int k = 0; k < 2; ++k)
	{
		if(r[k] >= 0)
		{
			s += load_u32_coherent(&size[r[k]]);
			if(tcount) t += load_u32_coherent(&tcount[r[k]]);
		}
		else if(r[k] != YMX_REF_EMPTY) t += ymx_leaf_count(r[k]);
	}
	__hip_atomic_store(&size[n], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	if(tcount) __hip_atomic_store(&tcount[n], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Internal nodes (size) and, when tcount is given, triangles per subtree,
// bottom-up: the last child to finish sums its parent (the arrival count of
// refit, over internal children only).
__global__ __launch_bounds__(kBuildBlock) void k_lay_size(const float4 *__restrict__ nodes, uint32_t n_int,
                                                          const int32_t *__restrict__ par,
                                                          const uint32_t *__restrict__ nint, uint32_t *arrive,
                                                          uint32_t *size, uint32_t *tcount, uint32_t *bad)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n_int || nint[i] != 0u) return;
	lay_sums(nodes, i, size, tcount);
	uint32_t node = i;
	for(uint32_t guard = 0; node != 0u; ++guard)
	{
		const int32_t p = par[node];
		if(p < 0) return;  // top of a subtree no ref reaches
		if(guard > n_int)
		{
			atomicOr(bad, 1u);
			return;
		}
		__threadfence();
		const uint32_t old = atomicAdd(&arrive[p], 1u);
		if(old + 1u < nint[p]) return;
		__threadfence();
		lay_sums(nodes, (uint32_t)p, size, tcount);
		node = (uint32_t)p;
	}
}

// Slot of node i: with B(n) the first slot of n's descendants (B(root) = 1),
// a node's internal children sit at B(n) and B(n) + 1, and
// B(child) = B(n) + nint(n) + (second child after an internal first one ?
// size(first) - 1 : 0).  Walking up from i sums these offsets.
__global__ __launch_bounds__(kBuildBlock) void k_lay_pos(const float4 *__restrict__ nodes, uint32_t n_int,
                                                         const int32_t *__restrict__ par,
                                                         const uint32_t *__restrict__ nint,
                                                         const uint32_t *__restrict__ size, uint32_t *__restrict__ pos,
                                                         uint32_t *bad)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n_int) return;
	if(i == 0u)
	{
		pos[0] = 0u;
		return;
	}
	uint32_t c = i, acc = 0u;
	bool own = true;
	for(uint32_t guard = 0; c != 0u; ++guard)
	{
		const int32_t p = par[c];
		if(p < 0)
		{
			pos[i] = kNoSlot;
			return;
		}
		if(guard > n_int)
		{
			atomicOr(bad, 1u);
			pos[i] = kNoSlot;
			return;
		}
		const float4 f3 = nodes[4 * (size_t)p + 3];
		const int32_t r0 = __float_as_int(f3.x), r1 = __float_as_int(f3.y);
		const bool after_internal = r1 == (int32_t)c && r0 != (int32_t)c && r0 >= 0;
		if(own) acc += after_internal ? 1u : 0u;
		else acc += nint[p] + (after_internal ? size[r0] - 1u : 0u);
		own = false;
		c = (uint32_t)p;
	}
	pos[i] = acc + 1u;
}

__global__ __launch_bounds__(kBuildBlock) void k_lay_write(const float4 *__restrict__ src, uint32_t n_int,
                                                           const uint32_t *__restrict__ pos, uint32_t n_out,
                                                           float4 *__restrict__ dst, uint32_t *bad)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n_int) return;
	const uint32_t d = pos[i];
	if(d == kNoSlot) return;
	if(d >= n_out)
	{
		atomicOr(bad, 1u);
		return;
	}
	float4 f3 = src[4 * (size_t)i + 3];
	int32_t r[2] = {__float_as_int(f3.x), __float_as_int(f3.y)};
	for(int k = 0; k < 2; ++k)
		if(r[k] >= 0)
		{
			const uint32_t t = pos[r[k]];
			if(t >= n_out)
			{
				atomicOr(bad, 1u);
				return;
			}
			r[k] = (int32_t)t;
		}
	f3.x = __int_as_float(r[0]);
	f3.y = __int_as_float(r[1]);
	dst[4 * (size_t)d + 0] = src[4 * (size_t)i + 0];
	dst[4 * (size_t)d + 1] = src[4 * (size_t)i + 1];
	dst[4 * (size_t)d + 2] = src[4 * (size_t)i + 2];
	dst[4 * (size_t)d + 3] = f3;
}

// PLOC leaves: triangles are rewritten in depth-first leaf order, so that
// every subtree covers a contiguous run of records, and a child subtree of at
// most max_leaf triangles becomes one leaf over its run.  With T(n) the first
// record of n's subtree (T(root) = 0), a node's second child starts after the
// triangles of its first: T(c1) = T(n) + tcount(c0).  New refs go to
// new_refs; the tree itself is read only.
__device__ __forceinline__ uint32_t ref_tris(int32_t r, const uint32_t *tcount)
{
	return r >= 0 ? tcount[r] : ymx_leaf_count(r);
}

__global__ __launch_bounds__(kBuildBlock) void k_ploc_leaves(const float4 *__restrict__ nodes, uint32_t n_int,
                                                             const int32_t *__restrict__ par,
                                                             const uint32_t *__restrict__ tcount,
                                                             const float4 *__restrict__ tris_in, uint32_t n_tris,
                                                             uint32_t max_leaf, float4 *__restrict__ tris_out,
                                                             int2 *__restrict__ new_refs, uint32_t *bad)
{
	const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
	if(n >= n_int) return;
	uint32_t first = 0u, c = n;
	for(uint32_t guard = 0; c != 0u; ++guard)
	{
		const int32_t p = par[c];
		if(p < 0 || guard > n_int)
		{
			atomicOr(bad, 1u);
			return;
		}
		const float4 f3 = nodes[4 * (size_t)p + 3];
		const int32_t r0 = __float_as_int(f3.x), r1 = __float_as_int(f3.y);
		if(r1 == (int32_t)c && r0 != (int32_t)c) first += ref_tris(r0, tcount);
		c = (uint32_t)p;
	}
	const float4 f3 = nodes[4 * (size_t)n + 3];
	int32_t r[2] = {__float_as_int(f3.x), __float_as_int(f3.y)};
	uint32_t off = first;
	for(int k = 0; k < 2; ++k)
	{
		const uint32_t cnt = ref_tris(r[k], tcount);
		if(off + cnt > n_tris)
		{
			atomicOr(bad, 1u);
			return;
		}
		if(r[k] < 0)
		{
			// a one-triangle leaf of the clustering: its record moves to slot off
			const uint32_t j = ymx_leaf_first(r[k]);
			if(j >= n_tris || cnt != 1u)
			{
				atomicOr(bad, 1u);
				return;
			}
			tris_out[3 * (size_t)off + 0] = tris_in[3 * (size_t)j + 0];
			tris_out[3 * (size_t)off + 1] = tris_in[3 * (size_t)j + 1];
			tris_out[3 * (size_t)off + 2] = tris_in[3 * (size_t)j + 2];
			r[k] = ymx_leaf_ref(off, 1u);
		}
		else if(cnt <= max_leaf) r[k] = ymx_leaf_ref(off, cnt);
		off += cnt;
	}
	new_refs[n] = make_int2(r[0], r[1]);
}

__global__ __launch_bounds__(kBuildBlock) void k_set_refs(float4 *__restrict__ nodes, uint32_t n_int,
                                                          const int2 *__restrict__ new_refs)
{
	const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
	if(n >= n_int) return;
	const int2 r = new_refs[n];
	float4 f3 = nodes[4 * (size_t)n + 3];
	f3.x = __int_as_float(r.x);
	f3.y = __int_as_float(r.y);
	nodes[4 * (size_t)n + 3] = f3;
}

__global__ void k_ploc_totals(const uint32_t *keep, const uint32_t *keep_pos, const uint32_t *merge,
                              const uint32_t *merge_pos, uint32_t count, uint32_t *totals)
{
	totals[0] = keep_pos[count - 1] + keep[count - 1];
	totals[1] = merge_pos[count - 1] + merge[count - 1];
}

__global__ __launch_bounds__(kBuildBlock) void k_ploc_merge(const Cluster *__restrict__ c, const int32_t *__restrict__ nn,
                                                            const uint32_t *__restrict__ keep,
                                                            const uint32_t *__restrict__ keep_pos,
                                                            const uint32_t *__restrict__ merge,
                                                            const uint32_t *__restrict__ merge_pos, uint32_t count,
                                                            int32_t next_node, float4 *__restrict__ nodes,
                                                            int32_t *__restrict__ parent_internal,
                                                            int32_t *__restrict__ parent_leaf, Cluster *__restrict__ out)
{
	const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if(i >= (int)count || !keep[i]) return;
	Cluster a = c[i];
	if(merge[i])
	{
		const Cluster b = c[nn[i]];
		const int32_t node = next_node - (int32_t)merge_pos[i];
		const int32_t refs[2] = {__float_as_int(a.lo.w), __float_as_int(b.lo.w)};
		const Cluster *kids[2] = {&a, &b};
		for(int s = 0; s < 2; ++s)
		{
			const float lo[3] = {kids[s]->lo.x, kids[s]->lo.y, kids[s]->lo.z};
			const float hi[3] = {kids[s]->hi.x, kids[s]->hi.y, kids[s]->hi.z};
			store_slot(nodes, node, s, lo, hi);
			if(refs[s] >= 0) parent_internal[refs[s]] = node;
			else parent_leaf[ymx_leaf_first(refs[s])] = node;
		}
		nodes[4 * (size_t)node + 3] = make_float4(__int_as_float(refs[0]), __int_as_float(refs[1]), 0.f, 0.f);
		a.lo = make_float4(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z), __int_as_float(node));
		a.hi = make_float4(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z), 0.f);
	}
	out[keep_pos[i]] = a;
}

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kBuildBlock - 1) / kBuildBlock); }

struct DevBuf
{
	void *p = nullptr;
	~DevBuf()
	{
		if(p) (void)hipFree(p);
	}
	hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
	template <typename T> T *as() const { return static_cast<T *>(p); }
	void *release()
	{
		void *r = p;
		p = nullptr;
		return r;
	}
};

struct EventPair
{
	hipEvent_t a = nullptr, b = nullptr;
	~EventPair()
	{
		if(a) (void)hipEventDestroy(a);
		if(b) (void)hipEventDestroy(b);
	}
};

// Tree bound of src/accelerator/accelerator_kdtree_original.cc:89-103: the
// union of primitive bounds, each axis grown by 0.001 x (float) extent
// evaluated in double and narrowed back to float.
void reference_tree_bound(const float lo_in[3], const float hi_in[3], uint32_t n_tris, float bmin[3], float bmax[3])
{
	for(int k = 0; k < 3; ++k)
	{
		float a = n_tris ? lo_in[k] : 0.f;
		float g = n_tris ? hi_in[k] : 0.f;
		const double offset = static_cast<double>(g - a) * 0.001;
		a -= static_cast<float>(offset);
		g += static_cast<float>(offset);
		bmin[k] = a;
		bmax[k] = g;
	}
}

// Depth-first relayout of a complete tree of n_tris - 1 internal nodes (root
// 0) into a new node buffer; with ploc_leaves, first the depth-first triangle
// order and the leaves of up to max_leaf triangles (k_ploc_leaves).
int relayout_tree(DevBuf &d_nodes, DevBuf &d_tris, uint32_t n_tris, bool ploc_leaves, uint32_t max_leaf,
                  hipStream_t stream, uint32_t &n_nodes)
{
	const uint32_t n_int = n_tris - 1;
	DevBuf d_par, d_nint, d_arrive, d_size, d_pos, d_bad, d_out;
	YMX_HIP_CHECK(d_par.alloc(sizeof(int32_t) * (size_t)n_int));
	YMX_HIP_CHECK(d_nint.alloc(sizeof(uint32_t) * (size_t)n_int));
	YMX_HIP_CHECK(d_arrive.alloc(sizeof(uint32_t) * (size_t)n_int));
	YMX_HIP_CHECK(d_size.alloc(sizeof(uint32_t) * (size_t)n_int));
	YMX_HIP_CHECK(d_pos.alloc(sizeof(uint32_t) * (size_t)n_int));
	YMX_HIP_CHECK(d_bad.alloc(sizeof(uint32_t)));
	auto sizes = [&](uint32_t *tcount) -> int {
		YMX_HIP_CHECK(hipMemsetAsync(d_par.p, 0xFF, sizeof(int32_t) * (size_t)n_int, stream));
		YMX_HIP_CHECK(hipMemsetAsync(d_arrive.p, 0, sizeof(uint32_t) * (size_t)n_int, stream));
		YMX_HIP_CHECK(hipMemsetAsync(d_size.p, 0, sizeof(uint32_t) * (size_t)n_int, stream));
		YMX_HIP_CHECK(hipMemsetAsync(d_bad.p, 0, sizeof(uint32_t), stream));
		k_lay_parents<<<grid_for(n_int), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), n_int, d_par.as<int32_t>(),
		                                                           d_nint.as<uint32_t>());
		YMX_HIP_CHECK(hipGetLastError());
		k_lay_size<<<grid_for(n_int), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), n_int, d_par.as<int32_t>(),
		                                                        d_nint.as<uint32_t>(), d_arrive.as<uint32_t>(),
		                                                        d_size.as<uint32_t>(), tcount, d_bad.as<uint32_t>());
		YMX_HIP_CHECK(hipGetLastError());
		return YMX_OK;
	};
	uint32_t hs[2] = {0u, 0u};
	if(ploc_leaves)
	{
		DevBuf d_tcount, d_tris_out, d_refs;
		YMX_HIP_CHECK(d_tcount.alloc(sizeof(uint32_t) * (size_t)n_int));
		YMX_HIP_CHECK(d_tris_out.alloc(sizeof(float4) * 3 * (size_t)n_tris));
		YMX_HIP_CHECK(d_refs.alloc(sizeof(int2) * (size_t)n_int));
		YMX_HIP_CHECK(hipMemsetAsync(d_tcount.p, 0, sizeof(uint32_t) * (size_t)n_int, stream));
		int rc = sizes(d_tcount.as<uint32_t>());
		if(rc != YMX_OK) return rc;
		k_ploc_leaves<<<grid_for(n_int), kBuildBlock, 0, stream>>>(
		    d_nodes.as<float4>(), n_int, d_par.as<int32_t>(), d_tcount.as<uint32_t>(), d_tris.as<float4>(), n_tris,
		    max_leaf, d_tris_out.as<float4>(), d_refs.as<int2>(), d_bad.as<uint32_t>());
		YMX_HIP_CHECK(hipGetLastError());
		YMX_HIP_CHECK(hipMemcpyAsync(&hs[0], d_tcount.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
		YMX_HIP_CHECK(hipMemcpyAsync(&hs[1], d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
		YMX_HIP_CHECK(hipStreamSynchronize(stream));
		if(hs[1] || hs[0] != n_tris)
		{
			ymx_set_error("ymx_accel_create: leaf collapse found an inconsistent tree (%u of %u triangles)", hs[0],
			              n_tris);
			return YMX_ERR_HIP;
		}
		k_set_refs<<<grid_for(n_int), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), n_int, d_refs.as<int2>());
		YMX_HIP_CHECK(hipGetLastError());
		std::swap(d_tris.p, d_tris_out.p);
	}
	int rc = sizes(nullptr);
	if(rc != YMX_OK) return rc;
	k_lay_pos<<<grid_for(n_int), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), n_int, d_par.as<int32_t>(),
	                                                       d_nint.as<uint32_t>(), d_size.as<uint32_t>(),
	                                                       d_pos.as<uint32_t>(), d_bad.as<uint32_t>());
	YMX_HIP_CHECK(hipGetLastError());
	YMX_HIP_CHECK(hipMemcpyAsync(&hs[0], d_size.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
	YMX_HIP_CHECK(hipMemcpyAsync(&hs[1], d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
	YMX_HIP_CHECK(hipStreamSynchronize(stream));
	if(hs[1] || hs[0] == 0u || hs[0] > n_int)
	{
		ymx_set_error("ymx_accel_create: node relayout found an inconsistent tree (%u of %u nodes)", hs[0], n_int);
		return YMX_ERR_HIP;
	}
	YMX_HIP_CHECK(d_out.alloc(sizeof(float4) * 4 * (size_t)hs[0]));
	k_lay_write<<<grid_for(n_int), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), n_int, d_pos.as<uint32_t>(), hs[0],
	                                                         d_out.as<float4>(), d_bad.as<uint32_t>());
	YMX_HIP_CHECK(hipGetLastError());
	YMX_HIP_CHECK(hipMemcpyAsync(&hs[1], d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
	YMX_HIP_CHECK(hipStreamSynchronize(stream));
	if(hs[1])
	{
		ymx_set_error("ymx_accel_create: node relayout wrote out of range");
		return YMX_ERR_HIP;
	}
	std::swap(d_nodes.p, d_out.p);
	n_nodes = hs[0];
	return YMX_OK;
}

}  // namespace

int ymx_build_device(ymx_accel *acc, const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t stride,
                     const uint32_t *quad_slot, const uint8_t *vis, uint32_t n_tris, int max_leaf, int builder,
                     int treelet_passes)
{
	acc->n_tris = n_tris;
	acc->build_ms = 0.f;
	hipStream_t stream = nullptr;

	if(n_tris < 2)
	{
		// A single node whose two children are the same one-triangle leaf
		// (testing it twice changes nothing).  With no triangle the leaf holds
		// a degenerate, invisible triangle that no ray can hit, so every query
		// is a miss.  (Slab tests take per-axis min/max, so an inverted "empty"
		// box would not be rejected; the tree therefore never has empty children.)
		float4 node[4];
		float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
		float4 tri[3] = {};
		const int32_t c0 = ymx_leaf_ref(0, 1);
		const bool quad = n_tris == 1 && quad_slot && quad_slot[0];
		float4 e3 = {};
		if(n_tris == 1)
		{
			float v[4][3];
			const int nv = quad ? 4 : 3;
			for(int c = 0; c < nv; ++c)
				for(int k = 0; k < 3; ++k) v[c][k] = verts[3 * (size_t)idx[c] + k];
			for(int k = 0; k < 3; ++k)
			{
				lo[k] = fminf(v[0][k], fminf(v[1][k], v[2][k]));
				hi[k] = fmaxf(v[0][k], fmaxf(v[1][k], v[2][k]));
				if(quad)
				{
					lo[k] = fminf(lo[k], v[3][k]);
					hi[k] = fmaxf(hi[k], v[3][k]);
				}
			}
			uint32_t zero = 0, vm = vis ? vis[0] : 3u, one = quad ? 1u : 0u;
			float fz, fv, fq;
			memcpy(&fz, &zero, 4);
			memcpy(&fv, &vm, 4);
			memcpy(&fq, &one, 4);
			tri[0] = make_float4(v[0][0], v[0][1], v[0][2], fz);
			tri[1] = make_float4(v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2], fv);
			tri[2] = make_float4(v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2], fq);
			if(quad) e3 = make_float4(v[3][0] - v[0][0], v[3][1] - v[0][1], v[3][2] - v[0][2], 0.f);
		}
		else
		{
			// degenerate triangle at the origin, visibility 0: never accepted
			tri[1] = make_float4(0.f, 0.f, 0.f, 0.f);
		}
		float pad = 0.f;
		if(n_tris == 1)
		{
			float ext = 0.f;
			for(int k = 0; k < 3; ++k) ext = fmaxf(ext, hi[k] - lo[k]);
			pad = ext * 0x1p-20f + 1e-30f;
		}
		float l0[3], h0[3];
		for(int k = 0; k < 3; ++k)
		{
			l0[k] = lo[k] - pad;
			h0[k] = hi[k] + pad;
		}
		float fc0;
		memcpy(&fc0, &c0, 4);
		node[0] = make_float4(l0[0], h0[0], l0[1], h0[1]);
		node[1] = node[0];
		node[2] = make_float4(l0[2], h0[2], l0[2], h0[2]);
		node[3] = make_float4(fc0, fc0, 0.f, 0.f);
		DevBuf dn, dt;
		YMX_HIP_CHECK(dn.alloc(sizeof node));
		YMX_HIP_CHECK(dt.alloc(sizeof tri));
		YMX_HIP_CHECK(hipMemcpy(dn.p, node, sizeof node, hipMemcpyHostToDevice));
		YMX_HIP_CHECK(hipMemcpy(dt.p, tri, sizeof tri, hipMemcpyHostToDevice));
		if(quad)
		{
			DevBuf dq;
			YMX_HIP_CHECK(dq.alloc(sizeof e3));
			YMX_HIP_CHECK(hipMemcpy(dq.p, &e3, sizeof e3, hipMemcpyHostToDevice));
			acc->d_quads = dq.as<float4>();
			acc->n_quads = 1;
			dq.release();
		}
		acc->d_nodes = dn.as<float4>();
		acc->d_tris = dt.as<float4>();
		dn.release();
		dt.release();
		acc->n_nodes = 1;
		acc->max_depth = 1;
		reference_tree_bound(lo, hi, n_tris, acc->bmin, acc->bmax);
		return YMX_OK;
	}

	DevBuf d_verts, d_idx, d_vis, d_tris_unsorted, d_prim_box, d_centroid, d_bounds, d_keys, d_keys_alt,
	    d_vals, d_vals_alt, d_nodes, d_tris, d_parent_int, d_parent_leaf, d_depth, d_temp, d_slot, d_quads;
	YMX_HIP_CHECK(d_verts.alloc(sizeof(float) * 3 * (size_t)n_verts));
	YMX_HIP_CHECK(d_idx.alloc(sizeof(uint32_t) * stride * (size_t)n_tris));
	if(quad_slot)
	{
		YMX_HIP_CHECK(d_slot.alloc(sizeof(uint32_t) * (size_t)n_tris));
		YMX_HIP_CHECK(d_quads.alloc(sizeof(float4) * (size_t)(acc->n_quads ? acc->n_quads : 1)));
		YMX_HIP_CHECK(hipMemcpy(d_slot.p, quad_slot, sizeof(uint32_t) * (size_t)n_tris, hipMemcpyHostToDevice));
	}
	if(vis) YMX_HIP_CHECK(d_vis.alloc(n_tris));
	YMX_HIP_CHECK(d_tris_unsorted.alloc(sizeof(float4) * 3 * (size_t)n_tris));
	YMX_HIP_CHECK(d_prim_box.alloc(sizeof(float) * 6 * (size_t)n_tris));
	YMX_HIP_CHECK(d_centroid.alloc(sizeof(float) * 3 * (size_t)n_tris));
	YMX_HIP_CHECK(d_bounds.alloc(sizeof(int) * 12));
	YMX_HIP_CHECK(d_keys.alloc(sizeof(uint64_t) * (size_t)n_tris));
	YMX_HIP_CHECK(d_keys_alt.alloc(sizeof(uint64_t) * (size_t)n_tris));
	YMX_HIP_CHECK(d_vals.alloc(sizeof(uint32_t) * (size_t)n_tris));
	YMX_HIP_CHECK(d_vals_alt.alloc(sizeof(uint32_t) * (size_t)n_tris));
	YMX_HIP_CHECK(d_nodes.alloc(sizeof(float4) * 4 * (size_t)(n_tris - 1)));
	YMX_HIP_CHECK(d_tris.alloc(sizeof(float4) * 3 * (size_t)n_tris));
	YMX_HIP_CHECK(d_parent_int.alloc(sizeof(int32_t) * (size_t)(n_tris - 1)));
	YMX_HIP_CHECK(d_parent_leaf.alloc(sizeof(int32_t) * (size_t)n_tris));
	YMX_HIP_CHECK(d_depth.alloc(sizeof(uint32_t)));

	YMX_HIP_CHECK(hipMemcpy(d_verts.p, verts, sizeof(float) * 3 * (size_t)n_verts, hipMemcpyHostToDevice));
	YMX_HIP_CHECK(hipMemcpy(d_idx.p, idx, sizeof(uint32_t) * stride * (size_t)n_tris, hipMemcpyHostToDevice));
	if(vis) YMX_HIP_CHECK(hipMemcpy(d_vis.p, vis, n_tris, hipMemcpyHostToDevice));

	EventPair ev;
	YMX_HIP_CHECK(hipEventCreate(&ev.a));
	YMX_HIP_CHECK(hipEventCreate(&ev.b));
	YMX_HIP_CHECK(hipEventRecord(ev.a, stream));

	k_init_bounds<<<1, 64, 0, stream>>>(d_bounds.as<int>());
	k_prep<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(
	    d_verts.as<float>(), d_idx.as<uint32_t>(), stride, quad_slot ? d_slot.as<uint32_t>() : nullptr,
	    vis ? d_vis.as<uint8_t>() : nullptr, n_tris, d_tris_unsorted.as<float4>(),
	    quad_slot ? d_quads.as<float4>() : nullptr, d_prim_box.as<float>(), d_centroid.as<float>(), d_bounds.as<int>());
	YMX_HIP_CHECK(hipGetLastError());
	int hb[12];
	YMX_HIP_CHECK(hipMemcpyAsync(hb, d_bounds.p, sizeof hb, hipMemcpyDeviceToHost, stream));
	YMX_HIP_CHECK(hipStreamSynchronize(stream));
	float scene_lo[3], scene_hi[3], c_lo[3], c_hi[3];
	for(int k = 0; k < 3; ++k)
	{
		scene_lo[k] = float_from_ordered(hb[k]);
		scene_hi[k] = float_from_ordered(hb[3 + k]);
		c_lo[k] = float_from_ordered(hb[6 + k]);
		c_hi[k] = float_from_ordered(hb[9 + k]);
	}
	float inv_ext[3];
	for(int k = 0; k < 3; ++k)
	{
		const float e = c_hi[k] - c_lo[k];
		inv_ext[k] = e > 0.f ? 1.f / e : 0.f;
	}
	k_morton<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(d_centroid.as<float>(), n_tris, c_lo[0], c_lo[1],
	                                                       c_lo[2], inv_ext[0], inv_ext[1], inv_ext[2],
	                                                       d_keys.as<uint64_t>(), d_vals.as<uint32_t>());
	YMX_HIP_CHECK(hipGetLastError());

	// the 63-bit Morton keys, ping-ponged between the two buffers: the sorted
	// run ends in whichever current() names
	rocprim::double_buffer<uint64_t> kb(d_keys.as<uint64_t>(), d_keys_alt.as<uint64_t>());
	rocprim::double_buffer<uint32_t> vb(d_vals.as<uint32_t>(), d_vals_alt.as<uint32_t>());
	size_t temp_bytes = 0;
	YMX_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, temp_bytes, kb, vb, (size_t)n_tris, 0u, 63u, stream));
	YMX_HIP_CHECK(d_temp.alloc(temp_bytes));
	YMX_HIP_CHECK(rocprim::radix_sort_pairs(d_temp.p, temp_bytes, kb, vb, (size_t)n_tris, 0u, 63u, stream));
	const uint64_t *keys = kb.current();
	const uint32_t *order = vb.current();

	k_reorder<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(d_tris_unsorted.as<float4>(), order, n_tris,
	                                                        d_tris.as<float4>());
	YMX_HIP_CHECK(hipGetLastError());
	YMX_HIP_CHECK(hipMemsetAsync(d_depth.p, 0, sizeof(uint32_t), stream));
	float ext = 0.f;
	for(int k = 0; k < 3; ++k) ext = fmaxf(ext, scene_hi[k] - scene_lo[k]);
	const float pad = ext * 0x1p-20f + 1e-30f;
	if(builder != YMX_BUILDER_PLOC)
	{
		DevBuf d_flags, d_range;
		YMX_HIP_CHECK(d_flags.alloc(sizeof(uint32_t) * (size_t)(n_tris - 1)));
		YMX_HIP_CHECK(d_range.alloc(sizeof(uint2) * (size_t)(n_tris - 1)));
		k_karras<<<grid_for(n_tris - 1), kBuildBlock, 0, stream>>>(keys, (int)n_tris, d_nodes.as<float4>(),
		                                                           d_parent_int.as<int32_t>(),
		                                                           d_parent_leaf.as<int32_t>(), d_range.as<uint2>());
		YMX_HIP_CHECK(hipGetLastError());
		YMX_HIP_CHECK(hipMemsetAsync(d_flags.p, 0, sizeof(uint32_t) * (size_t)(n_tris - 1), stream));
		k_refit<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), d_prim_box.as<float>(), order,
		                                                      n_tris, d_parent_int.as<int32_t>(),
		                                                      d_parent_leaf.as<int32_t>(), d_flags.as<uint32_t>(), pad);
		YMX_HIP_CHECK(hipGetLastError());
		if(max_leaf > 1 && treelet_passes == 0)
		{
			k_collapse_leaves<<<grid_for(n_tris - 1), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), d_range.as<uint2>(),
			                                                                    n_tris, (uint32_t)max_leaf);
			YMX_HIP_CHECK(hipGetLastError());
		}
	}
	else
	{
		YMX_HIP_CHECK(hipMemsetAsync(d_nodes.p, 0xFF, sizeof(float4) * 4 * (size_t)(n_tris - 1), stream));
		DevBuf d_ca, d_cb, d_nn, d_keep, d_merge, d_kpos, d_mpos, d_tot, d_scan;
		YMX_HIP_CHECK(d_ca.alloc(sizeof(Cluster) * (size_t)n_tris));
		YMX_HIP_CHECK(d_cb.alloc(sizeof(Cluster) * (size_t)n_tris));
		YMX_HIP_CHECK(d_nn.alloc(sizeof(int32_t) * (size_t)n_tris));
		YMX_HIP_CHECK(d_keep.alloc(sizeof(uint32_t) * (size_t)n_tris));
		YMX_HIP_CHECK(d_merge.alloc(sizeof(uint32_t) * (size_t)n_tris));
		YMX_HIP_CHECK(d_kpos.alloc(sizeof(uint32_t) * (size_t)n_tris));
		YMX_HIP_CHECK(d_mpos.alloc(sizeof(uint32_t) * (size_t)n_tris));
		YMX_HIP_CHECK(d_tot.alloc(sizeof(uint32_t) * 2));
		size_t scan_bytes = 0;
		YMX_HIP_CHECK(rocprim::exclusive_scan(nullptr, scan_bytes, d_keep.as<uint32_t>(), d_kpos.as<uint32_t>(), 0u,
		                                      (size_t)n_tris, rocprim::plus<uint32_t>(), stream));
		YMX_HIP_CHECK(d_scan.alloc(scan_bytes));
		k_ploc_init<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(d_prim_box.as<float>(), order, n_tris, pad,
		                                                          d_ca.as<Cluster>());
		YMX_HIP_CHECK(hipGetLastError());
		Cluster *cin = d_ca.as<Cluster>(), *cout = d_cb.as<Cluster>();
		uint32_t count = n_tris, iterations = 0;
		int32_t next_node = (int32_t)n_tris - 2;
		while(count > 1)
		{
			if(++iterations > 1000000u)
			{
				ymx_set_error("ymx_accel_create: BVH clustering does not converge");
				return YMX_ERR_HIP;
			}
			k_ploc_nn<<<grid_for(count), kBuildBlock, 0, stream>>>(cin, count, d_nn.as<int32_t>());
			YMX_HIP_CHECK(hipGetLastError());
			k_ploc_flags<<<grid_for(count), kBuildBlock, 0, stream>>>(d_nn.as<int32_t>(), count, d_keep.as<uint32_t>(),
			                                                          d_merge.as<uint32_t>());
			YMX_HIP_CHECK(hipGetLastError());
			size_t sb = scan_bytes;
			YMX_HIP_CHECK(rocprim::exclusive_scan(d_scan.p, sb, d_keep.as<uint32_t>(), d_kpos.as<uint32_t>(), 0u,
			                                      (size_t)count, rocprim::plus<uint32_t>(), stream));
			sb = scan_bytes;
			YMX_HIP_CHECK(rocprim::exclusive_scan(d_scan.p, sb, d_merge.as<uint32_t>(), d_mpos.as<uint32_t>(), 0u,
			                                      (size_t)count, rocprim::plus<uint32_t>(), stream));
			k_ploc_totals<<<1, 1, 0, stream>>>(d_keep.as<uint32_t>(), d_kpos.as<uint32_t>(), d_merge.as<uint32_t>(),
			                                   d_mpos.as<uint32_t>(), count, d_tot.as<uint32_t>());
			YMX_HIP_CHECK(hipGetLastError());
			k_ploc_merge<<<grid_for(count), kBuildBlock, 0, stream>>>(
			    cin, d_nn.as<int32_t>(), d_keep.as<uint32_t>(), d_kpos.as<uint32_t>(), d_merge.as<uint32_t>(),
			    d_mpos.as<uint32_t>(), count, next_node, d_nodes.as<float4>(), d_parent_int.as<int32_t>(),
			    d_parent_leaf.as<int32_t>(), cout);
			YMX_HIP_CHECK(hipGetLastError());
			uint32_t tot[2] = {0u, 0u};
			YMX_HIP_CHECK(hipMemcpyAsync(tot, d_tot.p, sizeof tot, hipMemcpyDeviceToHost, stream));
			YMX_HIP_CHECK(hipStreamSynchronize(stream));
			if(tot[1] == 0u || tot[1] > (uint32_t)(next_node + 1) || tot[0] + tot[1] != count)
			{
				ymx_set_error("ymx_accel_create: BVH clustering made no valid progress (%u clusters, %u merges)",
				              count, tot[1]);
				return YMX_ERR_HIP;
			}
			next_node -= (int32_t)tot[1];
			count = tot[0];
			Cluster *t = cin;
			cin = cout;
			cout = t;
		}
		// the tree must be complete before any traversal reads it
		YMX_HIP_CHECK(hipMemsetAsync(d_tot.p, 0, sizeof(uint32_t), stream));
		k_check_tree<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(d_nodes.as<float4>(), n_tris, d_tot.as<uint32_t>());
		YMX_HIP_CHECK(hipGetLastError());
		uint32_t bad = 0;
		YMX_HIP_CHECK(hipMemcpyAsync(&bad, d_tot.p, sizeof bad, hipMemcpyDeviceToHost, stream));
		YMX_HIP_CHECK(hipStreamSynchronize(stream));
		if(next_node != -1 || bad)
		{
			ymx_set_error("ymx_accel_create: BVH clustering left an incomplete tree");
			return YMX_ERR_HIP;
		}
	}
	if(treelet_passes > 0 && n_tris >= 3)
	{
		// SAH treelet restructuring of the one-triangle-leaf tree; leaves of
		// up to max_leaf triangles are then formed by the depth-first
		// triangle rewrite of relayout_tree (the LBVH key ranges no longer
		// describe the subtrees)
		DevBuf d_tflags, d_tcost, d_tcount, d_trew;
		YMX_HIP_CHECK(d_tflags.alloc(sizeof(uint32_t) * (size_t)(n_tris - 1)));
		YMX_HIP_CHECK(d_tcost.alloc(sizeof(float) * (size_t)(n_tris - 1)));
		YMX_HIP_CHECK(d_tcount.alloc(sizeof(uint32_t) * (size_t)(n_tris - 1)));
		YMX_HIP_CHECK(d_trew.alloc(2 * sizeof(uint32_t)));
		YMX_HIP_CHECK(hipMemsetAsync(d_trew.p, 0, 2 * sizeof(uint32_t), stream));
		for(int pass = 0; pass < treelet_passes; ++pass)
		{
			YMX_HIP_CHECK(hipMemsetAsync(d_tflags.p, 0, sizeof(uint32_t) * (size_t)(n_tris - 1), stream));
			k_treelets<<<(n_tris + kTreeletBlock - 1) / kTreeletBlock, kTreeletBlock, 0, stream>>>(
			    d_nodes.as<float4>(), d_parent_int.as<int32_t>(), d_parent_leaf.as<int32_t>(), n_tris,
			    d_tflags.as<uint32_t>(), d_tcost.as<float>(), d_tcount.as<uint32_t>(), (uint32_t)max_leaf + 1u,
			    d_trew.as<uint32_t>());
			YMX_HIP_CHECK(hipGetLastError());
		}
		uint32_t hr[2] = {0u, 0u};
		YMX_HIP_CHECK(hipMemcpyAsync(hr, d_trew.p, sizeof hr, hipMemcpyDeviceToHost, stream));
		YMX_HIP_CHECK(hipStreamSynchronize(stream));
		if(hr[1])
		{
			ymx_set_error("ymx_accel_create: treelet restructuring found an inconsistent tree");
			return YMX_ERR_HIP;
		}
		acc->treelets_rewritten = hr[0];
	}
	k_depth<<<grid_for(n_tris), kBuildBlock, 0, stream>>>(d_parent_int.as<int32_t>(), d_parent_leaf.as<int32_t>(),
	                                                      n_tris, d_depth.as<uint32_t>());
	YMX_HIP_CHECK(hipGetLastError());
	uint32_t n_nodes = n_tris - 1;
	// LBVH: the Karras node order puts a parent far from its children near the
	// root and leaves the nodes under collapsed leaves in place; the depth-first
	// relayout drops those (2.19M -> 0.70M nodes on the heightfield workload)
	// and traces 2 % faster.  PLOC: leaves of up to max_leaf triangles need the
	// triangles in depth-first order first; with one triangle per leaf the tree
	// is left as built (the relayout alone gains nothing on it).
	if(builder != YMX_BUILDER_PLOC || max_leaf > 1 || treelet_passes > 0)
	{
		const bool rewrite_leaves = builder == YMX_BUILDER_PLOC || treelet_passes > 0;
		const int rc = relayout_tree(d_nodes, d_tris, n_tris, rewrite_leaves, (uint32_t)max_leaf, stream, n_nodes);
		if(rc != YMX_OK) return rc;
	}
	YMX_HIP_CHECK(hipEventRecord(ev.b, stream));
	uint32_t depth = 0;
	YMX_HIP_CHECK(hipMemcpyAsync(&depth, d_depth.p, sizeof depth, hipMemcpyDeviceToHost, stream));
	YMX_HIP_CHECK(hipStreamSynchronize(stream));
	float ms = 0.f;
	YMX_HIP_CHECK(hipEventElapsedTime(&ms, ev.a, ev.b));

	acc->build_ms = ms;
	acc->n_nodes = n_nodes;
	acc->max_depth = depth;
	acc->d_nodes = d_nodes.as<float4>();
	acc->d_tris = d_tris.as<float4>();
	d_nodes.release();
	d_tris.release();
	if(quad_slot)
	{
		acc->d_quads = d_quads.as<float4>();
		d_quads.release();
	}
	reference_tree_bound(scene_lo, scene_hi, n_tris, acc->bmin, acc->bmax);
	return YMX_OK;
}
