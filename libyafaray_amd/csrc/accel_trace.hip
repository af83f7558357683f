// accel_trace.hip -- wave64 closest-hit and any-hit traversal of the LBVH.
//
// Persistent kernels: a grid that fills the device once; each wave takes
// 64 consecutive rays at a time from a device counter, so ray loads, hit
// stores and the any-hit occlusion ballot stay contiguous and coalesced.
// Per lane: one ray, a 16-entry traversal stack in LDS (lane-major so lanes
// never share a bank) spilling its oldest entries to a per-lane area in HBM
// when a path is deeper than 16.
//
// Bit-exactness with the reference (its FP32 x86 build, no FMA):
//   - this file is compiled with -ffp-contract=off and IEEE division;
//   - the tree-bound crossing, the dynamic ray bias and the triangle test are
//     evaluated in the reference's operation order;
//   - box tests only cull: boxes are padded, slab intervals are widened by
//     (1 + 2 gamma_3) (T. Ize, "Robust BVH Ray Traversal", JCGT 2013) and the
//     accepted [t_min, t_best] window is widened by 2^-12 relative, so every
//     triangle whose exact test would be accepted is tested.
#include "accel_internal.h"
#include "ray_tri.h"

#include <float.h>


namespace {

constexpr int kTraceBlock = 128;
constexpr int kStack = YMX_LDS_STACK;
constexpr float kBiasScale = 0.1f * 0.00005f;  // include/accelerator/accelerator.h:64,85
constexpr float kSlabWiden = ymx::kSlabWiden;
constexpr float kWindow = ymx::kBoxWindow;

struct RayState
{
	float o[3], d[3], idir[3];
	float tlo, thi;
};

using ymx::tree_cross;
using ymx::prim_test;

// Slab test of both children of one node.  Returns bit 0 / bit 1 for child
// 0 / 1 and their entry distances.
__device__ __forceinline__ int box2(const float4 f0, const float4 f1, const float4 f2, const RayState &r,
                                    float thi, float &tn0, float &tn1)
{
	const float a0x = (f0.x - r.o[0]) * r.idir[0], b0x = (f0.y - r.o[0]) * r.idir[0];
	const float a0y = (f0.z - r.o[1]) * r.idir[1], b0y = (f0.w - r.o[1]) * r.idir[1];
	const float a0z = (f2.x - r.o[2]) * r.idir[2], b0z = (f2.y - r.o[2]) * r.idir[2];
	const float a1x = (f1.x - r.o[0]) * r.idir[0], b1x = (f1.y - r.o[0]) * r.idir[0];
	const float a1y = (f1.z - r.o[1]) * r.idir[1], b1y = (f1.w - r.o[1]) * r.idir[1];
	const float a1z = (f2.z - r.o[2]) * r.idir[2], b1z = (f2.w - r.o[2]) * r.idir[2];
	const float n0 = fmaxf(fmaxf(fminf(a0x, b0x), fminf(a0y, b0y)), fmaxf(fminf(a0z, b0z), r.tlo));
	const float x0 = fminf(fminf(fmaxf(a0x, b0x), fmaxf(a0y, b0y)), fmaxf(a0z, b0z)) * kSlabWiden;
	const float n1 = fmaxf(fmaxf(fminf(a1x, b1x), fminf(a1y, b1y)), fmaxf(fminf(a1z, b1z), r.tlo));
	const float x1 = fminf(fminf(fmaxf(a1x, b1x), fmaxf(a1y, b1y)), fmaxf(a1z, b1z)) * kSlabWiden;
	tn0 = n0;
	tn1 = n1;
	int m = 0;
	if(n0 <= fminf(x0, thi)) m |= 1;
	if(n1 <= fminf(x1, thi)) m |= 2;
	return m;
}

struct LaneStack
{
	int32_t *lds;     // this lane's column: lds[k * kTraceBlock]
	int32_t *spill;   // this lane's spill column: spill[k * grid_threads]
	uint32_t stride;  // grid_threads
	uint32_t cap;     // spill capacity (entries)
	uint32_t sp;

	// The stack holds at most max_depth entries and kStack + cap is
	// max_depth + 1 (ymx_accel), so the capacity test never fails; it only
	// keeps every store inside the spill area.
	__device__ __forceinline__ void push(int32_t ref)
	{
		const uint32_t slot = sp & (kStack - 1);
		if(sp >= (uint32_t)kStack)
		{
			const uint32_t k = sp - kStack;
			if(k >= cap) return;
			spill[(size_t)k * stride] = lds[slot * kTraceBlock];
		}
		lds[slot * kTraceBlock] = ref;
		++sp;
	}
	__device__ __forceinline__ int32_t pop()
	{
		if(sp == 0) return YMX_REF_EMPTY;
		--sp;
		const uint32_t slot = sp & (kStack - 1);
		const int32_t ref = lds[slot * kTraceBlock];
		if(sp >= (uint32_t)kStack) lds[slot * kTraceBlock] = spill[(size_t)(sp - kStack) * stride];
		return ref;
	}
};

// ---- work distribution of the persistent grids.  One counter shared by the
// whole grid serialises: device-scope atomics on one address complete at about
// 90 M/s on this part (11 ns each), against about 16 G/s spread over many
// addresses, so 8,192 waves claiming 64 rays at a time queue at it, and even
// an empty launch costs each wave's claim and finish in turn (about 0.2 ms).
// The batch is cut into kSegments contiguous segments, each with its own
// claim counter a page apart from the others.  A wave starts at a home
// segment among the kSegments / 8 of its XCD's eighth of the batch (blocks
// are dealt to the 8 XCDs in turn, so one XCD's L2 holds the nodes that its
// contiguous share of rays visits).  When that one is used up, the wave reads
// every segment's counter at once (one load per lane) and moves to the next
// segment with rays left, so it claims only where claims can succeed.
//
// Work area words (a page, kWorkStride words, apart): slot s < kSegments:
// units claimed from segment s; slot kSegments + g: waves finished in block
// group g = blockIdx % kSegments; slot 2 kSegments: groups finished.  The
// last wave of the launch resets them all, so the next launch on the work
// area starts from zero.
constexpr uint32_t kSegments = 64;  // one per lane of the wave that scans them
static_assert(kTraceBlock % 64 == 0, "waves of 64 lanes");
static_assert(kSegments == kQueueSegments, "a segmented queue's segments are the claim segments");
constexpr uint32_t kXcds = 8;
constexpr uint32_t kWorkStride = 1024;
constexpr uint32_t kWorkSlots = 2 * kSegments + 1;
constexpr size_t kWorkWords = (size_t)kWorkSlots * kWorkStride;

__device__ __forceinline__ uint32_t seg_lo(uint32_t s, uint32_t n_units)
{
	return (uint32_t)(((uint64_t)n_units * s) / kSegments);
}

// What the segments hold, in claim units of 2^unit_shift rays: an even split
// of n_units, or the segments of a segmented queue (counts, cap_units each).
struct SegView
{
	uint32_t n_units;
	const uint32_t *counts;
	uint32_t cap_units;
	uint32_t unit_shift;
};

__device__ __forceinline__ SegView seg_view(const ymx_trace_args &args, uint32_t n_rays, uint32_t unit_shift)
{
	SegView v;
	v.counts = args.seg_counts;
	v.unit_shift = unit_shift;
	v.cap_units = args.seg_cap >> unit_shift;
	v.n_units = (uint32_t)(((uint64_t)n_rays + (1u << unit_shift) - 1u) >> unit_shift);
	return v;
}

// Segment s: first unit and unit count; ray_end, the end of its rays.
__device__ __forceinline__ void seg_bounds(const SegView &v, uint32_t s, uint32_t &lo, uint32_t &size,
                                           uint32_t &ray_end)
{
	if(v.counts)
	{
		uint32_t c = v.counts[(size_t)s * kQueueStride];
		const uint32_t cap = v.cap_units << v.unit_shift;
		if(c > cap) c = cap;
		lo = s * v.cap_units;
		size = (c + (1u << v.unit_shift) - 1u) >> v.unit_shift;
		ray_end = (lo << v.unit_shift) + c;
	}
	else
	{
		lo = seg_lo(s, v.n_units);
		size = seg_lo(s + 1u, v.n_units) - lo;
		ray_end = 0xFFFFFFFFu;  // only the batch's own end bounds the last unit
	}
}

struct WaveWork
{
	uint32_t seg;   // segment the wave claims from
	bool done;      // every segment used up
};

__device__ __forceinline__ WaveWork wave_work_start(const SegView &v)
{
	constexpr uint32_t per_xcd = kSegments / kXcds;
	const uint32_t xcd = blockIdx.x % kXcds;
	const uint32_t j = (blockIdx.x / kXcds) * (uint32_t)(kTraceBlock / 64) + (threadIdx.x >> 6);
	return WaveWork{xcd * per_xcd + j % per_xcd, !v.counts && v.n_units == 0u};
}

// After segment ww.seg ran out: lane k reads segment k's counter, and the
// wave moves to the first segment after ww.seg (cyclically) with units left,
// or sets done.  Every lane computes the same choice from the ballot.
__device__ __forceinline__ void next_segment(const uint32_t *work, WaveWork &ww, const SegView &v)
{
	const uint32_t k = __lane_id();
	const uint32_t claimed = __hip_atomic_load(work + (size_t)k * kWorkStride, __ATOMIC_RELAXED,
	                                           __HIP_MEMORY_SCOPE_AGENT);
	uint32_t lo, size, ray_end;
	seg_bounds(v, k, lo, size, ray_end);
	const bool left = claimed < size && k != ww.seg;
	const unsigned long long free_bits = __ballot(left);
	if(free_bits == 0ull)
	{
		ww.done = true;
		return;
	}
	const uint32_t start = (ww.seg + 1u) % kSegments;
	const unsigned long long rot = (free_bits >> start) | (start ? (free_bits << (64u - start)) : 0ull);
	ww.seg = (start + (uint32_t)(__ffsll((long long)rot) - 1)) % kSegments;
}

// Claims up to `want` consecutive units for the wave: returns the first and
// sets got (0 once every segment is used up) and ray_end (the end of the
// claimed segment's rays).  Called with every lane of the wave active; lane 0
// makes the atomic.
__device__ __forceinline__ uint32_t wave_claim(uint32_t *work, WaveWork &ww, const SegView &v, uint32_t want,
                                               uint32_t &got, uint32_t &ray_end)
{
	while(!ww.done)
	{
		uint32_t lo, size;
		seg_bounds(v, ww.seg, lo, size, ray_end);
		uint32_t b = size;
		if(__lane_id() == 0 && size > 0u) b = atomicAdd(work + (size_t)ww.seg * kWorkStride, want);
		b = __shfl(b, 0);
		if(b < size)
		{
			got = size - b < want ? size - b : want;
			if(b + want >= size) next_segment(work, ww, v);
			return lo + b;
		}
		next_segment(work, ww, v);
	}
	got = 0u;
	return 0u;
}

// End of a persistent wave.  Waves count themselves finished per block group
// (blockIdx % kSegments), the last of each group counts the group, and the
// last group's resets every counter.  A wave finishes after its last claim
// (an atomic whose result it waited for), so nothing of this launch claims
// after the reset.
__device__ __forceinline__ void finish_wave(uint32_t *work)
{
	if(__lane_id() == 0)
	{
		const uint32_t g = blockIdx.x % kSegments;
		const uint32_t blocks = gridDim.x / kSegments + (g < gridDim.x % kSegments ? 1u : 0u);
		const uint32_t waves = blocks * (uint32_t)(kTraceBlock / 64);
		if(atomicAdd(work + (size_t)(kSegments + g) * kWorkStride, 1u) == waves - 1u)
		{
			const uint32_t groups = gridDim.x < kSegments ? gridDim.x : kSegments;
			if(atomicAdd(work + (size_t)(2 * kSegments) * kWorkStride, 1u) == groups - 1u)
			{
				for(uint32_t k = 0; k < kWorkSlots; ++k) atomicExch(work + (size_t)k * kWorkStride, 0u);
			}
		}
	}
}

__device__ __forceinline__ void setup_ray(const ymx_ray &ray, RayState &r)
{
	for(int k = 0; k < 3; ++k)
	{
		r.o[k] = ray.org[k];
		r.d[k] = ray.dir[k];
		const float dk = (fabsf(ray.dir[k]) > 1e-30f) ? ray.dir[k] : copysignf(1e-30f, ray.dir[k]);
		r.idir[k] = 1.f / dk;
	}
}

__device__ __forceinline__ ymx_ray load_ray(const ymx_ray *rays, uint32_t i)
{
	const float4 *p = reinterpret_cast<const float4 *>(rays + i);
	const float4 a = p[0], b = p[1];
	ymx_ray r;
	r.org[0] = a.x;
	r.org[1] = a.y;
	r.org[2] = a.z;
	r.tmin = a.w;
	r.dir[0] = b.x;
	r.dir[1] = b.y;
	r.dir[2] = b.z;
	r.tmax = b.w;
	return r;
}


// kMode: kClosest (Accelerator::intersect), kShadowAny (intersectShadow) or
// kTransparent (intersectTransparentShadow).
enum : int { kClosest = 0, kShadowAny = 1, kTransparent = 2 };

template <int kMode, bool kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace(ymx_trace_args args)
{
	constexpr bool kShadow = kMode != kClosest;
	__shared__ int32_t lds_stack[kStack * kTraceBlock];
	const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
	LaneStack st;
	st.lds = lds_stack + threadIdx.x;
	st.spill = args.spill + gtid;
	st.stride = args.grid_threads;
	st.cap = args.spill_depth;
	unsigned long long n_fetch = 0, n_tests = 0;
	const uint32_t n_rays = args.n_dev ? *args.n_dev : args.n_rays;
	const SegView sv = seg_view(args, n_rays, 6u);
	WaveWork ww = wave_work_start(sv);

	for(;;)
	{
		// one group of 64 consecutive rays per claim
		uint32_t got = 0, ray_end = 0;
		const uint32_t group = wave_claim(args.work, ww, sv, 1u, got, ray_end);
		if(got == 0u) break;
		const uint32_t base = group * 64u;
		const uint32_t rid = base + __lane_id();
		const bool valid = rid < n_rays && rid < ray_end;

		ymx_ray ray = {};
		if(valid) ray = load_ray(args.rays, rid);
		RayState r;
		setup_ray(ray, r);
		const float t_max = (ray.tmax >= 0.f) ? ray.tmax : FLT_MAX;
		float enter = 0.f, leave = 0.f;
		bool live = valid && tree_cross(args.bmin, args.bmax, ray.org, ray.dir, t_max, enter, leave);
		// include/accelerator/accelerator_kdtree_common.h:139
		const float bias = kBiasScale * fabsf(leave - enter);
		const float t_min = (kMode == kShadowAny) ? bias : ((ray.tmin < bias) ? bias : ray.tmin);
		r.tlo = t_min - t_min * kWindow;
		float best_t = t_max;
		uint32_t best_prim = YMX_NO_HIT;
		float best_u = 0.f, best_v = 0.f;
		bool occluded = false;
		uint32_t n_transp = 0;  // transparent shadow casters kept (kTransparent)

		st.sp = 0;
		int32_t ref = live ? 0 : YMX_REF_EMPTY;
		for(;;)
		{
			// Descend internal nodes.  The first leaf met is postponed and the
			// descent goes on until every active lane of the wave has a leaf
			// to test (or no internal node left), so the lanes run the
			// triangle tests together (Aila & Laine 2009, speculative
			// traversal).  Every leaf reached is still tested exactly once.
			int32_t leaf = YMX_REF_EMPTY;
			while(ref >= 0)
			{
				const float4 *np = args.nodes + 4 * (size_t)ref;
				const float4 f0 = np[0], f1 = np[1], f2 = np[2], f3 = np[3];
				if(kStats) ++n_fetch;
				const float thi = best_t + best_t * kWindow;
				float tn0, tn1;
				const int m = box2(f0, f1, f2, r, thi, tn0, tn1);
				const int32_t c0 = __float_as_int(f3.x), c1 = __float_as_int(f3.y);
				// both children hit: the nearer next and the farther on the stack; one
				// hit: that one; none: the stack's top.  Selects for the next node, so
				// only the push and the pop stay branches.
				const bool swap = tn1 < tn0;
				int32_t next = m == 3 ? (swap ? c1 : c0) : (m == 1 ? c0 : c1);
				if(m == 3) st.push(swap ? c0 : c1);
				if(m == 0) next = st.pop();
				if(next < 0 && next != YMX_REF_EMPTY && leaf == YMX_REF_EMPTY)
				{
					leaf = next;
					next = st.pop();
				}
				ref = next;
				if(__all(leaf != YMX_REF_EMPTY || ref < 0)) break;
			}
			if(leaf == YMX_REF_EMPTY)
			{
				if(ref == YMX_REF_EMPTY) break;
				if(ref < 0)
				{
					leaf = ref;
					ref = st.pop();
				}
				else continue;  // the wave stopped while this lane still descends
			}
			const uint32_t first = ymx_leaf_first(leaf), count = ymx_leaf_count(leaf);
			for(uint32_t k = 0; k < count; ++k)
			{
				const float4 *tp = args.tris + 3 * (size_t)(first + k);
				const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
				if(kStats) ++n_tests;
				float u, v;
				const float t = prim_test(t0, t1, t2, args.quads, r.o, r.d, u, v);
				if(t <= 0.f || t < t_min) continue;
				const uint32_t prim = __float_as_uint(t0.w);
				const uint32_t vmask = __float_as_uint(t1.w);
				if(kMode == kShadowAny)
				{
					// include/accelerator/accelerator.h:135-145
					if(t >= t_max || !(vmask & YMX_VIS_CASTS_SHADOWS)) continue;
					occluded = true;
					best_prim = prim;
					break;
				}
				else if(kMode == kTransparent)
				{
					// include/accelerator/accelerator.h:147-168: an opaque caster
					// occludes; transparent ones are kept, once each (the
					// reference's `filtered` set: a one-triangle tree reaches its
					// leaf through both children), until more than max_depth of
					// them occlude.  The kept ones are stored sorted by t.
					if(t >= t_max || !(vmask & YMX_VIS_CASTS_SHADOWS)) continue;
					const bool transparent = (args.transp[prim >> 5] >> (prim & 31u)) & 1u;
					uint2 *list = args.t_hits + (size_t)rid * args.max_depth;
					bool seen = false;
					for(uint32_t j = 0; transparent && j < n_transp; ++j) seen |= list[j].x == prim;
					if(seen) continue;
					if(!transparent || n_transp >= args.max_depth)
					{
						occluded = true;
						best_prim = prim;
						break;
					}
					uint32_t k = n_transp++;
					while(k > 0 && __uint_as_float(list[k - 1].y) > t)
					{
						list[k] = list[k - 1];
						--k;
					}
					list[k] = make_uint2(prim, __float_as_uint(t));
				}
				else
				{
					// include/accelerator/accelerator.h:122-133; exact ties keep the lowest id
					if(!(vmask & YMX_VIS_VISIBLE)) continue;
					const bool closer = (t < best_t) || (t == best_t && best_prim != YMX_NO_HIT && prim < best_prim);
					if(!closer) continue;
					best_t = t;
					best_prim = prim;
					best_u = u;
					best_v = v;
				}
			}
			if(kShadow && occluded) break;
		}

		if(kShadow)
		{
			const unsigned long long mask = __ballot(occluded);
			if(__lane_id() == 0)
			{
				const uint32_t w = base >> 5;
				const uint32_t n_words = (n_rays + 31u) >> 5;
				args.occluded[w] = (uint32_t)mask;
				if(w + 1 < n_words) args.occluded[w + 1] = (uint32_t)(mask >> 32);
			}
			if(valid && args.occ_prim) args.occ_prim[rid] = occluded ? best_prim : YMX_NO_HIT;
			if(kMode == kTransparent && valid) args.t_count[rid] = occluded ? 0u : n_transp;
		}
		else if(valid)
		{
			ymx_hit h;
			if(best_prim != YMX_NO_HIT)
			{
				h.prim = best_prim;
				h.t = best_t;
				h.u = best_u;
				h.v = best_v;
			}
			else
			{
				h.prim = YMX_NO_HIT;
				h.t = 0.f;
				h.u = 0.f;
				h.v = 0.f;
			}
			reinterpret_cast<float4 *>(args.hits)[rid] =
			    make_float4(__uint_as_float(h.prim), h.t, h.u, h.v);
		}
	}
	finish_wave(args.work);
	if(kStats)
	{
		for(int off = 32; off > 0; off >>= 1)
		{
			n_fetch += __shfl_xor(n_fetch, off);
			n_tests += __shfl_xor(n_tests, off);
		}
		if(__lane_id() == 0)
		{
			atomicAdd(args.stats + 0, n_fetch);
			atomicAdd(args.stats + 1, n_tests);
		}
	}
}

// Closest hit and any hit with per-lane ray replacement (Aila & Laine 2009,
// persistent threads with dynamic ray fetch).  A lane whose ray is done takes the next
// unclaimed ray while the rest of the wave keeps its traversal state; the
// wave leaves its traversal loop to refill once fewer than kRefill lanes are
// still traversing.  Each ray is traversed exactly as by k_trace, so the
// result is the same; the any-hit occlusion bits are set by atomicOr into a
// cleared mask, since the rays of a wave are no longer one 64-ray group.  On
// the incoherent 8.4M-ray batch this is 17 % faster than k_trace for closest
// hit.  The path tracer's extension queues, whose lanes finish together, keep
// k_trace; its shadow queues use this kernel.
constexpr int kRefill = 12;
// Ray reordering's coherence test (below): neighbouring pairs sampled, and
// the count of them changing bucket from which a batch is reordered (1 in 8).
constexpr uint32_t kCoherenceSamples = 16384;
constexpr uint32_t kReorderChanges = kCoherenceSamples / 8u;

template <int kMode, bool kStats>
__global__ __launch_bounds__(kTraceBlock) void k_dyn(ymx_trace_args args)
{
	static_assert(kMode == kClosest || kMode == kShadowAny, "transparent shadows use k_trace");
	__shared__ int32_t lds_stack[kStack * kTraceBlock];
	const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
	const int lane = (int)__lane_id();
	LaneStack st;
	st.lds = lds_stack + threadIdx.x;
	st.spill = args.spill + gtid;
	st.stride = args.grid_threads;
	st.cap = args.spill_depth;
	st.sp = 0;
	unsigned long long n_fetch = 0, n_tests = 0;
	const uint32_t n_rays = args.n_dev ? *args.n_dev : args.n_rays;
	const bool reordered = args.order && *args.order_changes >= kReorderChanges;
	const SegView sv = seg_view(args, n_rays, 0u);
	WaveWork ww = wave_work_start(sv);

	bool alive = false, exhausted = false;
	uint32_t rid = 0;
	RayState r;
	float t_min = 0.f, t_max = 0.f, best_t = 0.f, best_u = 0.f, best_v = 0.f;
	uint32_t best_prim = YMX_NO_HIT;
	int32_t ref = YMX_REF_EMPTY;
	for(;;)
	{
		// lanes without a ray claim consecutive ray ids in lane order
		const bool need = !alive && !exhausted;
		const unsigned long long need_mask = __ballot(need);
		if(need_mask)
		{
			uint32_t got = 0, ray_end = 0;
			const uint32_t first = wave_claim(args.work, ww, sv, (uint32_t)__popcll(need_mask), got, ray_end);
			if(need)
			{
				const uint32_t rank = (uint32_t)__popcll(need_mask & ((1ull << lane) - 1ull));
				// nothing left anywhere: done; a short claim at a segment's end
				// leaves the lanes past it to claim again on the next pass
				if(got == 0u) exhausted = true;
				else if(rank < got)
				{
					rid = first + rank;
					// with ray reordering the k-th ray traced is the order[k]-th of the
					// batch; it reads and writes the caller's slot
					if(reordered) rid = args.order[rid];
					const ymx_ray ray = load_ray(args.rays, rid);
					setup_ray(ray, r);
					t_max = (ray.tmax >= 0.f) ? ray.tmax : FLT_MAX;
					float enter = 0.f, leave = 0.f;
					const bool live = tree_cross(args.bmin, args.bmax, ray.org, ray.dir, t_max, enter, leave);
					// include/accelerator/accelerator_kdtree_common.h:139
					const float bias = kBiasScale * fabsf(leave - enter);
					t_min = (kMode == kShadowAny) ? bias : ((ray.tmin < bias) ? bias : ray.tmin);
					r.tlo = t_min - t_min * kWindow;
					best_t = t_max;
					best_prim = YMX_NO_HIT;
					best_u = best_v = 0.f;
					st.sp = 0;
					ref = live ? 0 : YMX_REF_EMPTY;
					alive = true;
				}
			}
		}
		if(__all(exhausted)) break;
		if(alive)
		{
			for(;;)
			{
				// descent with the first leaf postponed (as in k_trace)
				int32_t leaf = YMX_REF_EMPTY;
				while(ref >= 0)
				{
					const float4 *np = args.nodes + 4 * (size_t)ref;
					const float4 f0 = np[0], f1 = np[1], f2 = np[2], f3 = np[3];
					if(kStats) ++n_fetch;
					const float thi = best_t + best_t * kWindow;
					float tn0, tn1;
					const int m = box2(f0, f1, f2, r, thi, tn0, tn1);
					const int32_t c0 = __float_as_int(f3.x), c1 = __float_as_int(f3.y);
					// both children hit: the nearer next and the farther on the stack; one
					// hit: that one; none: the stack's top.  Selects for the next node, so
					// only the push and the pop stay branches.
					const bool swap = tn1 < tn0;
					int32_t next = m == 3 ? (swap ? c1 : c0) : (m == 1 ? c0 : c1);
					if(m == 3) st.push(swap ? c0 : c1);
					if(m == 0) next = st.pop();
					if(next < 0 && next != YMX_REF_EMPTY && leaf == YMX_REF_EMPTY)
					{
						leaf = next;
						next = st.pop();
					}
					ref = next;
					if(__all(leaf != YMX_REF_EMPTY || ref < 0)) break;
				}
				if(leaf == YMX_REF_EMPTY && ref < 0 && ref != YMX_REF_EMPTY)
				{
					leaf = ref;
					ref = st.pop();
				}
				if(leaf != YMX_REF_EMPTY)
				{
					const uint32_t first = ymx_leaf_first(leaf), count = ymx_leaf_count(leaf);
					for(uint32_t k = 0; k < count; ++k)
					{
						const float4 *tp = args.tris + 3 * (size_t)(first + k);
						const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
						if(kStats) ++n_tests;
						float u, v;
						const float t = prim_test(t0, t1, t2, args.quads, r.o, r.d, u, v);
						if(t <= 0.f || t < t_min) continue;
						const uint32_t prim = __float_as_uint(t0.w);
						const uint32_t vmask = __float_as_uint(t1.w);
						if(kMode == kShadowAny)
						{
							// include/accelerator/accelerator.h:135-145
							if(t >= t_max || !(vmask & YMX_VIS_CASTS_SHADOWS)) continue;
							best_prim = prim;
							ref = YMX_REF_EMPTY;  // occluded: the ray is done
							break;
						}
						// include/accelerator/accelerator.h:122-133; exact ties keep the lowest id
						if(!(vmask & YMX_VIS_VISIBLE)) continue;
						const bool closer =
						    (t < best_t) || (t == best_t && best_prim != YMX_NO_HIT && prim < best_prim);
						if(!closer) continue;
						best_t = t;
						best_prim = prim;
						best_u = u;
						best_v = v;
					}
				}
				if(ref == YMX_REF_EMPTY)
				{
					// ray done
					const bool hit = best_prim != YMX_NO_HIT;
					if(kMode == kShadowAny)
					{
						if(hit) atomicOr(args.occluded + (rid >> 5), 1u << (rid & 31u));
						if(args.occ_prim) args.occ_prim[rid] = best_prim;
					}
					else
						reinterpret_cast<float4 *>(args.hits)[rid] = make_float4(
						    __uint_as_float(best_prim), hit ? best_t : 0.f, hit ? best_u : 0.f, hit ? best_v : 0.f);
					alive = false;
					break;
				}
				if(__popcll(__ballot(true)) < kRefill) break;
			}
		}
	}
	finish_wave(args.work);
	if(kStats)
	{
		for(int off = 32; off > 0; off >>= 1)
		{
			n_fetch += __shfl_xor(n_fetch, off);
			n_tests += __shfl_xor(n_tests, off);
		}
		if(__lane_id() == 0)
		{
			atomicAdd(args.stats + 0, n_fetch);
			atomicAdd(args.stats + 1, n_tests);
		}
	}
}

__global__ void k_clear_words(uint32_t *p, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n) p[i] = 0u;
}

// ---- ray reordering.  Rays that start near one another and head into the
// same octant take similar paths through the tree; traced in that order, a
// wave's lanes stay together longer.  The bucket of a ray: its direction
// octant (3 bits) above a Morton code of its origin's cell in the tree bound,
// kCellBits bits per axis (the sort granularity: 3 gives 4,096 buckets, 4
// gives 32,768).  One counting-sort pass over the buckets gives the order
// (ray indices bucket by bucket; within a bucket in no fixed order, which
// changes no result: every ray is traced on its own and writes its own slot),
// and the trace kernels read rays[order[k]].
//
// No host round trip decides whether to reorder: a first kernel counts how
// many of kCoherenceSamples neighbouring ray pairs change bucket, and every
// later kernel of the pass, the trace kernel included, reads that count and
// skips the reordering (or ignores the order) when the batch is coherent
// already (camera rays in pixel order).  Device work only, on the caller's
// stream.
constexpr uint32_t kReorderMin = 1u << 18;

template <int kCellBits>
struct Bucketing
{
	static constexpr uint32_t kBuckets = 8u << (3 * kCellBits);
	static constexpr uint32_t kBlock = kCellBits <= 3 ? 512u : 1024u;     // histogram and scatter blocks
	static constexpr uint32_t kTilesMax = kCellBits <= 3 ? 512u : 256u;   // their count per pass (at most)
};

struct KeyGrid
{
	float lo[3], scale[3];
};

template <int kCellBits>
__device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
	uint32_t out = 0u;
	for(int b = 0; b < kCellBits; ++b) out |= ((v >> b) & 1u) << (3 * b);
	return out;
}

// The origin's cell on a 64^3 grid over the tree bound: the bucket takes the
// octant and the top kCellBits bits of each coordinate; `fine` the Morton code
// of the remaining 6 - kCellBits bits, which orders the rays within a bucket.
constexpr int kOriginBits = 6;

template <int kCellBits>
__device__ __forceinline__ uint32_t ray_bucket(const ymx_ray &r, const KeyGrid &g, uint32_t &fine)
{
	constexpr int kFine = kOriginBits - kCellBits;
	uint32_t q[3];
	for(int k = 0; k < 3; ++k)
	{
		// NaN and out-of-bound origins clamp to the grid
		const float c = fminf(fmaxf((r.org[k] - g.lo[k]) * g.scale[k], 0.f), (float)((1 << kOriginBits) - 1));
		q[k] = c == c ? (uint32_t)c : 0u;
	}
	const uint32_t oct = (r.dir[0] < 0.f ? 1u : 0u) | (r.dir[1] < 0.f ? 2u : 0u) | (r.dir[2] < 0.f ? 4u : 0u);
	constexpr uint32_t low = (1u << kFine) - 1u;
	fine = spread_bits<kFine>(q[0] & low) | (spread_bits<kFine>(q[1] & low) << 1) | (spread_bits<kFine>(q[2] & low) << 2);
	return (oct << (3 * kCellBits)) | spread_bits<kCellBits>(q[0] >> kFine) |
	       (spread_bits<kCellBits>(q[1] >> kFine) << 1) | (spread_bits<kCellBits>(q[2] >> kFine) << 2);
}

// The work area's reordering buffers for one pass of n rays over `tiles`
// histogram blocks of `tile` rays.
struct ReorderBufs
{
	uint32_t *changes;   // 1 word: changing neighbour pairs among the samples (or kReorderChanges when forced)
	uint16_t *bucket;    // n: each ray's bucket
	uint16_t *fine;      // n: each ray's finer cell within its bucket
	uint16_t *fine_sorted;  // n: the fine cells in bucket order
	uint32_t *order;     // n: the ray indices in bucket order
	uint32_t *order2;    // n: ... and within each bucket by fine cell
	uint32_t *table;     // buckets x tiles, bucket-major: rays of bucket b in tile t, then its exclusive offset
	uint32_t *base;      // buckets: first slot of each bucket
	uint32_t n, tile, tiles;
};

__device__ __forceinline__ bool reorder_off(const uint32_t *changes)
{
	return *changes < kReorderChanges;
}

// Neighbouring pairs (i, i + 1) at positions of the golden-ratio (Weyl)
// sequence over the batch, so that no stride of the batch's own layout (an
// image row of camera rays) lines up with the samples.
template <int kCellBits>
__global__ __launch_bounds__(256) void k_ray_coherence(const ymx_ray *rays, uint32_t n, KeyGrid g, uint32_t *changes)
{
	const uint32_t s = blockIdx.x * 256u + threadIdx.x;
	uint32_t diff = 0u;
	if(s < kCoherenceSamples)
	{
		const uint32_t i = (uint32_t)(((uint64_t)(s * 2654435769u) * (uint64_t)(n - 1u)) >> 32);
		uint32_t f0, f1;
		diff = ray_bucket<kCellBits>(load_ray(rays, i), g, f0) != ray_bucket<kCellBits>(load_ray(rays, i + 1u), g, f1)
		           ? 1u
		           : 0u;
	}
	const int c = __syncthreads_count((int)diff);
	if(threadIdx.x == 0 && c) atomicAdd(changes, (uint32_t)c);
}

__global__ void k_set_word(uint32_t *p, uint32_t v)
{
	if(threadIdx.x == 0) *p = v;
}

// Per tile: each ray's bucket, and the tile's count of every bucket.
template <int kCellBits>
__global__ __launch_bounds__(Bucketing<kCellBits>::kBlock) void k_bucket_count(const ymx_ray *rays, KeyGrid g,
                                                                               ReorderBufs b)
{
	using B = Bucketing<kCellBits>;
	if(reorder_off(b.changes)) return;
	__shared__ uint32_t hist[B::kBuckets];
	for(uint32_t k = threadIdx.x; k < B::kBuckets; k += B::kBlock) hist[k] = 0u;
	__syncthreads();
	const uint64_t first = (uint64_t)blockIdx.x * b.tile;
	const uint64_t end = first + b.tile < b.n ? first + b.tile : b.n;
	for(uint64_t i = first + threadIdx.x; i < end; i += B::kBlock)
	{
		uint32_t fine;
		const uint32_t k = ray_bucket<kCellBits>(load_ray(rays, (uint32_t)i), g, fine);
		b.bucket[i] = (uint16_t)k;
		b.fine[i] = (uint16_t)fine;
		atomicAdd(&hist[k], 1u);
	}
	__syncthreads();
	for(uint32_t k = threadIdx.x; k < B::kBuckets; k += B::kBlock) b.table[k * b.tiles + blockIdx.x] = hist[k];
}

// One block per bucket: exclusive prefix of its counts over the tiles, in
// place, and the bucket's total into base[].
__global__ __launch_bounds__(256) void k_bucket_tiles(ReorderBufs b)
{
	if(reorder_off(b.changes)) return;
	__shared__ uint32_t wsum[4];
	uint32_t *row = b.table + (size_t)blockIdx.x * b.tiles;
	const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
	uint32_t carry = 0u;
	for(uint32_t c0 = 0; c0 < b.tiles; c0 += 256u)
	{
		const uint32_t t = c0 + threadIdx.x;
		const uint32_t v = t < b.tiles ? row[t] : 0u;
		uint32_t x = v;  // inclusive scan within the wave
		for(int off = 1; off < 64; off <<= 1)
		{
			const uint32_t y = __shfl_up(x, off);
			if(lane >= off) x += y;
		}
		if(lane == 63) wsum[wave] = x;
		__syncthreads();
		uint32_t before = carry;
		for(int w = 0; w < wave; ++w) before += wsum[w];
		if(t < b.tiles) row[t] = before + x - v;
		carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
		__syncthreads();
	}
	if(threadIdx.x == 0) b.base[blockIdx.x] = carry;
}

// One block: the buckets' totals into their exclusive prefix (first slots).
template <int kCellBits>
__global__ __launch_bounds__(1024) void k_bucket_base(ReorderBufs b)
{
	if(reorder_off(b.changes)) return;
	constexpr uint32_t kPer = Bucketing<kCellBits>::kBuckets / 1024u;
	__shared__ uint32_t wsum[16];
	const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
	uint32_t sum = 0u;
	for(uint32_t k = 0; k < kPer; ++k) sum += b.base[threadIdx.x * kPer + k];
	uint32_t x = sum;
	for(int off = 1; off < 64; off <<= 1)
	{
		const uint32_t y = __shfl_up(x, off);
		if(lane >= off) x += y;
	}
	if(lane == 63) wsum[wave] = x;
	__syncthreads();
	uint32_t run = x - sum;
	for(int w = 0; w < wave; ++w) run += wsum[w];
	for(uint32_t k = 0; k < kPer; ++k)
	{
		const uint32_t v = b.base[threadIdx.x * kPer + k];
		b.base[threadIdx.x * kPer + k] = run;
		run += v;
	}
}

// Per tile: each ray's slot is its bucket's first slot + the tile's offset in
// the bucket + its rank among the tile's rays of the bucket.
template <int kCellBits>
__global__ __launch_bounds__(Bucketing<kCellBits>::kBlock) void k_bucket_scatter(ReorderBufs b)
{
	using B = Bucketing<kCellBits>;
	if(reorder_off(b.changes)) return;
	__shared__ uint32_t next[B::kBuckets];
	for(uint32_t k = threadIdx.x; k < B::kBuckets; k += B::kBlock)
		next[k] = b.base[k] + b.table[k * b.tiles + blockIdx.x];
	__syncthreads();
	const uint64_t first = (uint64_t)blockIdx.x * b.tile;
	const uint64_t end = first + b.tile < b.n ? first + b.tile : b.n;
	for(uint64_t i = first + threadIdx.x; i < end; i += B::kBlock)
	{
		const uint32_t slot = atomicAdd(&next[b.bucket[i]], 1u);
		b.order[slot] = (uint32_t)i;
		b.fine_sorted[slot] = b.fine[i];
	}
}

// One block per bucket: its rays ordered by their fine cell (a counting sort
// in LDS; within a fine cell in no fixed order).
template <int kCellBits>
__global__ __launch_bounds__(256) void k_bucket_refine(ReorderBufs b)
{
	if(reorder_off(b.changes)) return;
	constexpr uint32_t kSub = 1u << (3 * (kOriginBits - kCellBits));
	__shared__ uint32_t next[kSub];
	__shared__ uint32_t wsum[4];
	const uint32_t lo = b.base[blockIdx.x];
	const uint32_t hi = blockIdx.x + 1u < Bucketing<kCellBits>::kBuckets ? b.base[blockIdx.x + 1u] : b.n;
	if(hi - lo < 2u)
	{
		if(hi > lo && threadIdx.x == 0) b.order2[lo] = b.order[lo];
		return;
	}
	for(uint32_t k = threadIdx.x; k < kSub; k += 256u) next[k] = 0u;
	__syncthreads();
	for(uint32_t i = lo + threadIdx.x; i < hi; i += 256u) atomicAdd(&next[b.fine_sorted[i]], 1u);
	__syncthreads();
	// exclusive prefix of the kSub counts, kSub / 256 consecutive ones per thread
	constexpr uint32_t kPer = kSub >= 256u ? kSub / 256u : 1u;
	const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
	uint32_t v[kPer], sum = 0u;
	for(uint32_t k = 0; k < kPer; ++k)
	{
		const uint32_t j = threadIdx.x * kPer + k;
		v[k] = j < kSub ? next[j] : 0u;
		sum += v[k];
	}
	uint32_t x = sum;
	for(int off = 1; off < 64; off <<= 1)
	{
		const uint32_t y = __shfl_up(x, off);
		if(lane >= off) x += y;
	}
	if(lane == 63) wsum[wave] = x;
	__syncthreads();
	uint32_t run = lo + x - sum;
	for(int w = 0; w < wave; ++w) run += wsum[w];
	for(uint32_t k = 0; k < kPer; ++k)
	{
		const uint32_t j = threadIdx.x * kPer + k;
		if(j < kSub) next[j] = run;
		run += v[k];
	}
	__syncthreads();
	for(uint32_t i = lo + threadIdx.x; i < hi; i += 256u)
	{
		const uint32_t slot = atomicAdd(&next[b.fine_sorted[i]], 1u);
		b.order2[slot] = b.order[i];
	}
}

template <int kCellBits>
int reorder_rays_bits(const ymx_accel *acc, ymx_work_area *wa, const ymx_ray *rays, uint32_t n, bool always,
                      hipStream_t stream, const uint32_t *&order, const uint32_t *&changes)
{
	using B = Bucketing<kCellBits>;
	ReorderBufs b;
	b.n = n;
	b.tiles = (n + 4095u) / 4096u;
	if(b.tiles > B::kTilesMax) b.tiles = B::kTilesMax;
	b.tile = (n + b.tiles - 1u) / b.tiles;
	b.tiles = (n + b.tile - 1u) / b.tile;
	auto up = [](size_t v) { return (v + 255u) / 256u * 256u; };
	const size_t o_bucket = 256, o_fine = o_bucket + up(2ull * n), o_fsort = o_fine + up(2ull * n);
	const size_t o_order = o_fsort + up(2ull * n), o_order2 = o_order + up(4ull * n), o_table = o_order2 + up(4ull * n);
	const size_t o_base = o_table + up(4ull * B::kBuckets * b.tiles), need = o_base + up(4ull * B::kBuckets);
	if(wa->sort_bytes < need)
	{
		// the previous launches on this area may still read the old buffer
		YMX_HIP_CHECK(hipStreamSynchronize(stream));
		if(wa->d_sort) YMX_HIP_CHECK(hipFree(wa->d_sort));
		wa->d_sort = nullptr;
		wa->sort_bytes = 0;
		YMX_HIP_CHECK(hipMalloc(&wa->d_sort, need));
		wa->sort_bytes = need;
	}
	char *base = static_cast<char *>(wa->d_sort);
	b.changes = reinterpret_cast<uint32_t *>(base);
	b.bucket = reinterpret_cast<uint16_t *>(base + o_bucket);
	b.fine = reinterpret_cast<uint16_t *>(base + o_fine);
	b.fine_sorted = reinterpret_cast<uint16_t *>(base + o_fsort);
	b.order = reinterpret_cast<uint32_t *>(base + o_order);
	b.order2 = reinterpret_cast<uint32_t *>(base + o_order2);
	b.table = reinterpret_cast<uint32_t *>(base + o_table);
	b.base = reinterpret_cast<uint32_t *>(base + o_base);
	KeyGrid g;
	for(int k = 0; k < 3; ++k)
	{
		const float e = acc->bmax[k] - acc->bmin[k];
		g.lo[k] = acc->bmin[k];
		g.scale[k] = e > 0.f ? (float)(1 << kOriginBits) / e : 0.f;
	}
	if(always || n < 2u) k_set_word<<<1, 64, 0, stream>>>(b.changes, always ? kReorderChanges : 0u);
	else
	{
		k_clear_words<<<1, 64, 0, stream>>>(b.changes, 1u);
		k_ray_coherence<kCellBits><<<kCoherenceSamples / 256u, 256, 0, stream>>>(rays, n, g, b.changes);
	}
	k_bucket_count<kCellBits><<<b.tiles, B::kBlock, 0, stream>>>(rays, g, b);
	k_bucket_tiles<<<B::kBuckets, 256, 0, stream>>>(b);
	k_bucket_base<kCellBits><<<1, 1024, 0, stream>>>(b);
	k_bucket_scatter<kCellBits><<<b.tiles, B::kBlock, 0, stream>>>(b);
	k_bucket_refine<kCellBits><<<B::kBuckets, 256, 0, stream>>>(b);
	YMX_HIP_CHECK(hipGetLastError());
	order = b.order2;
	changes = b.changes;
	return YMX_OK;
}

// Queues the reordering of one launch's rays on `stream` into the work area
// (grown on demand).  order receives the index array and changes the word the
// trace kernel reads to know whether the array was written.
int reorder_rays(const ymx_accel *acc, ymx_work_area *wa, const ymx_ray *rays, uint32_t n, bool always,
                 hipStream_t stream, const uint32_t *&order, const uint32_t *&changes)
{
	if(acc->reorder_cell_bits == 4) return reorder_rays_bits<4>(acc, wa, rays, n, always, stream, order, changes);
	return reorder_rays_bits<3>(acc, wa, rays, n, always, stream, order, changes);
}

template <int kMode, bool kStats, bool kDynamic = false>
int launch(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t *occ, uint32_t *prim, uint32_t n,
           hipStream_t stream, unsigned long long *stats, const uint32_t *n_dev = nullptr,
           const uint32_t *transp = nullptr, uint32_t max_depth = 0, uint32_t *t_count = nullptr,
           uint2 *t_hits = nullptr, const uint32_t *seg_counts = nullptr, uint32_t seg_cap = 0)
{
	constexpr bool kShadow = kMode != kClosest;
	if(n == 0) return YMX_OK;
	ymx_trace_args a;
	a.nodes = acc->d_nodes;
	a.tris = acc->d_tris;
	a.quads = acc->d_quads;
	a.rays = rays;
	a.n_rays = n;
	a.n_dev = n_dev;
	a.seg_counts = seg_counts;
	a.seg_cap = seg_cap;
	for(int k = 0; k < 3; ++k)
	{
		a.bmin[k] = acc->bmin[k];
		a.bmax[k] = acc->bmax[k];
	}
	std::lock_guard<std::mutex> lock(acc->work_mutex);
	ymx_work_area *wa = ymx_work_area_for(acc, stream);
	if(!wa) return YMX_ERR_OUT_OF_MEMORY;
	a.work = wa->d_work;
	a.spill = wa->d_spill;
	a.spill_depth = acc->spill_depth;
	a.grid_threads = kShadow ? acc->grid_threads_shadow : acc->grid_threads_closest;
	a.hits = hits;
	a.occluded = occ;
	a.occ_prim = prim;
	a.stats = stats;
	a.transp = transp;
	a.max_depth = max_depth;
	a.t_count = t_count;
	a.t_hits = t_hits;
	const uint32_t needed = ((n + 63u) / 64u) * 64u;
	uint32_t threads = a.grid_threads < needed ? a.grid_threads : needed;
	const uint32_t blocks = (threads + kTraceBlock - 1) / kTraceBlock;
	a.order = nullptr;
	a.order_changes = nullptr;
	if constexpr(kDynamic && kMode != kTransparent)
	{
		// automatic for closest hit only: an any-hit batch, whose rays stop at
		// their first occluder, traced 3.7 % slower reordered (2,675 -> 2,575
		// Mrays/s on the 8.4M-ray incoherent batch) once the grid claimed its
		// work XCD by XCD in contiguous segments
		const bool reorder = n_dev == nullptr && seg_counts == nullptr && !kStats &&
		                     (acc->ray_reorder > 0 ||
		                      (acc->ray_reorder == 0 && kMode == kClosest && n >= kReorderMin));
		if(reorder)
		{
			const int rc = reorder_rays(acc, wa, rays, n, acc->ray_reorder > 0, stream, a.order, a.order_changes);
			if(rc != YMX_OK) return rc;
		}
		if(kShadow)
		{
			// occlusion bits are OR-ed in: clear the mask (a kernel, not a
			// memset node, so the launch sequence can be graph-captured)
			const uint32_t words = (n + 31u) / 32u;
			k_clear_words<<<(words + 255u) / 256u, 256, 0, stream>>>(occ, words);
		}
		k_dyn<kMode, kStats><<<blocks, kTraceBlock, 0, stream>>>(a);
	}
	else k_trace<kMode, kStats><<<blocks, kTraceBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	YMX_HIP_CHECK(hipEventRecord(wa->done, stream));
	return YMX_OK;
}

}  // namespace

int ymx_trace_setup(ymx_accel *acc)
{
	int dev = 0;
	YMX_HIP_CHECK(hipGetDevice(&dev));
	hipDeviceProp_t prop;
	YMX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
	int per_cu_c = 0, per_cu_s = 0;
	YMX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_c, k_trace<kClosest, false>, kTraceBlock, 0));
	int per_cu_d = 0;
	YMX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_d, k_dyn<kClosest, false>, kTraceBlock, 0));
	if(per_cu_d < per_cu_c) per_cu_c = per_cu_d;  // the closest-hit grid serves both
	YMX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s, k_trace<kShadowAny, false>, kTraceBlock, 0));
	int per_cu_t = 0;
	YMX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_t, k_trace<kTransparent, false>, kTraceBlock, 0));
	if(per_cu_t < per_cu_s) per_cu_s = per_cu_t;  // the any-hit grid serves all three
	int per_cu_ds = 0;
	YMX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_ds, k_dyn<kShadowAny, false>, kTraceBlock, 0));
	if(per_cu_ds < per_cu_s) per_cu_s = per_cu_ds;
	if(per_cu_c < 1) per_cu_c = 1;
	if(per_cu_s < 1) per_cu_s = 1;
	acc->grid_threads_closest = (uint32_t)(prop.multiProcessorCount * per_cu_c * kTraceBlock);
	acc->grid_threads_shadow = (uint32_t)(prop.multiProcessorCount * per_cu_s * kTraceBlock);
	const uint32_t gt = acc->grid_threads_closest > acc->grid_threads_shadow ? acc->grid_threads_closest
	                                                                         : acc->grid_threads_shadow;
	acc->grid_threads_max = gt;
	acc->spill_depth = acc->max_depth > (uint32_t)kStack ? acc->max_depth - (uint32_t)kStack + 1u : 1u;
	// the allocation-free query's workspace on this device
	const int rc = ymx_noalloc_setup(acc->device);
	if(rc != YMX_OK) return rc;
	// the work area of the default stream, so its first query allocates nothing
	std::lock_guard<std::mutex> lock(acc->work_mutex);
	return ymx_work_area_for(acc, nullptr) ? YMX_OK : YMX_ERR_OUT_OF_MEMORY;
}

ymx_work_area *ymx_work_area_for(const ymx_accel *acc, hipStream_t stream)
{
	for(ymx_work_area &w : acc->work_areas)
		if(w.stream == stream) return &w;
	for(ymx_work_area &w : acc->work_areas)
	{
		if(hipEventQuery(w.done) == hipSuccess)
		{
			w.stream = stream;
			return &w;
		}
	}
	ymx_work_area w = {};
	w.stream = stream;
	const size_t spill_bytes = sizeof(int32_t) * (size_t)acc->spill_depth * acc->grid_threads_max;
	if(hipMalloc(&w.d_work, sizeof(uint32_t) * kWorkWords) != hipSuccess ||
	   hipMemset(w.d_work, 0, sizeof(uint32_t) * kWorkWords) != hipSuccess ||
	   hipMalloc(&w.d_spill, spill_bytes) != hipSuccess ||
	   hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess)
	{
		if(w.d_work) (void)hipFree(w.d_work);
		if(w.d_spill) (void)hipFree(w.d_spill);
		ymx_set_error("query work area: device allocation of %zu bytes failed",
		              spill_bytes + sizeof(uint32_t) * kWorkWords);
		return nullptr;
	}
	acc->work_areas.push_back(w);
	return &acc->work_areas.back();
}

void ymx_free_work_areas(ymx_accel *acc)
{
	for(ymx_work_area &w : acc->work_areas)
	{
		(void)hipFree(w.d_work);
		(void)hipFree(w.d_spill);
		if(w.d_sort) (void)hipFree(w.d_sort);
		(void)hipEventDestroy(w.done);
	}
	acc->work_areas.clear();
}

int ymx_launch_closest(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t n, hipStream_t stream,
                       unsigned long long *stats)
{
	if(stats) return launch<kClosest, true, true>(acc, rays, hits, nullptr, nullptr, n, stream, stats);
	return launch<kClosest, false, true>(acc, rays, hits, nullptr, nullptr, n, stream, nullptr);
}

int ymx_launch_closest_ordered(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t n,
                               hipStream_t stream)
{
	return launch<kClosest, false>(acc, rays, hits, nullptr, nullptr, n, stream, nullptr);
}

int ymx_launch_closest_queue(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, const uint32_t *n_dev,
                             uint32_t capacity, bool replacement, hipStream_t stream)
{
	if(replacement)
		return launch<kClosest, false, true>(acc, rays, hits, nullptr, nullptr, capacity, stream, nullptr, n_dev);
	return launch<kClosest, false>(acc, rays, hits, nullptr, nullptr, capacity, stream, nullptr, n_dev);
}

int ymx_launch_shadow_queue(const ymx_accel *acc, const ymx_ray *rays, uint32_t *occ, const uint32_t *n_dev,
                            uint32_t capacity, hipStream_t stream)
{
	return launch<kShadowAny, false>(acc, rays, nullptr, occ, nullptr, capacity, stream, nullptr, n_dev);
}

int ymx_launch_shadow(const ymx_accel *acc, const ymx_ray *rays, uint32_t *occ, uint32_t *prim, uint32_t n,
                      hipStream_t stream, unsigned long long *stats)
{
	if(stats) return launch<kShadowAny, true, true>(acc, rays, nullptr, occ, prim, n, stream, stats);
	return launch<kShadowAny, false, true>(acc, rays, nullptr, occ, prim, n, stream, nullptr);
}

int ymx_launch_transparent(const ymx_accel *acc, const ymx_ray *rays, const uint32_t *transp, uint32_t max_depth,
                           uint32_t *occ, uint32_t *counts, uint2 *hits, uint32_t n, const uint32_t *n_dev,
                           hipStream_t stream)
{
	return launch<kTransparent, false>(acc, rays, nullptr, occ, nullptr, n, stream, nullptr, n_dev, transp, max_depth,
	                                   counts, hits);
}

int ymx_launch_shadow_segmented(const ymx_accel *acc, const ymx_ray *rays, uint32_t *occ, const uint32_t *seg_counts,
                                uint32_t seg_cap, bool replacement, hipStream_t stream)
{
	if(seg_cap % 64u)
	{
		ymx_set_error("segmented shadow queue: %u rays per segment is not a multiple of 64", seg_cap);
		return YMX_ERR_INVALID_ARG;
	}
	if(replacement)
		return launch<kShadowAny, false, true>(acc, rays, nullptr, occ, nullptr, kQueueSegments * seg_cap, stream,
		                                       nullptr, nullptr, nullptr, 0, nullptr, nullptr, seg_counts, seg_cap);
	return launch<kShadowAny, false>(acc, rays, nullptr, occ, nullptr, kQueueSegments * seg_cap, stream, nullptr,
	                                 nullptr, nullptr, 0, nullptr, nullptr, seg_counts, seg_cap);
}

int ymx_launch_transparent_segmented(const ymx_accel *acc, const ymx_ray *rays, const uint32_t *transp,
                                     uint32_t max_depth, uint32_t *occ, uint32_t *counts, uint2 *hits,
                                     const uint32_t *seg_counts, uint32_t seg_cap, hipStream_t stream)
{
	if(seg_cap % 64u)
	{
		ymx_set_error("segmented shadow queue: %u rays per segment is not a multiple of 64", seg_cap);
		return YMX_ERR_INVALID_ARG;
	}
	return launch<kTransparent, false>(acc, rays, nullptr, occ, nullptr, kQueueSegments * seg_cap, stream, nullptr,
	                                   nullptr, transp, max_depth, counts, hits, seg_counts, seg_cap);
}

// ---------------------------------------------------------------------------
// Fixed workspace of the allocation-free closest-hit query: the work
// counters of the persistent grid in its first kScratchHead words, then the
// spill columns of the persistent lanes.  Every device holds its own copy
// from the load of this library until the process ends.  Host code must not
// pass `scratch` itself to a kernel (on the host it only names the symbol):
// ymx_noalloc_setup looks up each device's address once, and the launches use
// that address.
namespace {
constexpr size_t kScratchWords = (size_t)16 << 20;  // 64 MiB of int32
constexpr size_t kScratchHead = kWorkWords;         // the work counters (520 KiB)
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
int32_t *g_scratch_ptr[kScratchDevices];            // per device, set once
}  // namespace

static __device__ __align__(256) int32_t scratch[kScratchWords];

int ymx_noalloc_setup(int device)
{
	if(device < 0 || device >= kScratchDevices)
	{
		ymx_set_error("query workspace: device %d outside 0..%d", device, kScratchDevices - 1);
		return YMX_ERR_UNSUPPORTED;
	}
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if(g_scratch_ptr[device]) return YMX_OK;
	void *p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if(e != hipSuccess || !p)
	{
		ymx_set_error("query workspace: hipGetSymbolAddress on device %d failed: %s", device, hipGetErrorString(e));
		return YMX_ERR_HIP;
	}
	g_scratch_ptr[device] = static_cast<int32_t *>(p);
	return YMX_OK;
}

int ymx_launch_closest_noalloc(const ymx_accel *acc, const ymx_ray *rays, ymx_hit *hits, uint32_t n,
                               hipStream_t stream)
{
	if(n == 0) return YMX_OK;
	int32_t *ws = nullptr;
	{
		std::lock_guard<std::mutex> lock(g_scratch_mutex);
		if(acc->device >= 0 && acc->device < kScratchDevices) ws = g_scratch_ptr[acc->device];
	}
	if(!ws)
	{
		ymx_set_error("ymx_accel_intersect_noalloc: no query workspace on device %d", acc->device);
		return YMX_ERR_HIP;
	}
	// as many persistent lanes as the closest-hit grid has, or as many whole
	// blocks as the spill area holds columns of spill_depth entries
	const size_t fit = (kScratchWords - kScratchHead) / acc->spill_depth;
	uint32_t lanes = acc->grid_threads_closest;
	if(fit < lanes) lanes = (uint32_t)(fit / kTraceBlock) * (uint32_t)kTraceBlock;
	if(lanes < (uint32_t)kTraceBlock)
	{
		ymx_set_error("ymx_accel_intersect_noalloc: a tree of depth %u needs %zu bytes of workspace per %d-lane "
		              "block; the fixed workspace holds %zu",
		              acc->max_depth, sizeof(int32_t) * (size_t)acc->spill_depth * kTraceBlock, kTraceBlock,
		              sizeof(int32_t) * (kScratchWords - kScratchHead));
		return YMX_ERR_UNSUPPORTED;
	}
	ymx_trace_args a = {};
	a.nodes = acc->d_nodes;
	a.tris = acc->d_tris;
	a.quads = acc->d_quads;
	a.rays = rays;
	a.n_rays = n;
	for(int k = 0; k < 3; ++k)
	{
		a.bmin[k] = acc->bmin[k];
		a.bmax[k] = acc->bmax[k];
	}
	a.work = reinterpret_cast<uint32_t *>(ws);
	a.spill = ws + kScratchHead;
	a.spill_depth = acc->spill_depth;
	a.grid_threads = lanes;  // the spill columns' stride
	a.hits = hits;
	const uint32_t needed = ((n + 63u) / 64u) * 64u;
	const uint32_t threads = lanes < needed ? lanes : needed;
	const uint32_t blocks = (threads + kTraceBlock - 1) / kTraceBlock;
	// the counters start at zero: a fill kernel, not a memset node
	k_clear_words<<<(uint32_t)((kWorkWords + 255u) / 256u), 256, 0, stream>>>(a.work, (uint32_t)kWorkWords);
	k_dyn<kClosest, false><<<blocks, kTraceBlock, 0, stream>>>(a);
	YMX_HIP_CHECK(hipGetLastError());
	return YMX_OK;
}
