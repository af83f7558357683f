// accel_api.hip -- the extern "C" entry points declared in include/ymx_accel.h.
#include "accel_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <vector>

namespace {
thread_local char g_last_error[512] = "";

void free_accel(ymx_accel *acc)
{
	if(!acc) return;
	if(acc->d_nodes) (void)hipFree(acc->d_nodes);
	if(acc->d_tris) (void)hipFree(acc->d_tris);
	if(acc->d_quads) (void)hipFree(acc->d_quads);
	ymx_free_work_areas(acc);
	delete acc;
}
}  // namespace

void ymx_set_error(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
	va_end(ap);
}

extern "C" {

const char *ymx_last_error(void) { return g_last_error; }

static int create_accel(const char *fn, const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t stride,
                        const uint8_t *vis, uint32_t n_tris, const ymx_accel_params *params, ymx_accel **out)
{
	g_last_error[0] = '\0';
	if(!out)
	{
		ymx_set_error("%s: out is NULL", fn);
		return YMX_ERR_INVALID_ARG;
	}
	*out = nullptr;
	if(n_tris > 0 && (!verts || !idx || n_verts == 0))
	{
		ymx_set_error("%s: NULL geometry with %u primitives", fn, n_tris);
		return YMX_ERR_INVALID_ARG;
	}
	if(n_tris > YMX_LEAF_MAX_FIRST)
	{
		ymx_set_error("%s: %u primitives exceed the limit of %u", fn, n_tris, YMX_LEAF_MAX_FIRST);
		return YMX_ERR_UNSUPPORTED;
	}
	// 4 triangles per leaf: on the 2.19M-triangle heightfield +2 % incoherent
	// and +14 % coherent closest hit over 1 (LBVH; PLOC keeps 1)
	int max_leaf = 4, builder = YMX_BUILDER_LBVH, treelet_passes = YMX_TREELET_PASSES_DEFAULT, ray_reorder = 0;
	int cell_bits = YMX_REORDER_CELL_BITS_DEFAULT;
	if(params)
	{
		if(params->struct_size != sizeof(ymx_accel_params))
		{
			ymx_set_error("%s: params.struct_size %u != %zu", fn, params->struct_size, sizeof(ymx_accel_params));
			return YMX_ERR_INVALID_ARG;
		}
		if(params->max_leaf_size < 0 || params->max_leaf_size > 8)
		{
			ymx_set_error("%s: max_leaf_size %d outside 0..8", fn, params->max_leaf_size);
			return YMX_ERR_INVALID_ARG;
		}
		if(params->max_leaf_size > 0) max_leaf = params->max_leaf_size;
		if(params->builder != YMX_BUILDER_LBVH && params->builder != YMX_BUILDER_PLOC)
		{
			ymx_set_error("%s: unknown builder %d", fn, params->builder);
			return YMX_ERR_INVALID_ARG;
		}
		builder = params->builder;
		if(params->treelet_passes < -1 || params->treelet_passes > 8)
		{
			ymx_set_error("%s: treelet_passes %d outside -1..8", fn, params->treelet_passes);
			return YMX_ERR_INVALID_ARG;
		}
		if(params->treelet_passes != 0) treelet_passes = params->treelet_passes < 0 ? 0 : params->treelet_passes;
		if(params->ray_reorder < -1 || params->ray_reorder > 1)
		{
			ymx_set_error("%s: ray_reorder %d outside -1..1", fn, params->ray_reorder);
			return YMX_ERR_INVALID_ARG;
		}
		ray_reorder = params->ray_reorder;
		if(params->reorder_cell_bits != 0 && params->reorder_cell_bits != 3 && params->reorder_cell_bits != 4)
		{
			ymx_set_error("%s: reorder_cell_bits %d is not 0, 3 or 4", fn, params->reorder_cell_bits);
			return YMX_ERR_INVALID_ARG;
		}
		if(params->reorder_cell_bits != 0) cell_bits = params->reorder_cell_bits;
	}
	// a face's 4th index is YMX_NO_VERTEX for a triangle; quads get e3 slots
	std::vector<uint32_t> slot;
	uint32_t n_quads = 0;
	for(size_t i = 0; i < (size_t)n_tris; ++i)
	{
		const bool quad = stride == 4 && idx[4 * i + 3] != YMX_NO_VERTEX;
		for(uint32_t k = 0; k < (quad ? 4u : 3u); ++k)
		{
			if(idx[stride * i + k] >= n_verts)
			{
				ymx_set_error("%s: primitive %zu references vertex %u >= n_verts %u", fn, i, idx[stride * i + k],
				              n_verts);
				return YMX_ERR_INVALID_ARG;
			}
		}
		if(quad)
		{
			if(slot.empty()) slot.assign(n_tris, 0u);
			slot[i] = ++n_quads;
		}
	}
	ymx_accel *acc = new(std::nothrow) ymx_accel();
	if(!acc)
	{
		ymx_set_error("%s: host allocation failed", fn);
		return YMX_ERR_OUT_OF_MEMORY;
	}
	if(hipGetDevice(&acc->device) != hipSuccess)
	{
		ymx_set_error("%s: no HIP device", fn);
		delete acc;
		return YMX_ERR_HIP;
	}
	acc->n_quads = n_quads;
	acc->treelet_passes = (uint32_t)treelet_passes;
	acc->ray_reorder = ray_reorder;
	acc->reorder_cell_bits = cell_bits;
	int rc = ymx_build_device(acc, verts, n_verts, idx, stride, n_quads ? slot.data() : nullptr, vis, n_tris, max_leaf,
	                          builder, treelet_passes);
	if(rc == YMX_OK) rc = ymx_trace_setup(acc);
	if(rc != YMX_OK)
	{
		free_accel(acc);
		return rc;
	}
	*out = acc;
	return YMX_OK;
}

int ymx_accel_create(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, const uint8_t *vis,
                     uint32_t n_tris, const ymx_accel_params *params, ymx_accel **out)
{
	return create_accel("ymx_accel_create", verts, n_verts, tri_idx, 3, vis, n_tris, params, out);
}

int ymx_accel_create_faces(const float *verts, uint32_t n_verts, const uint32_t *face_idx, const uint8_t *vis,
                           uint32_t n_faces, const ymx_accel_params *params, ymx_accel **out)
{
	return create_accel("ymx_accel_create_faces", verts, n_verts, face_idx, 4, vis, n_faces, params, out);
}

int ymx_accel_destroy(ymx_accel *accel)
{
	if(!accel) return YMX_OK;
	(void)hipDeviceSynchronize();
	free_accel(accel);
	return YMX_OK;
}

int ymx_accel_bound(const ymx_accel *accel, float bound_min[3], float bound_max[3])
{
	if(!accel || !bound_min || !bound_max)
	{
		ymx_set_error("ymx_accel_bound: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	memcpy(bound_min, accel->bmin, 3 * sizeof(float));
	memcpy(bound_max, accel->bmax, 3 * sizeof(float));
	return YMX_OK;
}

int ymx_accel_export(const ymx_accel *accel, void *nodes, uint64_t node_bytes, void *tris, uint64_t tri_bytes)
{
	if(!accel || !nodes || !tris || node_bytes != 64ull * accel->n_nodes ||
	   tri_bytes != 48ull * (accel->n_tris ? accel->n_tris : 1))
	{
		ymx_set_error("ymx_accel_export: NULL buffer or byte counts other than ymx_accel_get_info's");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipMemcpy(nodes, accel->d_nodes, node_bytes, hipMemcpyDeviceToHost));
	YMX_HIP_CHECK(hipMemcpy(tris, accel->d_tris, tri_bytes, hipMemcpyDeviceToHost));
	return YMX_OK;
}

int ymx_accel_export_quads(const ymx_accel *accel, void *quads, uint64_t quad_bytes)
{
	if(!accel || quad_bytes != 16ull * accel->n_quads || (accel->n_quads && !quads))
	{
		ymx_set_error("ymx_accel_export_quads: NULL buffer or a byte count other than 16 x n_quads");
		return YMX_ERR_INVALID_ARG;
	}
	if(accel->n_quads) YMX_HIP_CHECK(hipMemcpy(quads, accel->d_quads, quad_bytes, hipMemcpyDeviceToHost));
	return YMX_OK;
}

int ymx_accel_get_info(const ymx_accel *accel, ymx_accel_info *info)
{
	if(!accel || !info || info->struct_size != sizeof(ymx_accel_info))
	{
		ymx_set_error("ymx_accel_get_info: NULL argument or struct_size mismatch");
		return YMX_ERR_INVALID_ARG;
	}
	info->n_tris = accel->n_tris;
	info->n_nodes = accel->n_nodes;
	info->max_depth = accel->max_depth;
	info->node_bytes = 64ull * accel->n_nodes;
	info->tri_bytes = 48ull * (accel->n_tris ? accel->n_tris : 1);
	info->build_ms = accel->build_ms;
	info->n_quads = accel->n_quads;
	info->treelet_passes = accel->treelet_passes;
	info->treelets_rewritten = accel->treelets_rewritten;
	memcpy(info->bound_min, accel->bmin, sizeof info->bound_min);
	memcpy(info->bound_max, accel->bmax, sizeof info->bound_max);
	return YMX_OK;
}

static int check_query(const ymx_accel *accel, const void *rays, const void *out, const char *fn)
{
	if(!accel)
	{
		ymx_set_error("%s: NULL accelerator", fn);
		return YMX_ERR_INVALID_ARG;
	}
	if(!rays || !out)
	{
		ymx_set_error("%s: NULL ray or result buffer", fn);
		return YMX_ERR_INVALID_ARG;
	}
	if(((uintptr_t)rays & 15u) || ((uintptr_t)out & 3u))
	{
		ymx_set_error("%s: ray buffer must be 16-byte aligned", fn);
		return YMX_ERR_INVALID_ARG;
	}
	return YMX_OK;
}

static int check_count(uint32_t n_rays, const char *fn)
{
	// the persistent kernels hand out rays in chunks of 64 from a 32-bit counter
	if(n_rays > 0x80000000u)
	{
		ymx_set_error("%s: %u rays exceed the per-call limit of 2^31", fn, n_rays);
		return YMX_ERR_INVALID_ARG;
	}
	return YMX_OK;
}

int ymx_accel_intersect(const ymx_accel *accel, const ymx_ray *d_rays, ymx_hit *d_hits, uint32_t n_rays,
                        void *stream)
{
	if(n_rays == 0) return YMX_OK;
	int rc = check_count(n_rays, "ymx_accel_intersect");
	if(rc == YMX_OK) rc = check_query(accel, d_rays, d_hits, "ymx_accel_intersect");
	if(rc != YMX_OK) return rc;
	if((uintptr_t)d_hits & 15u)
	{
		ymx_set_error("ymx_accel_intersect: hit buffer must be 16-byte aligned");
		return YMX_ERR_INVALID_ARG;
	}
	return ymx_launch_closest(accel, d_rays, d_hits, n_rays, (hipStream_t)stream, nullptr);
}

int ymx_accel_intersect_noalloc(const ymx_accel *accel, const ymx_ray *d_rays, ymx_hit *d_hits, uint32_t n_rays,
                                void *stream)
{
	if(n_rays == 0) return YMX_OK;
	int rc = check_count(n_rays, "ymx_accel_intersect_noalloc");
	if(rc == YMX_OK) rc = check_query(accel, d_rays, d_hits, "ymx_accel_intersect_noalloc");
	if(rc != YMX_OK) return rc;
	if((uintptr_t)d_hits & 15u)
	{
		ymx_set_error("ymx_accel_intersect_noalloc: hit buffer must be 16-byte aligned");
		return YMX_ERR_INVALID_ARG;
	}
	return ymx_launch_closest_noalloc(accel, d_rays, d_hits, n_rays, (hipStream_t)stream);
}

int ymx_accel_intersect_shadow(const ymx_accel *accel, const ymx_ray *d_rays, uint32_t *d_occluded,
                               uint32_t *d_prim, uint32_t n_rays, void *stream)
{
	if(n_rays == 0) return YMX_OK;
	int rc = check_count(n_rays, "ymx_accel_intersect_shadow");
	if(rc == YMX_OK) rc = check_query(accel, d_rays, d_occluded, "ymx_accel_intersect_shadow");
	if(rc != YMX_OK) return rc;
	return ymx_launch_shadow(accel, d_rays, d_occluded, d_prim, n_rays, (hipStream_t)stream, nullptr);
}

int ymx_accel_intersect_transparent(const ymx_accel *accel, const ymx_ray *d_rays, const uint32_t *d_transparent,
                                    uint32_t max_depth, uint32_t *d_occluded, uint32_t *d_counts, uint32_t *d_hits,
                                    uint32_t n_rays, void *stream)
{
	if(n_rays == 0) return YMX_OK;
	int rc = check_count(n_rays, "ymx_accel_intersect_transparent");
	if(rc == YMX_OK) rc = check_query(accel, d_rays, d_occluded, "ymx_accel_intersect_transparent");
	if(rc != YMX_OK) return rc;
	if(!d_transparent || !d_counts || (max_depth > 0 && !d_hits) || max_depth > 64)
	{
		ymx_set_error("ymx_accel_intersect_transparent: NULL buffer or max_depth %u above 64", max_depth);
		return YMX_ERR_INVALID_ARG;
	}
	return ymx_launch_transparent(accel, d_rays, d_transparent, max_depth, d_occluded, d_counts,
	                              reinterpret_cast<uint2 *>(d_hits), n_rays, nullptr, (hipStream_t)stream);
}

int ymx_accel_traversal_stats(const ymx_accel *accel, const ymx_ray *d_rays, uint32_t n_rays, int query,
                              uint64_t *node_fetches, uint64_t *tri_tests, void *stream)
{
	if(!node_fetches || !tri_tests || (query != 0 && query != 1))
	{
		ymx_set_error("ymx_accel_traversal_stats: bad arguments");
		return YMX_ERR_INVALID_ARG;
	}
	*node_fetches = 0;
	*tri_tests = 0;
	if(n_rays == 0) return YMX_OK;
	int rc = check_query(accel, d_rays, node_fetches, "ymx_accel_traversal_stats");
	if(rc != YMX_OK) return rc;
	hipStream_t s = (hipStream_t)stream;
	void *scratch_out = nullptr;
	const size_t out_bytes = query == 0 ? sizeof(ymx_hit) * (size_t)n_rays : sizeof(uint32_t) * (((size_t)n_rays + 31) / 32);
	// the counters live behind the results, in this call's own allocation
	const size_t stats_off = (out_bytes + 15u) & ~(size_t)15u;
	YMX_HIP_CHECK(hipMalloc(&scratch_out, stats_off + 2 * sizeof(unsigned long long)));
	unsigned long long *d_stats = reinterpret_cast<unsigned long long *>((char *)scratch_out + stats_off);
	unsigned long long host[2] = {0, 0};
	hipError_t e = hipMemsetAsync(d_stats, 0, sizeof host, s);
	if(e == hipSuccess)
	{
		rc = query == 0 ? ymx_launch_closest(accel, d_rays, (ymx_hit *)scratch_out, n_rays, s, d_stats)
		                : ymx_launch_shadow(accel, d_rays, (uint32_t *)scratch_out, nullptr, n_rays, s, d_stats);
		if(rc == YMX_OK) e = hipMemcpyAsync(host, d_stats, sizeof host, hipMemcpyDeviceToHost, s);
		if(rc == YMX_OK && e == hipSuccess) e = hipStreamSynchronize(s);
	}
	(void)hipFree(scratch_out);
	if(rc != YMX_OK) return rc;
	YMX_HIP_CHECK(e);
	*node_fetches = host[0];
	*tri_tests = host[1];
	return YMX_OK;
}

}  // extern "C"
