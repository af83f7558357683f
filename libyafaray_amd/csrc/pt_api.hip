// pt_api.hip -- host side of the path tracer: the extern "C" entry points of
// include/ymx_pt.h, scene constants, and the per-pass launch loop.
#include "pt_internal.h"

#include <float.h>
#include <stdio.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

using namespace ymx;

namespace {

// Faure's scrambling permutations (H. Faure, J. Number Theory 42, 1992):
// sigma_2 = (0 1); sigma_{2c}(i) = 2 sigma_c(i), sigma_{2c}(c + i) = 2 sigma_c(i) + 1;
// sigma_{2c+1}(c) = c, the rest is sigma_{2c} with values >= c moved up by one.
std::vector<std::vector<int>> faure_permutations(int max_base)
{
	std::vector<std::vector<int>> t(max_base + 1);
	t[1] = {0};
	t[2] = {0, 1};
	for(int b = 3; b <= max_base; ++b)
	{
		t[b].resize(b);
		if(b % 2 == 0)
		{
			const int c = b / 2;
			for(int i = 0; i < c; ++i)
			{
				t[b][i] = 2 * t[c][i];
				t[b][c + i] = 2 * t[c][i] + 1;
			}
		}
		else
		{
			const int c = (b - 1) / 2;
			for(int i = 0; i < b; ++i)
			{
				if(i == c)
				{
					t[b][i] = c;
					continue;
				}
				const int s = t[b - 1][i < c ? i : i - 1];
				t[b][i] = s >= c ? s + 1 : s;
			}
		}
	}
	return t;
}

// The 50 bases of Halton::lowDiscrepancySampling (src/sampler/halton.cc:401-403).
constexpr int kBases[kLdsDims] = {1, 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53,
                                  59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131,
                                  137, 139, 149, 151, 157, 163, 167, 173, 179, 181, 191, 193, 197, 199, 211, 223, 227};

void free_nodes(PtNodes &nd)
{
	void *ptrs[] = {nd.ray, nd.info, nd.dec, nd.col, nd.wt, nd.child};
	for(void *p : ptrs)
		if(p) (void)hipFree(p);
	nd.ray = nullptr;
	nd.info = nullptr;
	nd.dec = nullptr;
	nd.col = nullptr;
	nd.wt = nullptr;
	nd.child = nullptr;
	nd.cap = 0;
}

void free_guide_buffers(PtGuide &g)
{
	void *ptrs[] = {g.query, g.mix_node, g.mix_path, g.rec, g.base, g.mask, g.train_buf, g.batch_buf, g.cell_sum, g.train_count,
	                g.query_count};
	for(void *p : ptrs)
		if(p) (void)hipFree(p);
	memset(&g, 0, sizeof g);  // the selection table belongs to the guide
}

// Points the tracer's PtGuide at its guide's selection table (allocated on
// first use) for the learned selection and the fallback the tracer has
// switched on.  False when the table cannot be allocated.
bool attach_selection(ymx_pt *pt)
{
	PtGuide &g = pt->g;
	g.learn_alpha = 0u;
	g.fallback = 0u;
	g.alpha_theta = g.alpha_m = g.alpha_v = g.alpha_grad = g.fb_sum = g.fb_run = nullptr;
	if(!pt->guide || (!pt->alpha_learned && !pt->fallback_on)) return true;
	float *t = ymx_guide_selection_table(pt->guide, kSelectionTableFloats);
	if(!t) return false;
	const size_t cells = (size_t)kGuideCells * kGuideCells * kGuideCells;
	g.alpha_theta = t;
	g.alpha_m = t + cells;
	g.alpha_v = t + 2 * cells;
	g.alpha_grad = t + 3 * cells;
	g.fb_sum = t + 4 * cells;
	g.fb_run = t + 7 * cells;
	g.learn_alpha = pt->alpha_learned ? 1u : 0u;
	g.fallback = pt->fallback_on ? 1u : 0u;
	return true;
}

void free_pt(ymx_pt *pt)
{
	if(!pt) return;
	free_guide_buffers(pt->g);
	free_nodes(pt->nd);
	void *ptrs[] = {pt->d_prim, pt->d_materials, pt->d_lights, pt->d_ltris, pt->d_const, pt->d_faure, pt->d_inv_prims, pt->d_prims,
	                pt->d_film, pt->d_film_w, pt->st.path_col, pt->st.tp_next, pt->st.tp_vertex, pt->st.pwo,
	                pt->st.cam_hit, pt->st.dxy, pt->st.flags, pt->st.light, pt->st.vert_emit, pt->st.ext_rays[0],
	                pt->st.ext_rays[1], pt->st.ext_path[0], pt->st.ext_path[1], pt->st.ext_count, pt->st.ext_hits,
	                pt->st.slot_contrib, pt->st.slot_queue, pt->st.sh_rays, pt->st.sh_occ, pt->st.sh_count,
	                pt->st.sh_tcount, pt->st.sh_thits, pt->d_transp, pt->st.totals, pt->nd.spawn, pt->d_aa,
	                pt->d_cnorm};
	for(void *p : ptrs)
		if(p) (void)hipFree(p);
	if(pt->ev0) (void)hipEventDestroy(pt->ev0);
	if(pt->ev1) (void)hipEventDestroy(pt->ev1);
	for(ymx_accel *tree : pt->light_trees) (void)ymx_accel_destroy(tree);
	delete pt;
}

template <typename T> hipError_t dalloc(T *&p, size_t count) { return hipMalloc((void **)&p, sizeof(T) * (count ? count : 1)); }

// Node arrays for `cap` nodes, the first `keep` copied from `old`.
hipError_t alloc_nodes(PtNodes &nd, uint32_t cap, const PtNodes *old, uint32_t keep)
{
	hipError_t e;
	if((e = dalloc(nd.ray, cap)) != hipSuccess) return e;
	if((e = dalloc(nd.info, cap)) != hipSuccess) return e;
	if((e = dalloc(nd.dec, cap)) != hipSuccess) return e;
	if((e = dalloc(nd.col, cap)) != hipSuccess) return e;
	if((e = dalloc(nd.wt, cap)) != hipSuccess) return e;
	if((e = dalloc(nd.child, cap)) != hipSuccess) return e;
	nd.cap = cap;
	if(old && keep)
	{
		if((e = hipMemcpy(nd.ray, old->ray, sizeof(ymx_ray) * keep, hipMemcpyDeviceToDevice)) != hipSuccess) return e;
		if((e = hipMemcpy(nd.info, old->info, sizeof(uint4) * keep, hipMemcpyDeviceToDevice)) != hipSuccess) return e;
		if((e = hipMemcpy(nd.dec, old->dec, sizeof(float2) * keep, hipMemcpyDeviceToDevice)) != hipSuccess) return e;
		if((e = hipMemcpy(nd.col, old->col, sizeof(float4) * keep, hipMemcpyDeviceToDevice)) != hipSuccess) return e;
		if((e = hipMemcpy(nd.wt, old->wt, sizeof(float4) * keep, hipMemcpyDeviceToDevice)) != hipSuccess) return e;
		if((e = hipMemcpy(nd.child, old->child, sizeof(uint2) * keep, hipMemcpyDeviceToDevice)) != hipSuccess) return e;
	}
	return hipSuccess;
}

// Gauss filter table of ImageFilm (src/render/imagefilm.cc:231-264, include/math/filter.h:32-37).
void filter_table(float pixelwidth, PtConst &c)
{
	float fw = pixelwidth * 0.5f;
	fw *= 2.f;
	if(fw < 0.501f) fw = 0.501f;
	if(fw > 0.5f * 8.f) fw = 0.5f * 8.f;
	const float scale = 1.f / 16.f;
	int k = 0;
	for(int y = 0; y < 16; ++y)
		for(int x = 0; x < 16; ++x)
		{
			const float dx = ((float)x + .5f) * scale, dy = ((float)y + .5f) * scale;
			const float r_2 = dx * dx + dy * dy;
			const float v = fast_exp((float)-6 * r_2) - 0.00247875f;
			c.filter_table[k++] = v > 0.f ? v : 0.f;
		}
	c.filter_table_scale = 0.9999f * 16.f / fw;
	c.filter_width = fw;
}

// Rgb::clampProportionalRgb (include/color/color.h:399-431): the largest
// component clamped to max_value and the others scaled by the same factor;
// max_value <= 0 leaves the colour as it is.
Rgb clamp_proportional(Rgb c, float max_value)
{
	if(max_value > 0.f)
	{
		const float max_rgb = std::max(c.r, std::max(c.g, c.b));
		const float adj = max_value / max_rgb;
		if(max_rgb > max_value)
		{
			if(c.r >= max_rgb)
			{
				c.r = max_value;
				c.g *= adj;
				c.b *= adj;
			}
			else if(c.g >= max_rgb)
			{
				c.g = max_value;
				c.r *= adj;
				c.b *= adj;
			}
			else
			{
				c.b = max_value;
				c.r *= adj;
				c.g *= adj;
			}
		}
	}
	return c;
}

// ObjectLight::initIs (src/light/light_object_light.cc:87-104) for the
// triangles idx[0 .. 3 n): each triangle's area 0.5 |e1 x e2|
// (include/geometry/shape/shape_polygon.h:64-76), the Pdf1D cdf of the areas
// (include/sampler/sample_pdf1d.h:60-81: a double running sum of area / n,
// rounded to float, divided by the float total) and the tree bound of the
// light's own kd-tree (src/accelerator/accelerator_kdtree_original.cc:89-103:
// the vertex bound grown by 0.001 of its extent on each axis).  Appends the
// triangles to `out`; returns area_, the double sum of the float areas.
float object_light_tables(const float *verts, const uint32_t *idx, uint32_t n, const uint8_t *vis,
                          const float4 *face_normals, std::vector<PtLightTri> &out, float bound_a[3], float bound_g[3])
{
	std::vector<float> areas(n);
	double total = 0.0;
	float a[3] = {0.f, 0.f, 0.f}, g[3] = {0.f, 0.f, 0.f};
	for(uint32_t i = 0; i < n; ++i)
	{
		const V3 v[3] = {load3(verts + 3 * (size_t)idx[3 * (size_t)i]), load3(verts + 3 * (size_t)idx[3 * (size_t)i + 1]),
		                 load3(verts + 3 * (size_t)idx[3 * (size_t)i + 2])};
		const V3 cr = cross(sub(v[1], v[0]), sub(v[2], v[0]));
		areas[i] = 0.5f * sqrtf(dot(cr, cr));
		total += (double)areas[i];
		for(int k = 0; k < 3; ++k)
		{
			const float p[3] = {v[k].x, v[k].y, v[k].z};
			for(int ax = 0; ax < 3; ++ax)
			{
				if(i == 0 && k == 0) a[ax] = g[ax] = p[ax];
				else
				{
					if(p[ax] < a[ax]) a[ax] = p[ax];
					if(p[ax] > g[ax]) g[ax] = p[ax];
				}
			}
		}
		PtLightTri t;
		const uint32_t vm = vis ? (uint32_t)vis[i] : (uint32_t)(YMX_VIS_VISIBLE | YMX_VIS_CASTS_SHADOWS);
		float vf;
		memcpy(&vf, &vm, 4);
		t.v0_cdf = make_float4(v[0].x, v[0].y, v[0].z, 0.f);
		t.v1_vis = make_float4(v[1].x, v[1].y, v[1].z, vf);
		t.v2 = make_float4(v[2].x, v[2].y, v[2].z, 0.f);
		t.n = make_float4(face_normals[i].x, face_normals[i].y, face_normals[i].z, 0.f);
		out.push_back(t);
	}
	const size_t first = out.size() - n;
	const double delta = 1.0 / (double)n;
	double acc = 0.0;
	std::vector<float> cdf(n);
	for(uint32_t i = 0; i < n; ++i)
	{
		acc += (double)areas[i] * delta;
		cdf[i] = (float)acc;
	}
	const float integral = (float)acc;
	for(uint32_t i = 0; i < n; ++i) out[first + i].v0_cdf.w = cdf[i] / integral;
	for(int ax = 0; ax < 3; ++ax)
	{
		const double offset = (double)(g[ax] - a[ax]) * 0.001;
		bound_a[ax] = a[ax] - (float)offset;
		bound_g[ax] = g[ax] + (float)offset;
	}
	return (float)total;
}

}  // namespace

int ymx_pt_grow_nodes(ymx_pt *pt, uint64_t need, uint32_t keep)
{
	if(need > (1ull << 31))
	{
		ymx_set_error("ymx_pt_render: the recursive raytrace needs %llu ray nodes (at most 2^31)",
		              (unsigned long long)need);
		return YMX_ERR_UNSUPPORTED;
	}
	uint64_t cap = (uint64_t)pt->nd.cap * 2;
	if(cap < need) cap = need;
	if(cap > (1ull << 31)) cap = 1ull << 31;
	YMX_HIP_CHECK(hipDeviceSynchronize());  // the old arrays may still be in use
	PtNodes nd = pt->nd;
	nd.ray = nullptr;
	nd.info = nullptr;
	nd.dec = nullptr;
	nd.col = nullptr;
	nd.wt = nullptr;
	nd.child = nullptr;
	const hipError_t e = alloc_nodes(nd, (uint32_t)cap, &pt->nd, keep);
	if(e != hipSuccess)
	{
		free_nodes(nd);
		ymx_set_error("ymx_pt_render: growing the ray-node pool to %llu nodes: %s", (unsigned long long)cap,
		              hipGetErrorString(e));
		return e == hipErrorOutOfMemory ? YMX_ERR_OUT_OF_MEMORY : YMX_ERR_HIP;
	}
	free_nodes(pt->nd);
	nd.spawn = pt->nd.spawn;
	pt->nd = nd;
	return YMX_OK;
}

extern "C" {

int ymx_pt_create(const ymx_accel *accel, const float *verts, uint32_t n_verts, const uint32_t *tri_idx,
                  const uint32_t *mat_id, const uint8_t *vis, uint32_t n_tris, const ymx_material *materials,
                  uint32_t n_materials,
                  const ymx_light *lights, uint32_t n_lights, const ymx_pt_settings *st, ymx_pt **out)
{
	if(!out || !accel || !st || (n_tris && (!verts || !tri_idx || !mat_id)) || (n_materials && !materials) ||
	   (n_lights && !lights))
	{
		ymx_set_error("ymx_pt_create: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	*out = nullptr;
	ymx_accel_info info;
	info.struct_size = sizeof info;
	if(ymx_accel_get_info(accel, &info) != YMX_OK || info.n_tris != n_tris)
	{
		ymx_set_error("ymx_pt_create: accelerator holds %u triangles, %u given", info.n_tris, n_tris);
		return YMX_ERR_INVALID_ARG;
	}
	if(st->resx <= 0 || st->resy <= 0 || (int64_t)st->resx * st->resy > (1 << 26) || st->bounces < 1 ||
	   st->bounces > 12 || st->aa_samples < 1 || st->path_samples < 0 || n_lights > (uint32_t)kMaxLights ||
	   st->raydepth < 0 || st->raydepth > kMaxRayDepth ||
	   (st->transparent_shadows && (st->shadow_depth < 0 || st->shadow_depth > 64)))
	{
		ymx_set_error("ymx_pt_create: unsupported settings (res %dx%d, bounces %d, AA %d, lights %u, raydepth %d)",
		              st->resx, st->resy, st->bounces, st->aa_samples, n_lights, st->raydepth);
		return YMX_ERR_UNSUPPORTED;
	}
	for(uint32_t i = 0; i < 3 * (size_t)n_tris; ++i)
		if(tri_idx[i] >= n_verts)
		{
			ymx_set_error("ymx_pt_create: vertex index out of range");
			return YMX_ERR_INVALID_ARG;
		}
	for(uint32_t i = 0; i < n_tris; ++i)
		if(mat_id[i] >= n_materials)
		{
			ymx_set_error("ymx_pt_create: triangle %u uses material %u of %u", i, mat_id[i], n_materials);
			return YMX_ERR_INVALID_ARG;
		}
	for(uint32_t m = 0; m < n_materials; ++m)
		if(materials[m].type > YMX_MAT_LIGHT)
		{
			ymx_set_error("ymx_pt_create: material %u has unsupported type %u", m, materials[m].type);
			return YMX_ERR_UNSUPPORTED;
		}
	for(uint32_t l = 0; l < n_lights; ++l)
	{
		const ymx_light &s = lights[l];
		if(s.type > YMX_LIGHT_PORTAL || (s.type != YMX_LIGHT_POINT && (s.samples < 1 || s.samples > 64)))
		{
			ymx_set_error("ymx_pt_create: light %u unsupported (type %u, samples %d)", l, s.type, s.samples);
			return YMX_ERR_UNSUPPORTED;
		}
		const bool mesh = s.type == YMX_LIGHT_OBJECT || s.type == YMX_LIGHT_PORTAL;
		if(mesh && (s.n_tris == 0 || s.first_tri > n_tris || s.n_tris > n_tris - s.first_tri))
		{
			ymx_set_error("ymx_pt_create: %s light %u covers triangles [%u, %u + %u) of %u",
			              s.type == YMX_LIGHT_PORTAL ? "portal" : "object", l, s.first_tri, s.first_tri, s.n_tris,
			              n_tris);
			return YMX_ERR_INVALID_ARG;
		}
		// BackgroundPortalLight::init takes the scene's background
		// (src/light/light_background_portal.cc:96-99) and evaluates it in every
		// sample: a scene without one has nothing to send through the portal
		if(s.type == YMX_LIGHT_PORTAL && !st->has_background)
		{
			ymx_set_error("ymx_pt_create: portal light %u needs a background", l);
			return YMX_ERR_INVALID_ARG;
		}
	}

	const int32_t aa_passes = st->aa_passes > 1 ? st->aa_passes : 1;
	if(aa_passes > 1 && (aa_passes > 1000 || st->aa_inc_samples < 1 || !(st->aa_sample_multiplier_factor > 0.f) ||
	                     !(st->aa_light_sample_multiplier_factor > 0.f) || st->aa_dark_detection_type < 0 ||
	                     st->aa_dark_detection_type > 2 || st->aa_variance_edge_size < 0 ||
	                     st->aa_variance_edge_size > 256))
	{
		ymx_set_error("ymx_pt_create: unsupported AA pass settings (passes %d, inc samples %d, multipliers %g %g)",
		              st->aa_passes, st->aa_inc_samples, st->aa_sample_multiplier_factor,
		              st->aa_light_sample_multiplier_factor);
		return YMX_ERR_UNSUPPORTED;
	}
	// the pass sample counts and light multipliers of TiledIntegrator::render,
	// the most samples a pixel can receive (integrator_tiled.cc:258-262) and
	// the largest light multiplier (it sizes the shadow slots)
	float max_light_mult = 1.f;
	int32_t max_pass_samples = st->aa_samples;
	int32_t aa_total = st->aa_samples;
	{
		float sm = 1.f, lm = 1.f;
		for(int32_t i = 1; i < aa_passes; ++i)
		{
			sm *= st->aa_sample_multiplier_factor;
			lm *= st->aa_light_sample_multiplier_factor;
			const int32_t n = (int32_t)ceilf((float)st->aa_inc_samples * sm);
			max_pass_samples = n > max_pass_samples ? n : max_pass_samples;
			max_light_mult = lm > max_light_mult ? lm : max_light_mult;
			aa_total = (int32_t)((float)aa_total + ceilf((float)((double)st->aa_inc_samples *
			                                                     pow((double)st->aa_sample_multiplier_factor, i))));
		}
	}
	for(uint32_t l = 0; l < n_lights; ++l)
		if(lights[l].type != YMX_LIGHT_POINT && ceilf((float)lights[l].samples * max_light_mult) > 4096.f)
		{
			ymx_set_error("ymx_pt_create: light %u takes more than 4096 samples in an AA pass", l);
			return YMX_ERR_UNSUPPORTED;
		}

	ymx_pt *pt = new(std::nothrow) ymx_pt();
	if(!pt)
	{
		ymx_set_error("ymx_pt_create: host allocation failed");
		return YMX_ERR_OUT_OF_MEMORY;
	}
	memset(pt, 0, sizeof *pt);
	pt->settings = *st;
	pt->settings.aa_passes = aa_passes;
	pt->max_light_mult = max_light_mult;
	pt->guide_depth = -1;
	pt->accel = accel;
	pt->n_tris = n_tris;
	pt->n_materials = n_materials;
	(void)hipGetDevice(&pt->device);

#define PT_TRY(expr)                      \
	do                                    \
	{                                     \
		hipError_t e_ = (expr);           \
		if(e_ != hipSuccess)              \
		{                                 \
			ymx_set_error("%s: %s", #expr, hipGetErrorString(e_)); \
			free_pt(pt);                  \
			return e_ == hipErrorOutOfMemory ? YMX_ERR_OUT_OF_MEMORY : YMX_ERR_HIP; \
		}                                 \
	} while(0)

	// per-primitive face normal ((v1 - v0) ^ (v2 - v0)).normalize() and material
	// (include/geometry/shape/shape_polygon.h:58-62)
	std::vector<float4> prim(n_tris ? n_tris : 1);
	for(uint32_t i = 0; i < n_tris; ++i)
	{
		const V3 p0 = load3(verts + 3 * (size_t)tri_idx[3 * (size_t)i]);
		const V3 p1 = load3(verts + 3 * (size_t)tri_idx[3 * (size_t)i + 1]);
		const V3 p2 = load3(verts + 3 * (size_t)tri_idx[3 * (size_t)i + 2]);
		const V3 ng = normalize(cross(sub(p1, p0), sub(p2, p0)));
		uint32_t m = mat_id[i];
		float fm;
		memcpy(&fm, &m, 4);
		prim[i] = make_float4(ng.x, ng.y, ng.z, fm);
	}
	PT_TRY(dalloc(pt->d_prim, prim.size()));
	PT_TRY(hipMemcpy(pt->d_prim, prim.data(), sizeof(float4) * prim.size(), hipMemcpyHostToDevice));

	// materials: the values each reference material derives at construction
	std::vector<PtMaterial> mats(n_materials ? n_materials : 1);
	bool any_glossy_branches = false, any_specular = false;
	for(uint32_t m = 0; m < n_materials; ++m)
	{
		const ymx_material &s = materials[m];
		PtMaterial &d = mats[m];
		memset(&d, 0, sizeof d);
		d.type = s.type;
		d.receive_shadows = s.receive_shadows;
		d.color = rgb(s.color[0], s.color[1], s.color[2]);
		d.diffuse_reflect = s.diffuse_reflect;
		switch(s.type)
		{
			case YMX_MAT_GLOSSY:  // src/material/material_glossy.cc:130-150
				if(s.diffuse_reflect > 0.f) d.flags = kDiffuse | kReflect;
				d.flags |= s.as_diffuse ? (kDiffuse | kReflect) : (kGlossy | kReflect);
				d.as_diffuse = s.as_diffuse ? 1u : 0u;
				d.glossy_reflect = s.glossy_reflect;
				d.diffuse_color = rgb(s.diffuse_color[0], s.diffuse_color[1], s.diffuse_color[2]);
				d.exponent = s.exponent;
				d.p_diffuse = std_min(0.6f, 1.f - (s.glossy_reflect / (s.glossy_reflect + (1.f - s.glossy_reflect) *
				                                                                             s.diffuse_reflect)));
				break;
			case YMX_MAT_MIRROR:  // src/material/material_mirror.cc:64-68, include/material/material_mirror.h:64
				d.flags = kSpecular;
				d.ref_col = cscale(d.color, std_min(s.reflect, 1.f));
				break;
			case YMX_MAT_GLASS:  // src/material/material_glass.cc:157-166, include/material/material_glass.h:92
			{
				d.flags = kSpecular | kReflect | kTransmit;
				if(s.fake_shadows) d.flags |= kFilter;
				d.transmit_flags = s.fake_shadows ? (kFilter | kTransmit) : (kSpecular | kTransmit);
				d.ior = s.ior;
				const float tf = s.transmit_filter;
				d.filter_color = rgb(tf * s.filter_color[0] + (1.f - tf), tf * s.filter_color[1] + (1.f - tf),
				                     tf * s.filter_color[2] + (1.f - tf));
				d.mirror_color = rgb(s.mirror_color[0], s.mirror_color[1], s.mirror_color[2]);
				break;
			}
			case YMX_MAT_LIGHT:  // src/material/material_light.cc:69-73, include/material/material_light.h:69
				d.flags = kEmit;
				d.emit_color = cscale(d.color, s.power);
				d.double_sided = s.double_sided ? 1u : 0u;
				break;
			default:  // shinydiffuse, src/material/material_shiny_diffuse.cc:131-183
				d.emit_color = cscale(d.color, s.emit);
				if(s.diffuse_reflect > 0.00001f) d.flags |= kDiffuse | kReflect;
				if(s.emit > 0.f) d.flags |= kEmit;
		}
		if(d.flags & kGlossy) any_glossy_branches = true;
		if(d.flags & (kSpecular | kFilter)) any_specular = true;
	}
	pt->recursive = (any_glossy_branches || any_specular) && st->raydepth >= 1;
	pt->max_fan = any_glossy_branches ? kGlossyBranches : 2u;
	PT_TRY(dalloc(pt->d_materials, mats.size()));
	PT_TRY(hipMemcpy(pt->d_materials, mats.data(), sizeof(PtMaterial) * mats.size(), hipMemcpyHostToDevice));

	// lights (src/light/light_area.cc, include/light/light_area.h:70-80, include/light/light_point.h:64,
	// src/light/light_object_light.cc:87-104)
	std::vector<PtLight> lts(n_lights ? n_lights : 1);
	std::vector<PtLightTri> ltris;
	uint32_t slots_camera = 0, slots_path = 0;
	for(uint32_t l = 0; l < n_lights; ++l)
	{
		const ymx_light &s = lights[l];
		PtLight &d = lts[l];
		memset(&d, 0, sizeof d);
		d.type = s.type;
		d.cast_shadows = s.cast_shadows;
		d.samples = s.samples;
		const Rgb cp = cscale(rgb(s.color[0], s.color[1], s.color[2]), s.power);
		if(s.type == YMX_LIGHT_AREA)
		{
			const V3 c0 = load3(s.corner), p1 = load3(s.point1), p2 = load3(s.point2);
			d.color = cscale(cp, kPi);
			d.corner = c0;
			d.to_x = sub(p1, c0);
			d.to_y = sub(p2, c0);
			d.normal_flipped = cross(sub(p2, c0), sub(p1, c0));
			const V3 nn = normalize(d.normal_flipped);
			d.area = sqrtf(dot(nn, nn));
			d.quad[0] = c0;
			d.quad[1] = p1;
			d.quad[2] = sub(add(p1, p2), c0);
			d.quad[3] = p2;
		}
		else if(s.type == YMX_LIGHT_OBJECT || s.type == YMX_LIGHT_PORTAL)
		{
			if(s.type == YMX_LIGHT_OBJECT)
			{
				d.color = cscale(cp, kPi);
				d.double_sided = s.double_sided ? 1u : 0u;
			}
			else  // ConstantBackground::eval (src/background/background_constant.cc:63-66) * power
				d.color = cscale(rgb(st->background[0], st->background[1], st->background[2]), s.power);
			d.first = (uint32_t)ltris.size();
			d.count = s.n_tris;
			d.area = object_light_tables(verts, tri_idx + 3 * (size_t)s.first_tri, s.n_tris,
			                             vis ? vis + s.first_tri : nullptr, prim.data() + s.first_tri, ltris,
			                             d.bound_a, d.bound_g);
		}
		else
		{
			d.color = cp;
			d.from = load3(s.from);
		}
		// BackgroundPortalLight::intersect clamps its colour
		// (src/light/light_background_portal.cc:219-220); the other lights return theirs
		d.hit_color = s.type == YMX_LIGHT_PORTAL ? clamp_proportional(d.color, s.ibl_clamp_sampling) : d.color;
		const uint32_t k = s.type == YMX_LIGHT_POINT ? 1u : 2u * (uint32_t)ceilf((float)s.samples * max_light_mult);
		slots_camera += k;
		slots_path = k > slots_path ? k : slots_path;
	}
	// the lights' own BVHs (ObjectLight::initIs, BackgroundPortalLight::initIs:
	// src/light/light_object_light.cc:95-98, src/light/light_background_portal.cc:90-93)
	for(uint32_t l = 0; l < n_lights; ++l)
	{
		const ymx_light &s = lights[l];
		if((s.type != YMX_LIGHT_OBJECT && s.type != YMX_LIGHT_PORTAL) || s.n_tris <= kLightTreeMin) continue;
		ymx_accel_params ap;
		memset(&ap, 0, sizeof ap);
		ap.struct_size = sizeof ap;
		// the light's triangles on their own vertices, so the build reads no
		// more than the light's mesh
		std::vector<float> lv(9 * (size_t)s.n_tris);
		std::vector<uint32_t> li(3 * (size_t)s.n_tris);
		for(size_t k = 0; k < 3 * (size_t)s.n_tris; ++k)
		{
			memcpy(&lv[3 * k], verts + 3 * (size_t)tri_idx[3 * (size_t)s.first_tri + k], 3 * sizeof(float));
			li[k] = (uint32_t)k;
		}
		ymx_accel *tree = nullptr;
		const int rc = ymx_accel_create(lv.data(), 3u * s.n_tris, li.data(), vis ? vis + s.first_tri : nullptr,
		                                s.n_tris, &ap, &tree);
		if(rc != YMX_OK)
		{
			free_pt(pt);
			return rc;
		}
		pt->light_trees.push_back(tree);
		if(tree->max_depth < kLightStack)
		{
			lts[l].tree_nodes = tree->d_nodes;
			lts[l].tree_tris = tree->d_tris;
		}
	}
	PT_TRY(dalloc(pt->d_lights, lts.size()));
	PT_TRY(hipMemcpy(pt->d_lights, lts.data(), sizeof(PtLight) * lts.size(), hipMemcpyHostToDevice));
	if(ltris.empty()) ltris.resize(1);
	PT_TRY(dalloc(pt->d_ltris, ltris.size()));
	PT_TRY(hipMemcpy(pt->d_ltris, ltris.data(), sizeof(PtLightTri) * ltris.size(), hipMemcpyHostToDevice));

	// constants: camera (src/camera/camera.cc:89-104, camera_perspective.cc:97-111)
	PtConst &c = pt->host_const;
	memset(&c, 0, sizeof c);
	{
		const V3 from = load3(st->from);
		V3 cy = sub(load3(st->up), from);
		V3 cz = sub(load3(st->to), from);
		V3 cx = cross(cz, cy);
		cy = cross(cz, cx);
		cx = normalize(cx);
		cy = normalize(cy);
		cz = normalize(cz);
		const float aspect = st->aspect_factor * (float)st->resy / (float)st->resx;
		c.vright = cx;
		c.vup = mul(aspect, cy);
		c.vto = sub(mul(st->focal, cz), mul(0.5f, add(c.vup, c.vright)));
		c.vup = v3(c.vup.x / (float)st->resy, c.vup.y / (float)st->resy, c.vup.z / (float)st->resy);
		c.vright = v3(c.vright.x / (float)st->resx, c.vright.y / (float)st->resx, c.vright.z / (float)st->resx);
		c.cam_from = from;
	}
	c.resx = st->resx;
	c.resy = st->resy;
	c.path_samples = st->path_samples;
	c.bounces = st->bounces;
	c.rr_min_bounces = st->rr_min_bounces;
	c.aa_samples = st->aa_samples;
	c.aa_passes = aa_passes;
	c.aa_total = (uint32_t)aa_total;
	c.aa_clamp = st->aa_clamp_samples;
	c.raydepth = st->raydepth;
	c.transparent_shadows = st->transparent_shadows ? 1u : 0u;
	c.shadow_depth = st->transparent_shadows ? (uint32_t)st->shadow_depth : 0u;
	c.n_lights = n_lights;
	c.slots_camera = slots_camera;
	c.slots_path = slots_path;
	c.has_background = st->has_background ? 1u : 0u;
	c.background = rgb(st->background[0], st->background[1], st->background[2]);
	c.seed = st->seed;
	filter_table(st->aa_pixelwidth, c);
	PT_TRY(dalloc(pt->d_const, 1));
	PT_TRY(hipMemcpy(pt->d_const, &c, sizeof c, hipMemcpyHostToDevice));

	// low-discrepancy tables (src/sampler/halton.cc:25-440)
	{
		const auto perms = faure_permutations(kLdsMaxBase);
		std::vector<int32_t> faure((size_t)kLdsDims * kLdsMaxBase, 0);
		std::vector<double> inv(kLdsDims);
		std::vector<uint32_t> bases(kLdsDims);
		for(int d = 0; d < kLdsDims; ++d)
		{
			const int b = kBases[d];
			const int table = d <= 2 ? 3 : b;  // the first three dimensions share base 3's table
			for(int i = 0; i < (int)perms[table].size() && i < kLdsMaxBase; ++i) faure[(size_t)d * kLdsMaxBase + i] = perms[table][i];
			inv[d] = floor(1e9 / (double)b + 0.5) / 1e9;  // 1/p rounded to nine decimals
			bases[d] = (uint32_t)b;
		}
		PT_TRY(dalloc(pt->d_faure, faure.size()));
		PT_TRY(hipMemcpy(pt->d_faure, faure.data(), sizeof(int32_t) * faure.size(), hipMemcpyHostToDevice));
		PT_TRY(dalloc(pt->d_inv_prims, inv.size()));
		PT_TRY(hipMemcpy(pt->d_inv_prims, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice));
		PT_TRY(dalloc(pt->d_prims, bases.size()));
		PT_TRY(hipMemcpy(pt->d_prims, bases.data(), sizeof(uint32_t) * bases.size(), hipMemcpyHostToDevice));
	}

	// wavefront state: `batch` AA samples of every pixel per wavefront, as
	// many as the memory budget holds (at least one), at most 2^25 paths.
	// Larger wavefronts keep the deep bounces' ray queues full: on the
	// 1920 x 1080 terrain frame (65 lights, 72 slots per path) 7 samples per
	// wavefront instead of 1 took the frame from 4.51 to 2.50 s.
	const uint32_t npix = (uint32_t)(st->resx * st->resy);
	PtState &s = pt->st;
	s.slots = slots_camera > slots_path ? slots_camera : slots_path;
	if(s.slots == 0) s.slots = 1;
	uint64_t budget = (uint64_t)(st->wavefront_mb > 0 ? st->wavefront_mb : 0) << 20;
	if(budget == 0)
	{
		size_t free_b = 0, total_b = 0;
		if(hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
		budget = free_b / 4 < (64ull << 30) ? free_b / 4 : (64ull << 30);
	}
	// bytes per path: path state (112), ray nodes (96, twice with recursion),
	// extension queues (88), and per shadow slot its contribution, queue
	// index, queue ray and occlusion bit (53), plus the caster lists of
	// transparent shadows
	const uint64_t per_slot = 53u + (c.transparent_shadows ? 4u + 8u * (uint64_t)(c.shadow_depth ? c.shadow_depth : 1) : 0u);
	const uint64_t per_path = 112u + (pt->recursive ? 192u : 96u) + 88u + per_slot * s.slots;
	uint64_t max_paths = budget / per_path;
	if(max_paths > (1ull << 25)) max_paths = 1ull << 25;
	if(max_paths * s.slots > 0xF0000000ull) max_paths = 0xF0000000ull / s.slots;
	uint64_t batch = max_paths / npix;
	if(batch < 1) batch = 1;
	if(batch > (uint64_t)max_pass_samples) batch = (uint64_t)max_pass_samples;
	pt->batch = (uint32_t)batch;
	pt->batch_max = (uint32_t)batch;
	// kernel choices (ymx_pt_autotune re-measures them per scene): ray
	// replacement on the shadow queues was 2-5 % faster per frame on the
	// Cornell and terrain frames; the extension queues' lanes finish together
	// and keep the wave-ordered kernel.  Transparent shadows have only that one.
	pt->shadow_replacement = !c.transparent_shadows;
	pt->ext_replacement = false;
	const uint32_t n = (uint32_t)(npix * batch);
	s.capacity = n;
	const size_t slot_cap = (size_t)n * s.slots;
	if(slot_cap > 0xFFFFFFFFull)
	{
		ymx_set_error("ymx_pt_create: %zu shadow slots per wavefront exceed 2^32", slot_cap);
		free_pt(pt);
		return YMX_ERR_UNSUPPORTED;
	}
	PT_TRY(dalloc(pt->d_film, npix));
	PT_TRY(dalloc(pt->d_film_w, npix));
	PT_TRY(hipMemset(pt->d_film, 0, sizeof(float4) * npix));
	PT_TRY(hipMemset(pt->d_film_w, 0, sizeof(float) * npix));
	if(aa_passes > 1)
	{
		// flags, pair bits, pixel list, slots; per-block counts and the count word
		PT_TRY(dalloc(pt->d_aa, 4 * (size_t)npix + (npix + 255) / 256 + 1));
	}
	PT_TRY(dalloc(s.path_col, n));
	PT_TRY(dalloc(s.tp_next, n));
	PT_TRY(dalloc(s.tp_vertex, n));
	PT_TRY(dalloc(s.pwo, n));
	PT_TRY(dalloc(s.cam_hit, n));
	PT_TRY(dalloc(s.dxy, n));
	PT_TRY(dalloc(s.flags, n));
	PT_TRY(dalloc(s.light, n));
	PT_TRY(dalloc(s.vert_emit, n));
	// node pool: the camera samples, plus room for their children when materials recurse
	PT_TRY(alloc_nodes(pt->nd, pt->recursive ? 2u * n : n, nullptr, 0));
	PT_TRY(dalloc(pt->nd.spawn, 1));
	PT_TRY(hipMemset(pt->nd.spawn, 0, sizeof(uint32_t)));
	PT_TRY(dalloc(s.ext_rays[0], n));
	PT_TRY(dalloc(s.ext_rays[1], n));
	PT_TRY(dalloc(s.ext_path[0], n));
	PT_TRY(dalloc(s.ext_path[1], n));
	PT_TRY(dalloc(s.ext_count, 2));
	PT_TRY(hipMemset(s.ext_count, 0, 2 * sizeof(uint32_t)));
	PT_TRY(dalloc(s.ext_hits, n));
	PT_TRY(dalloc(s.slot_contrib, slot_cap));
	PT_TRY(dalloc(s.slot_queue, slot_cap));
	// the segmented shadow queue: a segment takes the appends of every
	// kQueueSegments-th wave of the shading grid (at most capacity paths, 64
	// per wave), each appending at most `slots` rays per path
	const uint64_t waves = ((uint64_t)n + 255u) / 256u * 4u;
	const uint64_t seg_cap = (waves + kQueueSegments - 1u) / kQueueSegments * 64u * s.slots;
	const uint64_t sh_cap = seg_cap * kQueueSegments;
	if(sh_cap > 0xFFFFFFC0ull)
	{
		ymx_set_error("ymx_pt_create: a %llu-ray shadow queue exceeds 2^32", (unsigned long long)sh_cap);
		free_pt(pt);
		return YMX_ERR_UNSUPPORTED;
	}
	s.sh_seg_cap = (uint32_t)seg_cap;
	PT_TRY(dalloc(s.sh_rays, sh_cap));
	PT_TRY(dalloc(s.sh_occ, (sh_cap + 31) / 32 + 2));
	PT_TRY(dalloc(s.sh_count, (size_t)kQueueSegments * kQueueStride));
	PT_TRY(hipMemset(s.sh_count, 0, sizeof(uint32_t) * kQueueSegments * kQueueStride));
	if(c.transparent_shadows)
	{
		// Material::isTransparent: glass with fake shadows (include/material/material_glass.h:82)
		std::vector<uint32_t> bits(((size_t)n_tris + 31) / 32 + 1, 0u);
		for(uint32_t i = 0; i < n_tris; ++i)
		{
			const ymx_material &m = materials[mat_id[i]];
			if(m.type == YMX_MAT_GLASS && m.fake_shadows) bits[i >> 5] |= 1u << (i & 31u);
		}
		PT_TRY(dalloc(pt->d_transp, bits.size()));
		PT_TRY(hipMemcpy(pt->d_transp, bits.data(), sizeof(uint32_t) * bits.size(), hipMemcpyHostToDevice));
		PT_TRY(dalloc(s.sh_tcount, sh_cap));
		PT_TRY(dalloc(s.sh_thits, sh_cap * (size_t)(c.shadow_depth ? c.shadow_depth : 1)));
	}
	PT_TRY(dalloc(s.totals, 3));
	PT_TRY(hipMemset(s.totals, 0, 3 * sizeof(unsigned long long)));
	PT_TRY(hipEventCreate(&pt->ev0));
	PT_TRY(hipEventCreate(&pt->ev1));
#undef PT_TRY
	*out = pt;
	return YMX_OK;
}

int ymx_pt_set_guide_depth(ymx_pt *pt, int32_t max_depth)
{
	if(!pt)
	{
		ymx_set_error("ymx_pt_set_guide_depth: NULL handle");
		return YMX_ERR_INVALID_ARG;
	}
	pt->guide_depth = max_depth < 0 ? -1 : max_depth;
	pt->g.max_depth = max_depth < 0 ? INT32_MAX : max_depth;
	return YMX_OK;
}

int ymx_pt_set_guide(ymx_pt *pt, ymx_guide *guide, float bsdf_prob, int32_t train, uint32_t train_steps)
{
	if(!pt)
	{
		ymx_set_error("ymx_pt_set_guide: NULL handle");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	free_guide_buffers(pt->g);
	pt->guide = guide;
	pt->train_steps = 0;
	if(!guide) return YMX_OK;
	PtGuide &g = pt->g;
	// the guide buffers cover the whole wavefront (2^24 paths at most: 7.5 GB
	// of queries and mixtures, 13 GB more of pending training records)
	const uint32_t cap = pt->st.capacity;
	pt->guide_cap = cap;
	const uint32_t bounces = (uint32_t)pt->host_const.bounces;
	hipError_t e = hipSuccess;
#define GB_TRY(expr)                                                          \
	do                                                                        \
	{                                                                         \
		if((e = (expr)) != hipSuccess)                                        \
		{                                                                     \
			ymx_set_error("ymx_pt_set_guide: %s: %s", #expr, hipGetErrorString(e)); \
			free_guide_buffers(g);                                            \
			pt->guide = nullptr;                                              \
			return e == hipErrorOutOfMemory ? YMX_ERR_OUT_OF_MEMORY : YMX_ERR_HIP; \
		}                                                                     \
	} while(0)
	GB_TRY(dalloc(g.query, cap));
	GB_TRY(dalloc(g.mix_node, (size_t)cap * YMX_GUIDE_MIX_FLOATS));
	GB_TRY(dalloc(g.mix_path, (size_t)cap * YMX_GUIDE_MIX_FLOATS));
	GB_TRY(dalloc(g.query_count, 1));
	g.bsdf_prob = bsdf_prob < kAlphaMin ? kAlphaMin : (bsdf_prob > 1.f ? 1.f : bsdf_prob);
	g.max_depth = pt->guide_depth < 0 ? INT32_MAX : pt->guide_depth;
	g.on = 1u;
	{
		// regions of the guide's bounds (training-weight normalisation, learned selection)
		float bmax[3];
		ymx_guide_bounds(guide, g.bmin, bmax);
		for(int k = 0; k < 3; ++k) g.scale[k] = (float)kGuideCells / (bmax[k] - g.bmin[k]);
	}
	if(!attach_selection(pt))
	{
		pt->alpha_learned = false;
		pt->fallback_on = false;
	}
	if(train)
	{
		// ring of recent samples: a power of two, at least 4 mini-batches, 1M by default
		uint64_t tcap = 1ull << 20;
		while(tcap < 4ull * ymx_guide_batch_limit(guide) && tcap < (1ull << 24)) tcap <<= 1;
		GB_TRY(dalloc(g.rec, (size_t)bounces * cap));
		GB_TRY(dalloc(g.base, (size_t)bounces * cap));
		GB_TRY(dalloc(g.mask, cap));
		GB_TRY(hipMemset(g.mask, 0, sizeof(uint32_t) * cap));
		GB_TRY(dalloc(g.train_buf, (size_t)tcap));
		GB_TRY(dalloc(g.batch_buf, (size_t)ymx_guide_batch_limit(guide)));
		const size_t cells = (size_t)kGuideCells * kGuideCells * kGuideCells;
		GB_TRY(dalloc(g.cell_sum, cells));
		GB_TRY(hipMemset(g.cell_sum, 0, sizeof(float2) * cells));
		GB_TRY(dalloc(g.train_count, 1));
		GB_TRY(hipMemset(g.train_count, 0, sizeof(uint32_t)));
		g.train_cap = (uint32_t)tcap;
		g.accept = 1.f;
		g.train = 1u;
		pt->train_steps = train_steps;
		pt->g_seen = 0;
		pt->g_filled = 0;
		pt->g_rng = 0x853C49E6748FEA9Bull;
	}
#undef GB_TRY
	return YMX_OK;
}

int ymx_pt_set_guide_selection(ymx_pt *pt, int32_t mode, float lr)
{
	if(!pt || !pt->guide || !(lr > 0.f) || lr > 1.f || mode < 0 || mode > 2)
	{
		ymx_set_error("ymx_pt_set_guide_selection: no guide set, mode outside 0..2 or lr outside (0, 1]");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	const size_t cells = (size_t)kGuideCells * kGuideCells * kGuideCells;
	PtGuide &g = pt->g;
	pt->alpha_lr = lr;
	pt->alpha_steps = 0;
	if(mode == 0)
	{
		pt->alpha_learned = false;
		attach_selection(pt);
		return YMX_OK;
	}
	float *t = ymx_guide_selection_table(pt->guide, kSelectionTableFloats);
	if(!t)
	{
		ymx_set_error("ymx_pt_set_guide_selection: no device memory for the selection table");
		return YMX_ERR_OUT_OF_MEMORY;
	}
	if(mode == 1)
	{
		// every region starts at the fixed probability:
		// theta = logit((bsdf_prob - min) / (1 - min)), moments zero
		const float s = (g.bsdf_prob - kAlphaMin) / (1.f - kAlphaMin);
		const float sc = s < 1e-4f ? 1e-4f : (s > 1.f - 1e-4f ? 1.f - 1e-4f : s);
		std::vector<float> init(4 * cells, 0.f);
		for(size_t i = 0; i < cells; ++i) init[i] = logf(sc / (1.f - sc));
		YMX_HIP_CHECK(hipMemcpy(t, init.data(), sizeof(float) * init.size(), hipMemcpyHostToDevice));
	}
	pt->alpha_learned = true;
	attach_selection(pt);
	return YMX_OK;
}

int ymx_pt_set_guide_fallback(ymx_pt *pt, int32_t mode)
{
	if(!pt || !pt->guide || mode < 0 || mode > 2)
	{
		ymx_set_error("ymx_pt_set_guide_fallback: no guide set or mode outside 0..2");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	pt->fallback_on = mode != 0;
	if(mode != 0)
	{
		float *t = ymx_guide_selection_table(pt->guide, kSelectionTableFloats);
		if(!t)
		{
			pt->fallback_on = false;
			ymx_set_error("ymx_pt_set_guide_fallback: no device memory for the selection table");
			return YMX_ERR_OUT_OF_MEMORY;
		}
		// mode 1: no region has fallen back yet: sums and running moments zero
		const size_t cells = (size_t)kGuideCells * kGuideCells * kGuideCells;
		if(mode == 1) YMX_HIP_CHECK(hipMemset(t + 4 * cells, 0, sizeof(float) * 6 * cells));
	}
	attach_selection(pt);
	return YMX_OK;
}

int ymx_pt_get_guide_fallback(const ymx_pt *pt, float *moments, uint32_t n)
{
	const size_t cells = (size_t)kGuideCells * kGuideCells * kGuideCells;
	if(!pt || !moments || n < 3 * cells || !pt->g.fallback)
	{
		ymx_set_error("ymx_pt_get_guide_fallback: NULL argument, fewer than %zu entries or no fallback set",
		              3 * cells);
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	YMX_HIP_CHECK(hipMemcpy(moments, pt->g.fb_run, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost));
	return YMX_OK;
}

int ymx_pt_get_guide_selection(const ymx_pt *pt, float *alpha, uint32_t n)
{
	const size_t cells = (size_t)kGuideCells * kGuideCells * kGuideCells;
	if(!pt || !alpha || n < cells)
	{
		ymx_set_error("ymx_pt_get_guide_selection: NULL argument or fewer than %zu entries", cells);
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	if(!pt->g.learn_alpha)
		for(size_t i = 0; i < cells; ++i) alpha[i] = pt->g.bsdf_prob;
	else
	{
		std::vector<float> theta(cells);
		YMX_HIP_CHECK(hipMemcpy(theta.data(), pt->g.alpha_theta, sizeof(float) * cells, hipMemcpyDeviceToHost));
		for(size_t i = 0; i < cells; ++i) alpha[i] = kAlphaMin + (1.f - kAlphaMin) / (1.f + expf(-theta[i]));
	}
	if(pt->g.fallback)
	{
		std::vector<float> run(3 * cells);
		YMX_HIP_CHECK(hipMemcpy(run.data(), pt->g.fb_run, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost));
		for(size_t i = 0; i < cells; ++i)
			if(run[2 * cells + i] >= kFallbackMinRecords && run[i] > run[cells + i]) alpha[i] = 1.f;
	}
	return YMX_OK;
}

int ymx_pt_set_normals(ymx_pt *pt, const float *corner_normals)
{
	if(!pt)
	{
		ymx_set_error("ymx_pt_set_normals: NULL handle");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	if(!corner_normals)
	{
		if(pt->d_cnorm) YMX_HIP_CHECK(hipFree(pt->d_cnorm));
		pt->d_cnorm = nullptr;
		return YMX_OK;
	}
	// the transparent-shadow query returns no barycentrics, so a smooth
	// fake-shadow glass caster's Fresnel term cannot use its shading normal
	if(pt->host_const.transparent_shadows && pt->d_transp)
	{
		std::vector<uint32_t> bits(((size_t)pt->n_tris + 31) / 32 + 1);
		YMX_HIP_CHECK(hipMemcpy(bits.data(), pt->d_transp, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost));
		for(uint32_t i = 0; i < pt->n_tris; ++i)
			if(((bits[i >> 5] >> (i & 31u)) & 1u) && corner_normals[12 * (size_t)i + 3] >= 0.f)
			{
				ymx_set_error("ymx_pt_set_normals: triangle %u is smooth fake-shadow glass under transparent shadows", i);
				return YMX_ERR_UNSUPPORTED;
			}
	}
	if(!pt->d_cnorm) YMX_HIP_CHECK(hipMalloc(&pt->d_cnorm, sizeof(float4) * 3 * (size_t)(pt->n_tris ? pt->n_tris : 1)));
	YMX_HIP_CHECK(hipMemcpy(pt->d_cnorm, corner_normals, sizeof(float) * 12 * (size_t)pt->n_tris,
	                        hipMemcpyHostToDevice));
	return YMX_OK;
}

int ymx_pt_destroy(ymx_pt *pt)
{
	if(!pt) return YMX_OK;
	(void)hipDeviceSynchronize();
	free_pt(pt);
	return YMX_OK;
}

int ymx_pt_clear_film(ymx_pt *pt, void *stream)
{
	if(!pt)
	{
		ymx_set_error("ymx_pt_clear_film: NULL handle");
		return YMX_ERR_INVALID_ARG;
	}
	const uint32_t n = (uint32_t)(pt->host_const.resx * pt->host_const.resy);  // film pixels
	hipStream_t s = (hipStream_t)stream;
	YMX_HIP_CHECK(hipMemsetAsync(pt->d_film, 0, sizeof(float4) * n, s));
	YMX_HIP_CHECK(hipMemsetAsync(pt->d_film_w, 0, sizeof(float) * n, s));
	return YMX_OK;
}

// Samples [first, first + n) of n_active pixels (pix_list / pix_slot NULL:
// every pixel) in wavefronts of as many samples as the buffers hold.
static int render_range(ymx_pt *pt, int32_t first, int32_t n, const uint32_t *pix_list, const uint32_t *pix_slot,
                        uint32_t n_active, float light_mult, int32_t pass_samples, hipStream_t s)
{
	uint64_t per = pt->st.capacity / n_active;
	if(per > pt->nd.cap / n_active) per = pt->nd.cap / n_active;
	if(per < 1) per = 1;
	for(int32_t k = 0; k < n; k += (int32_t)per)
	{
		const int32_t count = (uint64_t)(n - k) < per ? (n - k) : (int32_t)per;
		const int rc = ymx_pt_trace_subset(pt, first + k, count, pix_list, pix_slot, n_active, light_mult, pass_samples, s);
		if(rc != YMX_OK) return rc;
	}
	return YMX_OK;
}

static int begin_render(ymx_pt *pt, hipStream_t s)
{
	YMX_HIP_CHECK(hipMemsetAsync(pt->st.totals, 0, 3 * sizeof(unsigned long long), s));
	pt->stats.guide_samples = 0;
	pt->stats.guide_steps = 0;
	YMX_HIP_CHECK(hipEventRecord(pt->ev0, s));
	return YMX_OK;
}

static int end_render(ymx_pt *pt, hipStream_t s)
{
	YMX_HIP_CHECK(hipEventRecord(pt->ev1, s));
	pt->stats.struct_size = sizeof(ymx_pt_stats);
	return YMX_OK;
}

int ymx_pt_render(ymx_pt *pt, int32_t first_sample, int32_t n_samples, void *stream)
{
	if(!pt || first_sample < 0 || n_samples < 0)
	{
		ymx_set_error("ymx_pt_render: bad arguments");
		return YMX_ERR_INVALID_ARG;
	}
	hipStream_t s = (hipStream_t)stream;
	int rc = begin_render(pt, s);
	if(rc != YMX_OK) return rc;
	const uint32_t npix = (uint32_t)(pt->host_const.resx * pt->host_const.resy);
	for(int32_t k = 0; k < n_samples; k += (int32_t)pt->batch)
	{
		const int32_t count = (n_samples - k) < (int32_t)pt->batch ? (n_samples - k) : (int32_t)pt->batch;
		rc = ymx_pt_trace_subset(pt, first_sample + k, count, nullptr, nullptr, npix, 1.f, pt->host_const.aa_samples,
		                         s);
		if(rc != YMX_OK) return rc;
	}
	return end_render(pt, s);
}

int ymx_pt_render_passes(ymx_pt *pt, int32_t *resampled, void *stream)
{
	if(!pt)
	{
		ymx_set_error("ymx_pt_render_passes: NULL handle");
		return YMX_ERR_INVALID_ARG;
	}
	hipStream_t s = (hipStream_t)stream;
	const ymx_pt_settings &st = pt->settings;
	const uint32_t npix = (uint32_t)(pt->host_const.resx * pt->host_const.resy);
	int rc = begin_render(pt, s);
	if(rc != YMX_OK) return rc;
	// TiledIntegrator::render, src/integrator/surface/integrator_tiled.cc:160-214
	rc = render_range(pt, 0, st.aa_samples, nullptr, nullptr, npix, 1.f, st.aa_samples, s);
	if(rc != YMX_OK) return rc;
	if(resampled) resampled[0] = (int32_t)npix;
	const int32_t floor_pixels = (int32_t)floorf(st.aa_resampled_floor * (float)npix / 100.f);
	float sample_mult = 1.f, light_mult = 1.f, threshold = st.aa_threshold;
	int32_t acum = st.aa_samples, res = 0;
	bool threshold_changed = true;
	const uint32_t *list = pt->d_aa ? pt->d_aa + 2 * (size_t)npix : nullptr;
	const uint32_t *slot = pt->d_aa ? pt->d_aa + 3 * (size_t)npix : nullptr;
	const uint32_t *count = pt->d_aa ? pt->d_aa + 4 * (size_t)npix + (npix + 255) / 256 : nullptr;
	for(int32_t i = 1; i < st.aa_passes; ++i)
	{
		sample_mult *= st.aa_sample_multiplier_factor;
		light_mult *= st.aa_light_sample_multiplier_factor;
		bool all = false;
		if(res <= 0 && !threshold_changed) res = 0;  // ImageFilm::nextPass skipped
		else if(threshold <= 0.f)
		{
			// ImageFilm::doMoreSamples: every pixel
			res = (int32_t)npix;
			all = true;
			threshold_changed = false;
		}
		else
		{
			rc = ymx_pt_aa_flags(pt, threshold, s);
			if(rc != YMX_OK) return rc;
			uint32_t hc = 0;
			YMX_HIP_CHECK(hipMemcpyAsync(&hc, count, sizeof hc, hipMemcpyDeviceToHost, s));
			YMX_HIP_CHECK(hipStreamSynchronize(s));
			res = (int32_t)hc;
			threshold_changed = false;
		}
		const int32_t n = (int32_t)ceilf((float)st.aa_inc_samples * sample_mult);
		if(res > 0)
		{
			rc = all ? render_range(pt, acum, n, nullptr, nullptr, npix, light_mult, n, s)
			         : render_range(pt, acum, n, list, slot, (uint32_t)res, light_mult, n, s);
			if(rc != YMX_OK) return rc;
		}
		if(resampled) resampled[i] = res;
		acum += n;
		if(res < floor_pixels)
		{
			const float ratio = fminf(8.f, (float)floor_pixels / (float)res);
			threshold = threshold * (1.f - 0.1f * ratio);
			if(st.aa_threshold > 0.f) threshold_changed = true;
		}
	}
	return end_render(pt, s);
}

int ymx_pt_read_film(ymx_pt *pt, float *rgba, float *weight)
{
	if(!pt || !rgba || !weight)
	{
		ymx_set_error("ymx_pt_read_film: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	const uint32_t n = (uint32_t)(pt->host_const.resx * pt->host_const.resy);  // film pixels
	YMX_HIP_CHECK(hipMemcpy(rgba, pt->d_film, sizeof(float4) * n, hipMemcpyDeviceToHost));
	YMX_HIP_CHECK(hipMemcpy(weight, pt->d_film_w, sizeof(float) * n, hipMemcpyDeviceToHost));
	return YMX_OK;
}

int ymx_pt_save_film(ymx_pt *pt, const char *path, int32_t computer_node, int32_t base_sampling_offset,
                     int32_t sampling_offset)
{
	if(!pt || !path)
	{
		ymx_set_error("ymx_pt_save_film: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	const int32_t w = pt->host_const.resx, h = pt->host_const.resy;
	const size_t n = (size_t)w * (size_t)h;
	std::vector<float> rgba(4 * n), weight(n);
	const int rc = ymx_pt_read_film(pt, rgba.data(), weight.data());
	if(rc != YMX_OK) return rc;
	FILE *f = fopen(path, "wb");
	if(!f)
	{
		ymx_set_error("ymx_pt_save_film: cannot open '%s' for writing", path);
		return YMX_ERR_INVALID_ARG;
	}
	static const char kTag[] = "YAF_FILMv4_0_0";  // written with its NUL
	const int32_t head[10] = {computer_node, base_sampling_offset, sampling_offset, w, h, 0, w - 1, 0, h - 1, 1};
	bool ok = fwrite(kTag, 1, sizeof kTag, f) == sizeof kTag;
	ok = ok && fwrite(head, sizeof(int32_t), 10, f) == 10;
	ok = ok && fwrite(weight.data(), sizeof(float), n, f) == n;
	ok = ok && fwrite(rgba.data(), sizeof(float), 4 * n, f) == 4 * n;
	ok = (fclose(f) == 0) && ok;
	if(!ok)
	{
		ymx_set_error("ymx_pt_save_film: write to '%s' failed", path);
		return YMX_ERR_HIP;
	}
	return YMX_OK;
}

int ymx_pt_load_film(ymx_pt *pt, const char *path, int32_t add, int32_t *computer_node,
                     int32_t *base_sampling_offset, int32_t *sampling_offset)
{
	if(!pt || !path)
	{
		ymx_set_error("ymx_pt_load_film: NULL argument");
		return YMX_ERR_INVALID_ARG;
	}
	const int32_t w = pt->host_const.resx, h = pt->host_const.resy;
	const size_t n = (size_t)w * (size_t)h;
	FILE *f = fopen(path, "rb");
	if(!f)
	{
		ymx_set_error("ymx_pt_load_film: cannot open '%s'", path);
		return YMX_ERR_INVALID_ARG;
	}
	static const char kTag[] = "YAF_FILMv4_0_0";
	char tag[sizeof kTag];
	int32_t head[10];
	std::vector<float> rgba(4 * n), weight(n);
	bool ok = fread(tag, 1, sizeof tag, f) == sizeof tag && memcmp(tag, kTag, sizeof kTag) == 0;
	ok = ok && fread(head, sizeof(int32_t), 10, f) == 10;
	const bool shape = ok && head[3] == w && head[4] == h && head[5] == 0 && head[6] == w - 1 && head[7] == 0 &&
	                   head[8] == h - 1 && head[9] == 1;
	ok = shape && fread(weight.data(), sizeof(float), n, f) == n && fread(rgba.data(), sizeof(float), 4 * n, f) == 4 * n;
	fclose(f);
	if(!ok)
	{
		ymx_set_error("ymx_pt_load_film: '%s' is not a %dx%d single-layer YAF_FILMv4_0_0 film", path, w, h);
		return YMX_ERR_INVALID_ARG;
	}
	if(add)
	{
		std::vector<float> cur_rgba(4 * n), cur_w(n);
		const int rc = ymx_pt_read_film(pt, cur_rgba.data(), cur_w.data());
		if(rc != YMX_OK) return rc;
		for(size_t i = 0; i < n; ++i) weight[i] = cur_w[i] + weight[i];
		for(size_t i = 0; i < 4 * n; ++i) rgba[i] = cur_rgba[i] + rgba[i];
	}
	YMX_HIP_CHECK(hipDeviceSynchronize());
	YMX_HIP_CHECK(hipMemcpy(pt->d_film, rgba.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
	YMX_HIP_CHECK(hipMemcpy(pt->d_film_w, weight.data(), sizeof(float) * n, hipMemcpyHostToDevice));
	if(computer_node) *computer_node = head[0];
	if(base_sampling_offset) *base_sampling_offset = head[1];
	if(sampling_offset) *sampling_offset = head[2];
	return YMX_OK;
}

int ymx_pt_get_tuning(const ymx_pt *pt, ymx_pt_tuning *t)
{
	if(!pt || !t || t->struct_size != sizeof(ymx_pt_tuning))
	{
		ymx_set_error("ymx_pt_get_tuning: NULL argument or struct_size mismatch");
		return YMX_ERR_INVALID_ARG;
	}
	t->samples_per_wavefront = (int32_t)pt->batch;
	t->shadow_kernel = pt->shadow_replacement ? YMX_TRACE_REPLACEMENT : YMX_TRACE_WAVE_ORDERED;
	t->extension_kernel = pt->ext_replacement ? YMX_TRACE_REPLACEMENT : YMX_TRACE_WAVE_ORDERED;
	return YMX_OK;
}

int ymx_pt_set_tuning(ymx_pt *pt, const ymx_pt_tuning *t)
{
	if(!pt || !t || t->struct_size != sizeof(ymx_pt_tuning))
	{
		ymx_set_error("ymx_pt_set_tuning: NULL argument or struct_size mismatch");
		return YMX_ERR_INVALID_ARG;
	}
	const bool kernels_ok = (t->shadow_kernel == YMX_TRACE_WAVE_ORDERED || t->shadow_kernel == YMX_TRACE_REPLACEMENT) &&
	                        (t->extension_kernel == YMX_TRACE_WAVE_ORDERED ||
	                         t->extension_kernel == YMX_TRACE_REPLACEMENT);
	if(t->samples_per_wavefront < 1 || (uint32_t)t->samples_per_wavefront > pt->batch_max || !kernels_ok)
	{
		ymx_set_error("ymx_pt_set_tuning: %d samples per wavefront (1..%u) or an unknown kernel",
		              t->samples_per_wavefront, pt->batch_max);
		return YMX_ERR_INVALID_ARG;
	}
	if(t->shadow_kernel == YMX_TRACE_REPLACEMENT && pt->host_const.transparent_shadows)
	{
		ymx_set_error("ymx_pt_set_tuning: transparent shadows are traced wave-ordered only");
		return YMX_ERR_UNSUPPORTED;
	}
	pt->batch = (uint32_t)t->samples_per_wavefront;
	pt->shadow_replacement = t->shadow_kernel == YMX_TRACE_REPLACEMENT;
	pt->ext_replacement = t->extension_kernel == YMX_TRACE_REPLACEMENT;
	return YMX_OK;
}

int ymx_pt_autotune(ymx_pt *pt, int32_t trial_samples, ymx_pt_tuning *result, void *stream)
{
	if(!pt || !result || result->struct_size != sizeof(ymx_pt_tuning) || trial_samples < 0)
	{
		ymx_set_error("ymx_pt_autotune: NULL argument, struct_size mismatch or negative trial_samples");
		return YMX_ERR_INVALID_ARG;
	}
	hipStream_t s = (hipStream_t)stream;
	const uint32_t npix = (uint32_t)(pt->host_const.resx * pt->host_const.resy);
	const int32_t trial = trial_samples > 0 ? trial_samples : (int32_t)pt->batch_max;
	// the film is kept aside and restored; guide training pauses
	float4 *keep_film = nullptr;
	float *keep_w = nullptr;
	YMX_HIP_CHECK(hipMalloc(&keep_film, sizeof(float4) * npix));
	if(hipMalloc(&keep_w, sizeof(float) * npix) != hipSuccess)
	{
		(void)hipFree(keep_film);
		ymx_set_error("ymx_pt_autotune: no device memory to keep the film");
		return YMX_ERR_OUT_OF_MEMORY;
	}
	int rc = YMX_OK;
	const uint32_t train = pt->g.train;
	pt->g.train = 0u;
	const uint32_t batch0 = pt->batch;
	const bool sh0 = pt->shadow_replacement, ext0 = pt->ext_replacement;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	ymx_pt_tuning out = *result;
	out.n_trials = 0;
	float best_ms = 0.f;
	struct Cfg
	{
		uint32_t batch;
		bool sh, ext;
	} best{batch0, sh0, ext0};
	auto measure = [&](const Cfg &cfg, float &ms_per_sample) -> int {
		pt->batch = cfg.batch;
		pt->shadow_replacement = cfg.sh;
		pt->ext_replacement = cfg.ext;
		if(hipEventRecord(e0, s) != hipSuccess) return YMX_ERR_HIP;
		int r = ymx_pt_render(pt, 0, trial, stream);
		if(r != YMX_OK) return r;
		if(hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return YMX_ERR_HIP;
		float ms = 0.f;
		if(hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return YMX_ERR_HIP;
		ms_per_sample = ms / (float)trial;
		if(out.n_trials < YMX_PT_TUNING_TRIALS)
		{
			const int k = out.n_trials++;
			out.trial_samples_per_wavefront[k] = (int32_t)cfg.batch;
			out.trial_shadow_kernel[k] = cfg.sh ? YMX_TRACE_REPLACEMENT : YMX_TRACE_WAVE_ORDERED;
			out.trial_extension_kernel[k] = cfg.ext ? YMX_TRACE_REPLACEMENT : YMX_TRACE_WAVE_ORDERED;
			out.trial_ms_per_sample[k] = ms_per_sample;
		}
		if(best_ms == 0.f || ms_per_sample < best_ms)
		{
			best_ms = ms_per_sample;
			best = cfg;
		}
		return YMX_OK;
	};
	do
	{
		if(hipMemcpyAsync(keep_film, pt->d_film, sizeof(float4) * npix, hipMemcpyDeviceToDevice, s) != hipSuccess ||
		   hipMemcpyAsync(keep_w, pt->d_film_w, sizeof(float) * npix, hipMemcpyDeviceToDevice, s) != hipSuccess ||
		   hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
		{
			ymx_set_error("ymx_pt_autotune: event or copy failed");
			rc = YMX_ERR_HIP;
			break;
		}
		// a warm-up render (first launches, first touch of the wavefront state)
		pt->batch = pt->batch_max;
		if((rc = ymx_pt_render(pt, 0, trial < (int32_t)pt->batch_max ? trial : (int32_t)pt->batch_max, stream)) !=
		   YMX_OK)
			break;
		float ms = 0.f;
		const bool transparent = pt->host_const.transparent_shadows != 0;
		for(int sh = 0; sh < (transparent ? 1 : 2) && rc == YMX_OK; ++sh)
			for(int ext = 0; ext < 2 && rc == YMX_OK; ++ext)
				rc = measure(Cfg{pt->batch_max, sh == 1, ext == 1}, ms);
		const Cfg kernels = best;
		for(uint32_t b = pt->batch_max / 2u; b >= 1u && rc == YMX_OK && b * 4u >= pt->batch_max; b /= 2u)
			rc = measure(Cfg{b, kernels.sh, kernels.ext}, ms);
	} while(false);
	if(e0) (void)hipEventDestroy(e0);
	if(e1) (void)hipEventDestroy(e1);
	// restore the film whatever happened
	(void)hipMemcpyAsync(pt->d_film, keep_film, sizeof(float4) * npix, hipMemcpyDeviceToDevice, s);
	(void)hipMemcpyAsync(pt->d_film_w, keep_w, sizeof(float) * npix, hipMemcpyDeviceToDevice, s);
	(void)hipStreamSynchronize(s);
	(void)hipFree(keep_film);
	(void)hipFree(keep_w);
	pt->g.train = train;
	if(rc != YMX_OK)
	{
		pt->batch = batch0;
		pt->shadow_replacement = sh0;
		pt->ext_replacement = ext0;
		return rc;
	}
	pt->batch = best.batch;
	pt->shadow_replacement = best.sh;
	pt->ext_replacement = best.ext;
	out.samples_per_wavefront = (int32_t)best.batch;
	out.shadow_kernel = best.sh ? YMX_TRACE_REPLACEMENT : YMX_TRACE_WAVE_ORDERED;
	out.extension_kernel = best.ext ? YMX_TRACE_REPLACEMENT : YMX_TRACE_WAVE_ORDERED;
	*result = out;
	return YMX_OK;
}

int ymx_pt_get_stats(const ymx_pt *pt, ymx_pt_stats *stats)
{
	if(!pt || !stats || stats->struct_size != sizeof(ymx_pt_stats))
	{
		ymx_set_error("ymx_pt_get_stats: NULL argument or struct_size mismatch");
		return YMX_ERR_INVALID_ARG;
	}
	YMX_HIP_CHECK(hipEventSynchronize(pt->ev1));
	unsigned long long t[3] = {0, 0, 0};
	YMX_HIP_CHECK(hipMemcpy(t, pt->st.totals, sizeof t, hipMemcpyDeviceToHost));
	float ms = 0.f;
	YMX_HIP_CHECK(hipEventElapsedTime(&ms, pt->ev0, pt->ev1));
	stats->camera_rays = t[0];
	stats->extension_rays = t[1];
	stats->shadow_rays = t[2];
	stats->render_ms = ms;
	stats->guide_samples = pt->stats.guide_samples;
	stats->light_trees = 0;
	for(const ymx_accel *tree : pt->light_trees) stats->light_trees += tree->max_depth < kLightStack ? 1u : 0u;
	stats->guide_steps = pt->stats.guide_steps;
	return YMX_OK;
}

}  // extern "C"
