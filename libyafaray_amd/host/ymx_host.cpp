// ymx_host.cpp -- the C++ host layer (ymx_host.hpp) over libymx.so's C ABI.
#include "ymx_host.hpp"

#include <cmath>
#include <cstring>

#include "../csrc/pt_math.h"
#include "../csrc/ray_tri.h"

namespace ymx_host {

// ---------------------------------------------------------------- ParamMap

namespace {

template <typename T>
T get_as(const std::map<std::string, Param> &m, const std::string &name, const T &def, bool *ok)
{
	const auto it = m.find(name);
	if(it == m.end()) return def;
	if(const T *v = std::get_if<T>(&it->second)) return *v;
	if(ok) *ok = false;
	return def;
}

}  // namespace

bool ParamMap::getBool(const std::string &name, bool def, bool *ok) const { return get_as(map_, name, def, ok); }
int ParamMap::getInt(const std::string &name, int def, bool *ok) const { return get_as(map_, name, def, ok); }
double ParamMap::getFloat(const std::string &name, double def, bool *ok) const
{
	// an int param where a float is expected is read as a float, as the reference does
	const auto it = map_.find(name);
	if(it != map_.end())
		if(const int *i = std::get_if<int>(&it->second)) return (double)*i;
	return get_as(map_, name, def, ok);
}
std::string ParamMap::getString(const std::string &name, const std::string &def, bool *ok) const
{
	return get_as(map_, name, def, ok);
}
Color ParamMap::getColor(const std::string &name, const Color &def, bool *ok) const { return get_as(map_, name, def, ok); }
Vector ParamMap::getVector(const std::string &name, const Vector &def, bool *ok) const
{
	return get_as(map_, name, def, ok);
}

// ---------------------------------------------------------------- Accelerator

namespace {

constexpr float kBiasScale = 0.1f * 0.00005f;  // include/accelerator/accelerator.h:64,85
constexpr float kSlabWiden = 1.0000004f;       // the traversal kernels' conservative slab widening
constexpr float kWindow = 0x1p-12f;
constexpr int32_t kRefEmpty = (int32_t)0xFFFFFFFF;

inline float bits_f(int32_t i)
{
	float f;
	memcpy(&f, &i, 4);
	return f;
}
inline int32_t f_bits(float f)
{
	int32_t i;
	memcpy(&i, &f, 4);
	return i;
}

struct HostRay
{
	float o[3], d[3], idir[3];
	float tlo;
};

// The traversal kernels' two-child slab test (accel_trace.hip, box2), evaluated
// the same way so that the host visits every leaf the device visits.
inline int box2(const float *n, const HostRay &r, float thi, float &tn0, float &tn1)
{
	const float a0x = (n[0] - r.o[0]) * r.idir[0], b0x = (n[1] - r.o[0]) * r.idir[0];
	const float a0y = (n[2] - r.o[1]) * r.idir[1], b0y = (n[3] - r.o[1]) * r.idir[1];
	const float a0z = (n[8] - r.o[2]) * r.idir[2], b0z = (n[9] - r.o[2]) * r.idir[2];
	const float a1x = (n[4] - r.o[0]) * r.idir[0], b1x = (n[5] - r.o[0]) * r.idir[0];
	const float a1y = (n[6] - r.o[1]) * r.idir[1], b1y = (n[7] - r.o[1]) * r.idir[1];
	const float a1z = (n[10] - r.o[2]) * r.idir[2], b1z = (n[11] - r.o[2]) * r.idir[2];
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

}  // namespace

std::unique_ptr<Accelerator> Accelerator::factory(const ParamMap &params, const float *verts, uint32_t n_verts,
                                                  const uint32_t *tri_idx, const uint8_t *vis, uint32_t n_tris,
                                                  std::string *error)
{
	return create(params, verts, n_verts, tri_idx, false, vis, n_tris, error);
}

std::unique_ptr<Accelerator> Accelerator::factoryFaces(const ParamMap &params, const float *verts, uint32_t n_verts,
                                                       const uint32_t *face_idx, const uint8_t *vis, uint32_t n_faces,
                                                       std::string *error)
{
	return create(params, verts, n_verts, face_idx, true, vis, n_faces, error);
}

std::unique_ptr<Accelerator> Accelerator::create(const ParamMap &params, const float *verts, uint32_t n_verts,
                                                 const uint32_t *tri_idx, bool faces, const uint8_t *vis,
                                                 uint32_t n_tris, std::string *error)
{
	const std::string type = params.getString("type", "");
	if(type != "mi355x-lbvh")
	{
		if(error) *error = "Accelerator: type '" + type + "' is not built here (only 'mi355x-lbvh')";
		return nullptr;
	}
	ymx_accel_params p{};
	p.struct_size = sizeof p;
	p.max_leaf_size = params.getInt("max_leaf_size", 0);
	const std::string builder = params.getString("builder", "lbvh");
	if(builder != "lbvh" && builder != "ploc")
	{
		if(error) *error = "Accelerator: builder '" + builder + "' is not one of 'lbvh', 'ploc'";
		return nullptr;
	}
	p.builder = builder == "ploc" ? YMX_BUILDER_PLOC : YMX_BUILDER_LBVH;
	p.treelet_passes = params.getInt("treelet_passes", 0);
	p.ray_reorder = params.getInt("ray_reorder", 0);
	p.reorder_cell_bits = params.getInt("reorder_cell_bits", 0);
	std::unique_ptr<Accelerator> a(new Accelerator());
	const int rc = faces ? ymx_accel_create_faces(verts, n_verts, tri_idx, vis, n_tris, &p, &a->acc_)
	                     : ymx_accel_create(verts, n_verts, tri_idx, vis, n_tris, &p, &a->acc_);
	if(rc != YMX_OK)
	{
		if(error) *error = std::string("Accelerator: ") + ymx_last_error();
		return nullptr;
	}
	ymx_accel_info info{};
	info.struct_size = sizeof info;
	if(ymx_accel_get_info(a->acc_, &info) != YMX_OK)
	{
		if(error) *error = std::string("Accelerator: ") + ymx_last_error();
		return nullptr;
	}
	a->n_tris_ = n_tris;
	a->nodes_.resize(info.node_bytes / 4);
	a->tris_.resize(info.tri_bytes / 4);
	a->quads_.resize(4 * (size_t)info.n_quads);
	if(ymx_accel_export(a->acc_, a->nodes_.data(), info.node_bytes, a->tris_.data(), info.tri_bytes) != YMX_OK ||
	   ymx_accel_export_quads(a->acc_, a->quads_.data(), 16ull * info.n_quads) != YMX_OK)
	{
		if(error) *error = std::string("Accelerator: ") + ymx_last_error();
		return nullptr;
	}
	for(int k = 0; k < 3; ++k)
	{
		a->bound_.a[k] = info.bound_min[k];
		a->bound_.g[k] = info.bound_max[k];
	}
	a->max_depth_ = info.max_depth;
	return a;
}

Accelerator::~Accelerator()
{
	if(acc_) ymx_accel_destroy(acc_);
}

IntersectData Accelerator::traverse(const Ray &ray, float t_max, Mode mode,
                                    const std::function<bool(uint32_t, float, float, float)> &any) const
{
	IntersectData out;
	float enter = 0.f, leave = 0.f;
	if(!ymx::tree_cross(bound_.a, bound_.g, ray.from, ray.dir, t_max, enter, leave)) return out;
	// include/accelerator/accelerator_kdtree_common.h:139
	const float bias = kBiasScale * fabsf(leave - enter);
	const float t_min = (mode == Mode::Shadow) ? bias : ((ray.tmin < bias) ? bias : ray.tmin);
	HostRay r;
	for(int k = 0; k < 3; ++k)
	{
		r.o[k] = ray.from[k];
		r.d[k] = ray.dir[k];
		const float dk = (fabsf(ray.dir[k]) > 1e-30f) ? ray.dir[k] : copysignf(1e-30f, ray.dir[k]);
		r.idir[k] = 1.f / dk;
	}
	r.tlo = t_min - t_min * kWindow;
	float best_t = t_max;
	uint32_t best_prim = kNoHit;
	float best_u = 0.f, best_v = 0.f;
	std::vector<int32_t> stack;
	stack.reserve(64);
	int32_t ref = 0;
	while(ref != kRefEmpty)
	{
		while(ref >= 0)
		{
			const float *n = nodes_.data() + 16 * (size_t)ref;
			const float thi = best_t + best_t * kWindow;
			float tn0, tn1;
			const int m = box2(n, r, thi, tn0, tn1);
			int32_t c0 = f_bits(n[12]), c1 = f_bits(n[13]);
			if(m == 3)
			{
				if(tn1 < tn0) std::swap(c0, c1);
				stack.push_back(c1);
				const uint32_t used = (uint32_t)stack.size();
				uint32_t seen = max_stack_.load(std::memory_order_relaxed);
				while(used > seen && !max_stack_.compare_exchange_weak(seen, used, std::memory_order_relaxed))
				{
				}
				ref = c0;
			}
			else if(m == 1) ref = c0;
			else if(m == 2) ref = c1;
			else
			{
				ref = stack.empty() ? kRefEmpty : stack.back();
				if(!stack.empty()) stack.pop_back();
			}
		}
		if(ref == kRefEmpty) break;
		const uint32_t first = ((uint32_t)ref & 0x7FFFFFFFu) >> 4, count = ((uint32_t)ref & 15u) + 1u;
		for(uint32_t k = 0; k < count; ++k)
		{
			const float *t = tris_.data() + 12 * (size_t)(first + k);
			const float4 t0{t[0], t[1], t[2], t[3]}, t1{t[4], t[5], t[6], t[7]}, t2{t[8], t[9], t[10], t[11]};
			float u, v;
			const float th = ymx::prim_test(t0, t1, t2, reinterpret_cast<const float4 *>(quads_.data()), r.o, r.d, u, v);
			if(th <= 0.f || th < t_min) continue;
			const uint32_t prim = (uint32_t)f_bits(t0.w);
			const uint32_t vmask = (uint32_t)f_bits(t1.w);
			if(mode == Mode::Closest)
			{
				// include/accelerator/accelerator.h:122-133; exact ties keep the lowest id
				if(!(vmask & YMX_VIS_VISIBLE)) continue;
				const bool closer = (th < best_t) || (th == best_t && best_prim != kNoHit && prim < best_prim);
				if(!closer) continue;
				best_t = th;
				best_prim = prim;
				best_u = u;
				best_v = v;
			}
			else
			{
				// include/accelerator/accelerator.h:135-168
				if(th >= t_max || !(vmask & YMX_VIS_CASTS_SHADOWS)) continue;
				if(any(prim, th, u, v))
				{
					out.t_hit = out.t_max = th;
					out.u = u;
					out.v = v;
					out.primitive = prim;
					return out;
				}
			}
		}
		ref = stack.empty() ? kRefEmpty : stack.back();
		if(!stack.empty()) stack.pop_back();
	}
	if(mode == Mode::Closest && best_prim != kNoHit)
	{
		out.t_hit = out.t_max = best_t;
		out.u = best_u;
		out.v = best_v;
		out.primitive = best_prim;
	}
	return out;
}

IntersectData Accelerator::intersect(const Ray &ray, float t_max) const
{
	return traverse(ray, t_max, Mode::Closest, nullptr);
}

IntersectData Accelerator::intersectShadow(const Ray &ray, float t_max) const
{
	return traverse(ray, t_max, Mode::Shadow, [](uint32_t, float, float, float) { return true; });
}

int Accelerator::intersect(const ymx_ray *d_rays, ymx_hit *d_hits, uint32_t n, void *stream) const
{
	return ymx_accel_intersect(acc_, d_rays, d_hits, n, stream);
}

int Accelerator::intersectShadow(const ymx_ray *d_rays, uint32_t *d_occluded, uint32_t n, void *stream) const
{
	return ymx_accel_intersect_shadow(acc_, d_rays, d_occluded, nullptr, n, stream);
}

// ---------------------------------------------------------------- Scene

namespace {

void set3(float dst[3], const Color &c)
{
	for(int k = 0; k < 3; ++k) dst[k] = (float)c[k];
}
void set3(float dst[3], const Vector &v)
{
	for(int k = 0; k < 3; ++k) dst[k] = (float)v[k];
}

// include/common/visibility.h:31-37 ("normal" and unknown names: both bits)
uint8_t visibility_mask(const std::string &vis)
{
	if(vis == "invisible") return 0;
	if(vis == "shadow_only") return YMX_VIS_CASTS_SHADOWS;
	if(vis == "no_shadows") return YMX_VIS_VISIBLE;
	return YMX_VIS_VISIBLE | YMX_VIS_CASTS_SHADOWS;
}

}  // namespace

size_t Scene::createObject(const std::string &name, const ParamMap &params)
{
	Object o;
	o.name = name;
	o.visibility = visibility_mask(params.getString("visibility", "normal"));
	o.base = params.getBool("is_base_object", false);
	objects_.push_back(o);
	modified_ |= ModifiedObjects;
	return objects_.size() - 1;
}

size_t Scene::addVertex(size_t object_id, double x, double y, double z)
{
	if(object_id >= objects_.size()) return (size_t)-1;
	std::vector<float> &v = objects_[object_id].verts;
	modified_ |= ModifiedObjects;
	v.push_back((float)x);
	v.push_back((float)y);
	v.push_back((float)z);
	return v.size() / 3 - 1;
}

bool Scene::addNormal(size_t object_id, double nx, double ny, double nz)
{
	if(object_id >= objects_.size()) return false;
	std::vector<float> &n = objects_[object_id].normals;
	n.insert(n.end(), {(float)nx, (float)ny, (float)nz});
	modified_ |= ModifiedObjects;
	return true;
}

bool Scene::addTriangle(size_t object_id, size_t a, size_t b, size_t c, size_t material_id)
{
	if(object_id >= objects_.size() || material_id >= materials_.size()) return false;
	Object &o = objects_[object_id];
	const size_t nv = o.verts.size() / 3;
	if(a >= nv || b >= nv || c >= nv) return false;
	o.tris.insert(o.tris.end(), {(uint32_t)a, (uint32_t)b, (uint32_t)c});
	o.mats.push_back((uint32_t)material_id);
	// a mesh with vertex normals: the corners take their vertices' normals
	// (src/geometry/object/object_mesh.cc:110)
	const bool has_normals = !o.normals.empty();
	for(size_t v : {a, b, c}) o.corner_normal.push_back(has_normals ? (int32_t)v : -1);
	modified_ |= ModifiedObjects;
	return true;
}

namespace {

using ymx::V3;

V3 load_v3(const std::vector<float> &a, size_t i) { return ymx::v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

// Vec::length, sinFromVectors and math::asin (include/geometry/vector.h:63,342-349,
// include/math/math.h:280-286)
float length(V3 a) { return std::sqrt(ymx::dot(a, a)); }
float sin_from_vectors(V3 a, V3 b)
{
	const float div = (length(a) * length(b)) * 0.99999f + 0.00001f;
	float arg = (length(ymx::cross(a, b)) / div) * 0.99999f;
	if(arg > 1.f) arg = 1.f;
	if(arg <= -1.f) return -1.5707963267948966f;
	if(arg >= 1.f) return 1.5707963267948966f;
	return std::asin(arg);
}

}  // namespace

void Scene::appendCornerNormals(const Object &o, const float *m, std::vector<float> &out)
{
	// PrimitivePolygon::getSurfacePolygon shades a smooth mesh, or one with
	// vertex normals, with its corners' normals (the face normal for a corner
	// without one); an instance's normals are transformed by its matrix and
	// normalised (FacePrimitive::getVertexNormal, include/geometry/matrix.h:122-129)
	const bool shaded = o.smooth || !o.normals.empty();
	for(size_t t = 0; t < o.mats.size(); ++t)
		for(int k = 0; k < 3; ++k)
		{
			const int32_t idx = shaded ? o.corner_normal[3 * t + k] : -1;
			float c[4] = {0.f, 0.f, 0.f, 0.f};
			if(!shaded && k == 0) c[3] = -1.f;
			if(idx >= 0 && 3 * (size_t)idx + 2 < o.normals.size())
			{
				V3 n = load_v3(o.normals, (size_t)idx);
				if(m)
				{
					float r[3] = {0.f, 0.f, 0.f};
					for(int i = 0; i < 3; ++i)
						for(int j = 0; j < 3; ++j) r[i] += m[4 * i + j] * (&n.x)[j];
					n = ymx::normalize(ymx::v3(r[0], r[1], r[2]));
				}
				c[0] = n.x;
				c[1] = n.y;
				c[2] = n.z;
				c[3] = 1.f;
			}
			out.insert(out.end(), c, c + 4);
		}
}

std::vector<float> Scene::cornerNormals(size_t object_id) const
{
	std::vector<float> out;
	if(object_id < objects_.size()) appendCornerNormals(objects_[object_id], nullptr, out);
	return out;
}

bool Scene::smoothMesh(size_t object_id, double angle_d)
{
	if(object_id >= objects_.size()) return false;
	Object &o = objects_[object_id];
	const float angle = (float)angle_d;
	const size_t nv = o.verts.size() / 3, nt = o.tris.size() / 3;
	if(o.normals.size() < 3 * nv) o.normals.resize(3 * nv, 0.f);
	// MeshObject::getAngleSine of corner k of triangle t, and the face normals
	// (ShapePolygon::calculateFaceNormal)
	auto corner_sine = [&](size_t t, int k) {
		const uint32_t *ix = &o.tris[3 * t];
		const V3 p0 = load_v3(o.verts, ix[k]), p1 = load_v3(o.verts, ix[(k + 1) % 3]), p2 = load_v3(o.verts, ix[(k + 2) % 3]);
		return sin_from_vectors(ymx::sub(p1, p0), ymx::sub(p2, p0));
	};
	std::vector<V3> face_n(nt);
	for(size_t t = 0; t < nt; ++t)
	{
		const V3 p0 = load_v3(o.verts, o.tris[3 * t]), p1 = load_v3(o.verts, o.tris[3 * t + 1]),
		         p2 = load_v3(o.verts, o.tris[3 * t + 2]);
		face_n[t] = ymx::normalize(ymx::cross(ymx::sub(p1, p0), ymx::sub(p2, p0)));
	}
	if(angle >= 180.f)
	{
		for(size_t t = 0; t < nt; ++t)
			for(int k = 0; k < 3; ++k)
			{
				const uint32_t v = o.tris[3 * t + k];
				const V3 add = ymx::mul(corner_sine(t, k), face_n[t]);
				for(int c = 0; c < 3; ++c) o.normals[3 * v + c] += (&add.x)[c];
				o.corner_normal[3 * t + k] = (int32_t)v;
			}
		for(size_t v = 0; v < o.normals.size() / 3; ++v)
		{
			const V3 n = ymx::normalize(load_v3(o.normals, v));
			o.normals[3 * v] = n.x;
			o.normals[3 * v + 1] = n.y;
			o.normals[3 * v + 2] = n.z;
		}
	}
	else if(angle > 0.1f)
	{
		// math::cos(math::degToRad(angle)) with the reference's fast cosine
		const float threshold = ymx::fast_cos(angle * 0.017453292519943295f);
		std::vector<std::vector<std::pair<size_t, int>>> point_faces(nv);  // (triangle, corner)
		std::vector<std::vector<float>> point_sines(nv);
		for(size_t t = 0; t < nt; ++t)
			for(int k = 0; k < 3; ++k)
			{
				const uint32_t v = o.tris[3 * t + k];
				point_sines[v].push_back(corner_sine(t, k));
				point_faces[v].emplace_back(t, k);
			}
		for(size_t v = 0; v < nv; ++v)
		{
			std::vector<V3> made;
			std::vector<int32_t> made_idx;
			for(size_t j = 0; j < point_faces[v].size(); ++j)
			{
				const size_t t = point_faces[v][j].first;
				bool smooth = false;
				V3 vn = ymx::mul(point_sines[v][j], face_n[t]);
				for(size_t k = 0; k < point_faces[v].size(); ++k)
				{
					if(k == j) continue;
					const V3 n2 = face_n[point_faces[v][k].first];
					if(ymx::dot(face_n[t], n2) > threshold)
					{
						smooth = true;
						vn = ymx::add(vn, ymx::mul(point_sines[v][k], n2));
					}
				}
				int32_t idx = -1;
				if(smooth)
				{
					vn = ymx::normalize(vn);
					for(size_t m = 0; m < made.size(); ++m)
						if(ymx::dot(vn, made[m]) > 0.999f)
						{
							idx = made_idx[m];
							break;
						}
					if(idx < 0)
					{
						idx = (int32_t)(o.normals.size() / 3);
						made.push_back(vn);
						made_idx.push_back(idx);
						o.normals.insert(o.normals.end(), {vn.x, vn.y, vn.z});
					}
				}
				o.corner_normal[3 * t + point_faces[v][j].second] = idx;
			}
		}
	}
	o.smooth = true;
	modified_ |= ModifiedObjects;
	return true;
}

bool Scene::createMaterial(size_t *id, const std::string &name, const ParamMap &p, std::string *error)
{
	// include/material/material.h:151-166 type names, each type's param defaults
	const std::string type = p.getString("type", "");
	ymx_material m{};
	if(type == "shinydiffusemat") m.type = YMX_MAT_SHINYDIFFUSE;
	else if(type == "glossy") m.type = YMX_MAT_GLOSSY;
	else if(type == "mirror") m.type = YMX_MAT_MIRROR;
	else if(type == "glass") m.type = YMX_MAT_GLASS;
	else if(type == "light_mat") m.type = YMX_MAT_LIGHT;
	else
	{
		if(error) *error = "Scene: material '" + name + "' has unsupported type '" + type + "'";
		return false;
	}
	bool ok = true;
	set3(m.color, p.getColor("color", {1, 1, 1, 1}, &ok));
	m.diffuse_reflect = (float)p.getFloat("diffuse_reflect", type == "glossy" ? 0.0 : 1.0, &ok);
	m.emit = (float)p.getFloat("emit", 0.0, &ok);
	m.receive_shadows = p.getBool("receive_shadows", true, &ok) ? 1u : 0u;
	set3(m.diffuse_color, p.getColor("diffuse_color", {1, 1, 1, 1}, &ok));
	m.glossy_reflect = (float)p.getFloat("glossy_reflect", 1.0, &ok);
	m.exponent = (float)p.getFloat("exponent", 50.0, &ok);
	m.as_diffuse = p.getBool("as_diffuse", true, &ok) ? 1u : 0u;
	m.reflect = (float)p.getFloat("reflect", 1.0, &ok);
	m.ior = (float)p.getFloat("IOR", 1.4, &ok);
	set3(m.filter_color, p.getColor("filter_color", {1, 1, 1, 1}, &ok));
	m.transmit_filter = (float)p.getFloat("transmit_filter", 0.0, &ok);
	set3(m.mirror_color, p.getColor("mirror_color", {1, 1, 1, 1}, &ok));
	m.fake_shadows = p.getBool("fake_shadows", false, &ok) ? 1u : 0u;
	m.power = (float)p.getFloat("power", 1.0, &ok);
	m.double_sided = p.getBool("double_sided", false, &ok) ? 1u : 0u;
	if(!ok)
	{
		if(error) *error = "Scene: material '" + name + "' has a param of the wrong type";
		return false;
	}
	materials_.push_back(m);
	material_names_.push_back(name);
	// Material "visibility" (include/material/material.h:175), ANDed with the
	// object's in every query (include/accelerator/accelerator.h:126-127,138-139)
	material_vis_.push_back(visibility_mask(p.getString("visibility", "normal")));
	if(id) *id = materials_.size() - 1;
	modified_ |= ModifiedMaterials;
	return true;
}

bool Scene::createLight(const std::string &name, const ParamMap &p, std::string *error)
{
	// include/light/light.h:117-128 type names
	const std::string type = p.getString("type", "");
	ymx_light l{};
	bool ok = true;
	if(type == "arealight")
	{
		l.type = YMX_LIGHT_AREA;
		set3(l.corner, p.getVector("corner", {0, 0, 0}, &ok));
		set3(l.point1, p.getVector("point1", {0, 0, 0}, &ok));
		set3(l.point2, p.getVector("point2", {0, 0, 0}, &ok));
	}
	else if(type == "pointlight")
	{
		l.type = YMX_LIGHT_POINT;
		set3(l.from, p.getVector("from", {0, 0, 0}, &ok));
	}
	else if(type == "objectlight")
	{
		l.type = YMX_LIGHT_OBJECT;
		l.double_sided = p.getBool("double_sided", false, &ok) ? 1u : 0u;
	}
	else if(type == "bgPortalLight")  // include/light/light_background_portal.h:58-61
	{
		l.type = YMX_LIGHT_PORTAL;
		l.ibl_clamp_sampling = (float)p.getFloat("ibl_clamp_sampling", 0.0, &ok);
	}
	else
	{
		if(error) *error = "Scene: light '" + name + "' has unsupported type '" + type + "'";
		return false;
	}
	set3(l.color, p.getColor("color", {1, 1, 1, 1}, &ok));
	l.power = (float)p.getFloat("power", 1.0, &ok);
	l.samples = p.getInt("samples", l.type == YMX_LIGHT_PORTAL ? 16 : 4, &ok);
	l.cast_shadows = p.getBool("cast_shadows", true, &ok) ? 1u : 0u;
	const std::string object = p.getString("object_name", "", &ok);
	if(!ok || ((l.type == YMX_LIGHT_OBJECT || l.type == YMX_LIGHT_PORTAL) && object.empty()))
	{
		if(error) *error = "Scene: light '" + name + "' has a missing or mistyped param";
		return false;
	}
	lights_.emplace_back(l, object);
	modified_ |= ModifiedLights;
	return true;
}

// ---------------------------------------------------------------- PathIntegrator

size_t Scene::createInstance()
{
	instances_.emplace_back();
	modified_ |= ModifiedObjects;
	return instances_.size() - 1;
}

bool Scene::addInstanceObject(size_t instance_id, size_t base_object_id)
{
	if(instance_id >= instances_.size() || base_object_id >= objects_.size()) return false;
	instances_[instance_id].bases.push_back({base_object_id, false});
	modified_ |= ModifiedObjects;
	return true;
}

// Scene::addInstanceOfInstance (src/scene/scene.cc:259-264): no check of the
// base beyond its range; an instance of itself is refused at preprocess.
bool Scene::addInstanceOfInstance(size_t instance_id, size_t base_instance_id)
{
	if(instance_id >= instances_.size() || base_instance_id >= instances_.size()) return false;
	instances_[instance_id].bases.push_back({base_instance_id, true});
	modified_ |= ModifiedObjects;
	return true;
}

bool Scene::addInstanceMatrix(size_t instance_id, const double m[16], float time)
{
	(void)time;
	if(instance_id >= instances_.size() || !m) return false;
	std::array<float, 16> f;
	for(int k = 0; k < 16; ++k) f[k] = (float)m[k];  // Matrix4f of the reference
	instances_[instance_id].matrices.push_back(f);
	modified_ |= ModifiedObjects;
	return true;
}

bool Scene::preprocess(uint32_t flags, std::string *error)
{
	if(accel_ && !(flags & (ModifiedObjects | ModifiedAcceleratorParams))) return true;
	flat_verts_.clear();
	flat_tris_.clear();
	flat_mats_.clear();
	flat_vis_.clear();
	ranges_.clear();
	// Scene::preprocess skips base objects (src/scene/scene.cc:322-326); they
	// enter through their instances
	flat_cnorm_.clear();
	bool any_smooth = false;
	for(const Object &o : objects_) any_smooth |= o.smooth || !o.normals.empty();
	auto append = [&](const Object &o, const float *m) {
		if(any_smooth) appendCornerNormals(o, m, flat_cnorm_);
		const uint32_t v0 = (uint32_t)(flat_verts_.size() / 3);
		if(!m) flat_verts_.insert(flat_verts_.end(), o.verts.begin(), o.verts.end());
		else
			for(size_t v = 0; v + 2 < o.verts.size(); v += 3)
				for(int i = 0; i < 3; ++i)
				{
					// include/geometry/matrix.h:131-143: sum_j m[i][j] p[j], then + m[i][3]
					float a = 0.f;
					for(int j = 0; j < 3; ++j) a += m[4 * i + j] * o.verts[v + j];
					a += m[4 * i + 3];
					flat_verts_.push_back(a);
				}
		for(uint32_t i : o.tris) flat_tris_.push_back(v0 + i);
		flat_mats_.insert(flat_mats_.end(), o.mats.begin(), o.mats.end());
		for(uint32_t mat : o.mats) flat_vis_.push_back((uint8_t)(o.visibility & material_vis_[mat]));
	};
	// Objects in creation order, then instances: the primitive order of the
	// reference's list (src/scene/scene.cc:320-341), which leaves out objects
	// of visibility "invisible" (:323).  Their triangles are appended after
	// all others with a zero mask, so no query returns them and the ids before
	// them are the reference's; an object light still finds its mesh there.
	for(const Object &o : objects_)
	{
		if(o.base || o.visibility == 0) continue;
		const uint32_t t0 = (uint32_t)flat_mats_.size();
		append(o, nullptr);
		ranges_[o.name] = {t0, (uint32_t)o.mats.size()};
	}
	// Instance::updatePrimitives in id order (src/geometry/instance.cc,
	// src/scene/scene.cc:327-341): a base instance contributes the list it
	// holds now, updated earlier in this pass when its id is lower, else at the
	// previous preprocess (empty before the first); an instance that names
	// itself is not updated fully and adds nothing to the scene.  A copy's
	// vertices take the product of its chain's matrices, outermost first, in
	// Matrix4f's float order (include/geometry/matrix.h:76-92).
	auto mul4 = [](const std::array<float, 16> &a, const std::array<float, 16> &b) {
		std::array<float, 16> r;
		for(int i = 0; i < 4; ++i)
			for(int k = 0; k < 4; ++k)
			{
				float s = 0.f;
				for(int j = 0; j < 4; ++j) s += a[4 * i + j] * b[4 * j + k];
				r[4 * i + k] = s;
			}
		return r;
	};
	for(size_t id = 0; id < instances_.size(); ++id)
	{
		Instance &in = instances_[id];
		if(in.matrices.size() > 2)
		{
			if(error) *error = "Scene: an instance with three matrices or more (motion blur) is not built";
			return false;
		}
		std::vector<std::pair<size_t, std::vector<size_t>>> prims;
		bool updated = true;
		for(const Instance::Base &b : in.bases)
		{
			if(!b.instance)
			{
				prims.push_back({b.id, {id}});
				continue;
			}
			if(b.id == id)
			{
				updated = false;
				continue;
			}
			for(const auto &e : instances_[b.id].prims)
			{
				std::vector<size_t> chain{id};
				chain.insert(chain.end(), e.second.begin(), e.second.end());
				prims.push_back({e.first, chain});
			}
		}
		in.prims = std::move(prims);
		if(!updated) continue;
		for(const auto &[oid, chain] : in.prims)
		{
			bool have = true;
			for(size_t c : chain) have = have && !instances_[c].matrices.empty();
			if(!have) continue;
			std::array<float, 16> m = instances_[chain[0]].matrices[0];
			for(size_t k = 1; k < chain.size(); ++k) m = mul4(m, instances_[chain[k]].matrices[0]);
			append(objects_[oid], m.data());
		}
	}
	for(const Object &o : objects_)
	{
		if(o.base || o.visibility != 0) continue;
		const uint32_t t0 = (uint32_t)flat_mats_.size();
		append(o, nullptr);
		ranges_[o.name] = {t0, (uint32_t)o.mats.size()};
	}
	ParamMap p = accel_params_;
	if(!p.has("type")) p.setString("type", "mi355x-lbvh");
	// the clustered tree traces the path tracer's rays 4-7 % faster
	if(!p.has("builder")) p.setString("builder", "ploc");
	// with leaves of 2 triangles: the path tracer's best over its frames
	// (occluded room 34.6 ms per 8 spp against 36.4 with 4; Cornell frame
	// 11.9 ms against 12.1 with 1)
	if(!p.has("max_leaf_size")) p.setInt("max_leaf_size", 2);
	std::string aerr;
	std::shared_ptr<const Accelerator> a = Accelerator::factory(p, flat_verts_.data(), (uint32_t)(flat_verts_.size() / 3),
	                                                            flat_tris_.data(), flat_vis_.data(),
	                                                            (uint32_t)flat_mats_.size(), &aerr);
	if(!a)
	{
		if(error) *error = "Scene: " + aerr;
		return false;
	}
	accel_ = std::move(a);
	return true;
}

std::unique_ptr<PathIntegrator> PathIntegrator::create(Scene &scene, const ParamMap &integrator,
                                                       const ParamMap &camera, const ParamMap &film,
                                                       const ParamMap *background, std::string *error)
{
	auto fail = [&](const std::string &msg) -> std::unique_ptr<PathIntegrator> {
		if(error) *error = "PathIntegrator: " + msg;
		return nullptr;
	};
	if(integrator.getString("type", "pathtracing") != "pathtracing") return fail("only the 'pathtracing' type is built");
	if(camera.getString("type", "perspective") != "perspective") return fail("only the 'perspective' camera is built");
	std::string perr;
	if(!scene.preprocess(scene.checkAndClearModifiedFlags(), &perr)) return fail(perr);
	const std::vector<float> &verts = scene.flat_verts_;
	const std::vector<uint32_t> &tris = scene.flat_tris_, &mats = scene.flat_mats_;
	const std::vector<uint8_t> &vis = scene.flat_vis_;
	const auto &ranges = scene.ranges_;
	// an object or portal light covers its object's triangle range
	std::vector<ymx_light> lights;
	for(const auto &[l, object] : scene.lights_)
	{
		ymx_light x = l;
		if(x.type == YMX_LIGHT_OBJECT || x.type == YMX_LIGHT_PORTAL)
		{
			const auto it = ranges.find(object);
			if(it == ranges.end() || it->second.second == 0) return fail("light's object '" + object + "' not found");
			x.first_tri = it->second.first;
			x.n_tris = it->second.second;
		}
		lights.push_back(x);
	}
	std::unique_ptr<PathIntegrator> pi(new PathIntegrator());
	pi->accel_ = scene.accelerator();
	bool ok = true;
	ymx_pt_settings &s = pi->settings_;
	set3(s.from, camera.getVector("from", {0, 1, 0}, &ok));
	set3(s.to, camera.getVector("to", {0, 0, 0}, &ok));
	set3(s.up, camera.getVector("up", {0, 1, 1}, &ok));
	s.focal = (float)camera.getFloat("focal", 1.0, &ok);
	s.aspect_factor = (float)camera.getFloat("aspect_ratio_factor", 1.0, &ok);
	s.resx = camera.getInt("resx", 320, &ok);
	s.resy = camera.getInt("resy", 200, &ok);
	s.path_samples = integrator.getInt("path_samples", 32, &ok);
	s.bounces = integrator.getInt("bounces", 3, &ok);
	s.rr_min_bounces = integrator.getInt("russian_roulette_min_bounces", 0, &ok);
	s.raydepth = integrator.getInt("raydepth", 5, &ok);
	s.transparent_shadows = integrator.getBool("transpShad", false, &ok) ? 1 : 0;
	s.shadow_depth = integrator.getInt("shadowDepth", 4, &ok);
	s.seed = (uint32_t)integrator.getInt("seed", 0, &ok);
	s.aa_samples = film.getInt("AA_minsamples", 1, &ok);
	s.aa_pixelwidth = (float)film.getFloat("AA_pixelwidth", 1.5, &ok);
	// the AA pass params of ImageFilm (include/render/imagefilm.h:212-226)
	s.aa_passes = film.getInt("AA_passes", 1, &ok);
	s.aa_inc_samples = film.getInt("AA_inc_samples", 1, &ok);
	s.aa_threshold = (float)film.getFloat("AA_threshold", 0.05, &ok);
	s.aa_resampled_floor = (float)film.getFloat("AA_resampled_floor", 0.0, &ok);
	s.aa_sample_multiplier_factor = (float)film.getFloat("AA_sample_multiplier_factor", 1.0, &ok);
	s.aa_light_sample_multiplier_factor = (float)film.getFloat("AA_light_sample_multiplier_factor", 1.0, &ok);
	s.aa_detect_color_noise = film.getBool("AA_detect_color_noise", false, &ok) ? 1 : 0;
	const std::string dark = film.getString("AA_dark_detection_type", "none", &ok);
	s.aa_dark_detection_type = dark == "linear" ? 1 : (dark == "curve" ? 2 : 0);
	s.aa_dark_threshold_factor = (float)film.getFloat("AA_dark_threshold_factor", 0.0, &ok);
	s.aa_variance_edge_size = film.getInt("AA_variance_edge_size", 10, &ok);
	s.aa_variance_pixels = film.getInt("AA_variance_pixels", 0, &ok);
	s.aa_clamp_samples = (float)film.getFloat("AA_clamp_samples", 0.0, &ok);
	s.wavefront_mb = integrator.getInt("wavefront_mb", 0, &ok);
	if(background)
	{
		// ConstantBackground: color * power (include/background/background_constant.h:55,
		// include/background/background.h:79)
		s.has_background = 1;
		set3(s.background, background->getColor("color", {0, 0, 0, 1}, &ok));
		const float power = (float)background->getFloat("power", 1.0, &ok);
		for(int k = 0; k < 3; ++k) s.background[k] *= power;
	}
	if(!ok) return fail("a param has the wrong type");
	if(ymx_pt_create(pi->accel_->handle(), verts.data(), (uint32_t)(verts.size() / 3), tris.data(), mats.data(),
	                 vis.data(), (uint32_t)mats.size(), scene.materials_.data(), (uint32_t)scene.materials_.size(),
	                 lights.data(), (uint32_t)lights.size(), &s, &pi->pt_) != YMX_OK)
		return fail(ymx_last_error());
	if(!scene.flat_cnorm_.empty() && ymx_pt_set_normals(pi->pt_, scene.flat_cnorm_.data()) != YMX_OK)
		return fail(ymx_last_error());
	return pi;
}

PathIntegrator::~PathIntegrator()
{
	if(pt_) ymx_pt_destroy(pt_);
}

int PathIntegrator::render(int first, int n)
{
	const int rc = ymx_pt_render(pt_, first, n > 0 ? n : settings_.aa_samples, nullptr);
	if(rc != YMX_OK) error_ = ymx_last_error();
	return rc;
}

int PathIntegrator::renderPasses(std::vector<int> *resampled)
{
	std::vector<int32_t> res((size_t)(settings_.aa_passes > 1 ? settings_.aa_passes : 1), 0);
	const int rc = ymx_pt_render_passes(pt_, res.data(), nullptr);
	if(rc != YMX_OK) error_ = ymx_last_error();
	if(resampled) resampled->assign(res.begin(), res.end());
	return rc;
}

int PathIntegrator::autotune(int trial_samples, ymx_pt_tuning *result)
{
	ymx_pt_tuning t{};
	t.struct_size = sizeof t;
	const int rc = ymx_pt_autotune(pt_, trial_samples, &t, nullptr);
	if(rc != YMX_OK) error_ = ymx_last_error();
	else if(result) *result = t;
	return rc;
}

int PathIntegrator::clearFilm()
{
	const int rc = ymx_pt_clear_film(pt_, nullptr);
	if(rc != YMX_OK) error_ = ymx_last_error();
	return rc;
}

int PathIntegrator::film(std::vector<float> &rgba, std::vector<float> &weight) const
{
	const size_t n = (size_t)settings_.resx * settings_.resy;
	rgba.assign(4 * n, 0.f);
	weight.assign(n, 0.f);
	return ymx_pt_read_film(pt_, rgba.data(), weight.data());
}

int PathIntegrator::saveFilm(const std::string &path, int computer_node, int base_sampling_offset, int sampling_offset)
{
	const int rc = ymx_pt_save_film(pt_, path.c_str(), computer_node, base_sampling_offset, sampling_offset);
	if(rc != YMX_OK) error_ = ymx_last_error();
	return rc;
}

int PathIntegrator::loadFilm(const std::string &path, bool add, int *computer_node, int *base_sampling_offset,
                             int *sampling_offset)
{
	int32_t node = 0, base = 0, off = 0;
	const int rc = ymx_pt_load_film(pt_, path.c_str(), add ? 1 : 0, &node, &base, &off);
	if(rc != YMX_OK)
	{
		error_ = ymx_last_error();
		return rc;
	}
	if(computer_node) *computer_node = node;
	if(base_sampling_offset) *base_sampling_offset = base;
	if(sampling_offset) *sampling_offset = off;
	return YMX_OK;
}

ymx_pt_stats PathIntegrator::stats() const
{
	ymx_pt_stats st{};
	st.struct_size = sizeof st;
	(void)ymx_pt_get_stats(pt_, &st);
	return st;
}

}  // namespace ymx_host
