// ymx_host.hpp -- C++ host layer over the C ABI of libymx.so (include/ymx_*.h).
//
// It mirrors the reference's interfaces for the hot path, so that a
// maintainer's adapter in libYafaRay is a thin forwarding layer and the
// parity tests drive the library the way the reference drives its own code:
//   * ParamMap: the reference's string-keyed parameter maps
//     (include/param/param.h, set through yafaray_setParamMap* in
//     include/public_api/yafaray_c_api.h:108-117);
//   * Accelerator: the "mi355x-lbvh" accelerator type with the reference's
//     three queries intersect / intersectShadow / intersectTransparentShadow and
//     getBound (include/accelerator/accelerator.h:37-87).  Single-ray queries
//     traverse a host copy of the GPU-built tree with the same arithmetic, so
//     they return what the GPU kernels return; batched queries run on the GPU;
//   * Scene + PathIntegrator: a triangle scene built object by object with the
//     reference's material, light, camera and integrator param names, rendered
//     by the wavefront path tracer (the "pathtracing" SurfaceIntegrator,
//     src/integrator/surface/integrator_path_tracer.cc).
// Errors follow the reference's pattern of a result flag plus a logged
// message: factories return nullptr and set `error`, queries never throw.
#pragma once

#include <algorithm>
#include <array>
#include <cfloat>
#include <cstdint>
#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <variant>
#include <vector>

#include "../../include/ymx_accel.h"
#include "../../include/ymx_guide.h"
#include "../../include/ymx_pt.h"

namespace ymx_host {

// ---------------------------------------------------------------- parameters

using Color = std::array<double, 4>;
using Vector = std::array<double, 3>;
using Param = std::variant<bool, int, double, std::string, Color, Vector>;

class ParamMap
{
	public:
		void setBool(const std::string &name, bool b) { map_[name] = b; }
		void setInt(const std::string &name, int i) { map_[name] = i; }
		void setFloat(const std::string &name, double f) { map_[name] = f; }
		void setString(const std::string &name, const std::string &s) { map_[name] = s; }
		void setColor(const std::string &name, double r, double g, double b, double a = 1.0) { map_[name] = Color{r, g, b, a}; }
		void setVector(const std::string &name, double x, double y, double z) { map_[name] = Vector{x, y, z}; }
		void clear() { map_.clear(); }
		bool has(const std::string &name) const { return map_.count(name) != 0; }
		// Typed reads with the reference's defaults; a param of another type is
		// an error reported through `ok` (the reference's ParamResult).
		bool getBool(const std::string &name, bool def, bool *ok = nullptr) const;
		int getInt(const std::string &name, int def, bool *ok = nullptr) const;
		double getFloat(const std::string &name, double def, bool *ok = nullptr) const;
		std::string getString(const std::string &name, const std::string &def, bool *ok = nullptr) const;
		Color getColor(const std::string &name, const Color &def, bool *ok = nullptr) const;
		Vector getVector(const std::string &name, const Vector &def, bool *ok = nullptr) const;

	private:
		std::map<std::string, Param> map_;
};

// ---------------------------------------------------------------- accelerator

// include/geometry/ray.h:30-50: tmax < 0 means unbounded.
struct Ray
{
	float from[3] = {0.f, 0.f, 0.f};
	float dir[3] = {0.f, 0.f, 1.f};
	float tmin = 0.f;
	float tmax = -1.f;
};

constexpr uint32_t kNoHit = YMX_NO_HIT;

// include/accelerator/intersect_data.h:30-39, with the primitive's index in the
// scene's triangle list in place of the Primitive pointer.
struct IntersectData
{
	float t_hit = 0.f;
	float u = 0.f, v = 0.f;
	float t_max = FLT_MAX;
	uint32_t primitive = kNoHit;
	float color[3] = {1.f, 1.f, 1.f};  // transparent shadows: product of the casters' transparencies
	bool isHit() const { return t_hit > 0.f; }
};

struct Bound
{
	float a[3], g[3];  // include/geometry/bound.h: a = min corner, g = max corner
};

class Accelerator
{
	public:
		// Accelerator::factory (src/accelerator/accelerator.cc:44-55) for type
		// "mi355x-lbvh" (params "builder": "lbvh" default or "ploc",
		// "max_leaf_size"): builds the tree on the current GPU over n_tris
		// triangles (verts: 3 floats each; tri_idx: 3 indices each; vis: one
		// visibility mask each or nullptr).  nullptr and `error` on failure.
		static std::unique_ptr<Accelerator> factory(const ParamMap &params, const float *verts, uint32_t n_verts,
		                                            const uint32_t *tri_idx, const uint8_t *vis, uint32_t n_tris,
		                                            std::string *error);
		// The same over faces of 4 indices each, face_idx[4 i + 3] ==
		// YMX_NO_VERTEX for a triangle, else a quad (one primitive per face;
		// ymx_accel_create_faces).
		static std::unique_ptr<Accelerator> factoryFaces(const ParamMap &params, const float *verts, uint32_t n_verts,
		                                                 const uint32_t *face_idx, const uint8_t *vis,
		                                                 uint32_t n_faces, std::string *error);
		~Accelerator();
		Accelerator(const Accelerator &) = delete;
		Accelerator &operator=(const Accelerator &) = delete;

		// include/accelerator/accelerator.h:50-53, one ray on the host copy.
		IntersectData intersect(const Ray &ray, float t_max) const;
		IntersectData intersectShadow(const Ray &ray, float t_max) const;
		// transparent[p]: primitive p's material isTransparent(); transparency
		// (p, dir, t) its getTransparency at the caster's hit ray.from + t * dir,
		// multiplied in the order met front to back.  A hit is an occlusion.
		template <typename F>
		IntersectData intersectTransparentShadow(const Ray &ray, int max_depth, float t_max,
		                                         const std::vector<uint8_t> &transparent, F transparency) const;
		Bound getBound() const { return bound_; }

		// The same queries for device arrays of rays (the wavefront form).
		int intersect(const ymx_ray *d_rays, ymx_hit *d_hits, uint32_t n, void *stream) const;
		int intersectShadow(const ymx_ray *d_rays, uint32_t *d_occluded, uint32_t n, void *stream) const;

		const ymx_accel *handle() const { return acc_; }
		uint32_t numPrimitives() const { return n_tris_; }
		uint32_t maxDepth() const { return max_depth_; }
		// the deepest traversal stack the host queries have used so far (never
		// above maxDepth(), which sizes the GPU kernels' stacks)
		uint32_t maxStackUsed() const { return max_stack_.load(); }

	private:
		Accelerator() = default;
		static std::unique_ptr<Accelerator> create(const ParamMap &params, const float *verts, uint32_t n_verts,
		                                           const uint32_t *idx, bool faces, const uint8_t *vis, uint32_t n,
		                                           std::string *error);
		enum class Mode { Closest, Shadow, Transparent };
		// Traversal of the host copy with the GPU kernels' arithmetic and
		// culling.  Closest: returns the nearest visible hit in [t_min, t_max)
		// (lowest id on exact ties).  Shadow / Transparent: calls `any` for
		// every shadow-casting hit in the window until it returns true.
		IntersectData traverse(const Ray &ray, float t_max, Mode mode,
		                       const std::function<bool(uint32_t, float, float, float)> &any) const;
		ymx_accel *acc_ = nullptr;
		uint32_t n_tris_ = 0;
		std::vector<float> nodes_;  // 16 floats per node
		std::vector<float> tris_;   // 12 floats per triangle record
		std::vector<float> quads_;  // 4 floats (e3, 0) per quad record
		Bound bound_{};
		uint32_t max_depth_ = 0;
		mutable std::atomic<uint32_t> max_stack_{0};
};

// ---------------------------------------------------------------- path tracer

// A triangle scene as the reference's C API builds it: objects of vertices and
// triangles (yafaray_createObject / addVertex / addTriangle), materials and
// lights from param maps (yafaray_createMaterial / yafaray_createLight).
class Scene
{
	public:
		// Returns the object id; params: "visibility" ("normal", "no_shadows",
		// "shadow_only", "invisible"), "is_base_object" (only rendered through
		// instances), as src/geometry/object/object.cc.
		size_t createObject(const std::string &name, const ParamMap &params);
		// Instances (yafaray_createInstance / yafaray_addInstanceObject /
		// yafaray_addInstanceOfInstance / yafaray_addInstanceMatrix,
		// include/public_api/yafaray_c_api.h:227-231): copies of base objects and
		// of other instances' primitives under an object-to-world matrix
		// (row-major m[4][4]; the first one is used, as Instance does without
		// motion blur).  preprocess flattens each copy into the triangle list,
		// transforming its vertices as Matrix4f * Point3f does under the product
		// of the chain's matrices, outermost first (src/geometry/primitive/primitive_face.h,
		// include/geometry/primitive/primitive_instance.h:83-91,
		// include/geometry/matrix.h:76-92,131-143), so hits equal the reference's.
		// Not built: motion blur (three matrices or more).
		size_t createInstance();
		bool addInstanceObject(size_t instance_id, size_t base_object_id);
		bool addInstanceOfInstance(size_t instance_id, size_t base_instance_id);
		bool addInstanceMatrix(size_t instance_id, const double m[16], float time);
		size_t addVertex(size_t object_id, double x, double y, double z);
		// yafaray_addNormal (include/public_api/yafaray_c_api.h:217): the normal
		// of the object's next vertex; triangles added while the object has
		// normals take them at their corners (FacePrimitive::generateInitialVerticesNormalsIndices).
		bool addNormal(size_t object_id, double nx, double ny, double nz);
		bool addTriangle(size_t object_id, size_t a, size_t b, size_t c, size_t material_id);
		// yafaray_smoothObjectMesh (include/public_api/yafaray_c_api.h:224),
		// MeshObject::smoothVerticesNormals (src/geometry/object/object_mesh.cc:150-264):
		// angle >= 180 gives every vertex the sine-weighted mean of its faces'
		// normals; 0.1 < angle < 180 averages, per face corner, the faces whose
		// normals lie within `angle` of the face's (the others keep the face
		// normal at that corner).  The object is then shaded smooth.
		bool smoothMesh(size_t object_id, double angle);
		// The object's shading normals at its triangles' corners in the form of
		// ymx_pt_set_normals (12 floats per triangle; w < 0 in corner 0 for a
		// flat object, w == 0 for a corner without a normal).
		std::vector<float> cornerNormals(size_t object_id) const;
		// Material params (type "shinydiffusemat", "glossy", "mirror", "glass",
		// "light_mat") and light params (type "arealight", "pointlight",
		// "objectlight" with "object_name") use the reference's names.
		bool createMaterial(size_t *id, const std::string &name, const ParamMap &params, std::string *error);
		bool createLight(const std::string &name, const ParamMap &params, std::string *error);
		// "type" "mi355x-lbvh" (the only GPU type; others are refused).
		void setAcceleratorParams(const ParamMap &params)
		{
			accel_params_ = params;
			modified_ |= ModifiedAcceleratorParams;
		}

		// Modified flags, as yafaray_SceneModifiedFlags
		// (include/public_api/yafaray_c_api.h:47-57): the kinds of change since
		// the last checkAndClearModifiedFlags (Scene::checkAndClearSceneModifiedFlags).
		enum : uint32_t
		{
			ModifiedNothing = 0,
			ModifiedObjects = 1u << 0,
			ModifiedLights = 1u << 1,
			ModifiedMaterials = 1u << 2,
			ModifiedAcceleratorParams = 1u << 6,
			ModifiedAll = ModifiedObjects | ModifiedLights | ModifiedMaterials | ModifiedAcceleratorParams
		};
		uint32_t checkAndClearModifiedFlags()
		{
			const uint32_t m = modified_;
			modified_ = ModifiedNothing;
			return m;
		}
		// Scene::preprocess (src/scene/scene.cc:314-362): with modified objects or
		// accelerator params (or no accelerator yet), flatten the objects into the
		// device arrays and rebuild the accelerator (builder "ploc" unless the
		// params name one); otherwise keep both.  false and `error` on failure.
		bool preprocess(uint32_t flags, std::string *error);
		std::shared_ptr<const Accelerator> accelerator() const { return accel_; }

	private:
		friend class PathIntegrator;
		struct Object
		{
			std::string name;
			uint8_t visibility = YMX_VIS_VISIBLE | YMX_VIS_CASTS_SHADOWS;
			bool base = false;
			std::vector<float> verts;
			std::vector<uint32_t> tris, mats;  // local vertex indices
			std::vector<float> normals;        // vertex normals, 3 floats each (may outgrow verts when smoothed)
			std::vector<int32_t> corner_normal;  // per triangle corner: index into normals, -1 for none
			bool smooth = false;               // isSmooth(): shaded with interpolated normals
		};
		std::vector<Object> objects_;
		struct Instance
		{
			struct Base
			{
				size_t id;
				bool instance;  // id of another instance, else of an object
			};
			std::vector<Base> bases;  // in the order they were added
			std::vector<std::array<float, 16>> matrices;
			// Instance::updatePrimitives's list at the last preprocess: per entry a
			// base object and the chain of instances it is copied through, this one first
			std::vector<std::pair<size_t, std::vector<size_t>>> prims;
		};
		std::vector<Instance> instances_;
		std::vector<ymx_material> materials_;
		std::vector<std::string> material_names_;
		std::vector<uint8_t> material_vis_;  // "visibility" of each material
		std::vector<std::pair<ymx_light, std::string>> lights_;  // object lights keep their object name
		ParamMap accel_params_;
		uint32_t modified_ = ModifiedAll;
		// state of the last preprocess
		std::shared_ptr<const Accelerator> accel_;
		std::vector<float> flat_verts_;
		std::vector<uint32_t> flat_tris_, flat_mats_;
		std::vector<uint8_t> flat_vis_;
		std::vector<float> flat_cnorm_;  // 12 floats per triangle (ymx_pt_set_normals), empty if none is smooth
		std::map<std::string, std::pair<uint32_t, uint32_t>> ranges_;  // object name -> (first triangle, count)
		// appends the corner normals of o (instanced under the row-major matrix m, or nullptr)
		static void appendCornerNormals(const Object &o, const float *m, std::vector<float> &out);
};

class PathIntegrator
{
	public:
		// "pathtracing" integrator params ("bounces", "path_samples",
		// "russian_roulette_min_bounces", "raydepth", "transpShad",
		// "shadowDepth"), camera params ("type" "perspective", "from", "to",
		// "up", "resx", "resy", "focal", "aspect_ratio_factor"), film params
		// ("AA_minsamples", "AA_pixelwidth") and an optional constant
		// background "color".  Builds the accelerator and the wavefront state.
		// Preprocesses the scene with its modified flags first (the accelerator is
		// shared with the scene and rebuilt only when its objects changed).
		static std::unique_ptr<PathIntegrator> create(Scene &scene, const ParamMap &integrator,
		                                              const ParamMap &camera, const ParamMap &film,
		                                              const ParamMap *background, std::string *error);
		~PathIntegrator();
		PathIntegrator(const PathIntegrator &) = delete;
		PathIntegrator &operator=(const PathIntegrator &) = delete;

		// Renders AA samples [first, first + n) of every pixel into the film
		// (n <= 0: all of "AA_minsamples"); synchronous.
		int render(int first = 0, int n = 0);
		// TiledIntegrator::render: pass 0 of "AA_minsamples", then the adaptive
		// passes of the film params "AA_passes", "AA_inc_samples", "AA_threshold",
		// "AA_resampled_floor", "AA_sample_multiplier_factor",
		// "AA_light_sample_multiplier_factor", "AA_detect_color_noise",
		// "AA_dark_detection_type", "AA_dark_threshold_factor",
		// "AA_variance_edge_size", "AA_variance_pixels" (and "AA_clamp_samples"
		// for every sample); resampled: the pixels of each pass.  Returns once
		// the last pass is queued (film() waits for it).
		int renderPasses(std::vector<int> *resampled = nullptr);
		// ymx_pt_autotune: times short renders under each launch configuration
		// and keeps the fastest (the film is left as it was); trial_samples 0:
		// the samples one wavefront holds.  Synchronous.
		int autotune(int trial_samples = 0, ymx_pt_tuning *result = nullptr);
		int clearFilm();
		// The film as ImageFilm holds it: colour and alpha sums weighted by the
		// filter, and the weight sums; pixel (x, y) at index y * resx + x.
		int film(std::vector<float> &rgba, std::vector<float> &weight) const;
		// ImageFilm::imageFilmSave / imageFilmLoad (src/render/imagefilm.cc:910-1176):
		// the YAF_FILMv4_0_0 film file; load with add merges a render node's film.
		int saveFilm(const std::string &path, int computer_node = 0, int base_sampling_offset = 0,
		             int sampling_offset = 0);
		int loadFilm(const std::string &path, bool add, int *computer_node = nullptr,
		             int *base_sampling_offset = nullptr, int *sampling_offset = nullptr);
		int resx() const { return settings_.resx; }
		int resy() const { return settings_.resy; }
		ymx_pt_stats stats() const;
		const std::string &lastError() const { return error_; }

	private:
		PathIntegrator() = default;
		std::shared_ptr<const Accelerator> accel_;
		ymx_pt *pt_ = nullptr;
		ymx_pt_settings settings_{};
		std::string error_;
};

// ---------------------------------------------------------------- inline parts

template <typename F>
IntersectData Accelerator::intersectTransparentShadow(const Ray &ray, int max_depth, float t_max,
                                                      const std::vector<uint8_t> &transparent, F transparency) const
{
	IntersectData d;
	std::vector<std::pair<float, uint32_t>> kept;
	bool occluded = false;
	(void)traverse(ray, t_max, Mode::Transparent, [&](uint32_t prim, float t, float u, float v) {
		if(prim < transparent.size() && transparent[prim])
		{
			for(const auto &k : kept)
				if(k.second == prim) return false;
			if((int)kept.size() < max_depth)
			{
				kept.emplace_back(t, prim);
				return false;
			}
		}
		occluded = true;
		d.t_hit = t;
		d.t_max = t;
		d.u = u;
		d.v = v;
		d.primitive = prim;
		return true;  // stop
	});
	if(occluded) return d;
	std::sort(kept.begin(), kept.end());
	for(const auto &k : kept)
	{
		const std::array<float, 3> c = transparency(k.second, ray.dir, k.first);
		for(int i = 0; i < 3; ++i) d.color[i] *= c[i];
	}
	return d;
}

}  // namespace ymx_host
