// test_host.cpp -- the C++ host layer (libyafaray_amd/host/ymx_host.hpp)
// driven the way the reference's clients drive libYafaRay: param maps for the
// accelerator, materials, lights, camera and integrator.  Checks against the
// CPU oracle (oracle/), which it links as the checker only:
//   accel_closest / accel_shadow: single-ray host queries, the GPU batch
//     queries and the oracle's brute force agree bit for bit;
//   accel_transparent: host, GPU and the kd-tree port agree on occlusion and
//     on the transparent casters passed;
//   pathtracer: a scene built object by object renders the oracle's film;
//   errors: unknown types and missing objects are refused with a message.
// Usage: test_host [--cpu]   (--cpu: only the checks that need no GPU)
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <array>
#include <algorithm>
#include <cstdlib>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../../libyafaray_amd/csrc/ray_tri.h"
#include "../../libyafaray_amd/host/ymx_host.hpp"
#include "../../oracle/ora_pt.h"
#include "../../oracle/oracle.h"

using namespace ymx_host;

static int g_fail = 0;
#define CHECK(cond, ...)                   \
	do                                     \
	{                                      \
		if(!(cond))                        \
		{                                  \
			printf("  FAIL: " __VA_ARGS__); \
			printf("\n");                  \
			++g_fail;                      \
		}                                  \
	} while(0)

struct Soup
{
	std::vector<float> verts;
	std::vector<uint32_t> tris;
};

static Soup random_soup(uint32_t n, uint32_t seed)
{
	std::mt19937 rng(seed);
	std::uniform_real_distribution<float> u(0.f, 1.f);
	Soup s;
	for(uint32_t i = 0; i < n; ++i)
	{
		const float c[3] = {u(rng), u(rng), u(rng)};
		const float e = 0.02f + 0.1f * u(rng);
		for(int k = 0; k < 3; ++k)
		{
			for(int a = 0; a < 3; ++a) s.verts.push_back(c[a] + e * (u(rng) - 0.5f));
			s.tris.push_back(3 * i + k);
		}
	}
	return s;
}

static std::vector<Ray> random_rays(uint32_t n, uint32_t seed, bool shadow)
{
	std::mt19937 rng(seed);
	std::uniform_real_distribution<float> u(0.f, 1.f);
	std::normal_distribution<float> g;
	std::vector<Ray> rays(n);
	for(Ray &r : rays)
	{
		float d[3] = {g(rng), g(rng), g(rng)};
		const float l = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
		for(int k = 0; k < 3; ++k)
		{
			r.from[k] = -0.2f + 1.4f * u(rng);
			r.dir[k] = d[k] / l;
		}
		r.tmin = 1e-3f * u(rng);
		r.tmax = shadow ? 0.05f + u(rng) : (u(rng) < 0.5f ? -1.f : 0.1f + 2.f * u(rng));
	}
	return rays;
}

static float tmax_of(const Ray &r) { return r.tmax >= 0.f ? r.tmax : FLT_MAX; }

static void to_device(const std::vector<Ray> &rays, ymx_ray **d)
{
	std::vector<ymx_ray> h(rays.size());
	for(size_t i = 0; i < rays.size(); ++i)
	{
		memcpy(h[i].org, rays[i].from, 12);
		memcpy(h[i].dir, rays[i].dir, 12);
		h[i].tmin = rays[i].tmin;
		h[i].tmax = rays[i].tmax;
	}
	(void)hipMalloc(d, sizeof(ymx_ray) * h.size());
	(void)hipMemcpy(*d, h.data(), sizeof(ymx_ray) * h.size(), hipMemcpyHostToDevice);
}

static std::vector<ora_ray> to_oracle(const std::vector<Ray> &rays)
{
	std::vector<ora_ray> o(rays.size());
	for(size_t i = 0; i < rays.size(); ++i)
	{
		memcpy(o[i].org, rays[i].from, 12);
		memcpy(o[i].dir, rays[i].dir, 12);
		o[i].tmin = rays[i].tmin;
		o[i].tmax = rays[i].tmax;
	}
	return o;
}

static const char *g_builder = "lbvh";

static std::unique_ptr<Accelerator> build(const Soup &s, const std::vector<uint8_t> &vis)
{
	ParamMap p;
	p.setString("type", "mi355x-lbvh");
	p.setString("builder", g_builder);
	std::string err;
	auto a = Accelerator::factory(p, s.verts.data(), (uint32_t)(s.verts.size() / 3), s.tris.data(), vis.data(),
	                              (uint32_t)(s.tris.size() / 3), &err);
	CHECK(a != nullptr, "factory: %s", err.c_str());
	return a;
}

static void test_accel_closest()
{
	const Soup s = random_soup(4000, 1);
	const uint32_t nt = (uint32_t)(s.tris.size() / 3);
	std::vector<uint8_t> vis(nt, 3);
	for(uint32_t i = 0; i < nt; i += 7) vis[i] = 2;  // shadow_only: not visible to closest hit
	auto a = build(s, vis);
	if(!a) return;
	const std::vector<Ray> rays = random_rays(20000, 2, false);
	std::vector<ymx_hit> gpu(rays.size());
	ymx_ray *d_rays;
	ymx_hit *d_hits;
	to_device(rays, &d_rays);
	(void)hipMalloc(&d_hits, sizeof(ymx_hit) * rays.size());
	CHECK(a->intersect(d_rays, d_hits, (uint32_t)rays.size(), nullptr) == YMX_OK, "batch intersect");
	(void)hipMemcpy(gpu.data(), d_hits, sizeof(ymx_hit) * rays.size(), hipMemcpyDeviceToHost);
	std::vector<ora_hit> ora(rays.size());
	const std::vector<ora_ray> orays = to_oracle(rays);
	ora_closest_brute(s.verts.data(), s.tris.data(), vis.data(), nt, orays.data(), (uint32_t)rays.size(), ora.data(), 8);
	size_t hits = 0, bad_host = 0, bad_ora = 0;
	for(size_t i = 0; i < rays.size(); ++i)
	{
		const IntersectData h = a->intersect(rays[i], tmax_of(rays[i]));
		hits += h.isHit();
		if(h.primitive != gpu[i].prim || (h.isHit() && (h.t_hit != gpu[i].t || h.u != gpu[i].u || h.v != gpu[i].v)))
			++bad_host;
		if(h.primitive != ora[i].prim || (h.isHit() && h.t_hit != ora[i].t)) ++bad_ora;
	}
	printf("  closest: %zu rays, %zu hits, host != gpu: %zu, host != oracle: %zu\n", rays.size(), hits, bad_host, bad_ora);
	CHECK(hits > rays.size() / 10 && bad_host == 0 && bad_ora == 0, "closest-hit mismatch");
	(void)hipFree(d_rays);
	(void)hipFree(d_hits);
}

static void test_accel_shadow()
{
	const Soup s = random_soup(4000, 3);
	const uint32_t nt = (uint32_t)(s.tris.size() / 3);
	std::vector<uint8_t> vis(nt, 3);
	for(uint32_t i = 0; i < nt; i += 5) vis[i] = 1;  // no_shadows
	auto a = build(s, vis);
	if(!a) return;
	const std::vector<Ray> rays = random_rays(20000, 4, true);
	ymx_ray *d_rays;
	uint32_t *d_occ;
	to_device(rays, &d_rays);
	const size_t words = (rays.size() + 31) / 32;
	(void)hipMalloc(&d_occ, 4 * words);
	CHECK(a->intersectShadow(d_rays, d_occ, (uint32_t)rays.size(), nullptr) == YMX_OK, "batch shadow");
	std::vector<uint32_t> occ(words);
	(void)hipMemcpy(occ.data(), d_occ, 4 * words, hipMemcpyDeviceToHost);
	std::vector<uint8_t> ora(rays.size());
	const std::vector<ora_ray> orays = to_oracle(rays);
	ora_shadow_brute(s.verts.data(), s.tris.data(), vis.data(), nt, orays.data(), (uint32_t)rays.size(), ora.data(), 8);
	size_t n_occ = 0, bad = 0;
	for(size_t i = 0; i < rays.size(); ++i)
	{
		const bool h = a->intersectShadow(rays[i], tmax_of(rays[i])).isHit();
		const bool g = (occ[i >> 5] >> (i & 31)) & 1u;
		n_occ += h;
		bad += (h != g) || (h != (bool)ora[i]);
	}
	printf("  shadow: %zu rays, %zu occluded, mismatches %zu\n", rays.size(), n_occ, bad);
	CHECK(n_occ > rays.size() / 20 && n_occ < rays.size() && bad == 0, "shadow mismatch");
	(void)hipFree(d_rays);
	(void)hipFree(d_occ);
}

static void test_accel_transparent()
{
	const Soup s = random_soup(3000, 5);
	const uint32_t nt = (uint32_t)(s.tris.size() / 3);
	std::vector<uint8_t> vis(nt, 3), transp(nt, 0);
	std::mt19937 rng(6);
	for(uint32_t i = 0; i < nt; ++i) transp[i] = (rng() % 10) < 7;
	auto a = build(s, vis);
	if(!a) return;
	const int depth = 3;
	const std::vector<Ray> rays = random_rays(8000, 7, true);
	// GPU query
	ymx_ray *d_rays;
	to_device(rays, &d_rays);
	std::vector<uint32_t> mask((nt + 31) / 32 + 1, 0u);
	for(uint32_t i = 0; i < nt; ++i)
		if(transp[i]) mask[i >> 5] |= 1u << (i & 31);
	uint32_t *d_mask, *d_occ, *d_cnt, *d_hits;
	(void)hipMalloc(&d_mask, 4 * mask.size());
	(void)hipMemcpy(d_mask, mask.data(), 4 * mask.size(), hipMemcpyHostToDevice);
	const size_t words = (rays.size() + 31) / 32;
	(void)hipMalloc(&d_occ, 4 * words);
	(void)hipMalloc(&d_cnt, 4 * rays.size());
	(void)hipMalloc(&d_hits, 8 * depth * rays.size());
	CHECK(ymx_accel_intersect_transparent(a->handle(), d_rays, d_mask, depth, d_occ, d_cnt, d_hits,
	                                      (uint32_t)rays.size(), nullptr) == YMX_OK, "batch transparent");
	std::vector<uint32_t> occ(words), cnt(rays.size()), hits(2 * depth * rays.size());
	(void)hipMemcpy(occ.data(), d_occ, 4 * words, hipMemcpyDeviceToHost);
	(void)hipMemcpy(cnt.data(), d_cnt, 4 * rays.size(), hipMemcpyDeviceToHost);
	(void)hipMemcpy(hits.data(), d_hits, 8 * depth * rays.size(), hipMemcpyDeviceToHost);
	// kd-tree port
	ora_kdtree *kd = ora_kd_build(s.verts.data(), s.tris.data(), vis.data(), nt);
	const std::vector<ora_ray> orays = to_oracle(rays);
	size_t bad = 0, passed = 0, n_occ = 0;
	for(size_t i = 0; i < rays.size(); ++i)
	{
		std::set<uint32_t> met;
		auto tr = [&](uint32_t p, const float *, float) {
			met.insert(p);
			return std::array<float, 3>{0.5f, 0.75f, 1.f};
		};
		const IntersectData h = a->intersectTransparentShadow(rays[i], depth, tmax_of(rays[i]), transp, tr);
		const bool g = (occ[i >> 5] >> (i & 31)) & 1u;
		uint32_t prims[64];
		int pc = 0;
		const int o = ora_kd_transparent1(kd, &orays[i], tmax_of(rays[i]), transp.data(), depth, prims, &pc);
		n_occ += h.isHit();
		if(h.isHit() != g || h.isHit() != (bool)o)
		{
			++bad;
			continue;
		}
		if(h.isHit()) continue;
		std::set<uint32_t> gset, oset(prims, prims + pc);
		for(uint32_t k = 0; k < cnt[i]; ++k) gset.insert(hits[2 * (depth * i + k)]);
		if(met != gset || met != oset) ++bad;
		if(!met.empty())
		{
			++passed;
			const float expect = powf(0.5f, (float)met.size());
			if(fabsf(h.color[0] - expect) > 1e-6f * expect) ++bad;
		}
	}
	ora_kd_free(kd);
	printf("  transparent: %zu rays, %zu occluded, %zu pass casters, mismatches %zu\n", rays.size(), n_occ, passed, bad);
	CHECK(bad == 0 && passed > 50 && n_occ > 50, "transparent-shadow mismatch");
	(void)hipFree(d_rays);
	(void)hipFree(d_mask);
	(void)hipFree(d_occ);
	(void)hipFree(d_cnt);
	(void)hipFree(d_hits);
}

// A Cornell-style box built through the Scene API, with the oracle fed the
// same flattened arrays.
static void quad(Scene &sc, size_t obj, size_t mat, const float c[3], const float e1[3], const float e2[3],
                 std::vector<float> &verts, std::vector<uint32_t> &tris, std::vector<uint32_t> &mats)
{
	const float p[4][3] = {{c[0], c[1], c[2]},
	                       {c[0] + e1[0], c[1] + e1[1], c[2] + e1[2]},
	                       {c[0] + e1[0] + e2[0], c[1] + e1[1] + e2[1], c[2] + e1[2] + e2[2]},
	                       {c[0] + e2[0], c[1] + e2[1], c[2] + e2[2]}};
	size_t id[4];
	const uint32_t base = (uint32_t)(verts.size() / 3);
	for(int k = 0; k < 4; ++k)
	{
		id[k] = sc.addVertex(obj, p[k][0], p[k][1], p[k][2]);
		verts.insert(verts.end(), p[k], p[k] + 3);
	}
	(void)id;
	const size_t nv = verts.size() / 3;
	const size_t local0 = id[0];
	sc.addTriangle(obj, local0, local0 + 1, local0 + 2, mat);
	sc.addTriangle(obj, local0, local0 + 2, local0 + 3, mat);
	tris.insert(tris.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
	mats.insert(mats.end(), {(uint32_t)mat, (uint32_t)mat});
	(void)nv;
}

static void test_pathtracer()
{
	Scene sc;
	std::string err;
	std::vector<ora_material> omats;
	auto material = [&](const std::string &name, double r, double g, double b) {
		ParamMap p;
		p.setString("type", "shinydiffusemat");
		p.setColor("color", r, g, b);
		size_t id = 0;
		CHECK(sc.createMaterial(&id, name, p, &err), "material: %s", err.c_str());
		ora_material m{};
		m.type = ORA_MAT_SHINYDIFFUSE;
		m.color[0] = (float)r;
		m.color[1] = (float)g;
		m.color[2] = (float)b;
		m.diffuse_reflect = 1.f;
		m.receive_shadows = 1;
		m.diffuse_color[0] = m.diffuse_color[1] = m.diffuse_color[2] = 1.f;
		m.glossy_reflect = 1.f;
		m.exponent = 50.f;
		m.as_diffuse = 1;
		m.reflect = 1.f;
		m.ior = 1.4f;
		m.filter_color[0] = m.filter_color[1] = m.filter_color[2] = 1.f;
		m.mirror_color[0] = m.mirror_color[1] = m.mirror_color[2] = 1.f;
		m.power = 1.f;
		omats.push_back(m);
		return id;
	};
	const size_t white = material("white", 0.7, 0.7, 0.7), red = material("red", 0.7, 0.1, 0.1),
	             green = material("green", 0.1, 0.7, 0.1), grey = material("grey", 0.4, 0.4, 0.4);
	std::vector<float> verts;
	std::vector<uint32_t> tris, mats;
	ParamMap normal;
	const size_t walls = sc.createObject("walls", normal);
	const float o[3] = {0, 0, 0}, x[3] = {1, 0, 0}, y[3] = {0, 1, 0}, z[3] = {0, 0, 1};
	const float top[3] = {0, 0, 1}, back[3] = {0, 1, 0}, right[3] = {1, 0, 0};
	quad(sc, walls, white, o, y, x, verts, tris, mats);       // floor
	quad(sc, walls, white, top, x, y, verts, tris, mats);     // ceiling
	quad(sc, walls, white, back, z, x, verts, tris, mats);    // back wall
	quad(sc, walls, red, o, z, y, verts, tris, mats);         // left wall
	quad(sc, walls, green, right, y, z, verts, tris, mats);   // right wall
	const size_t box = sc.createObject("box", normal);
	const float b0[3] = {0.3f, 0.4f, 1.f / 512}, bx[3] = {0.3f, 0, 0}, by[3] = {0, 0.3f, 0}, bz[3] = {0, 0, 0.4f};
	const float b1[3] = {0.3f, 0.4f, 0.4f + 1.f / 512};
	quad(sc, box, grey, b1, bx, by, verts, tris, mats);  // box top
	quad(sc, box, grey, b0, bz, bx, verts, tris, mats);  // box front
	quad(sc, box, grey, b0, by, bz, verts, tris, mats);  // box left
	ParamMap al;
	al.setString("type", "arealight");
	al.setVector("corner", 0.4, 0.4, 0.999);
	al.setVector("point1", 0.4, 0.6, 0.999);
	al.setVector("point2", 0.6, 0.4, 0.999);
	al.setColor("color", 1, 1, 1);
	al.setFloat("power", 10.0 / M_PI);
	al.setInt("samples", 4);
	CHECK(sc.createLight("area", al, &err), "light: %s", err.c_str());
	ParamMap pl;
	pl.setString("type", "pointlight");
	pl.setVector("from", 0.8, 0.2, 0.6);
	pl.setColor("color", 1, 0.9, 0.8);
	pl.setFloat("power", 0.05);
	CHECK(sc.createLight("point", pl, &err), "light: %s", err.c_str());
	ParamMap integ, cam, film;
	integ.setString("type", "pathtracing");
	integ.setInt("bounces", 4);
	integ.setInt("path_samples", 1);
	integ.setInt("russian_roulette_min_bounces", 2);
	cam.setString("type", "perspective");
	cam.setVector("from", 0.5, -1.6, 0.5);
	cam.setVector("to", 0.5, 0.5, 0.5);
	cam.setVector("up", 0.5, -1.6, 1.5);
	cam.setFloat("focal", 1.0 / (2.0 * tan(20.0 * M_PI / 180.0)));
	cam.setInt("resx", 40);
	cam.setInt("resy", 32);
	film.setInt("AA_minsamples", 4);
	auto pi = PathIntegrator::create(sc, integ, cam, film, nullptr, &err);
	CHECK(pi != nullptr, "create: %s", err.c_str());
	if(!pi) return;
	CHECK(pi->clearFilm() == YMX_OK && pi->render() == YMX_OK, "render: %s", pi->lastError().c_str());
	std::vector<float> rgba, w;
	CHECK(pi->film(rgba, w) == YMX_OK, "film");
	// the oracle on the same arrays
	ora_light lights[2] = {};
	const float corner[3] = {0.4f, 0.4f, 0.999f}, p1[3] = {0.4f, 0.6f, 0.999f}, p2[3] = {0.6f, 0.4f, 0.999f};
	lights[0].type = ORA_LIGHT_AREA;
	memcpy(lights[0].corner, corner, 12);
	memcpy(lights[0].point1, p1, 12);
	memcpy(lights[0].point2, p2, 12);
	lights[0].color[0] = lights[0].color[1] = lights[0].color[2] = 1.f;
	lights[0].power = (float)(10.0 / M_PI);
	lights[0].samples = 4;
	lights[0].cast_shadows = 1;
	lights[1].type = ORA_LIGHT_POINT;
	lights[1].from[0] = 0.8f;
	lights[1].from[1] = 0.2f;
	lights[1].from[2] = 0.6f;
	lights[1].color[0] = 1.f;
	lights[1].color[1] = 0.9f;
	lights[1].color[2] = 0.8f;
	lights[1].power = 0.05f;
	lights[1].samples = 4;
	lights[1].cast_shadows = 1;
	ora_pt_settings st{};
	const float from[3] = {0.5f, -1.6f, 0.5f}, to[3] = {0.5f, 0.5f, 0.5f}, up[3] = {0.5f, -1.6f, 1.5f};
	memcpy(st.from, from, 12);
	memcpy(st.to, to, 12);
	memcpy(st.up, up, 12);
	st.focal = (float)(1.0 / (2.0 * tan(20.0 * M_PI / 180.0)));
	st.aspect_factor = 1.f;
	st.resx = 40;
	st.resy = 32;
	st.path_samples = 1;
	st.bounces = 4;
	st.rr_min_bounces = 2;
	st.aa_samples = 4;
	st.aa_pixelwidth = 1.5f;
	st.raydepth = 5;
	st.shadow_depth = 4;
	std::vector<uint8_t> vis(mats.size(), 3);
	ora_kdtree *kd = ora_kd_build(verts.data(), tris.data(), vis.data(), (uint32_t)mats.size());
	std::vector<float> orgba(4 * 40 * 32, 0.f), ow(40 * 32, 0.f);
	ora_pt_render(kd, verts.data(), tris.data(), mats.data(), vis.data(), omats.data(), lights, 2, &st, orgba.data(),
	              ow.data(), 8);
	ora_kd_free(kd);
	size_t bad_w = 0, bad_c = 0;
	double mean = 0.0;
	for(size_t i = 0; i < w.size(); ++i)
	{
		bad_w += w[i] != ow[i];
		for(int k = 0; k < 3; ++k)
		{
			const float g = rgba[4 * i + k] / w[i], r = orgba[4 * i + k] / ow[i];
			mean += r;
			if(fabsf(g - r) > 1e-4f * fabsf(r) + 1e-6f) ++bad_c;
		}
	}
	printf("  pathtracer: 40x32x4, mean %.5f, weight mismatches %zu, colour values off %zu\n", mean / (3.0 * w.size()),
	       bad_w, bad_c);
	CHECK(bad_w == 0 && bad_c <= 3 && mean > 0.0, "film differs from the oracle");
	// the film file: save, then load into a second integrator (replace, then merge)
	char path[] = "/tmp/ymx_host_film_XXXXXX";
	const int fd = mkstemp(path);
	CHECK(fd >= 0, "temporary film file");
	if(fd < 0) return;
	close(fd);
	CHECK(pi->saveFilm(path, 1, 2, 4) == YMX_OK, "saveFilm: %s", pi->lastError().c_str());
	auto pi2 = PathIntegrator::create(sc, integ, cam, film, nullptr, &err);
	CHECK(pi2 != nullptr, "second integrator: %s", err.c_str());
	if(pi2)
	{
		int node = -1, base = -1, off = -1;
		CHECK(pi2->clearFilm() == YMX_OK && pi2->loadFilm(path, false, &node, &base, &off) == YMX_OK,
		      "loadFilm: %s", pi2->lastError().c_str());
		std::vector<float> rgba2, w2;
		CHECK(pi2->film(rgba2, w2) == YMX_OK && rgba2 == rgba && w2 == w && node == 1 && base == 2 && off == 4,
		      "loaded film differs");
		CHECK(pi2->loadFilm(path, true) == YMX_OK && pi2->film(rgba2, w2) == YMX_OK, "merge");
		bool doubled = true;
		for(size_t i = 0; i < w.size(); ++i) doubled = doubled && w2[i] == w[i] + w[i];
		CHECK(doubled, "merged weights are not the sum");
	}
	unlink(path);
}

// BackgroundPortalLight through the host layer (type "bgPortalLight",
// src/light/light_background_portal.cc): an open-backed box whose back is an
// invisible "window" object, the portal's mesh, under a constant background of
// colour (0.9, 0.95, 1) and power 2; film against the oracle on the same arrays.
static void test_portal_light()
{
	Scene sc;
	std::string err;
	ParamMap mp;
	mp.setString("type", "shinydiffusemat");
	mp.setColor("color", 0.7, 0.7, 0.7);
	size_t white = 0;
	CHECK(sc.createMaterial(&white, "white", mp, &err), "material: %s", err.c_str());
	ora_material om{};
	om.type = ORA_MAT_SHINYDIFFUSE;
	om.color[0] = om.color[1] = om.color[2] = 0.7f;
	om.diffuse_reflect = 1.f;
	om.receive_shadows = 1;
	om.diffuse_color[0] = om.diffuse_color[1] = om.diffuse_color[2] = 1.f;
	om.glossy_reflect = 1.f;
	om.exponent = 50.f;
	om.as_diffuse = 1;
	om.reflect = 1.f;
	om.ior = 1.4f;
	om.filter_color[0] = om.filter_color[1] = om.filter_color[2] = 1.f;
	om.mirror_color[0] = om.mirror_color[1] = om.mirror_color[2] = 1.f;
	om.power = 1.f;
	std::vector<float> verts;
	std::vector<uint32_t> tris, mats;
	ParamMap normal;
	const size_t walls = sc.createObject("walls", normal);
	const float o[3] = {0, 0, 0}, x[3] = {1, 0, 0}, y[3] = {0, 1, 0}, z[3] = {0, 0, 1};
	const float top[3] = {0, 0, 1}, back[3] = {0, 1, 0}, right[3] = {1, 0, 0};
	quad(sc, walls, white, o, y, x, verts, tris, mats);      // floor
	quad(sc, walls, white, top, x, y, verts, tris, mats);    // ceiling
	quad(sc, walls, white, o, z, y, verts, tris, mats);      // left wall
	quad(sc, walls, white, right, y, z, verts, tris, mats);  // right wall
	ParamMap inv;
	inv.setString("visibility", "invisible");
	const size_t window = sc.createObject("window", inv);
	quad(sc, window, white, back, x, z, verts, tris, mats);  // the open back, face normal -y
	ParamMap pl;
	pl.setString("type", "bgPortalLight");
	pl.setString("object_name", "window");
	pl.setFloat("power", 2.0);
	pl.setInt("samples", 4);
	CHECK(sc.createLight("portal", pl, &err), "light: %s", err.c_str());
	ParamMap integ, cam, film, bg;
	integ.setString("type", "pathtracing");
	integ.setInt("bounces", 3);
	integ.setInt("path_samples", 1);
	integ.setInt("russian_roulette_min_bounces", 2);
	cam.setString("type", "perspective");
	cam.setVector("from", 0.5, -1.6, 0.5);
	cam.setVector("to", 0.5, 0.5, 0.5);
	cam.setVector("up", 0.5, -1.6, 1.5);
	cam.setFloat("focal", 1.0 / (2.0 * tan(20.0 * M_PI / 180.0)));
	cam.setInt("resx", 32);
	cam.setInt("resy", 24);
	film.setInt("AA_minsamples", 4);
	bg.setString("type", "constant");
	bg.setColor("color", 0.45, 0.475, 0.5);
	bg.setFloat("power", 2.0);
	CHECK(PathIntegrator::create(sc, integ, cam, film, nullptr, &err) == nullptr && !err.empty(),
	      "portal light without a background refused");
	err.clear();
	auto pi = PathIntegrator::create(sc, integ, cam, film, &bg, &err);
	CHECK(pi != nullptr, "create: %s", err.c_str());
	if(!pi) return;
	CHECK(pi->clearFilm() == YMX_OK && pi->render() == YMX_OK, "render: %s", pi->lastError().c_str());
	std::vector<float> rgba, w;
	CHECK(pi->film(rgba, w) == YMX_OK, "film");
	ora_light light{};
	light.type = ORA_LIGHT_PORTAL;
	light.power = 2.f;
	light.samples = 4;
	light.cast_shadows = 1;
	light.first_tri = 8;  // the invisible window follows the walls' 8 triangles
	light.n_tris = 2;
	ora_pt_settings st{};
	const float from[3] = {0.5f, -1.6f, 0.5f}, to[3] = {0.5f, 0.5f, 0.5f}, up[3] = {0.5f, -1.6f, 1.5f};
	memcpy(st.from, from, 12);
	memcpy(st.to, to, 12);
	memcpy(st.up, up, 12);
	st.focal = (float)(1.0 / (2.0 * tan(20.0 * M_PI / 180.0)));
	st.aspect_factor = 1.f;
	st.resx = 32;
	st.resy = 24;
	st.path_samples = 1;
	st.bounces = 3;
	st.rr_min_bounces = 2;
	st.aa_samples = 4;
	st.aa_pixelwidth = 1.5f;
	st.raydepth = 5;
	st.shadow_depth = 4;
	st.has_background = 1;
	st.background[0] = 0.45f * 2.f;
	st.background[1] = 0.475f * 2.f;
	st.background[2] = 0.5f * 2.f;
	std::vector<uint8_t> vis(mats.size(), 3);
	vis[8] = vis[9] = 0;
	std::vector<ora_material> omats(1, om);
	ora_kdtree *kd = ora_kd_build(verts.data(), tris.data(), vis.data(), (uint32_t)mats.size());
	std::vector<float> orgba(4 * 32 * 24, 0.f), ow(32 * 24, 0.f);
	const int ok = ora_pt_render(kd, verts.data(), tris.data(), mats.data(), vis.data(), omats.data(), &light, 1, &st,
	                             orgba.data(), ow.data(), 8);
	ora_kd_free(kd);
	CHECK(ok == 1, "oracle render");
	size_t bad_w = 0, bad_c = 0;
	double mean = 0.0;
	for(size_t i = 0; i < w.size(); ++i)
	{
		bad_w += w[i] != ow[i];
		for(int k = 0; k < 3; ++k)
		{
			const float g = rgba[4 * i + k] / w[i], r = orgba[4 * i + k] / ow[i];
			mean += r;
			if(fabsf(g - r) > 1e-4f * fabsf(r) + 1e-6f) ++bad_c;
		}
	}
	printf("  portal light: 32x24x4, mean %.5f, weight mismatches %zu, colour values off %zu\n",
	       mean / (3.0 * w.size()), bad_w, bad_c);
	CHECK(bad_w == 0 && bad_c == 0 && mean > 0.0, "film differs from the oracle");
}

// Scene::preprocess keeps the accelerator unless objects or accelerator
// params changed (src/scene/scene.cc:314-362).
static void test_scene_preprocess()
{
	Scene sc;
	std::string err;
	ParamMap m;
	m.setString("type", "shinydiffusemat");
	size_t mat = 0;
	CHECK(sc.createMaterial(&mat, "grey", m, &err), "material: %s", err.c_str());
	ParamMap normal;
	const size_t obj = sc.createObject("quad", normal);
	const double q[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
	for(const auto &v : q) sc.addVertex(obj, v[0], v[1], v[2]);
	sc.addTriangle(obj, 0, 1, 2, mat);
	sc.addTriangle(obj, 0, 2, 3, mat);
	uint32_t flags = sc.checkAndClearModifiedFlags();
	CHECK((flags & Scene::ModifiedObjects) && (flags & Scene::ModifiedMaterials), "creation flags %x", flags);
	CHECK(sc.checkAndClearModifiedFlags() == Scene::ModifiedNothing, "flags cleared");
	CHECK(sc.preprocess(flags, &err), "preprocess: %s", err.c_str());
	const auto a0 = sc.accelerator();
	CHECK(a0 && a0->numPrimitives() == 2, "two primitives");
	size_t mat2 = 0;
	CHECK(sc.createMaterial(&mat2, "red", m, &err), "material: %s", err.c_str());
	flags = sc.checkAndClearModifiedFlags();
	CHECK(flags == Scene::ModifiedMaterials, "material flag %x", flags);
	CHECK(sc.preprocess(flags, &err) && sc.accelerator() == a0, "a material change keeps the accelerator");
	sc.addVertex(obj, 0.5, 0.5, 1.0);
	sc.addTriangle(obj, 0, 1, 4, mat2);
	flags = sc.checkAndClearModifiedFlags();
	CHECK(flags == Scene::ModifiedObjects, "object flag %x", flags);
	CHECK(sc.preprocess(flags, &err), "preprocess: %s", err.c_str());
	const auto a1 = sc.accelerator();
	CHECK(a1 && a1 != a0 && a1->numPrimitives() == 3, "an object change rebuilds the accelerator");
	ParamMap ap;
	ap.setString("type", "mi355x-lbvh");
	ap.setString("builder", "lbvh");
	sc.setAcceleratorParams(ap);
	flags = sc.checkAndClearModifiedFlags();
	CHECK(flags == Scene::ModifiedAcceleratorParams && sc.preprocess(flags, &err) && sc.accelerator() != a1,
	      "new accelerator params rebuild it");
	Ray r;
	r.from[0] = 0.5f;
	r.from[1] = 0.25f;
	r.from[2] = 2.f;  // the new triangle lies on z = 2y, above the quad at z = 0
	r.dir[2] = -1.f;
	const IntersectData h = sc.accelerator()->intersect(r, FLT_MAX);
	CHECK(h.isHit() && h.primitive == 2 && fabsf(h.t_hit - 1.5f) < 1e-6f, "the added triangle is hit first (%u, %g)",
	      h.primitive, h.t_hit);
}

// Visibility: an object of visibility "invisible" is left out of the
// primitive list (src/scene/scene.cc:323), so the next object's triangles keep
// the reference's ids; a material's "visibility" is ANDed with the object's in
// every query (include/accelerator/accelerator.h:126-127,138-139).
static void test_visibility_order()
{
	Scene sc;
	std::string err;
	ParamMap m;
	m.setString("type", "shinydiffusemat");
	size_t grey = 0, shadow_only = 0, ghost = 0;
	CHECK(sc.createMaterial(&grey, "grey", m, &err), "material: %s", err.c_str());
	m.setString("visibility", "shadow_only");
	CHECK(sc.createMaterial(&shadow_only, "shadow_only", m, &err), "material: %s", err.c_str());
	m.setString("visibility", "invisible");
	CHECK(sc.createMaterial(&ghost, "ghost", m, &err), "material: %s", err.c_str());
	ParamMap inv;
	inv.setString("visibility", "invisible");
	const size_t hidden = sc.createObject("hidden", inv);  // above everything, created first
	sc.addVertex(hidden, 0, 0, 0.5);
	sc.addVertex(hidden, 4, 0, 0.5);
	sc.addVertex(hidden, 0, 4, 0.5);
	sc.addTriangle(hidden, 0, 1, 2, grey);
	ParamMap normal;
	const size_t obj = sc.createObject("shown", normal);
	auto tri = [&](double x, double y, double z, size_t mat) {
		const size_t v = sc.addVertex(obj, x, y, z);
		sc.addVertex(obj, x + 1, y, z);
		sc.addVertex(obj, x, y + 1, z);
		sc.addTriangle(obj, v, v + 1, v + 2, mat);
	};
	tri(0, 0, 0.0, grey);         // id 0
	tri(2, 0, 0.3, shadow_only);  // id 1
	tri(0, 2, 0.2, ghost);        // id 2
	CHECK(sc.preprocess(sc.checkAndClearModifiedFlags(), &err), "preprocess: %s", err.c_str());
	const auto a = sc.accelerator();
	if(!a) return;
	auto down = [](float x, float y) {
		Ray r;
		r.from[0] = x;
		r.from[1] = y;
		r.from[2] = 3.f;
		r.dir[2] = -1.f;
		return r;
	};
	const IntersectData c0 = a->intersect(down(0.2f, 0.2f), FLT_MAX);
	CHECK(c0.isHit() && c0.primitive == 0 && c0.t_hit == 3.f,
	      "the invisible object is skipped and the next object's first triangle is id 0 (%u, %g)", c0.primitive,
	      c0.t_hit);
	const IntersectData s0 = a->intersectShadow(down(0.2f, 0.2f), 10.f);
	CHECK(s0.isHit() && s0.primitive == 0, "the invisible object casts no shadow (%u)", s0.primitive);
	CHECK(!a->intersect(down(2.2f, 0.2f), FLT_MAX).isHit(), "a shadow_only material is not seen");
	const IntersectData s1 = a->intersectShadow(down(2.2f, 0.2f), 10.f);
	CHECK(s1.isHit() && s1.primitive == 1, "a shadow_only material casts shadows (%u)", s1.primitive);
	CHECK(!a->intersect(down(0.2f, 2.2f), FLT_MAX).isHit(), "an invisible material is not seen");
	CHECK(!a->intersectShadow(down(0.2f, 2.2f), 10.f).isHit(), "an invisible material casts no shadow");
	// the batched GPU query agrees with the single-ray host walk
	std::vector<Ray> rays = {down(0.2f, 0.2f), down(2.2f, 0.2f), down(0.2f, 2.2f), down(1.5f, 1.5f)};
	ymx_ray *d_rays = nullptr;
	to_device(rays, &d_rays);
	ymx_hit *d_hits = nullptr;
	CHECK(hipMalloc(&d_hits, sizeof(ymx_hit) * rays.size()) == hipSuccess, "hipMalloc");
	CHECK(a->intersect(d_rays, d_hits, (uint32_t)rays.size(), nullptr) == YMX_OK, "batched query");
	std::vector<ymx_hit> hits(rays.size());
	CHECK(hipMemcpy(hits.data(), d_hits, sizeof(ymx_hit) * hits.size(), hipMemcpyDeviceToHost) == hipSuccess, "copy");
	CHECK(hits[0].prim == 0 && hits[1].prim == YMX_NO_HIT && hits[2].prim == YMX_NO_HIT && hits[3].prim == YMX_NO_HIT,
	      "batched ids %u %u %u %u", hits[0].prim, hits[1].prim, hits[2].prim, hits[3].prim);
	(void)hipFree(d_rays);
	(void)hipFree(d_hits);
}

// Instances: base objects rendered only through instances, under their
// object-to-world matrices.
static void test_instances()
{
	Scene sc;
	std::string err;
	ParamMap m;
	m.setString("type", "shinydiffusemat");
	size_t mat = 0;
	CHECK(sc.createMaterial(&mat, "grey", m, &err), "material: %s", err.c_str());
	ParamMap base;
	base.setBool("is_base_object", true);
	const size_t obj = sc.createObject("tri", base);
	sc.addVertex(obj, 0, 0, 0);
	sc.addVertex(obj, 1, 0, 0);
	sc.addVertex(obj, 0, 1, 0);
	sc.addTriangle(obj, 0, 1, 2, mat);
	const double up[16] = {2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1};  // x scaled by 2, lifted to z = 1
	const double side[16] = {1, 0, 0, 5, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // moved to x = 5
	const size_t i0 = sc.createInstance(), i1 = sc.createInstance();
	CHECK(sc.addInstanceObject(i0, obj) && sc.addInstanceMatrix(i0, up, 0.f), "instance 0");
	CHECK(sc.addInstanceObject(i1, obj) && sc.addInstanceMatrix(i1, side, 0.f), "instance 1");
	CHECK(!sc.addInstanceObject(9, obj) && !sc.addInstanceMatrix(9, up, 0.f), "unknown instance refused");
	CHECK(sc.preprocess(sc.checkAndClearModifiedFlags(), &err), "preprocess: %s", err.c_str());
	const auto a = sc.accelerator();
	CHECK(a && a->numPrimitives() == 2, "two instanced triangles, the base object left out");
	if(!a) return;
	auto down = [&](float x, float y) {
		Ray r;
		r.from[0] = x;
		r.from[1] = y;
		r.from[2] = 3.f;
		r.dir[2] = -1.f;
		return a->intersect(r, FLT_MAX);
	};
	const IntersectData h0 = down(1.5f, 0.2f), h1 = down(5.2f, 0.2f), h2 = down(0.2f, 0.9f), h3 = down(2.5f, 0.2f);
	CHECK(h0.isHit() && h0.primitive == 0 && h0.t_hit == 2.f, "scaled instance hit at z = 1 (%u, %g)", h0.primitive, h0.t_hit);
	CHECK(h1.isHit() && h1.primitive == 1 && h1.t_hit == 3.f, "moved instance hit at z = 0 (%u, %g)", h1.primitive, h1.t_hit);
	CHECK(h2.isHit() && h2.primitive == 0, "inside the scaled copy");
	CHECK(!h3.isHit(), "outside both copies");
	// a second matrix changes nothing (Instance::getObjToWorldMatrixAtTime uses
	// the first below three time steps); a third is motion blur, refused
	const double two[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
	sc.addInstanceMatrix(i1, two, 1.f);
	CHECK(sc.preprocess(sc.checkAndClearModifiedFlags(), &err), "two matrices: %s", err.c_str());
	const auto a2 = sc.accelerator();
	CHECK(a2 && a2->numPrimitives() == 2, "two matrices: the same two triangles");
	if(a2)
	{
		const IntersectData h = a2->intersect(
		    [] { Ray r; r.from[0] = 5.2f; r.from[1] = 0.2f; r.from[2] = 3.f; r.dir[2] = -1.f; return r; }(), FLT_MAX);
		CHECK(h.isHit() && h.primitive == 1 && h.t_hit == 3.f, "first matrix used (%u, %g)", h.primitive, h.t_hit);
	}
	sc.addInstanceMatrix(i1, two, 2.f);
	err.clear();
	CHECK(!sc.preprocess(sc.checkAndClearModifiedFlags(), &err) && !err.empty(), "motion blur refused");
}

// Instances of instances (Scene::addInstanceOfInstance, src/geometry/instance.cc):
// a copy takes the product of its chain's matrices, outermost first; a base
// instance contributes the list it held when last updated (so an outer
// instance of lower id sees its base's primitives only from the second
// preprocess on); an instance of itself adds nothing.
static void test_nested_instances()
{
	Scene sc;
	std::string err;
	ParamMap m;
	m.setString("type", "shinydiffusemat");
	size_t mat = 0;
	CHECK(sc.createMaterial(&mat, "m", m, &err), "material: %s", err.c_str());
	ParamMap base;
	base.setBool("is_base_object", true);
	const size_t obj = sc.createObject("tri", base);
	sc.addVertex(obj, 0, 0, 0);
	sc.addVertex(obj, 1, 0, 0);
	sc.addVertex(obj, 0, 1, 0);
	sc.addTriangle(obj, 0, 1, 2, mat);
	const double lift[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1};        // z + 1
	const double shift[16] = {2, 0, 0, 5, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};       // x * 2 + 5
	const double far_[16] = {1, 0, 0, 20, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};       // x + 20
	const size_t outer = sc.createInstance();   // 0: an instance of instance 2, created first
	const size_t inner = sc.createInstance();   // 1: the triangle lifted to z = 1
	const size_t nest = sc.createInstance();    // 2: instance 1 scaled by 2 in x and moved to x + 5
	const size_t self = sc.createInstance();    // 3: names itself
	CHECK(sc.addInstanceObject(inner, obj) && sc.addInstanceMatrix(inner, lift, 0.f), "inner");
	CHECK(sc.addInstanceOfInstance(nest, inner) && sc.addInstanceMatrix(nest, shift, 0.f), "nest");
	CHECK(sc.addInstanceOfInstance(outer, nest) && sc.addInstanceMatrix(outer, far_, 0.f), "outer");
	CHECK(sc.addInstanceObject(self, obj) && sc.addInstanceOfInstance(self, self) &&
	      sc.addInstanceMatrix(self, far_, 0.f), "self");
	CHECK(!sc.addInstanceOfInstance(9, inner) && !sc.addInstanceOfInstance(inner, 9), "unknown instance refused");
	auto down = [&](const std::shared_ptr<const Accelerator> &a, float x, float y) {
		Ray r;
		r.from[0] = x;
		r.from[1] = y;
		r.from[2] = 3.f;
		r.dir[2] = -1.f;
		return a->intersect(r, FLT_MAX);
	};
	CHECK(sc.preprocess(sc.checkAndClearModifiedFlags(), &err), "preprocess: %s", err.c_str());
	auto a = sc.accelerator();
	// first pass: instance 0 saw instance 2 before its update (nothing), 1 and 2 one triangle each
	CHECK(a && a->numPrimitives() == 2, "first pass: two triangles (%u)", a ? a->numPrimitives() : 0u);
	if(!a) return;
	const IntersectData h1 = down(a, 0.2f, 0.2f), h2 = down(a, 6.5f, 0.2f), h3 = down(a, 26.5f, 0.2f);
	CHECK(h1.isHit() && h1.primitive == 0 && h1.t_hit == 2.f, "inner copy at z = 1 (%u, %g)", h1.primitive, h1.t_hit);
	CHECK(h2.isHit() && h2.primitive == 1 && h2.t_hit == 2.f, "nested copy: x * 2 + 5 after z + 1 (%u, %g)",
	      h2.primitive, h2.t_hit);
	CHECK(!h3.isHit(), "outer copy not there yet");
	// second pass (any object change): instance 0 now copies instance 2's list, first in the order
	sc.addInstanceMatrix(inner, lift, 1.f);
	CHECK(sc.preprocess(sc.checkAndClearModifiedFlags(), &err), "preprocess: %s", err.c_str());
	a = sc.accelerator();
	CHECK(a && a->numPrimitives() == 3, "second pass: three triangles (%u)", a ? a->numPrimitives() : 0u);
	if(!a) return;
	const IntersectData g3 = down(a, 26.5f, 0.2f), g1 = down(a, 0.2f, 0.2f);
	CHECK(g3.isHit() && g3.primitive == 0 && g3.t_hit == 2.f, "outer copy: x + 20 after the nest (%u, %g)",
	      g3.primitive, g3.t_hit);
	CHECK(g1.isHit() && g1.primitive == 1, "inner copy now second");
	CHECK(!down(a, 20.2f, 0.2f).isHit(), "the self-naming instance adds nothing");
}

static void test_errors()
{
	std::string err;
	ParamMap p;
	p.setString("type", "yafaray-kdtree-original");
	const float v[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
	const uint32_t t[3] = {0, 1, 2};
	CHECK(Accelerator::factory(p, v, 3, t, nullptr, 1, &err) == nullptr && !err.empty(), "other accelerator types refused");
	p.setString("type", "mi355x-lbvh");
	p.setString("builder", "sah-sweep");
	err.clear();
	CHECK(Accelerator::factory(p, v, 3, t, nullptr, 1, &err) == nullptr && !err.empty(), "unknown builder refused");
	Scene sc;
	ParamMap m;
	m.setString("type", "roughglass");
	size_t id;
	err.clear();
	CHECK(!sc.createMaterial(&id, "x", m, &err) && !err.empty(), "unsupported material refused");
	m.setString("type", "glossy");
	m.setString("exponent", "fifty");
	err.clear();
	CHECK(!sc.createMaterial(&id, "x", m, &err) && !err.empty(), "mistyped param refused");
	ParamMap l;
	l.setString("type", "objectlight");
	err.clear();
	CHECK(!sc.createLight("l", l, &err) && !err.empty(), "object light without object_name refused");
	l.setString("object_name", "nowhere");
	CHECK(sc.createLight("l", l, &err), "object light accepted at creation: %s", err.c_str());
	ParamMap portal;
	portal.setString("type", "bgPortalLight");
	err.clear();
	CHECK(!sc.createLight("p", portal, &err) && !err.empty(), "portal light without object_name refused");
	ParamMap integ, cam, film;
	err.clear();
	CHECK(PathIntegrator::create(sc, integ, cam, film, nullptr, &err) == nullptr && !err.empty(),
	      "object light with a missing object refused at create");
	integ.setString("type", "photonmapping");
	err.clear();
	CHECK(PathIntegrator::create(sc, integ, cam, film, nullptr, &err) == nullptr && !err.empty(),
	      "other integrators refused");
	ParamMap q;
	q.setInt("n", 3);
	bool ok = true;
	CHECK(q.getFloat("n", 0.0, &ok) == 3.0 && ok, "int read as float");
	CHECK(q.getString("n", "d", &ok) == "d" && !ok, "mistyped read flagged");
}

// The tree-bound slab test the kernels and the host traversal share
// (ray_tri.h), on the NaN cases of include/geometry/bound.h:183-188: an origin
// on a face with a denormal direction component gives -0 * inf = NaN, which
// std::max(ltmin, lmin) overwrites when it comes from the first axis and keeps
// when it comes from a later one.
static void test_slab_nan_order()
{
	const float a[3] = {0.f, 0.f, 0.f}, g[3] = {1.f, 1.f, 1.f};
	float e = 0.f, l = 0.f;
	{
		const float o[3] = {0.f, 0.5f, 0.5f}, d[3] = {1e-40f, 1.f, 0.f};
		const bool hit = ymx::tree_cross(a, g, o, d, FLT_MAX, e, l);
		CHECK(hit && e == -0.5f && l == 0.5f, "NaN on x replaced by the y slab (hit %d, %g, %g)", hit, e, l);
	}
	{
		const float o[3] = {0.5f, 0.f, 0.5f}, d[3] = {1.f, 1e-40f, 0.f};
		CHECK(!ymx::tree_cross(a, g, o, d, FLT_MAX, e, l), "NaN on y kept: no crossing");
	}
	{
		const float o[3] = {0.5f, 0.5f, 1.f}, d[3] = {1.f, 0.f, -1e-40f};
		CHECK(!ymx::tree_cross(a, g, o, d, FLT_MAX, e, l), "NaN on z kept: no crossing");
	}
	{
		const float o[3] = {-1.f, 0.5f, 0.5f}, d[3] = {1.f, 0.f, 0.f};
		const bool hit = ymx::tree_cross(a, g, o, d, FLT_MAX, e, l);
		CHECK(hit && e == 1.f && l == 2.f, "axis ray enters at 1 and leaves at 2");
	}
}

// MeshObject::smoothVerticesNormals (src/geometry/object/object_mesh.cc:150-264)
// against closed forms: a flat grid keeps its face normal, a cube corner's
// normal averages its three faces equally (each face meets the corner with a
// total angle of pi/2, one right-angled corner or two 45-degree ones), a cube
// smoothed below 90 degrees keeps face normals at corners with a coplanar
// partner and none elsewhere, and an icosphere's normals are radial.
static size_t add_cube(Scene &sc, size_t mat, const float lo[3] = nullptr, float size = 1.f)
{
	ParamMap normal;
	const size_t obj = sc.createObject("cube", normal);
	const float o[3] = {lo ? lo[0] : 0.f, lo ? lo[1] : 0.f, lo ? lo[2] : 0.f};
	for(int i = 0; i < 8; ++i)
		sc.addVertex(obj, o[0] + size * (float)(i & 1), o[1] + size * (float)((i >> 1) & 1),
		             o[2] + size * (float)((i >> 2) & 1));
	const int faces[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
	for(const auto &f : faces)
	{
		sc.addTriangle(obj, f[0], f[1], f[2], mat);
		sc.addTriangle(obj, f[0], f[2], f[3], mat);
	}
	return obj;
}

static void test_smooth_cpu()
{
	Scene sc;
	std::string err;
	ParamMap m;
	m.setString("type", "shinydiffusemat");
	size_t mat = 0;
	CHECK(sc.createMaterial(&mat, "m", m, &err), "material");
	ParamMap normal;
	// flat grid
	const size_t grid = sc.createObject("grid", normal);
	for(int j = 0; j < 4; ++j)
		for(int i = 0; i < 4; ++i) sc.addVertex(grid, i * 0.3, j * 0.25, 0.5);
	for(int j = 0; j < 3; ++j)
		for(int i = 0; i < 3; ++i)
		{
			const size_t a = 4 * j + i;
			sc.addTriangle(grid, a, a + 1, a + 5, mat);
			sc.addTriangle(grid, a, a + 5, a + 4, mat);
		}
	CHECK(sc.cornerNormals(grid).size() == 18 * 12 && sc.cornerNormals(grid)[3] == -1.f, "unsmoothed grid is flat");
	CHECK(sc.smoothMesh(grid, 180.0), "smooth grid");
	const std::vector<float> gn = sc.cornerNormals(grid);
	bool flat_ok = true;
	for(size_t c = 0; c < gn.size() / 4; ++c)
		flat_ok = flat_ok && gn[4 * c + 3] == 1.f && fabsf(gn[4 * c]) < 1e-7f && fabsf(gn[4 * c + 1]) < 1e-7f &&
		          fabsf(gn[4 * c + 2] - 1.f) < 1e-7f;
	CHECK(flat_ok, "a flat grid smooths to its face normal");
	// cube at 180: corner normals along the cube's diagonals
	const size_t cube = add_cube(sc, mat);
	CHECK(sc.smoothMesh(cube, 180.0), "smooth cube");
	const std::vector<float> cn = sc.cornerNormals(cube);
	double worst = 0.0;
	const int faces[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
	for(int f = 0; f < 6; ++f)
		for(int t = 0; t < 2; ++t)
			for(int k = 0; k < 3; ++k)
			{
				const int v = faces[f][t == 0 ? k : (k == 0 ? 0 : k + 1)];
				const double e[3] = {(v & 1) ? 1.0 : -1.0, ((v >> 1) & 1) ? 1.0 : -1.0, ((v >> 2) & 1) ? 1.0 : -1.0};
				const float *c = &cn[12 * (2 * f + t) + 4 * k];
				double d = 0.0;
				for(int i = 0; i < 3; ++i) d += c[i] * e[i] / sqrt(3.0);
				worst = std::max(worst, 1.0 - d);
			}
	printf("  cube corners: 1 - cos(error) at most %.2e\n", worst);
	CHECK(worst < 1e-4, "smoothed cube corners point along the diagonals");
	// cube at 60 degrees: neighbouring faces are 90 degrees apart
	Scene sc2;
	size_t mat2 = 0;
	sc2.createMaterial(&mat2, "m", m, &err);
	const size_t cube2 = add_cube(sc2, mat2);
	CHECK(sc2.smoothMesh(cube2, 60.0), "smooth cube at 60");
	const std::vector<float> c2 = sc2.cornerNormals(cube2);
	const float fn[6][3] = {{0, 0, -1}, {0, 0, 1}, {0, -1, 0}, {0, 1, 0}, {-1, 0, 0}, {1, 0, 0}};
	int with = 0, without = 0;
	bool sharp_ok = true;
	for(int f = 0; f < 6; ++f)
		for(int t = 0; t < 2; ++t)
			for(int k = 0; k < 3; ++k)
			{
				const float *c = &c2[12 * (2 * f + t) + 4 * k];
				if(c[3] == 0.f) ++without;
				else
				{
					++with;
					for(int i = 0; i < 3; ++i) sharp_ok = sharp_ok && fabsf(c[i] - fn[f][i]) < 1e-6f;
				}
			}
	printf("  cube at 60 degrees: %d corners with the face normal, %d without\n", with, without);
	CHECK(sharp_ok && with == 24 && without == 12, "sharp cube edges keep face normals");
	// icosphere: radial normals
	Scene sc3;
	size_t mat3 = 0;
	sc3.createMaterial(&mat3, "m", m, &err);
	const size_t sph = sc3.createObject("sphere", normal);
	std::vector<float> sv;
	std::vector<uint32_t> st;
	{
		// octahedron subdivided twice
		std::vector<std::array<double, 3>> v = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
		std::vector<std::array<int, 3>> f = {{0, 2, 4}, {2, 1, 4}, {1, 3, 4}, {3, 0, 4}, {2, 0, 5}, {1, 2, 5}, {3, 1, 5}, {0, 3, 5}};
		for(int s_ = 0; s_ < 2; ++s_)
		{
			std::map<std::pair<int, int>, int> mid;
			auto m_ = [&](int a, int b) {
				const auto key = std::make_pair(std::min(a, b), std::max(a, b));
				auto it = mid.find(key);
				if(it != mid.end()) return it->second;
				std::array<double, 3> p = {v[a][0] + v[b][0], v[a][1] + v[b][1], v[a][2] + v[b][2]};
				const double l = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
				for(double &x : p) x /= l;
				v.push_back(p);
				return mid[key] = (int)v.size() - 1;
			};
			std::vector<std::array<int, 3>> nf;
			for(const auto &t : f)
			{
				const int ab = m_(t[0], t[1]), bc = m_(t[1], t[2]), ca = m_(t[2], t[0]);
				nf.push_back({t[0], ab, ca});
				nf.push_back({ab, t[1], bc});
				nf.push_back({ca, bc, t[2]});
				nf.push_back({ab, bc, ca});
			}
			f = nf;
		}
		for(const auto &p : v) sc3.addVertex(sph, p[0], p[1], p[2]);
		for(const auto &t : f) sc3.addTriangle(sph, t[0], t[1], t[2], mat3);
		CHECK(sc3.smoothMesh(sph, 180.0), "smooth sphere");
		const std::vector<float> c3 = sc3.cornerNormals(sph);
		double worst3 = 0.0;
		for(size_t t = 0; t < f.size(); ++t)
			for(int k = 0; k < 3; ++k)
			{
				const auto &p = v[f[t][k]];
				const float *c = &c3[12 * t + 4 * k];
				worst3 = std::max(worst3, 1.0 - (c[0] * p[0] + c[1] * p[1] + c[2] * p[2]));
			}
		printf("  icosphere: 1 - cos(error) at most %.2e\n", worst3);
		CHECK(worst3 < 2e-3, "smoothed sphere normals are radial");
	}
	// explicit vertex normals (yafaray_addNormal): the corners take them
	Scene sc4;
	size_t mat4 = 0;
	sc4.createMaterial(&mat4, "m", m, &err);
	const size_t tri = sc4.createObject("tri", normal);
	for(int i = 0; i < 3; ++i)
	{
		sc4.addVertex(tri, i == 1, i == 2, 0);
		sc4.addNormal(tri, 0.1 * i, 0.0, 1.0);
	}
	sc4.addTriangle(tri, 0, 1, 2, mat4);
	const std::vector<float> c4 = sc4.cornerNormals(tri);
	CHECK(c4.size() == 12 && c4[3] == 1.f && c4[4] == 0.1f && c4[8] == 0.2f && c4[11] == 1.f,
	      "explicit vertex normals at the corners");
}

// A scene built through the host layer with a smoothed sphere and multi-pass
// adaptive AA renders the oracle's film bit for bit.
static void test_pathtracer_smooth_aa()
{
	Scene sc;
	std::string err;
	ParamMap white, glossy;
	white.setString("type", "shinydiffusemat");
	white.setColor("color", 0.7, 0.7, 0.7);
	glossy.setString("type", "glossy");
	glossy.setColor("color", 0.8, 0.7, 0.6);
	glossy.setFloat("diffuse_reflect", 0.3);
	glossy.setFloat("exponent", 40.0);
	size_t mw = 0, mg = 0;
	CHECK(sc.createMaterial(&mw, "white", white, &err) && sc.createMaterial(&mg, "glossy", glossy, &err), "materials");
	ParamMap normal;
	const size_t walls = sc.createObject("walls", normal);
	std::vector<float> verts;
	std::vector<uint32_t> tris, mats;
	const float o[3] = {0, 0, 0}, x[3] = {1, 0, 0}, y[3] = {0, 1, 0}, z[3] = {0, 0, 1};
	const float top[3] = {0, 0, 1}, back[3] = {0, 1, 0}, right[3] = {1, 0, 0};
	quad(sc, walls, mw, o, y, x, verts, tris, mats);
	quad(sc, walls, mw, top, x, y, verts, tris, mats);
	quad(sc, walls, mw, back, z, x, verts, tris, mats);
	quad(sc, walls, mw, o, z, y, verts, tris, mats);
	quad(sc, walls, mw, right, y, z, verts, tris, mats);
	// a glossy cube standing in the box, shaded smooth (its corner normals
	// along the diagonals)
	const float cube_lo[3] = {0.35f, 0.35f, 0.0625f}, cube_size = 0.3f;
	const size_t cube = add_cube(sc, mg, cube_lo, cube_size);
	{
		// the oracle's copy of the cube's vertices, as add_cube added them
		const uint32_t v0 = (uint32_t)(verts.size() / 3);
		for(int i = 0; i < 8; ++i)
		{
			verts.push_back((float)((double)cube_lo[0] + (double)(cube_size * (float)(i & 1))));
			verts.push_back((float)((double)cube_lo[1] + (double)(cube_size * (float)((i >> 1) & 1))));
			verts.push_back((float)((double)cube_lo[2] + (double)(cube_size * (float)((i >> 2) & 1))));
		}
		const int faces[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
		for(const auto &f : faces)
		{
			tris.insert(tris.end(), {v0 + f[0], v0 + f[1], v0 + f[2], v0 + f[0], v0 + f[2], v0 + f[3]});
			mats.insert(mats.end(), {(uint32_t)mg, (uint32_t)mg});
		}
	}
	CHECK(sc.smoothMesh(cube, 180.0), "smooth");
	ParamMap al;
	al.setString("type", "arealight");
	al.setVector("corner", 0.4, 0.4, 0.999);
	al.setVector("point1", 0.4, 0.6, 0.999);
	al.setVector("point2", 0.6, 0.4, 0.999);
	al.setFloat("power", 10.0 / M_PI);
	CHECK(sc.createLight("area", al, &err), "light");
	ParamMap integ, cam, film;
	integ.setString("type", "pathtracing");
	integ.setInt("bounces", 3);
	integ.setInt("path_samples", 1);
	cam.setString("type", "perspective");
	cam.setVector("from", 0.5, -1.6, 0.5);
	cam.setVector("to", 0.5, 0.5, 0.5);
	cam.setVector("up", 0.5, -1.6, 1.5);
	cam.setFloat("focal", 1.0 / (2.0 * tan(20.0 * M_PI / 180.0)));
	cam.setInt("resx", 32);
	cam.setInt("resy", 24);
	film.setInt("AA_minsamples", 2);
	film.setInt("AA_passes", 3);
	film.setInt("AA_inc_samples", 2);
	film.setFloat("AA_threshold", 0.02);
	film.setFloat("AA_light_sample_multiplier_factor", 1.5);
	auto pi = PathIntegrator::create(sc, integ, cam, film, nullptr, &err);
	CHECK(pi != nullptr, "create: %s", err.c_str());
	if(!pi) return;
	std::vector<int> res;
	// the autotuner leaves the film alone and changes no pixel of a render
	ymx_pt_tuning tune{};
	CHECK(pi->autotune(1, &tune) == YMX_OK && tune.n_trials >= 4, "autotune: %s", pi->lastError().c_str());
	CHECK(pi->clearFilm() == YMX_OK && pi->renderPasses(&res) == YMX_OK, "render: %s", pi->lastError().c_str());
	std::vector<float> rgba, w;
	CHECK(pi->film(rgba, w) == YMX_OK, "film");
	printf("  resampled per pass: %d %d %d\n", res.size() > 0 ? res[0] : -1, res.size() > 1 ? res[1] : -1,
	       res.size() > 2 ? res[2] : -1);
	// the oracle with the same arrays, corner normals and AA params
	std::vector<float> cnorm = sc.cornerNormals(walls);
	const std::vector<float> cc = sc.cornerNormals(cube);
	cnorm.insert(cnorm.end(), cc.begin(), cc.end());
	std::vector<ora_material> omats(2);
	for(auto &mm : omats)
	{
		mm = ora_material{};
		mm.diffuse_reflect = 1.f;
		mm.receive_shadows = 1;
		mm.diffuse_color[0] = mm.diffuse_color[1] = mm.diffuse_color[2] = 1.f;
		mm.glossy_reflect = 1.f;
		mm.exponent = 50.f;
		mm.as_diffuse = 1;
		mm.reflect = 1.f;
		mm.ior = 1.4f;
		mm.filter_color[0] = mm.filter_color[1] = mm.filter_color[2] = 1.f;
		mm.mirror_color[0] = mm.mirror_color[1] = mm.mirror_color[2] = 1.f;
		mm.power = 1.f;
	}
	omats[0].type = ORA_MAT_SHINYDIFFUSE;
	omats[0].color[0] = omats[0].color[1] = omats[0].color[2] = 0.7f;
	omats[1].type = ORA_MAT_GLOSSY;
	omats[1].color[0] = 0.8f;
	omats[1].color[1] = 0.7f;
	omats[1].color[2] = 0.6f;
	omats[1].diffuse_reflect = 0.3f;
	omats[1].exponent = 40.f;
	ora_light light{};
	light.type = ORA_LIGHT_AREA;
	const float corner[3] = {0.4f, 0.4f, 0.999f}, p1[3] = {0.4f, 0.6f, 0.999f}, p2[3] = {0.6f, 0.4f, 0.999f};
	memcpy(light.corner, corner, 12);
	memcpy(light.point1, p1, 12);
	memcpy(light.point2, p2, 12);
	light.color[0] = light.color[1] = light.color[2] = 1.f;
	light.power = (float)(10.0 / M_PI);
	light.samples = 4;
	light.cast_shadows = 1;
	ora_pt_settings st{};
	const float from[3] = {0.5f, -1.6f, 0.5f}, to[3] = {0.5f, 0.5f, 0.5f}, up[3] = {0.5f, -1.6f, 1.5f};
	memcpy(st.from, from, 12);
	memcpy(st.to, to, 12);
	memcpy(st.up, up, 12);
	st.focal = (float)(1.0 / (2.0 * tan(20.0 * M_PI / 180.0)));
	st.aspect_factor = 1.f;
	st.resx = 32;
	st.resy = 24;
	st.path_samples = 1;
	st.bounces = 3;
	st.aa_samples = 2;
	st.aa_pixelwidth = 1.5f;
	st.raydepth = 5;
	st.shadow_depth = 4;
	st.aa_passes = 3;
	st.aa_inc_samples = 2;
	st.aa_threshold = 0.02f;
	st.aa_sample_multiplier_factor = 1.f;
	st.aa_light_sample_multiplier_factor = 1.5f;
	st.aa_variance_edge_size = 10;
	st.corner_normals = cnorm.data();
	st.trace_brute = 1;  // ties in t to the lowest id, as the GPU
	std::vector<uint8_t> vis(mats.size(), 3);
	ora_kdtree *kd = ora_kd_build(verts.data(), tris.data(), vis.data(), (uint32_t)mats.size());
	std::vector<float> orgba(4 * 32 * 24, 0.f), ow(32 * 24, 0.f);
	ora_pt_render(kd, verts.data(), tris.data(), mats.data(), vis.data(), omats.data(), &light, 1, &st, orgba.data(),
	              ow.data(), 8);
	ora_kd_free(kd);
	size_t bad = 0;
	for(size_t i = 0; i < rgba.size(); ++i) bad += memcmp(&rgba[i], &orgba[i], 4) != 0;
	for(size_t i = 0; i < w.size(); ++i) bad += memcmp(&w[i], &ow[i], 4) != 0;
	printf("  smooth + AA passes: %zu film values differ from the oracle\n", bad);
	CHECK(bad == 0 && res.size() == 3 && res[1] > 0 && res[1] < 32 * 24, "film differs from the oracle");
}

int main(int argc, char **argv)
{
	const bool cpu_only = argc > 1 && strcmp(argv[1], "--cpu") == 0;
	struct
	{
		const char *name;
		void (*fn)();
		bool gpu;
	} tests[] = {{"errors", test_errors, false},
	             {"slab_nan_order", test_slab_nan_order, false},
	             {"smooth_cpu", test_smooth_cpu, false},
	             {"pathtracer_smooth_aa", test_pathtracer_smooth_aa, true},
	             {"accel_closest", test_accel_closest, true},
	             {"accel_closest_ploc", [] { g_builder = "ploc"; test_accel_closest(); g_builder = "lbvh"; }, true},
	             {"accel_shadow_ploc", [] { g_builder = "ploc"; test_accel_shadow(); g_builder = "lbvh"; }, true},
	             {"accel_shadow", test_accel_shadow, true},
	             {"accel_transparent", test_accel_transparent, true},
	             {"scene_preprocess", test_scene_preprocess, true},
	             {"instances", test_instances, true},
	             {"nested_instances", test_nested_instances, true},
	             {"visibility_order", test_visibility_order, true},
	             {"pathtracer", test_pathtracer, true},
	             {"portal_light", test_portal_light, true}};
	for(const auto &t : tests)
	{
		if(cpu_only && t.gpu) continue;
		printf("%s\n", t.name);
		fflush(stdout);
		const int before = g_fail;
		t.fn();
		printf("%s %s\n", g_fail == before ? "PASS" : "FAIL", t.name);
		fflush(stdout);
	}
	printf("%s\n", g_fail ? "FAILED" : "ALL PASSED");
	return g_fail ? 1 : 0;
}
