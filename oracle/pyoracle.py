"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.  The shipped package never
does.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
NO_HIT = 0xFFFFFFFF

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        fp = C.POINTER(C.c_float)
        vp = C.c_void_p
        L.ora_triangle_intersect.argtypes = [fp, fp, fp, fp, fp, fp, fp, fp]
        L.ora_triangle_intersect.restype = C.c_int
        L.ora_quad_intersect.argtypes = [fp, fp, fp, fp, fp, fp, fp, fp, fp]
        L.ora_quad_intersect.restype = C.c_int
        L.ora_tree_bound_faces.argtypes = [vp, vp, C.c_uint32, C.c_uint32, fp, fp]
        L.ora_closest_brute_faces.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_int]
        L.ora_shadow_brute_faces.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_int]
        L.ora_closest_brute_in_bound.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, fp, fp, vp, C.c_uint32, vp,
                                                 C.c_int]
        L.ora_retest_faces.argtypes = [vp, vp, C.c_uint32, vp, vp, C.c_uint32, vp]
        L.ora_kd_build_faces.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32]
        L.ora_kd_build_faces.restype = vp
        L.ora_bound_cross.argtypes = [fp, fp, fp, fp, C.c_float, fp, fp]
        L.ora_bound_cross.restype = C.c_int
        L.ora_tree_bound.argtypes = [vp, vp, C.c_uint32, fp, fp]
        L.ora_dynamic_bias.argtypes = [C.c_float, C.c_float]
        L.ora_dynamic_bias.restype = C.c_float
        L.ora_closest_brute.argtypes = [vp, vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_int]
        L.ora_shadow_brute.argtypes = [vp, vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_int]
        L.ora_retest.argtypes = [vp, vp, vp, vp, C.c_uint32, vp]
        L.ora_kd_build.argtypes = [vp, vp, vp, C.c_uint32]
        L.ora_kd_build.restype = vp
        L.ora_kd_free.argtypes = [vp]
        L.ora_kd_bound.argtypes = [vp, fp, fp]
        L.ora_kd_num_nodes.argtypes = [vp]
        L.ora_kd_num_nodes.restype = C.c_uint32
        L.ora_kd_closest.argtypes = [vp, vp, C.c_uint32, vp, C.c_int]
        L.ora_kd_shadow.argtypes = [vp, vp, C.c_uint32, vp, vp, C.c_int]
        for name in ("ora_fast_log2", "ora_fast_exp2", "ora_fast_log", "ora_fast_sin", "ora_fast_cos"):
            getattr(L, name).argtypes = [C.c_float]
            getattr(L, name).restype = C.c_float
        L.ora_fast_pow.argtypes = [C.c_float, C.c_float]
        L.ora_fast_pow.restype = C.c_float
        _lib = L
    return _lib


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _f3(a) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in a])


def triangle_intersect(v0, v1, v2, org, dir):
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = lib().ora_triangle_intersect(_f3(v0), _f3(v1), _f3(v2), _f3(org), _f3(dir),
                                       C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def quad_intersect(v0, v1, v2, v3, org, dir):
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = lib().ora_quad_intersect(_f3(v0), _f3(v1), _f3(v2), _f3(v3), _f3(org), _f3(dir),
                                   C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def bound_cross(a, g, org, dir, t_max):
    e, l = C.c_float(), C.c_float()
    hit = lib().ora_bound_cross(_f3(a), _f3(g), _f3(org), _f3(dir), C.c_float(t_max), C.byref(e), C.byref(l))
    return bool(hit), e.value, l.value


def tree_bound(verts, tris):
    """tris: (n, 3) triangles, or (n, 4) faces whose 4th index is NO_HIT for a
    triangle (every function here takes either)."""
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    a, g = (C.c_float * 3)(), (C.c_float * 3)()
    lib().ora_tree_bound_faces(verts.ctypes.data, tris.ctypes.data, tris.shape[1], tris.shape[0], a, g)
    return np.array(a[:], dtype=np.float32), np.array(g[:], dtype=np.float32)


def dynamic_bias(enter, leave) -> float:
    return lib().ora_dynamic_bias(C.c_float(enter), C.c_float(leave))


def _prep(verts, tris, vis):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    if tris.ndim != 2 or tris.shape[1] not in (3, 4):
        raise ValueError("faces must be (n, 3) or (n, 4) vertex indices")
    if vis is None:
        vis = np.full(tris.shape[0], 3, dtype=np.uint8)
    vis = np.ascontiguousarray(vis, dtype=np.uint8)
    return verts, tris, vis


def closest_brute(verts, tris, rays, vis=None, threads=None) -> dict:
    verts, tris, vis = _prep(verts, tris, vis)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    out = np.zeros((rays.shape[0], 4), dtype=np.uint32)
    lib().ora_closest_brute_faces(verts.ctypes.data, tris.ctypes.data, tris.shape[1], vis.ctypes.data,
                                  tris.shape[0], rays.ctypes.data, rays.shape[0], out.ctypes.data,
                                  threads or default_threads())
    return _hits(out)


def closest_brute_in_bound(verts, tris, rays, bound, vis=None, threads=None) -> dict:
    """closest_brute over the faces given, clipped to and biased by the
    scene's tree bound (a, g) = tree_bound(all faces) instead of theirs."""
    verts, tris, vis = _prep(verts, tris, vis)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    out = np.zeros((rays.shape[0], 4), dtype=np.uint32)
    a = (C.c_float * 3)(*[float(x) for x in bound[0]])
    g = (C.c_float * 3)(*[float(x) for x in bound[1]])
    lib().ora_closest_brute_in_bound(verts.ctypes.data, tris.ctypes.data, tris.shape[1], vis.ctypes.data,
                                     tris.shape[0], a, g, rays.ctypes.data, rays.shape[0], out.ctypes.data,
                                     threads or default_threads())
    return _hits(out)


def shadow_brute(verts, tris, rays, vis=None, threads=None) -> np.ndarray:
    verts, tris, vis = _prep(verts, tris, vis)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    occ = np.zeros(rays.shape[0], dtype=np.uint8)
    lib().ora_shadow_brute_faces(verts.ctypes.data, tris.ctypes.data, tris.shape[1], vis.ctypes.data,
                                 tris.shape[0], rays.ctypes.data, rays.shape[0], occ.ctypes.data,
                                 threads or default_threads())
    return occ


def retest(verts, tris, rays, prim) -> dict:
    """Re-run the triangle test of each ray against its reported primitive."""
    verts, tris, _ = _prep(verts, tris, None)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    prim = np.ascontiguousarray(prim, dtype=np.uint32)
    out = np.zeros((rays.shape[0], 4), dtype=np.uint32)
    lib().ora_retest_faces(verts.ctypes.data, tris.ctypes.data, tris.shape[1], rays.ctypes.data, prim.ctypes.data,
                           rays.shape[0], out.ctypes.data)
    return _hits(out)


def _hits(out: np.ndarray) -> dict:
    return {"prim": out[:, 0].copy(), "t": out[:, 1].view(np.float32).copy(),
            "u": out[:, 2].view(np.float32).copy(), "v": out[:, 3].view(np.float32).copy()}


class KdTree:
    """Port of the reference's SAH kd-tree (CPU baseline and large-batch oracle)."""

    def __init__(self, verts, tris, vis=None):
        self.verts, self.tris, self.vis = _prep(verts, tris, vis)
        self._h = lib().ora_kd_build_faces(self.verts.ctypes.data, self.tris.ctypes.data, self.tris.shape[1],
                                           self.vis.ctypes.data, self.tris.shape[0])
        if not self._h:
            raise MemoryError("ora_kd_build failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_kd_free(self._h)
            self._h = None

    def bound(self):
        a, g = (C.c_float * 3)(), (C.c_float * 3)()
        lib().ora_kd_bound(self._h, a, g)
        return np.array(a[:], dtype=np.float32), np.array(g[:], dtype=np.float32)

    def num_nodes(self) -> int:
        return int(lib().ora_kd_num_nodes(self._h))

    def closest(self, rays, threads=None) -> dict:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 4), dtype=np.uint32)
        lib().ora_kd_closest(self._h, rays.ctypes.data, rays.shape[0], out.ctypes.data,
                             threads or default_threads())
        return _hits(out)

    def transparent(self, rays, transparent, max_depth):
        """TransparentShadow query per ray (accelerator.h:147-168): occluded flags,
        counts and the transparent casters met, in the kd-tree's order."""
        L = lib()
        if not getattr(L, "_tr_ready", False):
            L.ora_kd_transparent1.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_void_p,
                                              C.POINTER(C.c_int)]
            L.ora_kd_transparent1.restype = C.c_int
            L._tr_ready = True
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        flags = np.ascontiguousarray(transparent, dtype=np.uint8)
        n = rays.shape[0]
        occ = np.zeros(n, bool)
        counts = np.zeros(n, np.uint32)
        prims = np.zeros((n, max(1, max_depth)), np.uint32)
        buf = np.zeros(max(1, max_depth), np.uint32)
        cnt = C.c_int(0)
        for i in range(n):
            t_max = float(rays[i, 7]) if rays[i, 7] >= 0 else float(np.finfo(np.float32).max)
            occ[i] = bool(L.ora_kd_transparent1(self._h, rays[i].ctypes.data, t_max, flags.ctypes.data, int(max_depth),
                                                buf.ctypes.data, C.byref(cnt)))
            counts[i] = 0 if occ[i] else cnt.value
            prims[i, :cnt.value] = buf[:cnt.value]
        return occ, counts, prims

    def shadow(self, rays, threads=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], dtype=np.uint8)
        prim = np.zeros(rays.shape[0], dtype=np.uint32)
        lib().ora_kd_shadow(self._h, rays.ctypes.data, rays.shape[0], occ.ctypes.data, prim.ctypes.data,
                            threads or default_threads())
        return occ, prim


def fast_pow(a, b) -> float:
    return lib().ora_fast_pow(C.c_float(a), C.c_float(b))


# ---------------------------------------------------------------- path tracer
class OraMaterial(C.Structure):
    _fields_ = [("type", C.c_uint32), ("color", C.c_float * 3), ("diffuse_reflect", C.c_float),
                ("emit", C.c_float), ("receive_shadows", C.c_uint32), ("diffuse_color", C.c_float * 3),
                ("glossy_reflect", C.c_float), ("exponent", C.c_float), ("as_diffuse", C.c_uint32),
                ("reflect", C.c_float), ("ior", C.c_float), ("filter_color", C.c_float * 3),
                ("transmit_filter", C.c_float), ("mirror_color", C.c_float * 3), ("fake_shadows", C.c_uint32),
                ("power", C.c_float), ("double_sided", C.c_uint32)]


MAT_TYPES = {"shinydiffuse": 0, "shinydiffusemat": 0, "glossy": 1, "mirror": 2, "glass": 3, "light": 4,
             "light_mat": 4}


def fill_material(dst, m):
    """Material dict (the reference's param names, its defaults) -> record."""
    kind = m.get("type", "shinydiffuse")
    dst.type = MAT_TYPES[kind]
    dst.color = _f3v(m.get("color", (1, 1, 1)))
    dst.diffuse_reflect = float(m.get("diffuse_reflect", 0.0 if kind == "glossy" else 1.0))
    dst.emit = float(m.get("emit", 0.0))
    dst.receive_shadows = int(bool(m.get("receive_shadows", True)))
    dst.diffuse_color = _f3v(m.get("diffuse_color", (1, 1, 1)))
    dst.glossy_reflect = float(m.get("glossy_reflect", 1.0))
    dst.exponent = float(m.get("exponent", 50.0))
    dst.as_diffuse = int(bool(m.get("as_diffuse", True)))
    dst.reflect = float(m.get("reflect", 1.0))
    dst.ior = float(m.get("IOR", 1.4))
    dst.filter_color = _f3v(m.get("filter_color", (1, 1, 1)))
    dst.transmit_filter = float(m.get("transmit_filter", 0.0))
    dst.mirror_color = _f3v(m.get("mirror_color", (1, 1, 1)))
    dst.fake_shadows = int(bool(m.get("fake_shadows", False)))
    dst.power = float(m.get("power", 1.0))
    dst.double_sided = int(bool(m.get("double_sided", False)))


class OraLight(C.Structure):
    _fields_ = [("type", C.c_uint32), ("corner", C.c_float * 3), ("point1", C.c_float * 3),
                ("point2", C.c_float * 3), ("from_", C.c_float * 3), ("color", C.c_float * 3),
                ("power", C.c_float), ("samples", C.c_int32), ("cast_shadows", C.c_uint32),
                ("first_tri", C.c_uint32), ("n_tris", C.c_uint32), ("double_sided", C.c_uint32),
                ("ibl_clamp_sampling", C.c_float)]


# light "type" param values of the reference (src/light/light.cc factory), as ORA_LIGHT_*
LIGHT_TYPES = {"area": 0, "arealight": 0, "point": 1, "pointlight": 1, "object": 2, "objectlight": 2,
               "portal": 3, "bgPortalLight": 3}


def fill_light(dst, l):
    if l["type"] not in LIGHT_TYPES:
        raise ValueError(f"unsupported light type {l['type']!r}")
    dst.type = LIGHT_TYPES[l["type"]]
    for k in ("corner", "point1", "point2"):
        setattr(dst, k, _f3v(l.get(k, (0, 0, 0))))
    dst.from_ = _f3v(l.get("from", (0, 0, 0)))
    dst.color = _f3v(l.get("color", (1, 1, 1)))
    dst.power = float(l.get("power", 1.0))
    dst.samples = int(l.get("samples", 16 if dst.type == 3 else 4))
    dst.cast_shadows = int(bool(l.get("cast_shadows", True)))
    dst.first_tri = int(l.get("first_tri", 0))
    dst.n_tris = int(l.get("n_tris", 0))
    dst.double_sided = int(bool(l.get("double_sided", False)))
    dst.ibl_clamp_sampling = float(l.get("ibl_clamp_sampling", 0.0))


class OraPtSettings(C.Structure):
    _fields_ = [("from_", C.c_float * 3), ("to", C.c_float * 3), ("up", C.c_float * 3), ("focal", C.c_float),
                ("aspect_factor", C.c_float), ("resx", C.c_int32), ("resy", C.c_int32),
                ("path_samples", C.c_int32), ("bounces", C.c_int32), ("rr_min_bounces", C.c_int32),
                ("aa_samples", C.c_int32), ("aa_pixelwidth", C.c_float), ("has_background", C.c_int32),
                ("background", C.c_float * 3), ("seed", C.c_uint32), ("raydepth", C.c_int32),
                ("transparent_shadows", C.c_int32), ("shadow_depth", C.c_int32),
                ("aa_passes", C.c_int32), ("aa_inc_samples", C.c_int32), ("aa_threshold", C.c_float),
                ("aa_resampled_floor", C.c_float), ("aa_sample_multiplier_factor", C.c_float),
                ("aa_light_sample_multiplier_factor", C.c_float), ("aa_detect_color_noise", C.c_int32),
                ("aa_dark_detection_type", C.c_int32), ("aa_dark_threshold_factor", C.c_float),
                ("aa_variance_edge_size", C.c_int32), ("aa_variance_pixels", C.c_int32),
                ("aa_clamp_samples", C.c_float), ("corner_normals", C.c_void_p), ("trace_brute", C.c_int32)]


# ImageFilm's multi-pass AA params (include/render/imagefilm.h:212-226): the
# settings key, the struct field, the reference's default
AA_PARAMS = [("AA_passes", "aa_passes", 1), ("AA_inc_samples", "aa_inc_samples", 1),
             ("AA_threshold", "aa_threshold", 0.05), ("AA_resampled_floor", "aa_resampled_floor", 0.0),
             ("AA_sample_multiplier_factor", "aa_sample_multiplier_factor", 1.0),
             ("AA_light_sample_multiplier_factor", "aa_light_sample_multiplier_factor", 1.0),
             ("AA_detect_color_noise", "aa_detect_color_noise", 0),
             ("AA_dark_detection_type", "aa_dark_detection_type", "none"),
             ("AA_dark_threshold_factor", "aa_dark_threshold_factor", 0.0),
             ("AA_variance_edge_size", "aa_variance_edge_size", 10),
             ("AA_variance_pixels", "aa_variance_pixels", 0), ("AA_clamp_samples", "aa_clamp_samples", 0.0)]
DARK_DETECTION = {"none": 0, "linear": 1, "curve": 2}


def fill_aa(st, s):
    """The AA pass params of a scene's settings dict into a settings struct
    (ora_pt_settings or the library's ymx_pt_settings, same field names)."""
    for key, field, default in AA_PARAMS:
        v = s.get(key, default)
        if field == "aa_dark_detection_type":
            v = DARK_DETECTION[v] if isinstance(v, str) else int(v)
        setattr(st, field, float(v) if isinstance(default, float) else int(v))


def _f3v(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def pt_structs(scene):
    mats = (OraMaterial * len(scene["materials"]))()
    for i, m in enumerate(scene["materials"]):
        fill_material(mats[i], m)
    lights = (OraLight * max(1, len(scene["lights"])))()
    for i, l in enumerate(scene["lights"]):
        fill_light(lights[i], l)
    cam, s = scene["camera"], scene["settings"]
    st = OraPtSettings()
    st.from_ = _f3v(cam["from_"])
    st.to = _f3v(cam["to"])
    st.up = _f3v(cam["up"])
    st.focal = cam["focal"]
    st.aspect_factor = 1.0
    st.resx, st.resy = cam["resx"], cam["resy"]
    st.path_samples, st.bounces, st.rr_min_bounces = s["path_samples"], s["bounces"], s["rr_min_bounces"]
    st.aa_samples = s["aa_samples"]
    st.aa_pixelwidth = s["aa_pixelwidth"]
    bg = s.get("background")
    st.has_background = 0 if bg is None else 1
    st.background = _f3v(bg if bg is not None else (0, 0, 0))
    st.seed = int(s.get("seed", 0))
    st.raydepth = int(s.get("raydepth", 5))
    st.transparent_shadows = int(bool(s.get("transpShad", False)))
    st.shadow_depth = int(s.get("shadowDepth", 4))
    fill_aa(st, s)
    return mats, lights, st


def corner_normals_array(scene):
    """The scene's per-corner normals as the float32 (n_tris, 12) array the
    path tracers take, or None."""
    cn = scene.get("corner_normals")
    if cn is None:
        return None
    cn = np.ascontiguousarray(cn, dtype=np.float32).reshape(-1, 12)
    if cn.shape[0] != len(scene["tris"]):
        raise ValueError("one row of corner normals per triangle is required")
    return cn


def _pt_lib():
    L = lib()
    if not getattr(L, "_pt_ready", False):
        vp = C.c_void_p
        L.ora_pt_render.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_uint32, vp, vp, vp, C.c_int]
        L.ora_pt_render_pixels.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, C.c_int]
        L.ora_pt_render.restype = C.c_int
        L.ora_pt_render_pixels.restype = C.c_int
        L.ora_pt_pixel_sample.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_uint32, vp, C.c_int, C.c_int, C.c_int,
                                          vp, vp]
        L.ora_object_light_setup.argtypes = [vp, vp, vp, vp]
        L.ora_object_light_setup.restype = C.c_float
        L.ora_object_light_sample.argtypes = [vp, vp, vp, vp, C.c_float, C.c_float, vp]
        L.ora_object_light_sample.restype = C.c_int
        L.ora_object_light_intersect.argtypes = [vp, vp, vp, vp, vp, vp, C.c_float, C.POINTER(C.c_float)]
        L.ora_object_light_intersect.restype = C.c_int
        L.ora_ri_vdc.argtypes = [C.c_uint32, C.c_uint32]
        L.ora_ri_vdc.restype = C.c_float
        L.ora_ri_s.argtypes = [C.c_uint32, C.c_uint32]
        L.ora_ri_s.restype = C.c_float
        L.ora_ri_lp.argtypes = [C.c_uint32, C.c_uint32]
        L.ora_ri_lp.restype = C.c_float
        L.ora_fnv32a.argtypes = [C.c_uint32]
        L.ora_fnv32a.restype = C.c_uint32
        L.ora_low_discrepancy.argtypes = [C.c_int, C.c_uint32]
        L.ora_low_discrepancy.restype = C.c_double
        L.ora_halton_next.argtypes = [C.c_int, C.c_uint32]
        L.ora_halton_next.restype = C.c_float
        L.ora_faure_perm.argtypes = [C.c_int, C.c_int]
        L.ora_faure_perm.restype = C.c_int
        L.ora_rr_random.argtypes = [C.c_uint32] * 4
        L.ora_rr_random.restype = C.c_float
        L.ora_filter_table.argtypes = [C.c_float, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.ora_aa_next_pass.argtypes = [vp, C.c_float, vp, vp, vp]
        L.ora_aa_next_pass.restype = C.c_int
        L.ora_aa_max_samples.argtypes = [vp]
        L.ora_aa_max_samples.restype = C.c_int
        L._pt_ready = True
    return L


class PathTracer:
    """CPU restatement of the reference's pathtracing integrator (oracle)."""

    def __init__(self, scene, kd: "KdTree | None" = None, brute: bool = False):
        """brute: closest hits by brute force (exact ties to the lowest id, as
        the GPU) instead of the kd-tree port; for small scenes."""
        self.scene = scene
        self.verts = np.ascontiguousarray(scene["verts"], dtype=np.float32)
        self.tris = np.ascontiguousarray(scene["tris"], dtype=np.uint32)
        self.mat_id = np.ascontiguousarray(scene["mat_id"], dtype=np.uint32)
        self.kd = kd or KdTree(self.verts, self.tris, scene.get("vis"))
        self.vis = self.kd.vis
        self.mats, self.lights, self.st = pt_structs(scene)
        self.n_lights = len(scene["lights"])
        self._cn = corner_normals_array(scene)
        self.st.corner_normals = None if self._cn is None else self._cn.ctypes.data
        self.st.trace_brute = 1 if brute else 0

    def render(self, threads=None, aa_samples=None):
        L = _pt_lib()
        st = self.st
        if aa_samples is not None:
            st.aa_samples = int(aa_samples)
        rgba = np.zeros((st.resy, st.resx, 4), dtype=np.float32)
        w = np.zeros((st.resy, st.resx), dtype=np.float32)
        ok = L.ora_pt_render(self.kd._h, self.verts.ctypes.data, self.tris.ctypes.data, self.mat_id.ctypes.data,
                             self.vis.ctypes.data, C.addressof(self.mats), C.addressof(self.lights), self.n_lights,
                             C.addressof(st), rgba.ctypes.data, w.ctypes.data, threads or default_threads())
        if not ok:
            raise RuntimeError("ora_pt_render: scene not supported (see ora_pt.h)")
        return rgba, w

    def render_pixels(self, flags, threads=None):
        """Pass 0 over the pixels where flags (resy, resx) is set, splatted
        into a full-size film (ora_pt_render_pixels)."""
        L = _pt_lib()
        st = self.st
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        assert flags.shape == (st.resy, st.resx)
        rgba = np.zeros((st.resy, st.resx, 4), dtype=np.float32)
        w = np.zeros((st.resy, st.resx), dtype=np.float32)
        ok = L.ora_pt_render_pixels(self.kd._h, self.verts.ctypes.data, self.tris.ctypes.data,
                                    self.mat_id.ctypes.data, self.vis.ctypes.data, C.addressof(self.mats),
                                    C.addressof(self.lights), self.n_lights, C.addressof(st), flags.ctypes.data,
                                    rgba.ctypes.data, w.ctypes.data, threads or default_threads())
        if not ok:
            raise RuntimeError("ora_pt_render_pixels: scene not supported (see ora_pt.h)")
        return rgba, w

    def pixel_sample(self, x, y, sample):
        L = _pt_lib()
        out = np.zeros(4, dtype=np.float32)
        dxdy = np.zeros(2, dtype=np.float32)
        L.ora_pt_pixel_sample(self.kd._h, self.verts.ctypes.data, self.tris.ctypes.data, self.mat_id.ctypes.data,
                              self.vis.ctypes.data, C.addressof(self.mats), C.addressof(self.lights), self.n_lights, C.addressof(self.st),
                              int(x), int(y), int(sample), out.ctypes.data, dxdy.ctypes.data)
        return out, dxdy


def aa_next_pass(settings_scene, threshold, rgba, w):
    """ImageFilm::nextPass's resampling flags for film sums rgba (h, w, 4) and
    weights (h, w): returns (flags uint8 (h, w), count)."""
    _, _, st = pt_structs(settings_scene)
    L = _pt_lib()
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    flags = np.zeros(w.shape, dtype=np.uint8)
    n = L.ora_aa_next_pass(C.addressof(st), C.c_float(threshold), rgba.ctypes.data, w.ctypes.data,
                           flags.ctypes.data)
    return flags, n


def aa_max_samples(settings_scene):
    _, _, st = pt_structs(settings_scene)
    return _pt_lib().ora_aa_max_samples(C.addressof(st))


def camera_dir(scene, px, py):
    _, _, st = pt_structs(scene)
    L = _pt_lib()
    L.ora_pt_camera_dir.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
    out = np.zeros(3, dtype=np.float32)
    L.ora_pt_camera_dir(C.addressof(st), float(px), float(py), out.ctypes.data)
    return out


def filter_table(pixelwidth=1.5):
    L = _pt_lib()
    t = np.zeros(256, dtype=np.float32)
    scale, width = C.c_float(), C.c_float()
    L.ora_filter_table(C.c_float(pixelwidth), t.ctypes.data, C.byref(scale), C.byref(width))
    return t.reshape(16, 16), scale.value, width.value


# ---------------------------------------------------------------- single material functions
F_SPECULAR, F_GLOSSY, F_DIFFUSE, F_DISPERSIVE, F_REFLECT, F_TRANSMIT, F_FILTER, F_EMIT = 1, 2, 4, 8, 16, 32, 64, 128
F_ALL = 127


def _mat_lib():
    L = _pt_lib()
    if not getattr(L, "_mat_ready", False):
        vp = C.c_void_p
        L.ora_mat_flags.argtypes = [vp]
        L.ora_mat_flags.restype = C.c_uint32
        L.ora_bsdf_eval.argtypes = [vp, vp, vp, vp, C.c_uint32, vp]
        L.ora_bsdf_pdf.argtypes = [vp, vp, vp, vp, C.c_uint32]
        L.ora_bsdf_pdf.restype = C.c_float
        L.ora_bsdf_sample.argtypes = [vp, vp, vp, C.c_float, C.c_float, C.c_uint32, vp]
        L.ora_specular.argtypes = [vp, vp, vp, C.c_int, vp]
        L.ora_fresnel.argtypes = [vp, vp, C.c_float, vp]
        L.ora_refract.argtypes = [vp, vp, C.c_float, vp]
        L.ora_refract.restype = C.c_int
        L.ora_mat_emit.argtypes = [vp, vp, vp, vp]
        L._mat_ready = True
    return L


def _v(x):
    return np.ascontiguousarray(x, dtype=np.float32)


class Material:
    """One material record, for single-function checks of the oracle."""

    def __init__(self, **params):
        self.rec = OraMaterial()
        fill_material(self.rec, params)
        self.L = _mat_lib()

    @property
    def flags(self):
        return self.L.ora_mat_flags(C.byref(self.rec))

    def eval(self, ng, wo, wl, flags=F_ALL):
        out = np.zeros(3, np.float32)
        ng, wo, wl = _v(ng), _v(wo), _v(wl)
        self.L.ora_bsdf_eval(C.byref(self.rec), ng.ctypes.data, wo.ctypes.data, wl.ctypes.data, flags,
                             out.ctypes.data)
        return out

    def pdf(self, ng, wo, wi, flags=F_GLOSSY | F_DIFFUSE | F_DISPERSIVE | F_REFLECT | F_TRANSMIT):
        ng, wo, wi = _v(ng), _v(wo), _v(wi)
        return self.L.ora_bsdf_pdf(C.byref(self.rec), ng.ctypes.data, wo.ctypes.data, wi.ctypes.data, flags)

    def sample(self, ng, wo, s1, s2, flags=F_ALL):
        out = np.zeros(9, np.float32)
        ng, wo = _v(ng), _v(wo)
        self.L.ora_bsdf_sample(C.byref(self.rec), ng.ctypes.data, wo.ctypes.data, float(s1), float(s2), flags,
                               out.ctypes.data)
        return dict(wi=out[0:3].copy(), col=out[3:6].copy(), w=float(out[6]), pdf=float(out[7]),
                    sampled=int(out[8]))

    def specular(self, ng, wo, ray_level=1):
        out = np.zeros(14, np.float32)
        ng, wo = _v(ng), _v(wo)
        self.L.ora_specular(C.byref(self.rec), ng.ctypes.data, wo.ctypes.data, int(ray_level), out.ctypes.data)
        r = dict(dir=out[1:4].copy(), col=out[4:7].copy()) if out[0] else None
        t = dict(dir=out[8:11].copy(), col=out[11:14].copy()) if out[7] else None
        return r, t

    def emit(self, ng, wo):
        out = np.zeros(3, np.float32)
        ng, wo = _v(ng), _v(wo)
        self.L.ora_mat_emit(C.byref(self.rec), ng.ctypes.data, wo.ctypes.data, out.ctypes.data)
        return out


def fresnel(i, n, ior):
    L = _mat_lib()
    out = np.zeros(2, np.float32)
    i, n = _v(i), _v(n)
    L.ora_fresnel(i.ctypes.data, n.ctypes.data, float(ior), out.ctypes.data)
    return float(out[0]), float(out[1])


def refract(n, wi, ior):
    L = _mat_lib()
    out = np.zeros(3, np.float32)
    n, wi = _v(n), _v(wi)
    ok = L.ora_refract(n.ctypes.data, wi.ctypes.data, float(ior), out.ctypes.data)
    return out if ok else None


class ObjectLight:
    """ObjectLight pieces of the oracle on one light dict (type "object") over a
    triangle soup: area_, the area cdf, illumSample and intersect."""

    def __init__(self, verts, tris, light, vis=None):
        self.verts, self.tris, self.vis = _prep(verts, tris, vis)
        self.l = OraLight()
        fill_light(self.l, light)

    def setup(self):
        cdf = np.zeros(max(1, self.l.n_tris), dtype=np.float32)
        area = _pt_lib().ora_object_light_setup(self.verts.ctypes.data, self.tris.ctypes.data, C.addressof(self.l),
                                                cdf.ctypes.data)
        return float(area), cdf[:self.l.n_tris]

    def sample(self, p, s1, s2):
        out = np.zeros(5, dtype=np.float32)
        p = np.ascontiguousarray(p, dtype=np.float32)
        hit = _pt_lib().ora_object_light_sample(self.verts.ctypes.data, self.tris.ctypes.data, C.addressof(self.l),
                                                p.ctypes.data, float(s1), float(s2), out.ctypes.data)
        return hit, out[:3], float(out[3]), float(out[4])

    def intersect(self, frm, d, tmin):
        ipdf = C.c_float(0.0)
        frm = np.ascontiguousarray(frm, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        hit = _pt_lib().ora_object_light_intersect(self.verts.ctypes.data, self.tris.ctypes.data,
                                                   self.vis.ctypes.data, C.addressof(self.l), frm.ctypes.data,
                                                   d.ctypes.data, float(tmin), C.byref(ipdf))
        return hit, float(ipdf.value)


# ---------------------------------------------------------------- film file

FILM_TAG = b"YAF_FILMv4_0_0"


def write_film(path, rgba, weight, computer_node=0, base_sampling_offset=0, sampling_offset=0):
    """ImageFilm::imageFilmSave (src/render/imagefilm.cc:1099-1176) for a film
    of one image layer covering the whole image: the tag through
    File::append(std::string), which writes its NUL (src/common/file.cc:187-191);
    the node, the two sampling offsets, width, height, x0, x1 = x0 + w - 1,
    y0, y1 = y0 + h - 1 and the layer count as int; the weights for y, then x;
    each layer's r, g, b, a for y, then x.  File::append<T> writes host-order
    bytes (include/common/file.h:90-94)."""
    import struct
    weight = np.asarray(weight, np.float32)
    rgba = np.asarray(rgba, np.float32)
    h, w = weight.shape
    with open(path, "wb") as f:
        f.write(FILM_TAG + b"\0")
        f.write(struct.pack("=10i", computer_node, base_sampling_offset, sampling_offset, w, h, 0, w - 1, 0, h - 1, 1))
        for y in range(h):
            for x in range(w):
                f.write(struct.pack("=f", float(weight[y, x])))
        for y in range(h):
            for x in range(w):
                f.write(struct.pack("=4f", *(float(v) for v in rgba[y, x, :4])))


def read_film(path, w, h):
    """ImageFilm::imageFilmLoad (src/render/imagefilm.cc:910-1018) for a
    one-layer film of w x h: None when the tag, size, borders or layer count
    differ, else (rgba [h][w][4], weight [h][w], (node, base offset, offset))."""
    import struct
    data = Path(path).read_bytes()
    end = data.find(b"\0")
    if end < 0 or data[:end] != FILM_TAG:
        return None
    pos = end + 1
    head = struct.unpack_from("=10i", data, pos)
    pos += 40
    if head[3:] != (w, h, 0, w - 1, 0, h - 1, 1):
        return None
    weight = np.empty((h, w), np.float32)
    for y in range(h):
        for x in range(w):
            weight[y, x] = struct.unpack_from("=f", data, pos)[0]
            pos += 4
    rgba = np.empty((h, w, 4), np.float32)
    for y in range(h):
        for x in range(w):
            rgba[y, x] = struct.unpack_from("=4f", data, pos)
            pos += 16
    return rgba, weight, head[:3]
