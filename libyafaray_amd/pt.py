"""Python handle on the wavefront path tracer (C ABI in include/ymx_pt.h).

Mirrors the render side of libYafaRay's `pathtracing` SurfaceIntegrator:
`render(first, n)` traces AA samples like TiledIntegrator::renderPass and
`film()` returns ImageFilm's colour and weight sums.  Scenes are the dicts
built by libyafaray_amd.scenes (geometry, materials, lights, camera,
settings with the reference's param names).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .accel import GpuAccelerator
from .native import check


class YmxMaterial(C.Structure):
    _fields_ = [("type", C.c_uint32), ("color", C.c_float * 3), ("diffuse_reflect", C.c_float),
                ("emit", C.c_float), ("receive_shadows", C.c_uint32), ("diffuse_color", C.c_float * 3),
                ("glossy_reflect", C.c_float), ("exponent", C.c_float), ("as_diffuse", C.c_uint32),
                ("reflect", C.c_float), ("ior", C.c_float), ("filter_color", C.c_float * 3),
                ("transmit_filter", C.c_float), ("mirror_color", C.c_float * 3), ("fake_shadows", C.c_uint32),
                ("power", C.c_float), ("double_sided", C.c_uint32)]


# material "type" param values of the reference (include/material/material.h:151-166), plus short aliases
MATERIAL_TYPES = {"shinydiffuse": 0, "shinydiffusemat": 0, "glossy": 1, "mirror": 2, "glass": 3,
                  "light": 4, "light_mat": 4}


class YmxLight(C.Structure):
    _fields_ = [("type", C.c_uint32), ("corner", C.c_float * 3), ("point1", C.c_float * 3),
                ("point2", C.c_float * 3), ("from_", C.c_float * 3), ("color", C.c_float * 3),
                ("power", C.c_float), ("samples", C.c_int32), ("cast_shadows", C.c_uint32),
                ("first_tri", C.c_uint32), ("n_tris", C.c_uint32), ("double_sided", C.c_uint32),
                ("ibl_clamp_sampling", C.c_float)]


# light "type" param values of the reference (include/light/light.h:117-128), plus short aliases
LIGHT_TYPES = {"area": 0, "arealight": 0, "point": 1, "pointlight": 1, "object": 2, "objectlight": 2,
               "portal": 3, "bgPortalLight": 3}

# object "visibility" param values (include/common/visibility.h:31-37) as YMX_VIS_* masks
VISIBILITY = {"normal": 3, "no_shadows": 1, "shadow_only": 2, "invisible": 0}


class YmxPtSettings(C.Structure):
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
                ("aa_clamp_samples", C.c_float), ("wavefront_mb", C.c_int32)]


# ImageFilm's multi-pass AA params (include/render/imagefilm.h:212-226):
# the settings key (the reference's param name), the struct field, the default
AA_PARAMS = [("AA_passes", "aa_passes", 1), ("AA_inc_samples", "aa_inc_samples", 1),
             ("AA_threshold", "aa_threshold", 0.05), ("AA_resampled_floor", "aa_resampled_floor", 0.0),
             ("AA_sample_multiplier_factor", "aa_sample_multiplier_factor", 1.0),
             ("AA_light_sample_multiplier_factor", "aa_light_sample_multiplier_factor", 1.0),
             ("AA_detect_color_noise", "aa_detect_color_noise", 0),
             ("AA_dark_detection_type", "aa_dark_detection_type", "none"),
             ("AA_dark_threshold_factor", "aa_dark_threshold_factor", 0.0),
             ("AA_variance_edge_size", "aa_variance_edge_size", 10),
             ("AA_variance_pixels", "aa_variance_pixels", 0), ("AA_clamp_samples", "aa_clamp_samples", 0.0)]
# "AA_dark_detection_type" values (include/common/aa_noise_params.h:30-37)
DARK_DETECTION = {"none": 0, "linear": 1, "curve": 2}

YmxPtStats = native.YmxPtStats


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def fill_material(dst, m):
    """Material dict with the reference's param names; absent params take the
    reference's defaults for that material type."""
    kind = m.get("type", "shinydiffuse")
    if kind not in MATERIAL_TYPES:
        raise ValueError(f"unsupported material type {kind!r}")
    dst.type = MATERIAL_TYPES[kind]
    dst.color = _f3(m.get("color", (1, 1, 1)))
    dst.diffuse_reflect = float(m.get("diffuse_reflect", 0.0 if kind == "glossy" else 1.0))
    dst.emit = float(m.get("emit", 0.0))
    dst.receive_shadows = 1 if m.get("receive_shadows", True) else 0
    dst.diffuse_color = _f3(m.get("diffuse_color", (1, 1, 1)))
    dst.glossy_reflect = float(m.get("glossy_reflect", 1.0))
    dst.exponent = float(m.get("exponent", 50.0))
    dst.as_diffuse = 1 if m.get("as_diffuse", True) else 0
    dst.reflect = float(m.get("reflect", 1.0))
    dst.ior = float(m.get("IOR", 1.4))
    dst.filter_color = _f3(m.get("filter_color", (1, 1, 1)))
    dst.transmit_filter = float(m.get("transmit_filter", 0.0))
    dst.mirror_color = _f3(m.get("mirror_color", (1, 1, 1)))
    dst.fake_shadows = 1 if m.get("fake_shadows", False) else 0
    dst.power = float(m.get("power", 1.0))
    dst.double_sided = 1 if m.get("double_sided", False) else 0


def build_structs(scene):
    mats = (YmxMaterial * max(1, len(scene["materials"])))()
    for i, m in enumerate(scene["materials"]):
        fill_material(mats[i], m)
    lights = (YmxLight * max(1, len(scene["lights"])))()
    for i, l in enumerate(scene["lights"]):
        if l["type"] not in LIGHT_TYPES:
            raise ValueError(f"unsupported light type {l['type']!r}")
        lights[i].type = LIGHT_TYPES[l["type"]]
        for k in ("corner", "point1", "point2"):
            setattr(lights[i], k, _f3(l.get(k, (0, 0, 0))))
        lights[i].from_ = _f3(l.get("from", (0, 0, 0)))
        lights[i].color = _f3(l.get("color", (1, 1, 1)))
        lights[i].power = float(l.get("power", 1.0))
        lights[i].samples = int(l.get("samples", 16 if LIGHT_TYPES[l["type"]] == 3 else 4))
        lights[i].cast_shadows = 1 if l.get("cast_shadows", True) else 0
        lights[i].first_tri = int(l.get("first_tri", 0))
        lights[i].n_tris = int(l.get("n_tris", 0))
        lights[i].double_sided = 1 if l.get("double_sided", False) else 0
        lights[i].ibl_clamp_sampling = float(l.get("ibl_clamp_sampling", 0.0))
    cam, s = scene["camera"], scene["settings"]
    st = YmxPtSettings()
    st.from_ = _f3(cam["from_"])
    st.to = _f3(cam["to"])
    st.up = _f3(cam["up"])
    st.focal = float(cam["focal"])
    st.aspect_factor = float(cam.get("aspect_factor", 1.0))
    st.resx, st.resy = int(cam["resx"]), int(cam["resy"])
    st.path_samples = int(s["path_samples"])
    st.bounces = int(s["bounces"])
    st.rr_min_bounces = int(s["rr_min_bounces"])
    st.aa_samples = int(s["aa_samples"])
    st.aa_pixelwidth = float(s.get("aa_pixelwidth", 1.5))
    bg = s.get("background")
    st.has_background = 0 if bg is None else 1
    st.background = _f3(bg if bg is not None else (0, 0, 0))
    st.seed = int(s.get("seed", 0))
    st.raydepth = int(s.get("raydepth", 5))
    st.transparent_shadows = 1 if s.get("transpShad", False) else 0
    st.shadow_depth = int(s.get("shadowDepth", 4))
    st.wavefront_mb = int(s.get("wavefront_mb", 0))
    for key, field, default in AA_PARAMS:
        v = s.get(key, default)
        if field == "aa_dark_detection_type":
            v = DARK_DETECTION[v] if isinstance(v, str) else int(v)
        setattr(st, field, float(v) if isinstance(default, float) else int(v))
    return mats, lights, st


class GpuPathTracer:
    """Path tracer over a GpuAccelerator of the same triangles (built with the
    clustered PLOC tree and leaves of up to 2 triangles unless `accel` is
    given: its rays trace 4-12 % faster than on the LBVH tree, with the same
    results; 4-triangle leaves are faster on the Cornell frame but 5 % slower
    on the occluded room)."""

    def __init__(self, scene, accel: GpuAccelerator | None = None, builder: str = "ploc", max_leaf_size: int = 2,
                 treelet_passes: int = 0):
        self._lib = native.lib()
        self.verts = np.ascontiguousarray(scene["verts"], dtype=np.float32)
        self.tris = np.ascontiguousarray(scene["tris"], dtype=np.uint32)
        self.mat_id = np.ascontiguousarray(scene["mat_id"], dtype=np.uint32)
        vis = scene.get("vis")
        self.vis = None if vis is None else np.ascontiguousarray(vis, dtype=np.uint8)
        self.accel = accel or GpuAccelerator(self.verts, self.tris, self.vis, max_leaf_size=max_leaf_size,
                                              builder=builder, treelet_passes=treelet_passes)
        self.mats, self.lights, self.st = build_structs(scene)
        self.resx, self.resy = self.st.resx, self.st.resy
        h = C.c_void_p()
        rc = self._lib.ymx_pt_create(self.accel.handle, self.verts.ctypes.data, self.verts.shape[0],
                                     self.tris.ctypes.data, self.mat_id.ctypes.data,
                                     None if self.vis is None else self.vis.ctypes.data, self.tris.shape[0],
                                     C.addressof(self.mats), len(scene["materials"]), C.addressof(self.lights),
                                     len(scene["lights"]), C.addressof(self.st), C.byref(h))
        check(rc, "ymx_pt_create")
        self._h = h
        cn = scene.get("corner_normals")
        if cn is not None:
            cn = np.ascontiguousarray(cn, dtype=np.float32).reshape(-1, 12)
            if cn.shape[0] != self.tris.shape[0]:
                raise ValueError("one row of corner normals per triangle is required")
            self.set_normals(cn)

    def set_normals(self, corner_normals):
        """Shading normals at the triangles' corners ((n_tris, 12) float32, see
        ymx_pt_set_normals), or None for flat triangles."""
        arr = None if corner_normals is None else np.ascontiguousarray(corner_normals, dtype=np.float32)
        check(self._lib.ymx_pt_set_normals(self._h, None if arr is None else arr.ctypes.data), "ymx_pt_set_normals")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ymx_pt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch
        return torch.cuda.current_stream().cuda_stream

    def clear(self):
        check(self._lib.ymx_pt_clear_film(self._h, self._stream()), "ymx_pt_clear_film")

    def set_guide(self, guide, bsdf_prob: float = 0.5, train: bool = False, train_steps: int = 4):
        """Mix a GpuGuide into the shinydiffuse next-direction sampling (one-sample
        MIS); with train, each wavefront's paths train it.  guide None: off."""
        h = guide.handle if guide is not None else None
        check(self._lib.ymx_pt_set_guide(self._h, h, float(bsdf_prob), int(bool(train)), int(train_steps)),
              "ymx_pt_set_guide")
        self._guide = guide  # keep it alive while set

    def save_film(self, path, computer_node: int = 0, base_sampling_offset: int = 0, sampling_offset: int = 0):
        """Write the film in ImageFilm's YAF_FILMv4_0_0 format."""
        check(self._lib.ymx_pt_save_film(self._h, str(path).encode(), int(computer_node), int(base_sampling_offset),
                                         int(sampling_offset)), "ymx_pt_save_film")

    def load_film(self, path, add: bool = False) -> dict:
        """Read a YAF_FILMv4_0_0 film into this tracer's film (replace, or add
        its sums); returns the offsets stored in the file."""
        node, base, off = C.c_int32(), C.c_int32(), C.c_int32()
        check(self._lib.ymx_pt_load_film(self._h, str(path).encode(), int(bool(add)), C.byref(node), C.byref(base),
                                         C.byref(off)), "ymx_pt_load_film")
        return {"computer_node": node.value, "base_sampling_offset": base.value, "sampling_offset": off.value}

    def set_guide_depth(self, max_depth: int | None):
        """Guide only the sub-path start vertex and the next `max_depth` path
        vertices (None: every vertex)."""
        check(self._lib.ymx_pt_set_guide_depth(self._h, -1 if max_depth is None else int(max_depth)),
              "ymx_pt_set_guide_depth")

    def render(self, first_sample: int = 0, n_samples: int | None = None):
        n = self.st.aa_samples if n_samples is None else n_samples
        check(self._lib.ymx_pt_render(self._h, int(first_sample), int(n), self._stream()), "ymx_pt_render")

    def render_passes(self) -> list[int]:
        """The whole multi-pass AA render (TiledIntegrator::render): returns the
        pixels resampled in each pass."""
        passes = max(1, int(self.st.aa_passes))
        res = (C.c_int32 * passes)()
        check(self._lib.ymx_pt_render_passes(self._h, res, self._stream()), "ymx_pt_render_passes")
        return list(res)

    def film(self):
        rgba = np.zeros((self.resy, self.resx, 4), dtype=np.float32)
        w = np.zeros((self.resy, self.resx), dtype=np.float32)
        check(self._lib.ymx_pt_read_film(self._h, rgba.ctypes.data, w.ctypes.data), "ymx_pt_read_film")
        return rgba, w

    SELECTION = {"fixed": 0, "learn": 1, "keep": 2}
    FALLBACK = {"off": 0, "start": 1, "keep": 2}

    def set_guide_selection(self, mode: str, lr: float = 0.05):
        """ymx_pt_set_guide_selection (after set_guide): "learn" the
        BSDF-sampling probability per region from the fixed one, "keep" the
        guide's learned table, or go back to "fixed"."""
        check(self._lib.ymx_pt_set_guide_selection(self._h, self.SELECTION[mode], C.c_float(lr)),
              "ymx_pt_set_guide_selection")

    def guide_selection(self) -> np.ndarray:
        """Selection probability per region, shape (16, 16, 16) as [z, y, x]
        (1 where the region has fallen back to BSDF sampling)."""
        a = np.zeros(16 ** 3, dtype=np.float32)
        check(self._lib.ymx_pt_get_guide_selection(self._h, a.ctypes.data, a.size), "ymx_pt_get_guide_selection")
        return a.reshape(16, 16, 16)

    def set_guide_fallback(self, mode: str = "start"):
        """ymx_pt_set_guide_fallback (after set_guide): regions whose running
        guided second moment exceeds BSDF sampling's sample the BSDF alone;
        "start" from no estimates, "keep" the guide's table as another tracer
        left it, or "off"."""
        check(self._lib.ymx_pt_set_guide_fallback(self._h, self.FALLBACK[mode]), "ymx_pt_set_guide_fallback")

    def guide_fallback(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """The fallback's running second moments per region (guided, BSDF;
        0 where no estimate yet) and the training records each region has
        seen, each of shape (16, 16, 16) as [z, y, x]."""
        m = np.zeros(3 * 16 ** 3, dtype=np.float32)
        check(self._lib.ymx_pt_get_guide_fallback(self._h, m.ctypes.data, m.size), "ymx_pt_get_guide_fallback")
        return tuple(m[k * 16 ** 3:(k + 1) * 16 ** 3].reshape(16, 16, 16) for k in range(3))

    KERNELS = {0: "wave-ordered", 1: "replacement"}

    def autotune(self, trial_samples: int = 0) -> dict:
        """ymx_pt_autotune: time short renders under each launch configuration
        and keep the fastest (the film is left as it was)."""
        t = native.YmxPtTuning()
        t.struct_size = C.sizeof(native.YmxPtTuning)
        check(self._lib.ymx_pt_autotune(self._h, int(trial_samples), C.byref(t), self._stream()), "ymx_pt_autotune")
        out = self._tuning_dict(t)
        out["trials"] = [{"samples_per_wavefront": t.trial_samples_per_wavefront[k],
                          "shadow_kernel": self.KERNELS[t.trial_shadow_kernel[k]],
                          "extension_kernel": self.KERNELS[t.trial_extension_kernel[k]],
                          "ms_per_sample": round(float(t.trial_ms_per_sample[k]), 4)} for k in range(t.n_trials)]
        return out

    def _tuning_dict(self, t) -> dict:
        return {"samples_per_wavefront": t.samples_per_wavefront, "shadow_kernel": self.KERNELS[t.shadow_kernel],
                "extension_kernel": self.KERNELS[t.extension_kernel]}

    def tuning(self) -> dict:
        t = native.YmxPtTuning()
        t.struct_size = C.sizeof(native.YmxPtTuning)
        check(self._lib.ymx_pt_get_tuning(self._h, C.byref(t)), "ymx_pt_get_tuning")
        return self._tuning_dict(t)

    def set_tuning(self, samples_per_wavefront: int, shadow_kernel: str, extension_kernel: str):
        inv = {v: k for k, v in self.KERNELS.items()}
        t = native.YmxPtTuning()
        t.struct_size = C.sizeof(native.YmxPtTuning)
        t.samples_per_wavefront = int(samples_per_wavefront)
        t.shadow_kernel = inv[shadow_kernel]
        t.extension_kernel = inv[extension_kernel]
        check(self._lib.ymx_pt_set_tuning(self._h, C.byref(t)), "ymx_pt_set_tuning")

    def stats(self) -> dict:
        s = YmxPtStats()
        s.struct_size = C.sizeof(YmxPtStats)
        check(self._lib.ymx_pt_get_stats(self._h, C.byref(s)), "ymx_pt_get_stats")
        return {"camera_rays": s.camera_rays, "extension_rays": s.extension_rays, "shadow_rays": s.shadow_rays,
                "render_ms": s.render_ms, "guide_samples": s.guide_samples, "guide_steps": s.guide_steps,
                "light_trees": s.light_trees}
