#!/usr/bin/env python3
"""Ray-cast microbenchmark of the mi355x-lbvh accelerator on one MI355X.

Workload: a 1024 x 1024 fractal heightfield (2,093,058 triangles) plus
100,000 random triangles above it; 8,388,608 coherent primary rays
(4096 x 2048 pinhole) and 8,388,608 incoherent rays (origins uniform in the
scene box, directions uniform on the sphere); shadow rays are the incoherent
rays with tmax uniform in (0, 1].

A step is one closest-hit pass over the 8,388,608 incoherent rays, with rays
and BVH already resident in HBM.  The coherent closest-hit and incoherent
any-hit rates are reported alongside.  The same JSON line carries, under
`extra`, the path-traced Cornell frame (256 x 256, 16 spp) beside the CPU
restatement of the reference's integrator, the mixed-material terrain frame
(1920 x 1080, 64 spp, 2.73M triangles), the guided vs plain time to
relMSE 1e-3 on the occluded room (1280 x 720), and the guiding network's
kernels on their own.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s incoherent closest-hit; time-to-relMSE 1e-3, guided vs plain GPU PT"
HBM_PEAK_GBS = 8000.0
RAY_BYTES = 32
HIT_BYTES = 16
NODE_BYTES = 64
TRI_BYTES = 48
OCC_BYTES = 4  # any-hit output per ray, padded (the kernel writes 1 bit)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--cpu-sample", type=int, default=2_097_152,
                   help="incoherent rays timed on the CPU kd-tree port (0: skip)")
    p.add_argument("--cpu-threads", type=int, default=0, help="0: the CPU share of one GPU, min(16, cores)")
    p.add_argument("--res", type=int, default=1024)
    p.add_argument("--n-random", type=int, default=100_000)
    p.add_argument("--quiet", action="store_true")
    p.add_argument("--treelet-passes", type=int, default=0,
                   help="SAH treelet restructuring passes of the ray-cast tree (0: the library default, -1: none)")
    p.add_argument("--builder", type=str, default="lbvh", help="tree builder of the ray-cast batches")
    p.add_argument("--reorder-cell-bits", type=int, default=0, help="sort granularity of ray reordering (3 or 4)")
    p.add_argument("--ray-reorder", type=int, default=0,
                   help="ray reordering of the ray-cast batches (0: library default, 1: always, -1: never)")
    p.add_argument("--pt-treelet-passes", type=int, default=0,
                   help="treelet passes of the path tracer's tree (0: the library default, -1: none)")
    p.add_argument("--no-counters", action="store_true",
                   help="skip the rocprofv3 counter passes (traffic, VALU busy, MFMA utilisation)")
    p.add_argument("--counter-pass", choices=["raycast", "pathtrace", "guide"], default=None,
                   help="only launch one workload's kernels a few times, for a profiler's counter pass")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run only a plain 1 GiB device-to-device copy (HBM counter calibration)")
    p.add_argument("--no-pt", action="store_true", help="skip the path-tracing measurements")
    p.add_argument("--no-terrain", action="store_true", help="skip the 1920x1080x64 terrain frame")
    p.add_argument("--terrain-only", action="store_true", help="only the 1920x1080x64 terrain frame")
    p.add_argument("--large-only", action="store_true", help="only the 3840x2160 large-terrain frame")
    p.add_argument("--large-spp", type=int, default=4, help="samples of the large-terrain frame")
    p.add_argument("--pt-only", action="store_true", help="only the path-tracing measurement")
    p.add_argument("--no-guided", action="store_true", help="skip the guided path-tracing measurement")
    p.add_argument("--guided-only", action="store_true", help="only the guided path-tracing measurement")
    p.add_argument("--guided-res", type=str, default="1280x720")
    p.add_argument("--guided-scene", choices=["room", "large"], default="room",
                   help="the occluded room, or the large terrain (at --guided-res) for its guided half")
    p.add_argument("--guided-ref-spp", type=int, default=4096)
    p.add_argument("--guided-spp", type=int, default=640,
                   help="samples of each tracer; near the 1e-3 target so that little is extrapolated")
    p.add_argument("--guided-bsdf-prob", type=float, default=0.5)
    p.add_argument("--guided-selection", choices=["learn", "fixed"], default="learn",
                   help="BSDF selection probability: learned per region, or --guided-bsdf-prob everywhere")
    p.add_argument("--guided-selection-lr", type=float, default=0.05)
    p.add_argument("--guided-fallback", action="store_true",
                   help="per-region fallback to BSDF sampling where guiding's running second moment is the larger")
    p.add_argument("--guided-steps", type=int, default=16, help="Adam steps after each training pass")
    p.add_argument("--guided-lr", type=float, default=1e-2)
    p.add_argument("--guided-depth", type=int, default=GUIDE_DEPTH,
                   help="path vertices after the camera vertex that are guided (-1: all)")
    p.add_argument("--guided-verify", action="store_true",
                   help="also render both tracers to the predicted sample counts and measure the relMSE reached")
    p.add_argument("--guided-train-passes", type=int, default=8,
                   help="training passes over the first 32 spp (32 / passes spp each)")
    return p.parse_args()


def measure_path_tracer(args, torch):
    """Cornell box + 10k random triangles, 256 x 256, 16 spp, 5 bounces: GPU
    wavefront render time and rays/s over all rays cast, beside the CPU
    restatement of the reference's integrator on the host cores."""
    from libyafaray_amd import scenes
    from libyafaray_amd.pt import GpuPathTracer
    sc = scenes.cornell_path_scene()
    pt = GpuPathTracer(sc, treelet_passes=args.pt_treelet_passes)
    pt.clear()
    pt.render(0, 16)  # warm-up (kernel loads, first-touch)
    torch.cuda.synchronize()
    times, st = [], None
    for _ in range(3):
        pt.clear()
        t0 = time.perf_counter()
        pt.render(0, 16)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        st = pt.stats()
    pt.close()
    rays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
    best = min(times)
    out = {"workload": "cornell box + 10000 random tris, 256x256, 16 spp, bounces 5, RR from depth 3",
           "render_ms": round(best * 1e3, 2), "rays": int(rays),
           "mrays_per_s": round(rays / best / 1e6, 1),
           "camera_rays": int(st["camera_rays"]), "extension_rays": int(st["extension_rays"]),
           "shadow_rays": int(st["shadow_rays"])}
    if args.cpu_sample > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or cpu_share()
        cpu = pyoracle.PathTracer(sc)
        t0 = time.perf_counter()
        cpu.render(threads=threads)
        cpu_s = time.perf_counter() - t0
        out["cpu_port_render_ms"] = round(cpu_s * 1e3, 1)
        out["cpu_port_threads"] = threads
        out["speedup_vs_cpu_port"] = round(cpu_s / best, 1)
    log(args, f"path tracer: {out}")
    return out


def measure_terrain_frame(args, torch):
    """The mixed-material terrain frame at its full size: 1024^2 heightfield
    + 500 icospheres of 1280 triangles, 1920 x 1080, 64 spp, depth 8, 64 point
    lights + an area light, background 0.1.  One warm-up frame at 1 spp, the
    autotuner's choice of launch configuration (its trial renders reported
    apart), then the best of two 64-spp frames: render time, samples/s and
    rays/s."""
    from libyafaray_amd import scenes
    from libyafaray_amd.pt import GpuPathTracer
    sc = scenes.terrain_spheres_scene()
    cam, st0 = sc["camera"], sc["settings"]
    tb = time.perf_counter()
    pt = GpuPathTracer(sc, treelet_passes=args.pt_treelet_passes)
    setup_s = time.perf_counter() - tb
    pt.clear()
    pt.render(0, 1)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    tuning = pt.autotune()  # launch configuration for this scene (not timed with the frame)
    tune_s = time.perf_counter() - ta
    times, st = [], None
    for _ in range(2):
        pt.clear()
        t0 = time.perf_counter()
        pt.render(0, st0["aa_samples"])
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        st = pt.stats()
    pt.close()
    best = min(times)
    rays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
    samples = cam["resx"] * cam["resy"] * st0["aa_samples"]
    out = {"workload": f"terrain 1024^2 + 500 icospheres ({len(sc['tris'])} tris), {cam['resx']}x{cam['resy']}, "
                       f"{st0['aa_samples']} spp, depth {st0['bounces']}, {len(sc['lights'])} lights",
           "render_ms": round(best * 1e3, 1), "msamples_per_s": round(samples / best / 1e6, 1),
           "rays": int(rays), "mrays_per_s": round(rays / best / 1e6, 1),
           "camera_rays": int(st["camera_rays"]), "extension_rays": int(st["extension_rays"]),
           "shadow_rays": int(st["shadow_rays"]), "setup_ms": round(setup_s * 1e3, 1),
           "autotune_ms": round(tune_s * 1e3, 1), "tuning": tuning}
    log(args, f"terrain frame: {out}")
    return out


def measure_large_frame(args, torch):
    """The large terrain at its full size (`scenes.large_terrain_scene`: 4096^2
    heightfield + 2000 icospheres, 36.1M triangles, 256 point lights + 2 area
    lights, depth 12) at 3840 x 2160 and args.large_spp samples: BVH build,
    wavefront state, one frame.  Not in the default run (the scene takes about
    half a minute to generate on the host)."""
    from libyafaray_amd import scenes
    from libyafaray_amd.pt import GpuPathTracer
    tg = time.perf_counter()
    sc = scenes.large_terrain_scene(spp=args.large_spp)
    gen_s = time.perf_counter() - tg
    tb = time.perf_counter()
    pt = GpuPathTracer(sc, treelet_passes=args.pt_treelet_passes)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - tb
    pt.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    st = {"camera_rays": 0, "extension_rays": 0, "shadow_rays": 0}
    while done < args.large_spp:  # in slices of 16 samples, each reported (stats are per render call)
        k = min(16, args.large_spp - done)
        pt.render(done, k)
        torch.cuda.synchronize()
        done += k
        s_k = pt.stats()
        for key in st:
            st[key] += s_k[key]
        log(args, f"large frame: {done} of {args.large_spp} spp, {time.perf_counter() - t0:.1f} s")
    frame_s = time.perf_counter() - t0
    tuning = pt.tuning()
    rgba, w = pt.film()
    pt.close()
    img = rgba[..., :3] / w[..., None]
    cam = sc["camera"]
    rays = st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
    samples = cam["resx"] * cam["resy"] * args.large_spp
    out = {"workload": f"large terrain 4096^2 + 2000 icospheres ({len(sc['tris'])} tris), {cam['resx']}x{cam['resy']}, "
                       f"{args.large_spp} spp, depth {sc['settings']['bounces']}, {len(sc['lights'])} lights",
           "scene_generation_s": round(gen_s, 1), "setup_s": round(setup_s, 2), "frame_s": round(frame_s, 3),
           "msamples_per_s": round(samples / frame_s / 1e6, 2), "mrays_per_s": round(rays / frame_s / 1e6, 1),
           "rays": int(rays), "shadow_rays": int(st["shadow_rays"]), "samples_per_wavefront": tuning["samples_per_wavefront"],
           "finite": bool(np.isfinite(img).all()), "mean": round(float(img.mean()), 5)}
    log(args, f"large frame: {out}")
    return out


GUIDE_N_MIN, GUIDE_N_MAX = 4, 32   # hash-grid resolutions of the guided occluded-room run
GUIDE_TRAIN_SPP = 32               # the guide trains on the first 32 spp of the render
GUIDE_DEPTH = 1                    # guided: the camera vertex and the next path vertex
_GUIDING_REFERENCE = {}            # (large scene, resx, resy, spp) -> (reference image, its own relMSE)


def measure_guiding(args, torch):
    """Occluded room (4 x 3 x 3 m, about 533k triangles, lit through a 0.1 m
    slit from an adjacent room; max depth 10), or with --guided-scene large the
    large terrain (36.1M triangles, depth 12) at --guided-res: per-pixel relMSE
    (I - R)^2 / (R^2 + 0.01) against a `guided_ref_spp` plain reference (two
    independent halves, whose own error is subtracted), and the time each
    tracer needs to reach relMSE 1e-3.

    Plain: `guided_spp` samples; relMSE falls as 1 / samples, so the time to
    1e-3 is t_plain * relMSE / 1e-3.

    Guided: the guide (16-level hash grid of 2^19 entries x 2 features, 3 x 64
    MLP, 8 vMF lobes, BF16 MFMA, Adam lr 1e-2) trains online on the first 32
    spp of the render (8 passes of 4 spp, 16 Adam steps after each), then the
    render continues with the trained guide up to `guided_spp`.  The camera
    vertex and the next path vertex are guided (deeper vertices add cost and
    no measurable variance reduction on this scene).  Every sample,
    training passes included, is a sample of the image (all are unbiased);
    guided sampling mixes the BSDF sampler (probability bsdf_prob) and the
    mixture by one-sample MIS.  With v_t and v_g the per-sample error of the
    training and post-training samples (relMSE x samples of each part, the
    post-training part taken from the film sums after minus before), an image of
    32 + m samples has relMSE (32 v_t + m v_g) / (32 + m)^2; the time to 1e-3 is
    the training time plus m times the post-training time per sample, for the
    smallest such m."""
    from libyafaray_amd import scenes
    from libyafaray_amd.guide import GpuGuide
    from libyafaray_amd.pt import GpuPathTracer
    resx, resy = (int(v) for v in args.guided_res.split("x"))
    large = getattr(args, "guided_scene", "room") == "large"
    if large:
        sc = scenes.large_terrain_scene(res=(resx, resy), spp=16)
        workload = (f"large terrain 4096^2 + 2000 icospheres, {len(sc['tris'])} tris, {resx}x{resy}, "
                    f"depth {sc['settings']['bounces']}, {len(sc['lights'])} lights")
    else:
        sc = scenes.occluded_room_scene(res=(resx, resy), spp=16)
        workload = f"occluded room, {len(sc['tris'])} tris, {resx}x{resy}, depth 10, slit 0.1 m"
    spp, ref_spp = args.guided_spp, args.guided_ref_spp
    n0 = GUIDE_TRAIN_SPP
    if spp <= n0:
        raise SystemExit(f"--guided-spp must exceed the {n0} training samples")

    def tracer(aa, seed):
        s = dict(sc)
        s["settings"] = dict(sc["settings"], aa_samples=aa, seed=seed)
        return GpuPathTracer(s, treelet_passes=getattr(args, "pt_treelet_passes", 0))

    def ratio(rgba, w):
        return rgba[..., :3].astype(np.float64) / w[..., None]

    def image(pt):
        return ratio(*pt.film())

    def relmse(a, b):
        return float(((a - b) ** 2 / (b.astype(np.float64) ** 2 + 1e-2)).mean())

    log(args, f"guided: {workload}")
    key = (large, resx, resy, ref_spp)
    if key not in _GUIDING_REFERENCE:
        pt = tracer(ref_spp // 2, 777)
        pt.clear()
        pt.render(0, ref_spp // 2)
        r1 = image(pt)
        pt.clear()
        pt.render(ref_spp // 2, ref_spp // 2)
        r2 = image(pt)
        pt.close()
        _GUIDING_REFERENCE[key] = ((r1 + r2) / 2, relmse(r1, r2) / 4)
        log(args, f"guided: {ref_spp}-spp reference rendered")
    ref, ref_err = _GUIDING_REFERENCE[key]

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    p = tracer(spp, 0)
    p.clear()
    p.render(0, 1)  # warm-up
    p.clear()
    t_plain = timed(lambda: p.render(0, spp))
    e_plain = relmse(image(p), ref) - ref_err
    bmin, bmax = sc["verts"].min(0) - 1e-3, sc["verts"].max(0) + 1e-3
    alpha, steps = args.guided_bsdf_prob, args.guided_steps
    n_min, n_max = getattr(args, "guided_n_min", GUIDE_N_MIN), getattr(args, "guided_n_max", GUIDE_N_MAX)
    g = GpuGuide(bmin, bmax, log2_table_size=19, n_min=n_min, n_max=n_max, max_batch=1 << 16,
                 lr=args.guided_lr, seed=1)
    depth = getattr(args, "guided_depth", GUIDE_DEPTH)
    passes = getattr(args, "guided_train_passes", 8)
    if passes < 1 or n0 % passes:
        raise SystemExit(f"--guided-train-passes must divide {n0}")
    p.set_guide_depth(None if depth < 0 else depth)
    p.set_guide(g, bsdf_prob=alpha, train=True, train_steps=steps)
    selection = getattr(args, "guided_selection", "learn")
    if selection == "learn":
        p.set_guide_selection("learn", lr=args.guided_selection_lr)
    fallback = getattr(args, "guided_fallback", False)
    if fallback:
        p.set_guide_fallback("start")
    p.clear()

    def train():
        for k in range(passes):
            p.render(k * (n0 // passes), n0 // passes)

    t_train = timed(train)
    rgba_t, w_t = p.film()
    e_train = relmse(ratio(rgba_t, w_t), ref) - ref_err
    p.set_guide(g, bsdf_prob=alpha, train=False)
    sel_stats = None
    if selection == "learn":
        p.set_guide_selection("keep")
        a_sel = p.guide_selection()
        sel_stats = {"min": round(float(a_sel.min()), 3), "median": round(float(np.median(a_sel)), 3),
                     "max": round(float(a_sel.max()), 3)}
    fb_stats = None
    if fallback:
        p.set_guide_fallback("keep")
        mg, mb, records = p.guide_fallback()
        seen = records >= 256
        fb_stats = {"regions_with_256_records": int(seen.sum()), "fallen_back": int((seen & (mg > mb)).sum())}
    m = spp - n0
    t_rest = timed(lambda: p.render(n0, m))
    rgba_c, w_c = p.film()
    e_all = relmse(ratio(rgba_c, w_c), ref) - ref_err
    e_rest = relmse(ratio(rgba_c.astype(np.float64) - rgba_t, w_c.astype(np.float64) - w_t), ref) - ref_err
    target = 1e-3
    verify = None
    if getattr(args, "guided_verify", False):
        # render each tracer to its predicted sample count and measure the relMSE it reaches
        v_t0, v_g0 = e_train * n0, e_rest * m
        a0, b0, c0 = target, 2 * target * n0 - v_g0, target * n0 * n0 - n0 * v_t0
        m_v = 0 if c0 >= 0 else int(np.ceil((-b0 + (b0 * b0 - 4 * a0 * c0) ** 0.5) / (2 * a0)))
        n_plain = int(np.ceil(spp * e_plain / target))
        pv = tracer(n_plain, 0)
        pv.clear()
        t_pv = timed(lambda: pv.render(0, n_plain))
        e_pv = relmse(image(pv), ref) - ref_err
        pv.close()
        g2 = GpuGuide(bmin, bmax, log2_table_size=19, n_min=n_min, n_max=n_max, max_batch=1 << 16,
                      lr=args.guided_lr, seed=1)
        pg = tracer(n0 + m_v, 0)  # a tracer whose AA sample count covers every sample rendered
        pg.set_guide_depth(None if depth < 0 else depth)
        pg.set_guide(g2, bsdf_prob=alpha, train=True, train_steps=steps)
        if selection == "learn":
            pg.set_guide_selection("learn", lr=args.guided_selection_lr)
        if fallback:
            pg.set_guide_fallback("start")
        pg.clear()

        def train_v():
            for k in range(passes):
                pg.render(k * (n0 // passes), n0 // passes)

        t_gv = timed(train_v)
        pg.set_guide(g2, bsdf_prob=alpha, train=False)
        if selection == "learn":
            pg.set_guide_selection("keep")
        if fallback:
            pg.set_guide_fallback("keep")
        t_gv += timed(lambda: pg.render(n0, m_v) if m_v > 0 else None)
        e_gv = relmse(image(pg), ref) - ref_err
        pg.set_guide(None)
        pg.close()
        g2.close()
        verify = {"plain_spp": n_plain, "plain_s": round(t_pv, 3), "plain_relmse": e_pv,
                  "guided_spp": n0 + m_v, "guided_s": round(t_gv, 3), "guided_relmse": e_gv,
                  "guided_over_plain_time": round(t_gv / t_pv, 3) if t_pv > 0 else None}
    p.set_guide(None)
    g.close()
    p.close()
    tt_plain = t_plain * e_plain / target
    v_t, v_g = e_train * n0, e_rest * m
    # smallest m with (n0 v_t + m v_g) / (n0 + m)^2 <= target (a quadratic in m)
    a_, b_, c_ = target, 2 * target * n0 - v_g, target * n0 * n0 - n0 * v_t
    m_need = 0.0 if c_ >= 0 else (-b_ + (b_ * b_ - 4 * a_ * c_) ** 0.5) / (2 * a_)
    tt_guided = t_train + m_need * t_rest / m
    # equal time: the guided image rendered in the plain tracer's time for `spp`
    m_eq = max(0.0, (t_plain - t_train) / (t_rest / m))
    e_guided_eq = (n0 * v_t + m_eq * v_g) / (n0 + m_eq) ** 2
    out = {"workload": workload,
           "reference_spp": ref_spp, "reference_relmse": ref_err, "spp": spp,
           "plain_ms": round(t_plain * 1e3, 1),
           "guided_ms": round((t_train + t_rest) * 1e3, 1),
           "train_ms": round(t_train * 1e3, 1),
           "train": f"first {n0} spp of the render, {passes} passes of {n0 // passes} spp, {steps} Adam steps after each",
           "guided_vertices": "all" if depth < 0 else f"camera vertex and the next {depth}",
           "post_train_ms_per_spp": round(t_rest / m * 1e3, 2), "plain_ms_per_spp": round(t_plain / spp * 1e3, 2),
           "bsdf_prob": alpha, "lr": args.guided_lr, "grid_levels": f"16 (resolution {n_min}..{n_max})",
           "selection": selection if selection == "fixed" else f"learned per region (lr {args.guided_selection_lr})",
           "learned_selection": sel_stats,
           "fallback": fb_stats,
           "equal_time_relmse_plain": e_plain, "equal_time_relmse_guided": e_guided_eq,
           "equal_time_gain": round(e_plain / e_guided_eq, 3) if e_guided_eq > 0 else None,
           "relmse_plain": e_plain, "relmse_guided": e_all,
           "relmse_training_part": e_train, "relmse_post_training_part": e_rest,
           "equal_spp_gain": round(e_plain / e_all, 3) if e_all > 0 else None,
           "per_sample_gain_after_training": round((e_plain * spp) / v_g, 3) if v_g > 0 else None,
           "time_to_relmse_1e-3_plain_s": round(tt_plain, 2), "time_to_relmse_1e-3_guided_s": round(tt_guided, 2),
           "guided_over_plain_time": round(tt_guided / tt_plain, 3) if tt_plain > 0 else None}
    if verify is not None:
        out["verify_at_predicted_spp"] = verify
    log(args, f"guided path tracing: {out}")
    return out


BF16_PEAK_TFLOPS = 2500.0
MLP_FLOPS = 2 * (3 * 64 * 64 + 48 * 64)   # one query through the 3 x 64 hidden layers and the 48-wide output


def measure_guide_kernels(args, torch):
    """The guiding network's kernels on their own: inference over 2^21 queries
    and one training step (forward, backward, weight gradients, Adam) over a
    65,536-sample batch, timed with HIP events on the stream they run on.
    MFMA utilisation = MLP FLOPs (inference: 30,720 per query; training: three
    times that, forward plus the input- and weight-gradient products) over
    kernel time, against the dense BF16 peak.  The hash-grid gathers and
    scatters are not counted as FLOPs."""
    from libyafaray_amd.guide import GpuGuide, pack_samples
    rng = np.random.default_rng(11)

    def unit(n):
        v = rng.normal(size=(n, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    n_inf, n_train = 1 << 21, 1 << 16
    # the grid of the guided occluded-room run (16 levels, resolution 4 .. 32)
    g = GpuGuide((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), log2_table_size=19, n_min=GUIDE_N_MIN, n_max=GUIDE_N_MAX,
                 max_batch=n_train, seed=2)
    q = torch.from_numpy(pack_samples(rng.uniform(0, 1, (n_inf, 3)).astype(np.float32), unit(n_inf),
                                      unit(n_inf))).cuda()
    smp = torch.from_numpy(pack_samples(rng.uniform(0, 1, (n_train, 3)).astype(np.float32), unit(n_train),
                                        unit(n_train), unit(n_train),
                                        rng.uniform(0.1, 2.0, n_train).astype(np.float32))).cuda()
    mix = torch.empty((n_inf, 48), dtype=torch.float32, device="cuda")
    _, inf_ms = timed_passes(lambda: g.infer(q, mix), 10, 2, torch)
    _, train_ms = timed_passes(lambda: g.train_step(smp), 10, 2, torch)
    g.close()
    inf_tflops = MLP_FLOPS * n_inf / (inf_ms * 1e-3) / 1e12
    train_tflops = 3 * MLP_FLOPS * n_train / (train_ms * 1e-3) / 1e12
    out = {"inference_queries": n_inf, "inference_ms": round(inf_ms, 3),
           "inference_mqueries_per_s": round(n_inf / (inf_ms * 1e-3) / 1e6, 1),
           "inference_mlp_tflops": round(inf_tflops, 2), "inference_mfma_frac": round(inf_tflops / BF16_PEAK_TFLOPS, 5),
           "train_batch": n_train, "train_step_ms": round(train_ms, 3),
           "train_mlp_tflops": round(train_tflops, 2), "train_mfma_frac": round(train_tflops / BF16_PEAK_TFLOPS, 5),
           "peak_bf16_tflops": BF16_PEAK_TFLOPS}
    log(args, f"guide kernels: {out}")
    return out


def calibrate_copy():
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters give the
    scale between the counters and bytes moved."""
    import torch
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": 4 * n, "launches": 3}), flush=True)


def cpu_share() -> int:
    """Host threads for the CPU baselines: the machine's CPU share of one GPU
    (16), or fewer when fewer CPUs are visible."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def counter_pass(args, torch):
    """The kernels a counter pass profiles, launched a few times each with
    nothing else on the device: the ray-cast batches (incoherent closest hit,
    coherent closest hit, incoherent any hit), one path-traced Cornell frame,
    or the guide's inference and training step."""
    from libyafaray_amd import scenes
    if args.counter_pass == "raycast":
        from libyafaray_amd.accel import GpuAccelerator
        verts, tris, coherent, incoherent, shadow = scenes.raycast_workload(args.res, args.n_random)
        acc = GpuAccelerator(verts, tris, builder=args.builder, treelet_passes=args.treelet_passes,
                             ray_reorder=args.ray_reorder)
        d_inc = torch.from_numpy(incoherent).cuda()
        d_coh = torch.from_numpy(coherent).cuda()
        d_sha = torch.from_numpy(shadow).cuda()
        hits = torch.empty((d_inc.shape[0], 4), dtype=torch.int32, device="cuda")
        occ = torch.empty(((d_inc.shape[0] + 31) // 32,), dtype=torch.int32, device="cuda")
        for _ in range(3):
            acc.intersect_device(d_inc, hits)
            acc.intersect_device(d_coh, hits)
            acc.intersect_shadow_device(d_sha, occ)
        torch.cuda.synchronize()
        acc.close()
        del d_inc, d_coh, d_sha, hits, occ
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE scale the counters to bytes
        a = torch.ones(COPY_FLOATS, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    elif args.counter_pass == "pathtrace":
        from libyafaray_amd.pt import GpuPathTracer
        pt = GpuPathTracer(scenes.cornell_path_scene(), treelet_passes=args.pt_treelet_passes)
        pt.clear()
        pt.render(0, 16)
        torch.cuda.synchronize()
        pt.close()
    else:
        measure_guide_kernels(args, torch)
    print(json.dumps({"counter_pass": args.counter_pass, "done": True}), flush=True)


COPY_FLOATS = 1 << 28  # 1 GiB of float32

# rocprofv3 counter passes, each within one run's hardware limits (at most 8
# SQ and 4 TCC counters: FETCH_SIZE takes 3 TCC, WRITE_SIZE 2).  Derived
# counters as rocprofv3 defines them: VALUBusy = 100 SQ_ACTIVE_INST_VALU /
# CU_NUM / GRBM_GUI_ACTIVE, VALUUtilization = 100 SQ_THREAD_CYCLES_VALU /
# (SQ_ACTIVE_INST_VALU x 64), MfmaUtil = 100 SQ_VALU_MFMA_BUSY_CYCLES /
# (GRBM_GUI_ACTIVE x SIMD_NUM).
COUNTER_PASSES = {
    "valu": ["VALUBusy", "VALUUtilization", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_WAVES", "SQ_INSTS_VMEM_RD",
             "SQ_INSTS_LDS"],
    "fetch": ["FETCH_SIZE", "TCC_HIT_sum", "MfmaUtil", "SQ_INSTS_VALU_MFMA_MOPS_BF16"],
    "write": ["WRITE_SIZE", "TCC_MISS_sum", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES"],
}


def counter_run(args, workload, counters, regex):
    """One rocprofv3 counter pass over `bench.py --counter-pass workload` in a
    child process; returns [(kernel name, {counter: value})] in dispatch order,
    or None when rocprofv3 is missing or the pass fails."""
    rocprof = shutil.which("rocprofv3")
    if not rocprof:
        return None
    out = tempfile.mkdtemp(prefix="ymx_pmc_", dir="/tmp")
    cmd = ["timeout", "-s", "KILL", "180", rocprof, "--pmc", *counters, "--kernel-include-regex", regex, "-d", out,
           "-o", "run", "--output-format", "csv", "--", sys.executable, str(Path(__file__).resolve()),
           "--counter-pass", workload, "--builder", args.builder, "--treelet-passes", str(args.treelet_passes),
           "--pt-treelet-passes", str(args.pt_treelet_passes), "--ray-reorder", str(args.ray_reorder),
           "--res", str(args.res), "--n-random",
           str(args.n_random), "--quiet"]
    env = dict(os.environ, TMPDIR="/tmp")
    res = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd="/tmp")
    path = Path(out) / "run_counter_collection.csv"
    if not path.exists():
        log(args, f"counter pass {workload} {counters[0]}...: no output (rc {res.returncode}) {res.stderr[-400:]}")
        return None
    rows = {}
    for r in csv.DictReader(path.open()):
        key = int(r["Dispatch_Id"])
        rows.setdefault(key, (r["Kernel_Name"], {}))[1][r["Counter_Name"]] = float(r["Counter_Value"])
    shutil.rmtree(out, ignore_errors=True)
    return [rows[k] for k in sorted(rows)]


def short_kernel(name):
    """'void (anonymous namespace)::k_dyn<0, false>(ymx_trace_args)' -> 'k_dyn<0, false>'."""
    n = name.replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    n = n.split("(")[0]
    base, _, tmpl = n.partition("<")
    base = base.split("::")[-1]
    return base + ("<" + tmpl[:40] if tmpl else "")


def profile_counters(args):
    """Counter passes of the ray-cast batches (with the 1 GiB calibration
    copy), the path-traced Cornell frame and the guide kernels; averages per
    kernel and batch over the dispatches of each pass."""
    out = {}

    def add(label, counters):
        slot = out.setdefault(label, {"dispatches": 0})
        for k, v in counters.items():
            slot.setdefault(k, []).append(v)

    # ray cast: dispatches repeat (incoherent closest, coherent closest, any hit), then the copies
    for pname in ("valu", "fetch", "write"):
        rows = counter_run(args, "raycast", COUNTER_PASSES[pname], "k_dyn|copy|elementwise")
        if rows is None:
            return None
        order = ["incoherent closest-hit k_dyn<0>", "coherent closest-hit k_dyn<0>", "incoherent any-hit k_dyn<1>"]
        k = 0
        for name, c in rows:
            s = short_kernel(name)
            if s.startswith("k_dyn"):
                add(order[k % 3], c)
                k += 1
            else:
                add("copy 1 GiB", c)
    for workload, regex, passes in (("pathtrace", "k_trace|k_dyn|k_shade|k_subpath|k_resolve|k_camera|k_film",
                                     ("valu",)),
                                    ("guide", "k_infer|k_forward|k_backward|k_grid|k_wgrad|k_adam", ("valu", "fetch"))):
        for pname in passes:
            rows = counter_run(args, workload, COUNTER_PASSES[pname], regex)
            if rows is None:
                continue
            for name, c in rows:
                add(f"{workload} {short_kernel(name)}", c)
    res = {}
    for label, slot in out.items():
        r = {}
        for k, v in slot.items():
            if k == "dispatches":
                continue
            r[k] = float(np.mean(v))
        r["dispatches"] = max(len(v) for k, v in slot.items() if k != "dispatches")
        res[label] = r
    copy = res.get("copy 1 GiB", {})
    fs = ws = None
    if copy.get("FETCH_SIZE") and copy.get("WRITE_SIZE"):
        fs = 4.0 * COPY_FLOATS / (copy["FETCH_SIZE"] * 1024.0)
        ws = 4.0 * COPY_FLOATS / (copy["WRITE_SIZE"] * 1024.0)
    for label, r in res.items():
        if fs and "FETCH_SIZE" in r:
            r["hbm_read_bytes"] = r["FETCH_SIZE"] * 1024.0 * fs
        if ws and "WRITE_SIZE" in r:
            r["hbm_write_bytes"] = r["WRITE_SIZE"] * 1024.0 * ws
        if "TCC_HIT_sum" in r and "TCC_MISS_sum" in r and r["TCC_HIT_sum"] + r["TCC_MISS_sum"] > 0:
            r["l2_hit_rate"] = r["TCC_HIT_sum"] / (r["TCC_HIT_sum"] + r["TCC_MISS_sum"])
    return {"kernels": res, "fetch_scale": fs, "write_scale": ws,
            "note": "per-dispatch means over rocprofv3 --pmc passes; bytes = counter KiB x 1024 x the factor that "
                    "maps a 1 GiB device copy's FETCH_SIZE / WRITE_SIZE to 1 GiB"}


def log(args, msg):
    if not args.quiet:
        print(msg, file=sys.stderr, flush=True)


def timed_passes(fn, steps, warmup, torch):
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps * 1e3, ev0.elapsed_time(ev1) / steps


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N > 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from libyafaray_amd import scenes
    from libyafaray_amd.accel import GpuAccelerator

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.counter_pass:
        counter_pass(args, torch)
        return
    if args.pt_only:
        print(json.dumps({"path_tracing": measure_path_tracer(args, torch)}), flush=True)
        return
    if args.terrain_only:
        print(json.dumps({"terrain_frame": measure_terrain_frame(args, torch)}), flush=True)
        return
    if args.large_only:
        print(json.dumps({"large_frame": measure_large_frame(args, torch)}), flush=True)
        return
    if args.guided_only:
        print(json.dumps({"guide_kernels": measure_guide_kernels(args, torch),
                          "guided_path_tracing": measure_guiding(args, torch)}), flush=True)
        return

    t0 = time.perf_counter()
    verts, tris, coherent, incoherent, shadow = scenes.raycast_workload(args.res, args.n_random)
    log(args, f"workload: {len(tris)} tris, {len(coherent)} coherent + {len(incoherent)} incoherent rays "
              f"({time.perf_counter() - t0:.1f}s)")
    acc = GpuAccelerator(verts, tris, builder=args.builder, treelet_passes=args.treelet_passes,
                         ray_reorder=args.ray_reorder, reorder_cell_bits=args.reorder_cell_bits)
    info = acc.info()
    log(args, f"{args.builder}: {info['n_nodes']} nodes, depth {info['max_depth']}, build {info['build_ms']:.1f} ms, "
              f"treelet passes {info['treelet_passes']} ({info['treelets_rewritten']} rewritten)")

    d_inc = torch.from_numpy(incoherent).cuda()
    d_coh = torch.from_numpy(coherent).cuda()
    d_sha = torch.from_numpy(shadow).cuda()
    n = d_inc.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    occ = torch.empty(((n + 31) // 32,), dtype=torch.int32, device="cuda")

    wall_ms, ev_ms = timed_passes(lambda: acc.intersect_device(d_inc, hits), args.steps, args.warmup, torch)
    value = n / (wall_ms * 1e-3) / 1e6
    log(args, f"incoherent closest: {ev_ms:.3f} ms/pass (events), {value:.1f} Mrays/s")
    _, coh_ms = timed_passes(lambda: acc.intersect_device(d_coh, hits), max(3, args.steps // 2), 2, torch)
    _, sha_ms = timed_passes(lambda: acc.intersect_shadow_device(d_sha, occ), max(3, args.steps // 2), 2, torch)

    # counting replay of the same kernels on the same batches
    nf_inc, nt_inc = acc.traversal_stats(d_inc, 0)
    nf_coh, nt_coh = acc.traversal_stats(d_coh, 0)
    nf_sha, nt_sha = acc.traversal_stats(d_sha, 1)
    nodes_per_ray = nf_inc / n
    tris_per_ray = nt_inc / n
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray
    achieved = bytes_per_ray * n / (ev_ms * 1e-3) / 1e9

    counters = None if args.no_counters else profile_counters(args)
    traffic = valu_busy = None
    if counters:
        k = counters["kernels"].get("incoherent closest-hit k_dyn<0>", {})
        if "hbm_read_bytes" in k and "hbm_write_bytes" in k:
            traffic = int(k["hbm_read_bytes"] + k["hbm_write_bytes"])
        valu_busy = k.get("VALUBusy")
        log(args, f"counters: traffic {traffic}, VALU busy {valu_busy}")

    cpu = None
    if args.cpu_sample > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or cpu_share()
        sample = incoherent[: args.cpu_sample]
        tb = time.perf_counter()
        kd = pyoracle.KdTree(verts, tris)
        build_s = time.perf_counter() - tb
        tq = time.perf_counter()
        kd.closest(sample, threads=threads)
        q_s = time.perf_counter() - tq
        ssample = shadow[: args.cpu_sample]
        ts = time.perf_counter()
        kd.shadow(ssample, threads=threads)
        s_s = time.perf_counter() - ts
        cpu = {"value": round(len(sample) / q_s / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} of the incoherent rays, closest-hit, C port of the reference kd-tree "
                         f"(SAH build {build_s:.1f}s not timed)",
               "anyhit_value": round(len(ssample) / s_s / 1e6, 4),
               "cpu_model": cpu_model(), "logical_cpus_visible": os.cpu_count(),
               "threads_note": "threads = the host-core share of one GPU on this machine (16), all of it used"}
        log(args, f"cpu kd-tree port: {cpu['value']} Mrays/s closest, {cpu['anyhit_value']} any hit on {threads} threads")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"raycast heightfield {args.res}x{args.res} + {args.n_random} random tris "
                        f"({len(tris)} tris); {n} incoherent closest-hit rays per step",
            "triangles": int(len(tris)),
            "rays_per_step": int(n),
            "query": "closest-hit",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "valu_busy_pct": None if valu_busy is None else round(valu_busy, 2),
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 2),
            "tris_per_ray": round(tris_per_ray, 2),
            "kernel_ms": round(ev_ms, 4),
        },
        "cpu_baseline": cpu,
        "extra": {
            "coherent_closest_mrays": round(len(coherent) / (coh_ms * 1e-3) / 1e6, 1),
            "incoherent_anyhit_mrays": round(n / (sha_ms * 1e-3) / 1e6, 1),
            "coherent_nodes_per_ray": round(nf_coh / len(coherent), 2),
            "coherent_tris_per_ray": round(nt_coh / len(coherent), 2),
            "anyhit_nodes_per_ray": round(nf_sha / n, 2),
            "anyhit_tris_per_ray": round(nt_sha / n, 2),
            "bvh_builder": args.builder,
            "bvh_treelet_passes": info["treelet_passes"],
            "bvh_build_ms": round(info["build_ms"], 2),
            "bvh_nodes": info["n_nodes"],
            "bvh_depth": info["max_depth"],
        },
    }
    if counters:
        out["extra"]["counters"] = counters
    if not args.no_pt:
        out["extra"]["path_tracing"] = measure_path_tracer(args, torch)
        if not args.no_terrain:
            out["extra"]["terrain_frame"] = measure_terrain_frame(args, torch)
    if not args.no_guided:
        out["extra"]["guided_path_tracing"] = measure_guiding(args, torch)
        out["extra"]["guide_kernels"] = measure_guide_kernels(args, torch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
