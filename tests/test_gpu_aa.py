"""Multi-pass adaptive anti-aliasing on the GPU (ymx_pt_render_passes) against
the oracle's restatement of TiledIntegrator::render and ImageFilm::nextPass:
films bit-exact (so every pass flagged the same pixels and traced the same
samples), for subsets of pixels, every pixel, the variance window, dark
detection, sample clamping and growing sample / light multipliers."""
import sys

import numpy as np
import pytest

from libyafaray_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def scene(res=(40, 30), spp=2, n_random=1500, seed=3, light_scale=1.0, **aa):
    sc = scenes.cornell_path_scene(n_random=n_random, seed=seed, res=res[0], spp=spp)
    sc["camera"] = dict(sc["camera"], resx=res[0], resy=res[1])
    sc["settings"] = dict(sc["settings"], **aa)
    sc["lights"] = [dict(l, power=l.get("power", 1.0) * light_scale) for l in sc["lights"]]
    return sc


CASES = {
    "subset": dict(AA_passes=3, AA_inc_samples=2, AA_threshold=0.02),
    "all_pixels": dict(AA_passes=2, AA_inc_samples=1, AA_threshold=0.0),
    "variance": dict(AA_passes=3, AA_inc_samples=1, AA_threshold=0.04, AA_variance_edge_size=6,
                     AA_variance_pixels=4),
    "dark_curve_clamp": dict(AA_passes=3, AA_inc_samples=1, AA_threshold=0.01, AA_dark_detection_type="curve",
                             AA_clamp_samples=3.0, light_scale=20.0),
    "multipliers_floor": dict(AA_passes=4, AA_inc_samples=2, AA_threshold=0.03, AA_sample_multiplier_factor=1.5,
                              AA_light_sample_multiplier_factor=2.0, AA_resampled_floor=20.0,
                              AA_detect_color_noise=True, AA_dark_detection_type="linear",
                              AA_dark_threshold_factor=0.5),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_passes_match_oracle(oracle, case):
    from libyafaray_amd.pt import GpuPathTracer
    sc = scene(**CASES[case])
    say(f"aa passes case={case} res={sc['camera']['resx']}x{sc['camera']['resy']} settings={CASES[case]}")
    pt = GpuPathTracer(sc)
    pt.clear()
    res = pt.render_passes()
    grgba, gw = pt.film()
    pt.close()
    # brute-force closest hits: exact ties in t resolve to the lowest id on
    # both sides (the kd-tree port breaks them in leaf order, and the scrambled
    # multi-pass offsets put samples on the box's edges)
    orgba, ow = oracle.PathTracer(sc, brute=True).render(threads=8)
    say(f"  resampled per pass {res}")
    assert res[0] == 40 * 30
    if case in ("subset", "variance", "dark_curve_clamp"):
        assert 0 < res[1] < 40 * 30
    if case == "all_pixels":
        assert res[1] == 40 * 30
    assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32)), f"{case}: filter weights differ"
    same = np.all(grgba.view(np.uint32) == orgba.view(np.uint32), axis=-1)
    bad = np.argwhere(~same)
    assert same.all(), f"{case}: {(~same).sum()} pixels differ, first (y, x) {bad[:6].tolist()}"


def test_single_pass_render_is_pass_zero(oracle):
    """ymx_pt_render_passes with one pass is ymx_pt_render of AA_minsamples."""
    from libyafaray_amd.pt import GpuPathTracer
    sc = scene(spp=3)
    a = GpuPathTracer(sc)
    a.clear()
    assert a.render_passes() == [40 * 30]
    fa = a.film()
    a.clear()
    a.render(0, 3)
    fb = a.film()
    a.close()
    assert np.array_equal(fa[0].view(np.uint32), fb[0].view(np.uint32))
    assert np.array_equal(fa[1].view(np.uint32), fb[1].view(np.uint32))
