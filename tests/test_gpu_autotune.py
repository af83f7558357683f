"""The wavefront's launch configuration (ymx_pt_autotune, ymx_pt_set_tuning):
every configuration -- wave-ordered or ray-replacement kernels on the shadow
and extension queues, fewer AA samples per wavefront -- gives the same film
bit for bit, the autotuner leaves the film as it was and keeps the fastest
trial, and invalid configurations are refused."""
import sys

import numpy as np
import pytest

from libyafaray_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def u32(x):
    return np.ascontiguousarray(x).view(np.uint32)


def render(pt, spp):
    pt.clear()
    pt.render(0, spp)
    return pt.film()


def same(a, b):
    return np.array_equal(u32(a[0]), u32(b[0])) and np.array_equal(u32(a[1]), u32(b[1]))


def test_every_configuration_gives_the_same_film():
    from libyafaray_amd.pt import GpuPathTracer
    sc = scenes.terrain_spheres_scene(res=(96, 54), spp=8, hf_res=64, n_spheres=40, subdiv=1, n_point_lights=8, seed=4)
    pt = GpuPathTracer(sc)
    t0 = pt.tuning()
    say(f"autotune scene {len(sc['tris'])} tris, default tuning {t0}")
    assert t0["samples_per_wavefront"] == 8
    ref = render(pt, 8)
    for spw in (8, 3, 1):
        for sh in ("wave-ordered", "replacement"):
            for ext in ("wave-ordered", "replacement"):
                pt.set_tuning(spw, sh, ext)
                assert pt.tuning() == {"samples_per_wavefront": spw, "shadow_kernel": sh, "extension_kernel": ext}
                assert same(render(pt, 8), ref), (spw, sh, ext)
    pt.close()


def test_autotune_keeps_the_film_and_the_fastest():
    from libyafaray_amd.pt import GpuPathTracer
    from libyafaray_amd.native import YmxError
    sc = scenes.cornell_path_scene(n_random=3000, res=64, spp=8, seed=5)
    pt = GpuPathTracer(sc)
    ref = render(pt, 8)
    before = pt.film()
    res = pt.autotune(trial_samples=4)
    say(f"  autotune: {res}")
    after = pt.film()
    assert same(before, after)
    trials = res["trials"]
    assert 4 <= len(trials) <= 8
    best = min(trials, key=lambda t: t["ms_per_sample"])
    assert {k: res[k] for k in ("samples_per_wavefront", "shadow_kernel", "extension_kernel")} == \
        {k: best[k] for k in ("samples_per_wavefront", "shadow_kernel", "extension_kernel")}
    assert pt.tuning() == {k: res[k] for k in ("samples_per_wavefront", "shadow_kernel", "extension_kernel")}
    assert same(render(pt, 8), ref)
    with pytest.raises(YmxError):
        pt.set_tuning(9, "wave-ordered", "wave-ordered")  # more than the allocated 8
    with pytest.raises(YmxError):
        pt.set_tuning(0, "wave-ordered", "wave-ordered")
    pt.close()


def test_transparent_shadows_stay_wave_ordered():
    from libyafaray_amd.pt import GpuPathTracer
    from libyafaray_amd.native import YmxError
    sc = scenes.cornell_materials_scene(n_random=200, res=32, spp=2, seed=7, raydepth=2, point_lights=1)
    sc["settings"] = dict(sc["settings"], transpShad=True, shadowDepth=3)
    for m in sc["materials"]:
        if m.get("type") == "glass":
            m["fake_shadows"] = True
    pt = GpuPathTracer(sc)
    assert pt.tuning()["shadow_kernel"] == "wave-ordered"
    ref = render(pt, 2)
    res = pt.autotune(trial_samples=1)
    assert res["shadow_kernel"] == "wave-ordered"
    assert all(t["shadow_kernel"] == "wave-ordered" for t in res["trials"])
    with pytest.raises(YmxError):
        pt.set_tuning(1, "replacement", "wave-ordered")
    assert same(render(pt, 2), ref)
    pt.close()
