"""Smooth shading normals on the GPU (ymx_pt_set_normals) against the oracle:
icospheres of every material with radial vertex normals, some with corners
that have none, in the Cornell box; films bit-exact, and different from the
flat-shaded film."""
import sys

import numpy as np
import pytest

from libyafaray_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_smooth_spheres_match_oracle(oracle):
    from libyafaray_amd.pt import GpuPathTracer
    sc = scenes.smooth_spheres_scene(res=64, spp=4)
    say(f"smooth spheres tris={len(sc['tris'])} res=64 spp=4")
    pt = GpuPathTracer(sc)
    pt.clear()
    pt.render(0, 4)
    g_rgba, g_w = pt.film()
    pt.set_normals(None)
    pt.clear()
    pt.render(0, 4)
    f_rgba, _ = pt.film()
    pt.close()
    o_rgba, o_w = oracle.PathTracer(sc, brute=True).render(threads=8)
    assert np.array_equal(g_w.view(np.uint32), o_w.view(np.uint32))
    same = np.all(g_rgba.view(np.uint32) == o_rgba.view(np.uint32), axis=-1)
    assert same.all(), f"{(~same).sum()} pixels differ"
    changed = np.any(f_rgba != g_rgba, axis=-1).mean()
    say(f"  pixels changed by the shading normals: {changed:.3f}")
    assert changed > 0.2


def test_smooth_normals_refused_on_fake_shadow_glass():
    from libyafaray_amd import native
    from libyafaray_amd.pt import GpuPathTracer
    sc = scenes.smooth_spheres_scene(res=16, spp=1)
    sc["materials"] = [dict(m, fake_shadows=True) if m.get("type") == "glass" else m for m in sc["materials"]]
    sc["settings"] = dict(sc["settings"], transpShad=True)
    with pytest.raises(native.YmxError, match="fake-shadow glass"):
        GpuPathTracer(sc)
