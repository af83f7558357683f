"""The allocation-free closest-hit entry point (ymx_accel_intersect_noalloc)
against ymx_accel_intersect, through the C ABI: bit-exact hit records on two
calls in a row, the second with a smaller batch (the shared workspace keeps
the first call's contents), and on a tree deep enough to spill its stack."""
import numpy as np
import pytest

from libyafaray_amd import native, scenes
from libyafaray_amd.accel import GpuAccelerator

pytestmark = pytest.mark.gpu


def _noalloc(acc, rays, hits, torch):
    rc = native.lib().ymx_accel_intersect_noalloc(acc.handle, rays.data_ptr(), hits.data_ptr(), rays.shape[0],
                                                  torch.cuda.current_stream().cuda_stream)
    native.check(rc, "ymx_accel_intersect_noalloc")
    return hits


def _scene(kind):
    if kind == "large":
        # 544k triangles: a tree deeper than the 16-entry LDS stack, so the
        # spill columns of the workspace are used
        return scenes.heightfield_scene(res=512, n_random=20000, seed=3)
    return scenes.heightfield_scene(res=32, n_random=200, seed=4)


@pytest.mark.parametrize("kind", ["large", "small"])
def test_noalloc_matches_intersect(kind):
    import torch
    verts, tris = _scene(kind)
    acc = GpuAccelerator(verts, tris)
    rng = np.random.default_rng(7)
    for n in (300_000, 70_001):  # second call smaller, odd-sized
        rays = torch.from_numpy(scenes.incoherent_rays(rng, n, verts.min(0), verts.max(0))).cuda()
        ref = acc.intersect_device(rays)
        got = torch.full_like(ref, -7)
        _noalloc(acc, rays, got, torch)
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        print(f"{kind}: n={n} depth={acc.info()['max_depth']} hits={(r[:, 0] != -1).sum()}", flush=True)
        assert np.array_equal(r, g), f"{kind}: noalloc hits differ from ymx_accel_intersect at n={n}"
    acc.close()


def test_noalloc_rejects_bad_arguments():
    import torch
    verts, tris = _scene("small")
    acc = GpuAccelerator(verts, tris)
    rays = torch.zeros((64, 8), dtype=torch.float32, device="cuda")
    hits = torch.zeros((64, 4), dtype=torch.int32, device="cuda")
    lib = native.lib()
    assert lib.ymx_accel_intersect_noalloc(None, rays.data_ptr(), hits.data_ptr(), 64, None) == -1
    assert lib.ymx_accel_intersect_noalloc(acc.handle, rays.data_ptr(), hits.data_ptr() + 4, 64, None) == -1
    assert b"16-byte aligned" in lib.ymx_last_error()
    assert lib.ymx_accel_intersect_noalloc(acc.handle, rays.data_ptr(), hits.data_ptr(), 0, None) == 0
    acc.close()
