"""Ray reordering (ymx_accel_params.ray_reorder): the batch is traced in the
order of a device counting sort by direction octant and origin cell, and
every result lands in the caller's slot.  Closest and any hit stay bit-exact
against the brute-force oracle (forced on for small and ragged batches), and
large batches (where it is on by default, or forced over more than the
sort's tile limit) equal the unreordered trace bit for bit, incoherent ones
(reordered) and coherent camera batches (left in their order) alike."""
import sys

import numpy as np
import pytest

from libyafaray_amd import scenes
from libyafaray_amd.accel import GpuAccelerator

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_reordered_small_batch_matches_oracle(oracle, builder):
    verts, tris = scenes.heightfield_scene(res=48, n_random=1500, seed=31)
    rng = np.random.default_rng(31)
    lo, hi = verts.min(0), verts.max(0)
    rays = np.concatenate([scenes.incoherent_rays(rng, 30000, lo, hi),
                           scenes.incoherent_rays(rng, 3000, lo - 2.0, hi + 2.0)])
    rays[5, 0:3] = np.nan  # a NaN origin sorts somewhere and still gets its answer
    say(f"reorder small builder={builder} rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, builder=builder, ray_reorder=1)
    got = acc.intersect(rays)
    exp = oracle.closest_brute(verts, tris, rays)
    assert np.array_equal(got["prim"], exp["prim"])
    for k in ("t", "u", "v"):
        assert np.array_equal(bits(got[k]), bits(exp[k]))
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(srays))
    occ, _ = acc.intersect_shadow(srays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, tris, srays))
    acc.close()


def test_reordered_large_batch_equals_unreordered():
    import torch
    verts, tris = scenes.heightfield_scene(res=256, n_random=20000, seed=32)
    rng = np.random.default_rng(32)
    n = (1 << 18) + 12345  # above the automatic threshold, ragged
    rays = torch.from_numpy(scenes.incoherent_rays(rng, n, verts.min(0), verts.max(0))).cuda()
    say(f"reorder large tris={len(tris)} rays={n}")
    a = GpuAccelerator(verts, tris)                  # automatic: reordered
    b = GpuAccelerator(verts, tris, ray_reorder=-1)  # never
    ha, hb = a.intersect_device(rays), b.intersect_device(rays)
    srays = rays.clone()
    srays[:, 7] = torch.from_numpy(scenes.shadow_tmax(rng, n)).cuda()
    oa, ob = a.intersect_shadow_device(srays), b.intersect_shadow_device(srays)
    torch.cuda.synchronize()
    assert torch.equal(ha, hb)
    assert torch.equal(oa, ob)
    a.close()
    b.close()


@pytest.mark.parametrize("cell_bits", [3, 4])
@pytest.mark.parametrize("n", [1, 2, 3, 4097, 8193])
def test_forced_reorder_ragged_sizes(oracle, n, cell_bits):
    verts, tris = scenes.heightfield_scene(res=32, n_random=500, seed=33)
    rng = np.random.default_rng(33 + n)
    rays = scenes.incoherent_rays(rng, n, verts.min(0), verts.max(0))
    say(f"reorder forced n={n} cell_bits={cell_bits}")
    acc = GpuAccelerator(verts, tris, ray_reorder=1, reorder_cell_bits=cell_bits)
    got = acc.intersect(rays)
    exp = oracle.closest_brute(verts, tris, rays)
    assert np.array_equal(got["prim"], exp["prim"])
    assert np.array_equal(bits(got["t"]), bits(exp["t"]))
    acc.close()


@pytest.mark.parametrize("cell_bits", [3, 4])
def test_reorder_over_tile_limit_and_coherent_batch(cell_bits):
    """2.2M incoherent rays forced through the sort (more than the tile limit
    of 4096-ray tiles), and a 1024 x 512 camera batch under the automatic rule."""
    import torch
    verts, tris = scenes.heightfield_scene(res=256, n_random=20000, seed=34)
    rng = np.random.default_rng(34)
    n = 2_200_001
    inc = torch.from_numpy(scenes.incoherent_rays(rng, n, verts.min(0), verts.max(0))).cuda()
    cam = torch.from_numpy(scenes.camera_rays(1024, 512, **scenes.RAYCAST_CAMERA)).cuda()
    say(f"reorder tile limit n={n}, camera {len(cam)}, cell_bits={cell_bits}")
    forced = GpuAccelerator(verts, tris, ray_reorder=1, reorder_cell_bits=cell_bits)
    auto = GpuAccelerator(verts, tris, reorder_cell_bits=cell_bits)
    plain = GpuAccelerator(verts, tris, ray_reorder=-1)
    h_forced, h_plain = forced.intersect_device(inc), plain.intersect_device(inc)
    c_auto, c_plain = auto.intersect_device(cam), plain.intersect_device(cam)
    torch.cuda.synchronize()
    assert torch.equal(h_forced, h_plain)
    assert torch.equal(c_auto, c_plain)
    for a in (forced, auto, plain):
        a.close()
