"""GPU parity of the mi355x-lbvh accelerator against the CPU oracle, through
the C ABI (libymx.so).

Bar: closest-hit prim ids, t, u, v bit-exact against the brute-force oracle
(both resolve exact ties in t to the lowest id); any-hit occlusion bits
bit-exact.  Against the kd-tree port (ties in kd leaf order, like the
reference) every differing id must be an exact tie in t.  At the full
benchmark size the properties are size-independent: every reported hit
re-tests bit-exactly, and shadow rays clipped at the closest hit agree with
it."""
import sys

import numpy as np
import pytest

from libyafaray_amd import scenes
from libyafaray_amd.accel import GpuAccelerator

pytestmark = pytest.mark.gpu
NO_HIT = 0xFFFFFFFF


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)


def assert_hits_equal(got, exp, what):
    for k in ("prim",):
        bad = np.nonzero(got[k] != exp[k])[0]
        assert bad.size == 0, f"{what}: {bad.size} prim mismatches, first {bad[:5]} got {got[k][bad[:5]]} exp {exp[k][bad[:5]]}"
    for k in ("t", "u", "v"):
        bad = np.nonzero(bits(got[k]) != bits(exp[k]))[0]
        assert bad.size == 0, f"{what}: {bad.size} {k} mismatches, first {bad[:5]}"


def small_scene(seed=0, res=64, n_random=2000):
    verts, tris = scenes.heightfield_scene(res=res, n_random=n_random, seed=seed)
    rng = np.random.default_rng(seed + 100)
    lo, hi = verts.min(0), verts.max(0)
    rays = np.concatenate([
        scenes.incoherent_rays(rng, 20000, lo, hi),
        scenes.camera_rays(160, 80, **scenes.RAYCAST_CAMERA),
        # origins outside the scene box, aimed anywhere
        scenes.incoherent_rays(rng, 3000, lo - 2.0, hi + 2.0),
    ])
    return verts, tris, rays, rng


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
@pytest.mark.parametrize("seed", [0, 1])
def test_closest_bit_exact_small(oracle, seed, builder):
    verts, tris, rays, _ = small_scene(seed)
    say(f"closest small seed={seed} tris={len(tris)} rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, builder=builder)
    a, g = acc.bound()
    ea, eg = oracle.tree_bound(verts, tris)
    assert np.array_equal(a, ea) and np.array_equal(g, eg)
    got = acc.intersect(rays)
    exp = oracle.closest_brute(verts, tris, rays)
    assert (exp["prim"] != NO_HIT).mean() > 0.3
    assert_hits_equal(got, exp, "closest")


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
@pytest.mark.parametrize("seed", [0, 1])
def test_shadow_bit_exact_small(oracle, seed, builder):
    verts, tris, rays, rng = small_scene(seed)
    rays = rays.copy()
    rays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    say(f"shadow small seed={seed} rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, builder=builder)
    occ, prim = acc.intersect_shadow(rays)
    exp = oracle.shadow_brute(verts, tris, rays)
    assert 0.05 < exp.mean() < 0.95
    assert np.array_equal(occ, exp)
    # every reported occluder really blocks the ray inside [bias, tmax)
    assert np.all(prim[occ == 0] == NO_HIT)
    r = oracle.retest(verts, tris, rays, np.where(occ == 1, prim, NO_HIT).astype(np.uint32))
    t = r["t"][occ == 1]
    assert np.all(r["prim"][occ == 1] == prim[occ == 1])
    assert np.all((t > 0) & (t < rays[occ == 1, 7]))


def test_visibility_masks(oracle):
    verts, tris, rays, rng = small_scene(2, res=40, n_random=800)
    vis = rng.integers(0, 4, size=len(tris)).astype(np.uint8)
    say(f"visibility rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, vis)
    assert_hits_equal(acc.intersect(rays), oracle.closest_brute(verts, tris, rays, vis=vis), "closest+vis")
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(rays)) * 2
    occ, _ = acc.intersect_shadow(srays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, tris, srays, vis=vis))


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_ties_on_grid_vertices(oracle, builder):
    """Vertical rays through heightfield vertices and edges meet several
    triangles at the same t; the lowest id must win, as in the oracle."""
    verts, tris = scenes.heightfield_scene(res=33, n_random=0, seed=3)
    g = np.linspace(0.0, 1.0, 65, dtype=np.float32)  # every vertex and edge midpoint
    xy = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(-1, 2)
    rays = np.zeros((len(xy), 8), np.float32)
    rays[:, 0:2] = xy
    rays[:, 2] = 1.0
    rays[:, 6] = -1.0
    rays[:, 7] = -1.0
    say(f"ties rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, builder=builder)
    got = acc.intersect(rays)
    exp = oracle.closest_brute(verts, tris, rays)
    assert_hits_equal(got, exp, "ties")


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
@pytest.mark.parametrize("n_tris", [0, 1, 2, 3])
def test_tiny_scenes(oracle, n_tris, builder):
    rng = np.random.default_rng(n_tris)
    tv = scenes.random_triangles(rng, max(n_tris, 1), (0.3, 0.3, 0.3), (0.7, 0.7, 0.7), edge=(0.1, 0.3))[:n_tris]
    verts, tris = scenes.soup(tv) if n_tris else (np.zeros((3, 3), np.float32), np.zeros((0, 3), np.uint32))
    rays = scenes.incoherent_rays(rng, 5000, (0.2, 0.2, 0.2), (0.8, 0.8, 0.8))
    say(f"tiny n_tris={n_tris}")
    acc = GpuAccelerator(verts, tris, builder=builder)
    got = acc.intersect(rays)
    if n_tris == 0:
        assert np.all(got["prim"] == NO_HIT)
        occ, _ = acc.intersect_shadow(rays)
        assert not occ.any()
        return
    assert_hits_equal(got, oracle.closest_brute(verts, tris, rays), f"tiny{n_tris}")
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    occ, _ = acc.intersect_shadow(srays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, tris, srays))


@pytest.mark.parametrize("n_rays", [1, 63, 64, 65, 1000, 4097])
def test_ragged_batches(oracle, n_rays):
    verts, tris, rays, rng = small_scene(4, res=24, n_random=200)
    rays = rays[:n_rays].copy()
    say(f"ragged n_rays={n_rays}")
    acc = GpuAccelerator(verts, tris)
    assert_hits_equal(acc.intersect(rays), oracle.closest_brute(verts, tris, rays), "ragged")
    rays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    occ, _ = acc.intersect_shadow(rays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, tris, rays))


def test_empty_batch_and_bad_args():
    import torch
    from libyafaray_amd import native
    verts, tris = scenes.heightfield_scene(res=8, n_random=0)
    acc = GpuAccelerator(verts, tris)
    r = torch.zeros((0, 8), dtype=torch.float32, device="cuda")
    assert acc.intersect_device(r).shape == (0, 4)
    with pytest.raises(native.YmxError):
        GpuAccelerator(verts, np.array([[0, 1, 10_000]], np.uint32))
    with pytest.raises(ValueError):
        acc.intersect_device(torch.zeros((4, 7), dtype=torch.float32, device="cuda"))


def test_axis_aligned_and_degenerate_rays(oracle):
    verts, tris, _, rng = small_scene(5, res=32, n_random=500)
    lo, hi = verts.min(0), verts.max(0)
    n = 6000
    rays = scenes.incoherent_rays(rng, n, lo, hi)
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    rays[:, 4:7] = 0.0
    rays[np.arange(n), 4 + axis] = sign
    # a third keep one extra non-zero component
    extra = rng.random(n) < 0.33
    rays[extra, 4 + (axis[extra] + 1) % 3] = 0.5
    say("axis-aligned rays")
    acc = GpuAccelerator(verts, tris)
    assert_hits_equal(acc.intersect(rays), oracle.closest_brute(verts, tris, rays), "axis")


def test_face_origin_denormal_direction(oracle):
    """Origins on a face of the tree bound with a denormal direction component
    along that face's axis: the slab test meets -0 * inf = NaN, which
    std::max(ltmin, lmin) drops when it comes from the first axis and keeps
    when it comes from a later one (include/geometry/bound.h:183-188).  Such a
    ray stays in the face plane, which no triangle reaches (the bound is grown
    by 0.1 %), so the order decides only whether the bound is crossed (pinned
    on the CPU in the oracle and the shared header); here every ray must come
    back as the oracle's miss, with no NaN-driven stray hit."""
    verts, tris, _, rng = small_scene(8, res=32, n_random=500)
    acc = GpuAccelerator(verts, tris)
    a, g = acc.bound()
    n = 3000
    den = np.float32(1e-40)
    assert den != 0 and np.isinf(np.float32(1) / den)

    def face_rays(axis, lo_face):
        r = np.zeros((n, 8), np.float32)
        r[:, 0:3] = rng.uniform(a, g, size=(n, 3))
        r[:, axis] = a[axis] if lo_face else g[axis]
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d[:, axis] = den if lo_face else -den
        r[:, 4:7] = d
        r[:, 7] = -1.0
        return r

    rays = np.concatenate([face_rays(0, True), face_rays(1, True), face_rays(2, False)])
    say("face-origin rays with denormal direction components")
    exp = oracle.closest_brute(verts, tris, rays)
    assert_hits_equal(acc.intersect(rays), exp, "face-origin rays")
    occ, _ = acc.intersect_shadow(rays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, tris, rays))
    acc.close()


def test_queries_on_several_streams(oracle):
    """Queries on one handle overlap across streams and host threads: each
    stream has its own work area (ray counter and stack spill), and every
    launch leaves its counters at zero for the next one on that stream."""
    import threading

    import torch
    verts, tris, rays, rng = small_scene(9, res=48, n_random=1500)
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    exp_c = oracle.closest_brute(verts, tris, rays)
    exp_s = oracle.shadow_brute(verts, tris, srays)
    acc = GpuAccelerator(verts, tris)
    d_rays = torch.from_numpy(rays).cuda()
    d_srays = torch.from_numpy(srays).cuda()
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(3)]
    say("interleaved launches on 3 streams")
    outs = []
    for rep in range(4):
        for k, s in enumerate(streams):
            if (rep + k) % 2 == 0:
                outs.append(("c", acc.intersect_device(d_rays, stream=s)))
            else:
                outs.append(("s", acc.intersect_shadow_device(d_srays, stream=s)))
    torch.cuda.synchronize()
    from libyafaray_amd.accel import unpack_bits, unpack_hits
    for kind, o in outs:
        if kind == "c":
            assert_hits_equal(unpack_hits(o.cpu().numpy()), exp_c, "closest on a side stream")
        else:
            assert np.array_equal(unpack_bits(o.cpu().numpy(), len(srays)), exp_s)
    say("two host threads, one stream each")
    results, errors = {}, []

    def worker(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                got = [acc.intersect_device(d_rays, stream=s) for _ in range(5)]
            s.synchronize()
            results[i] = [g.cpu().numpy() for g in got]
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(2):
        for h in results[i]:
            assert_hits_equal(unpack_hits(h), exp_c, f"thread {i}")
    acc.close()


def test_deterministic_repeat():
    verts, tris, rays, _ = small_scene(6)
    acc = GpuAccelerator(verts, tris)
    a = acc.intersect(rays)
    b = acc.intersect(rays)
    for k in a:
        assert np.array_equal(bits(a[k]) if k != "prim" else a[k], bits(b[k]) if k != "prim" else b[k])


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_medium_scene_against_kdtree_port(oracle, builder):
    """~140k triangles: GPU vs the reference-style kd-tree (ties listed)."""
    verts, tris = scenes.heightfield_scene(res=256, n_random=10_000, seed=7)
    rng = np.random.default_rng(70)
    lo, hi = verts.min(0), verts.max(0)
    rays = np.concatenate([scenes.incoherent_rays(rng, 150_000, lo, hi),
                           scenes.camera_rays(512, 256, **scenes.RAYCAST_CAMERA)])
    say(f"medium tris={len(tris)} rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, builder=builder)
    got = acc.intersect(rays)
    kd = oracle.KdTree(verts, tris)
    exp = kd.closest(rays)
    diff = np.nonzero(got["prim"] != exp["prim"])[0]
    say(f"  exact-tie ids (kd order vs lowest id): {diff.size}")
    assert np.array_equal(bits(got["t"]), bits(exp["t"]))
    assert diff.size <= 1e-3 * len(rays)
    if diff.size:
        # both candidates hit at the same t
        r = oracle.retest(verts, tris, rays[diff], exp["prim"][diff])
        assert np.array_equal(bits(r["t"]), bits(got["t"][diff]))
        assert np.all(got["prim"][diff] < exp["prim"][diff])
    # brute force on a subset settles the tie-break exactly
    sub = rng.choice(len(rays), 4000, replace=False)
    assert_hits_equal({k: v[sub] for k, v in got.items()}, oracle.closest_brute(verts, tris, rays[sub]), "medium")


@pytest.mark.timeout(600)
@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_full_raycast_workload_properties(oracle, builder):
    """The benchmark scene (~2.19M triangles) with all 16.8M rays."""
    import torch
    verts, tris, coherent, incoherent, shadow = scenes.raycast_workload()
    say(f"full tris={len(tris)} rays={len(coherent) + len(incoherent)}")
    acc = GpuAccelerator(verts, tris, builder=builder)
    info = acc.info()
    say(f"  nodes={info['n_nodes']} depth={info['max_depth']} build_ms={info['build_ms']:.1f}")
    for name, rays in (("coherent", coherent), ("incoherent", incoherent)):
        d_rays = torch.from_numpy(rays).cuda()
        hits = acc.intersect_device(d_rays)
        torch.cuda.synchronize()
        h = hits.cpu().numpy()
        prim = h[:, 0].view(np.uint32)
        t = h[:, 1].view(np.float32)
        hit = prim != NO_HIT
        say(f"  {name}: hit fraction {hit.mean():.4f}")
        # every reported hit re-tests bit-exactly
        r = oracle.retest(verts, tris, rays, prim)
        assert np.array_equal(r["prim"], prim)
        assert np.array_equal(bits(r["t"]), h[:, 1].view(np.uint32))
        assert np.array_equal(bits(r["u"]), h[:, 2].view(np.uint32))
        # shadow rays clipped at the closest hit: nothing lies strictly before it
        s = rays.copy()
        s[hit, 7] = t[hit]
        occ_at = acc.intersect_shadow_device(torch.from_numpy(s).cuda())
        s[hit, 7] = np.nextafter(t[hit], np.float32(np.inf))
        occ_after = acc.intersect_shadow_device(torch.from_numpy(s).cuda())
        torch.cuda.synchronize()
        from libyafaray_amd.accel import unpack_bits
        at = unpack_bits(occ_at.cpu().numpy(), len(rays))
        after = unpack_bits(occ_after.cpu().numpy(), len(rays))
        assert not at[hit].any()
        assert after[hit].all()
        assert not after[~hit].any()
        del d_rays, hits
    # subset against the kd-tree port
    rng = np.random.default_rng(9)
    sub = rng.choice(len(incoherent), 300_000, replace=False)
    kd = oracle.KdTree(verts, tris)
    exp = kd.closest(incoherent[sub])
    got = acc.intersect(incoherent[sub])
    diff = np.nonzero(got["prim"] != exp["prim"])[0]
    say(f"  kd-port subset: {diff.size} exact-tie ids of {len(sub)}")
    assert np.array_equal(bits(got["t"]), bits(exp["t"]))
    occ_g, _ = acc.intersect_shadow(shadow[sub])
    occ_k, _ = kd.shadow(shadow[sub])
    assert np.array_equal(occ_g, occ_k)


@pytest.mark.parametrize("max_depth", [0, 1, 4])
def test_transparent_shadow_query(oracle, max_depth):
    """intersectTransparentShadow against the kd-tree port's restatement: the
    same occlusion decision (an opaque caster, or more than max_depth
    transparent ones), and for unoccluded rays the same set of transparent
    casters (the GPU lists them by t, the kd-tree in its own order)."""
    verts, tris, rays, rng = small_scene(3, res=24, n_random=3000)
    rays = rays[:6000].copy()
    rays[:, 3] = rng.uniform(0.0, 0.01, len(rays)).astype(np.float32)   # tmin: max(tmin, bias) rule
    rays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    transparent = rng.uniform(size=len(tris)) < 0.7
    say(f"transparent shadows depth={max_depth} rays={len(rays)} tris={len(tris)}")
    acc = GpuAccelerator(verts, tris)
    occ, counts, hits = acc.intersect_transparent(rays, transparent, max_depth)
    kd = oracle.KdTree(verts, tris)
    eocc, ecounts, eprims = kd.transparent(rays, transparent, max_depth)
    say(f"  occluded {eocc.mean():.3f}, unoccluded with casters {((~eocc) & (ecounts > 0)).mean():.3f}")
    assert 0.05 < eocc.mean() < 0.95
    assert np.array_equal(occ.astype(bool), eocc)
    assert np.array_equal(counts, ecounts)
    with_casters = np.nonzero((~eocc) & (ecounts > 0))[0]
    if max_depth > 0:
        assert len(with_casters) > 100
    for i in with_casters:
        k = counts[i]
        assert set(hits[i, :k, 0].tolist()) == set(eprims[i, :k].tolist())
        t = hits[i, :k, 1].view(np.float32)
        assert np.all(np.diff(t) >= 0) and np.all(t > 0) and np.all(t < rays[i, 7])


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_exported_tree_structure(builder):
    """The exported tree is a tree over every triangle exactly once: node 0 is
    the root, every other node has exactly one parent, refs stay in range.
    With leaves of up to 4 triangles (the default) both builders store the
    nodes in depth-first order with a node's internal children in the next
    free adjacent slots, and keep no node under a collapsed leaf."""
    import ctypes as C
    from libyafaray_amd import native
    verts, tris = scenes.heightfield_scene(res=128, n_random=3000, seed=11)
    acc = GpuAccelerator(verts, tris, builder=builder)
    info = acc.info()
    n_nodes, n_tris = info["n_nodes"], len(tris)
    nodes = np.zeros((n_nodes, 16), dtype=np.float32)
    tri_rec = np.zeros((max(1, n_tris), 12), dtype=np.float32)
    native.check(native.lib().ymx_accel_export(acc.handle, nodes.ctypes.data, C.c_uint64(nodes.nbytes),
                                               tri_rec.ctypes.data, C.c_uint64(tri_rec.nbytes)), "ymx_accel_export")
    refs = nodes[:, 12:14].copy().view(np.int32)
    internal = refs >= 0
    assert np.all(refs[internal] < n_nodes)
    parents = np.bincount(refs[internal], minlength=n_nodes)
    assert parents[0] == 0 and np.all(parents[1:] == 1)
    leaf = refs[~internal].view(np.uint32)
    first, count = (leaf & 0x7FFFFFFF) >> 4, (leaf & 15) + 1
    covered = np.zeros(n_tris, dtype=np.int64)
    for f, c in zip(first, count):
        covered[f:f + c] += 1
    assert np.all(covered == 1)
    ids = tri_rec[:, 3].copy().view(np.uint32)
    assert np.array_equal(np.sort(ids), np.arange(n_tris, dtype=np.uint32))
    assert n_nodes < n_tris - 1  # collapsed subtrees dropped
    assert count.max() <= 4 and count.max() > 1
    nxt, stack = 1, [0]
    while stack:
        n = stack.pop()
        ch = [int(r) for r in refs[n] if r >= 0]
        assert ch == list(range(nxt, nxt + len(ch))), (n, ch, nxt)
        nxt += len(ch)
        stack.extend(reversed(ch))
    assert nxt == n_nodes


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
@pytest.mark.parametrize("max_leaf", [1, 2, 3, 8])
def test_leaf_sizes_bit_exact(oracle, builder, max_leaf):
    """Every leaf size gives the brute-force closest hits and occlusion bits
    (leaf collapse, triangle reordering and node relayout change no result)."""
    verts, tris, rays, rng = small_scene(4, res=48, n_random=1500)
    say(f"leaf sizes builder={builder} max_leaf={max_leaf} tris={len(tris)} rays={len(rays)}")
    acc = GpuAccelerator(verts, tris, builder=builder, max_leaf_size=max_leaf)
    assert_hits_equal(acc.intersect(rays), oracle.closest_brute(verts, tris, rays), f"closest leaf {max_leaf}")
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(srays))
    occ, _ = acc.intersect_shadow(srays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, tris, srays))


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
@pytest.mark.parametrize("max_leaf", [1, 4])
def test_quad_faces_bit_exact(oracle, builder, max_leaf):
    """Faces of triangles and quads (ymx_accel_create_faces): the heightfield
    as bent quads, planar and bent random quads, and triangles.  A quad is one
    primitive, tested as include/geometry/shape/shape_polygon.h:126-176 does;
    ids, t and both uv mappings bit-exact against the oracle."""
    verts, faces = scenes.quad_face_scene(res=64, n_random=3000, seed=12)
    rng = np.random.default_rng(12)
    lo, hi = verts.min(0), verts.max(0)
    g = np.linspace(0.0, 1.0, 127, dtype=np.float32)  # cell corners, edges and centres
    xy = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(-1, 2)
    vertical = np.zeros((len(xy), 8), np.float32)
    vertical[:, 0:2] = xy
    vertical[:, 2] = 1.0
    vertical[:, 6] = -1.0
    vertical[:, 7] = -1.0
    rays = np.concatenate([scenes.incoherent_rays(rng, 20000, lo, hi),
                           scenes.camera_rays(160, 80, **scenes.RAYCAST_CAMERA), vertical])
    say(f"quad faces builder={builder} max_leaf={max_leaf} faces={len(faces)} rays={len(rays)}")
    acc = GpuAccelerator(verts, faces, builder=builder, max_leaf_size=max_leaf)
    assert acc.info()["n_quads"] == int((faces[:, 3] != 0xFFFFFFFF).sum())
    a, gb = acc.bound()
    ea, eg = oracle.tree_bound(verts, faces)
    assert np.array_equal(a, ea) and np.array_equal(gb, eg)
    exp = oracle.closest_brute(verts, faces, rays)
    hit_quads = exp["prim"][exp["prim"] != NO_HIT]
    assert (faces[hit_quads, 3] != 0xFFFFFFFF).mean() > 0.5
    # both of the reference's uv mappings occur: second-triangle hits have v >= u
    assert ((exp["v"] > exp["u"]) & (exp["prim"] != NO_HIT)).any()
    assert_hits_equal(acc.intersect(rays), exp, "quad closest")
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    occ, prim = acc.intersect_shadow(srays)
    assert np.array_equal(occ, oracle.shadow_brute(verts, faces, srays))
    r = oracle.retest(verts, faces, srays, np.where(occ == 1, prim, NO_HIT).astype(np.uint32))
    assert np.all(r["prim"][occ == 1] == prim[occ == 1])
    acc.close()


def test_single_quad_and_mixed_tiny(oracle):
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0.2], [0, 1, 0], [0.5, 0.5, 0.5], [2, 0, 0.5], [2, 1, 0.5]],
                     np.float32)
    rng = np.random.default_rng(3)
    rays = scenes.incoherent_rays(rng, 4000, (-0.2, -0.2, -0.5), (1.2, 1.2, 1.0))
    for faces in (np.array([[0, 1, 2, 3]], np.uint32),
                  np.array([[0, 1, 2, 3], [4, 5, 6, 0xFFFFFFFF]], np.uint32)):
        say(f"tiny faces n={len(faces)}")
        acc = GpuAccelerator(verts, faces)
        assert_hits_equal(acc.intersect(rays), oracle.closest_brute(verts, faces, rays), f"{len(faces)} faces")
        acc.close()
