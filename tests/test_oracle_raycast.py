"""Pins the CPU oracle of the ray-cast path against closed forms, a float64
re-derivation, analytic geometry and brute force (no golden vectors exist in
the reference and its kd-tree cannot be compiled here without CMake-generated
headers).  Each check is chosen so that a plausible slip -- swapped u/v, a sign
error, farthest-instead-of-closest, a dropped bias -- fails it."""
import numpy as np
import pytest

from libyafaray_amd import scenes

F32 = np.float32
FMAX = float(np.finfo(np.float32).max)


# ---------------------------------------------------------------- triangle
TRI = ((0, 0, 0), (1, 0, 0), (0, 1, 0))


def test_triangle_closed_form_front(oracle):
    hit, t, u, v = oracle.triangle_intersect(*TRI, (0.125, 0.5, 1.0), (0, 0, -1))
    # u weighs v1 - v0 (x), v weighs v2 - v0 (y): P = v0 + u e1 + v e2
    assert hit and (t, u, v) == (1.0, 0.125, 0.5)


def test_triangle_double_sided(oracle):
    hit, t, u, v = oracle.triangle_intersect(*TRI, (0.125, 0.5, -1.0), (0, 0, 1))
    assert hit and (t, u, v) == (1.0, 0.125, 0.5)


@pytest.mark.parametrize("org,expect", [
    ((0.5, 0.5, 1.0), True),        # on the hypotenuse: u + v == 1 accepted
    ((0.0, 0.0, 1.0), True),        # vertex v0: u == v == 0 accepted
    ((1.0, 0.0, 1.0), True),        # vertex v1: u == 1 accepted
    ((0.5, 0.5000001, 1.0), False),  # just outside u + v <= 1
    ((-1e-7, 0.5, 1.0), False),     # u < 0
    ((0.5, -1e-7, 1.0), False),     # v < 0
])
def test_triangle_edges(oracle, org, expect):
    assert oracle.triangle_intersect(*TRI, org, (0, 0, -1))[0] is expect


def test_triangle_parallel_and_behind(oracle):
    assert not oracle.triangle_intersect(*TRI, (0.1, 0.1, 0.0), (1, 0, 0))[0]  # det == 0
    assert not oracle.triangle_intersect(*TRI, (0.125, 0.5, 1.0), (0, 0, 1))[0]  # t < 0
    hit, t, _, _ = oracle.triangle_intersect(*TRI, (0.125, 0.5, 0.0), (0, 0, -1))
    assert not hit and t == 0.0  # t == 0 is a miss


def _mt64(v0, v1, v2, o, d):
    """float64 Moller-Trumbore for reference values (not the oracle's formula
    order: Cramer's rule on [-d, e1, e2])."""
    e1 = v1 - v0
    e2 = v2 - v0
    m = np.stack([-d, e1, e2], axis=-1)
    det = np.linalg.det(m)
    ok = np.abs(det) > 1e-12
    rhs = o - v0
    sol = np.full(d.shape, np.nan)
    sol[ok] = np.linalg.solve(m[ok], rhs[ok][..., None])[..., 0]
    return sol[:, 0], sol[:, 1], sol[:, 2]


def test_triangle_matches_float64(oracle):
    rng = np.random.default_rng(3)
    n = 4000
    v = rng.uniform(-1, 1, size=(n, 3, 3)).astype(F32)
    o = rng.uniform(-2, 2, size=(n, 3)).astype(F32)
    target = (v[:, 0] + rng.uniform(-0.3, 1.0) * (v[:, 1] - v[:, 0]) * 0.6 +
              (v[:, 2] - v[:, 0]) * rng.uniform(-0.3, 1.0, size=(n, 1)) * 0.6)
    d = (target - o).astype(F32)
    t64, u64, v64 = _mt64(v[:, 0].astype(float), v[:, 1].astype(float), v[:, 2].astype(float),
                          o.astype(float), d.astype(float))
    margin = 1e-4
    checked = 0
    for i in range(n):
        if np.isnan(t64[i]):
            continue
        inside = u64[i] >= 0 and v64[i] >= 0 and u64[i] + v64[i] <= 1 and t64[i] > 0
        near = min(abs(u64[i]), abs(v64[i]), abs(1 - u64[i] - v64[i]), abs(t64[i])) < margin
        e1 = v[i, 1].astype(float) - v[i, 0]
        e2 = v[i, 2].astype(float) - v[i, 0]
        dd = d[i].astype(float)
        cond = abs(np.dot(np.cross(dd, e2), e1)) / (np.linalg.norm(dd) * np.linalg.norm(e1) * np.linalg.norm(e2))
        if near or cond < 0.05:  # ill-conditioned: float32 rounding dominates
            continue
        hit, t, u, vv = oracle.triangle_intersect(v[i, 0], v[i, 1], v[i, 2], o[i], d[i])
        assert hit == inside, i
        if hit:
            assert abs(t - t64[i]) <= 1e-4 * max(1.0, abs(t64[i]))
            assert abs(u - u64[i]) <= 1e-4 and abs(vv - v64[i]) <= 1e-4
        checked += 1
    assert checked > 2000


# ---------------------------------------------------------------- slab test
UNIT = ((0, 0, 0), (1, 1, 1))


@pytest.mark.parametrize("org,dir,tmax,expect", [
    ((-1, 0.5, 0.5), (1, 0, 0), FMAX, (True, 1.0, 2.0)),
    ((2, 0.5, 0.5), (-1, 0, 0), FMAX, (True, 1.0, 2.0)),
    ((-1, -1, -1), (1, 1, 1), FMAX, (True, 1.0, 2.0)),
    ((0.5, 0.5, 0.5), (0, 0, 1), FMAX, (True, -0.5, 0.5)),   # origin inside
    ((-1, 0.5, 0.5), (1, 0, 0), 0.5, (False, None, None)),   # lmin > t_max
    ((2, 0.5, 0.5), (1, 0, 0), FMAX, (False, None, None)),   # lmax < 0
    ((-1, 5.0, 0.5), (1, 0, 0), FMAX, (True, 1.0, 2.0)),     # zero dir.y: no y rejection (quirk)
    ((-1, 5.0, 0.5), (1, 1e-30, 0), FMAX, (False, None, None)),
    # origin on a face with a denormal direction component: 1/dir is inf and
    # -p * inv_dir is 0 * inf = NaN.  std::max(ltmin, lmin) keeps a NaN ltmin
    # and replaces a NaN lmin (include/geometry/bound.h:187), so a NaN on the
    # first axis is overwritten by the next axis ...
    ((0, 0.5, 0.5), (1e-40, 1, 0), FMAX, (True, -0.5, 0.5)),
    # ... and a NaN on a later axis survives and fails lmin <= lmax
    ((0.5, 0, 0.5), (1, 1e-40, 0), FMAX, (False, None, None)),
    ((0.5, 0.5, 1), (1, 0, -1e-40), FMAX, (False, None, None)),
])
def test_bound_cross_closed_form(oracle, org, dir, tmax, expect):
    hit, e, l = oracle.bound_cross(*UNIT, org, dir, tmax)
    assert hit is expect[0]
    if hit:
        assert (e, l) == expect[1:]


def test_tree_bound_growth(oracle):
    verts = np.array([[0, 0, 0], [1, 2, 4], [0.5, 0.5, 0.5]], dtype=F32)
    tris = np.array([[0, 1, 2]], dtype=np.uint32)
    a, g = oracle.tree_bound(verts, tris)
    # each axis grows by 0.1 % of its extent, rounded to float32
    assert list(a) == [-F32(0.001), -F32(0.002), -F32(0.004)]
    assert list(g) == [F32(1) + F32(0.001), F32(2) + F32(0.002), F32(4) + F32(0.004)]


def test_tree_bound_unused_vertex_ignored(oracle):
    verts = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [100, 100, 100]], dtype=F32)
    tris = np.array([[0, 1, 2]], dtype=np.uint32)
    a, g = oracle.tree_bound(verts, tris)
    assert g[0] < 2


def test_dynamic_bias_scale(oracle):
    for enter, leave in ((0.0, 1.0), (-3.0, 5.0), (2.0, 2.5)):
        b = oracle.dynamic_bias(enter, leave)
        assert abs(b - 5e-6 * abs(leave - enter)) <= 1e-6 * 5e-6 * abs(leave - enter) * 10


# ---------------------------------------------------------------- closest hit
def _rays(org, dir, tmin=0.0, tmax=-1.0):
    org = np.atleast_2d(np.asarray(org, dtype=F32))
    dir = np.atleast_2d(np.asarray(dir, dtype=F32))
    n = max(len(org), len(dir))
    r = np.zeros((n, 8), dtype=F32)
    r[:, 0:3] = org
    r[:, 3] = tmin
    r[:, 4:7] = dir
    r[:, 7] = tmax
    return r


def test_closest_on_analytic_heightfield(oracle):
    res = 17
    rng = np.random.default_rng(11)
    h = rng.uniform(0, 0.2, size=(res, res))
    verts, tris = scenes.grid_mesh(h)
    n = 3000
    xy = rng.uniform(0.001, 0.999, size=(n, 2))
    cell = xy * (res - 1)
    j = np.floor(cell[:, 0]).astype(int)
    i = np.floor(cell[:, 1]).astype(int)
    fx = cell[:, 0] - j
    fy = cell[:, 1] - i
    keep = (np.abs(fx - fy) > 1e-3) & (fx > 1e-3) & (fy > 1e-3) & (fx < 1 - 1e-3) & (fy < 1 - 1e-3)
    xy, i, j, fx, fy = xy[keep], i[keep], j[keep], fx[keep], fy[keep]
    upper = fy > fx  # second triangle of the cell (v00, v11, v10)
    prim = 2 * (i * (res - 1) + j) + upper
    hz = verts[:, 2].astype(float).reshape(res, res)
    h00, h01, h10, h11 = hz[i, j], hz[i, j + 1], hz[i + 1, j], hz[i + 1, j + 1]
    # first triangle: P = v00 + u (v01 - v00) + v (v11 - v00), (fx, fy) = (u + v, v)
    u1, v1 = fx - fy, fy
    # second triangle: P = v00 + u (v11 - v00) + v (v10 - v00), (fx, fy) = (u, u + v)
    u2, v2 = fx, fy - fx
    uu = np.where(upper, u2, u1)
    vv = np.where(upper, v2, v1)
    z = np.where(upper, h00 + u2 * (h11 - h00) + v2 * (h10 - h00), h00 + u1 * (h01 - h00) + v1 * (h11 - h00))
    org = np.column_stack([xy, np.ones(len(xy))])
    hits = oracle.closest_brute(verts, tris, _rays(org, (0, 0, -1)))
    assert np.array_equal(hits["prim"], prim.astype(np.uint32))
    assert np.allclose(hits["t"], 1.0 - z, atol=2e-6)
    assert np.allclose(hits["u"], uu, atol=2e-5) and np.allclose(hits["v"], vv, atol=2e-5)


def _planes(zs, size=4.0):
    """Stacked horizontal squares at heights zs (2 triangles each)."""
    verts, tris = [], []
    for k, z in enumerate(zs):
        b = 4 * k
        verts += [(-size, -size, z), (size, -size, z), (size, size, z), (-size, size, z)]
        tris += [(b, b + 1, b + 2), (b, b + 2, b + 3)]
    return np.array(verts, dtype=F32), np.array(tris, dtype=np.uint32)


def test_closest_is_nearest_not_first(oracle):
    verts, tris = _planes([0.1, 0.2, 0.3])
    down = oracle.closest_brute(verts, tris, _rays((0.3, 0.4, 1.0), (0, 0, -1)))
    up = oracle.closest_brute(verts, tris, _rays((0.3, 0.4, -1.0), (0, 0, 1)))
    assert down["prim"][0] // 2 == 2 and abs(down["t"][0] - 0.7) < 1e-6
    assert up["prim"][0] // 2 == 0 and abs(up["t"][0] - 1.1) < 1e-6


def test_closest_respects_tmin_tmax_and_bias(oracle):
    verts, tris = _planes([0.1, 0.2, 0.3])
    # tmin skips the first plane
    r = oracle.closest_brute(verts, tris, _rays((0.3, 0.4, 0.0), (0, 0, 1), tmin=0.15))
    assert r["prim"][0] // 2 == 1
    # tmax cuts below the first plane, and just above it
    assert oracle.closest_brute(verts, tris, _rays((0.3, 0.4, 0.05), (0, 0, 1), tmax=0.04))["prim"][0] == oracle.NO_HIT
    assert oracle.closest_brute(verts, tris, _rays((0.3, 0.4, 0.05), (0, 0, 1), tmax=0.06))["prim"][0] // 2 == 0
    # a hit closer than the dynamic bias (5e-6 x crossing length) is rejected
    z0 = np.nextafter(F32(0.2), F32(0), dtype=F32)
    r = oracle.closest_brute(verts, tris, _rays((0.3, 0.4, z0), (0, 0, 1)))
    assert r["prim"][0] // 2 == 2


def test_visibility_masks(oracle):
    verts, tris = _planes([0.1, 0.2, 0.3])
    vis = np.full(6, 3, dtype=np.uint8)
    vis[4:6] = 2  # top plane: shadow only (not visible to camera rays)
    r = oracle.closest_brute(verts, tris, _rays((0.3, 0.4, 1.0), (0, 0, -1)), vis=vis)
    assert r["prim"][0] // 2 == 1
    vis2 = np.full(6, 1, dtype=np.uint8)  # no plane casts shadows
    occ = oracle.shadow_brute(verts, tris, _rays((0.3, 0.4, 1.0), (0, 0, -1), tmax=5.0), vis=vis2)
    assert occ[0] == 0
    occ = oracle.shadow_brute(verts, tris, _rays((0.3, 0.4, 1.0), (0, 0, -1), tmax=5.0), vis=vis)
    assert occ[0] == 1


def test_shadow_tmax_window(oracle):
    verts, tris = _planes([0.5])
    rays = _rays(np.array([[0.1, 0.1, 0.0]] * 3), (0, 0, 1))
    rays[:, 7] = [0.49, 0.5, 0.51]
    occ = oracle.shadow_brute(verts, tris, rays)
    assert list(occ) == [0, 0, 1]  # t_hit >= t_max is not an occlusion


def test_exact_tie_keeps_lowest_id(oracle):
    verts, tris = _planes([0.3, 0.3])
    r = oracle.closest_brute(verts, tris, _rays((0.3, 0.4, 1.0), (0, 0, -1)))
    assert r["prim"][0] // 2 == 0  # the copy in the first (lower-id) plane


def test_brute_matches_float64_closest(oracle):
    rng = np.random.default_rng(5)
    tv = scenes.random_triangles(rng, 400, (0, 0, 0), (1, 1, 1), edge=(0.05, 0.2))
    verts, tris = scenes.soup(tv)
    rays = scenes.incoherent_rays(rng, 1500, (0, 0, 0), (1, 1, 1))
    got = oracle.closest_brute(verts, tris, rays)
    o = rays[:, 0:3].astype(float)
    d = rays[:, 4:7].astype(float)
    v = tv.astype(float)
    best = np.full(len(rays), np.inf)
    second = np.full(len(rays), np.inf)
    arg = np.full(len(rays), -1)
    for p in range(len(tv)):
        t, u, w = _mt64(np.broadcast_to(v[p, 0], o.shape), np.broadcast_to(v[p, 1], o.shape),
                        np.broadcast_to(v[p, 2], o.shape), o, d)
        ok = (t > 1e-4) & (u >= 0) & (w >= 0) & (u + w <= 1)
        t = np.where(ok, t, np.inf)
        closer = t < best
        second = np.where(closer, best, np.minimum(second, t))
        arg = np.where(closer, p, arg)
        best = np.where(closer, t, best)
    clear = np.isfinite(best) & ((second - best) > 1e-4) | ~np.isfinite(best)
    expect = np.where(np.isfinite(best), arg, -1).astype(np.int64)
    gotp = got["prim"].astype(np.int64)
    gotp[gotp == oracle.NO_HIT] = -1
    # ignore rays grazing triangle edges in float64 (tolerance band)
    assert clear.sum() > 1300
    assert np.mean(gotp[clear] == expect[clear]) > 0.995


# ---------------------------------------------------------------- kd-tree port
@pytest.mark.parametrize("seed", [0, 1])
def test_kdtree_port_equals_brute(oracle, seed):
    verts, tris = scenes.heightfield_scene(res=48, n_random=1500, seed=seed)
    rng = np.random.default_rng(seed + 7)
    lo, hi = verts.min(0), verts.max(0)
    rays = np.concatenate([
        scenes.incoherent_rays(rng, 6000, lo, hi),
        scenes.camera_rays(64, 32, **scenes.RAYCAST_CAMERA),
    ])
    kd = oracle.KdTree(verts, tris)
    assert np.array_equal(kd.bound()[0], oracle.tree_bound(verts, tris)[0])
    b = oracle.closest_brute(verts, tris, rays)
    k = kd.closest(rays)
    same = b["prim"] == k["prim"]
    # any difference must be an exact tie in t
    assert np.all(b["t"][~same] == k["t"][~same])
    assert np.array_equal(b["t"], k["t"])
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    occ_b = oracle.shadow_brute(verts, tris, srays)
    occ_k, _ = kd.shadow(srays)
    assert np.array_equal(occ_b, occ_k)


def test_brute_in_bound_equals_whole_scene(oracle):
    """Brute force over a culled subset in the whole scene's bound equals
    brute force over the whole scene for short segments the subset holds
    (ids mapped back).  The subset's own bound differs from the scene's, and
    the queries clip to the bound and scale their bias by its chord, so the
    scene's bound is the one to pass."""
    verts, tris = scenes.heightfield_scene(res=64, n_random=800, seed=5)
    rng = np.random.default_rng(5)
    # origins near the terrain surface of a small patch: centroids of its
    # triangles moved up or down by at most 0.02
    cen = verts[tris[: 2 * 63 * 63]].mean(1)
    patch = cen[(np.abs(cen[:, 0] - 0.42) < 0.03) & (np.abs(cen[:, 1] - 0.52) < 0.03)]
    org = patch[rng.integers(0, len(patch), 3000)] + np.stack(
        [np.zeros(3000), np.zeros(3000), rng.uniform(-0.02, 0.02, 3000)], 1)
    rays = scenes.incoherent_rays(rng, 3000, (0, 0, 0), (1, 1, 1))
    rays[:, 0:3] = org
    rays[:, 7] = rng.uniform(0.001, 0.05, 3000)
    full = oracle.closest_brute(verts, tris, rays)
    ends = rays[:, 0:3] + rays[:, 4:7] * rays[:, 7:8]
    lo = np.minimum(rays[:, 0:3], ends).min(0) - 1e-4
    hi = np.maximum(rays[:, 0:3], ends).max(0) + 1e-4
    tv = verts[tris]
    sub = np.nonzero(np.all((tv.max(1) >= lo) & (tv.min(1) <= hi), axis=1))[0]
    assert 0 < len(sub) < len(tris) // 4
    part = oracle.closest_brute_in_bound(verts, tris[sub], rays, oracle.tree_bound(verts, tris))
    hit = part["prim"] != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.95
    mapped = np.where(hit, sub[np.where(hit, part["prim"], 0)], 0xFFFFFFFF)
    assert np.array_equal(mapped, full["prim"])
    for k in ("t", "u", "v"):
        assert np.array_equal(part[k].view(np.uint32), full[k].view(np.uint32))
    # the subset's own bound is another bound (so closest_brute over the subset would clip and bias differently)
    a_sub, g_sub = oracle.tree_bound(verts, tris[sub])
    a_all, g_all = oracle.tree_bound(verts, tris)
    assert np.any(a_sub != a_all) and np.any(g_sub != g_all)


def test_retest_reproduces_hits(oracle):
    verts, tris = scenes.heightfield_scene(res=32, n_random=300, seed=2)
    rng = np.random.default_rng(2)
    rays = scenes.incoherent_rays(rng, 3000, verts.min(0), verts.max(0))
    h = oracle.closest_brute(verts, tris, rays)
    r = oracle.retest(verts, tris, rays, h["prim"])
    assert np.array_equal(r["prim"], h["prim"])
    assert np.array_equal(r["t"].view(np.uint32), h["t"].view(np.uint32))


# ---------------------------------------------------------------- fast math
def test_fast_math_known_values(oracle):
    L = oracle.lib()
    assert L.ora_fast_log2(1.0) == 0.0 and L.ora_fast_log2(2.0) == 1.0 and L.ora_fast_log2(0.5) == -1.0
    assert L.ora_fast_exp2(0.0) == F32(9.9999994e-1)   # polyexp(0)
    assert L.ora_fast_sin(0.0) == 0.0 and abs(L.ora_fast_sin(np.pi / 2) - 1.0) < 2e-7
    xs = np.linspace(-7, 7, 5001)
    assert all(-1.0 <= L.ora_fast_sin(float(x)) <= 1.0 for x in xs)  # clamped to [-1, 1]


def test_fast_math_error_bounds(oracle):
    L = oracle.lib()
    xs = np.linspace(-10, 10, 2001).astype(F32)
    assert max(abs(L.ora_fast_exp2(float(x)) / 2.0 ** float(x) - 1) for x in xs) < 3e-4
    xs = np.exp(np.linspace(-20, 20, 2001)).astype(F32)
    assert max(abs(L.ora_fast_log2(float(x)) - np.log2(float(x))) for x in xs) < 2e-5
    xs = np.linspace(-20, 20, 4001).astype(F32)
    assert max(abs(L.ora_fast_sin(float(x)) - np.sin(float(x))) for x in xs) < 1.2e-3
    assert max(abs(L.ora_fast_cos(float(x)) - np.cos(float(x))) for x in xs) < 1.2e-3


# ---------------------------------------------------------------- quads
# include/geometry/shape/shape_polygon.h:126-176 with N == 4.  On the unit
# square (v0, v1, v2, v3) = (0,0), (1,0), (1,1), (0,1) the first triangle has
# u = x - y, v = y and returns {u + v, v}; the second (v0, v2, v3) has u = x,
# v = y - x and returns {u, u + v}: both give uv == (x, y).
SQUARE = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))


@pytest.mark.parametrize("xy", [(0.75, 0.25), (0.25, 0.75), (0.5, 0.5), (1.0, 1.0), (0.0, 1.0)])
def test_quad_closed_form_uv(oracle, xy):
    hit, t, u, v = oracle.quad_intersect(*SQUARE, (xy[0], xy[1], 1.0), (0, 0, -1))
    assert hit and (t, u, v) == (1.0, xy[0], xy[1])


def test_quad_second_triangle_only_when_u_leaves_unit_interval(oracle):
    # v3 folded out: the point (1.2, 0.5) lies inside the second triangle
    # (u 0.36, v 0.56), but the first triangle's u = x - y = 0.7 is inside
    # [0, 1] and its v test fails (u + v = 1.2): the reference returns a miss
    quad = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (1.5, 0.25, 0))
    assert oracle.triangle_intersect(quad[0], quad[2], quad[3], (1.2, 0.5, 1.0), (0, 0, -1))[0]
    assert not oracle.quad_intersect(*quad, (1.2, 0.5, 1.0), (0, 0, -1))[0]
    # below the diagonal the first triangle's u = x - y = -0.25 < 0 hands over
    hit, t, u, v = oracle.quad_intersect(*quad, (0.5, 0.75, 1.0), (0, 0, -1))
    assert not hit  # outside the second triangle too
    # beyond the first triangle's u = 1 (x - y = 1.1) the second triangle
    # hits, at its (u, v) = (0.08, 0.88), returned as {u, u + v}
    hit, t, u, v = oracle.quad_intersect(*quad, (1.4, 0.3, 1.0), (0, 0, -1))
    assert hit and abs(t - 1.0) < 1e-6 and abs(u - 0.08) < 1e-6 and abs(v - 0.96) < 1e-6


def test_quad_first_triangle_parallel_is_a_miss(oracle):
    # non-planar quad: the ray runs parallel to the first triangle (det 0) but
    # crosses the second one at (0.2, 0.3, 0.1); the reference returns a miss
    quad = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 1))
    assert oracle.triangle_intersect(quad[0], quad[2], quad[3], (0.2, -1.0, 0.1), (0, 1, 0))[0]
    assert not oracle.quad_intersect(*quad, (0.2, -1.0, 0.1), (0, 1, 0))[0]


def test_quad_double_sided_and_behind(oracle):
    hit, t, u, v = oracle.quad_intersect(*SQUARE, (0.25, 0.75, -2.0), (0, 0, 1))
    assert hit and (t, u, v) == (2.0, 0.25, 0.75)
    assert not oracle.quad_intersect(*SQUARE, (0.25, 0.75, 1.0), (0, 0, 1))[0]
    assert not oracle.quad_intersect(*SQUARE, (1.25, 0.5, 1.0), (0, 0, -1))[0]


def test_quad_faces_in_brute_force_and_kdtree(oracle):
    """A face list mixing quads and triangles: every hit re-tests, the tree
    bound covers the fourth vertices, and the kd-tree port agrees."""
    rng = np.random.default_rng(11)
    n = 400
    c = rng.uniform(0.1, 0.9, size=(n, 3)).astype(F32)
    a = rng.normal(size=(n, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = np.cross(a, rng.normal(size=(n, 3)))
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    s = rng.uniform(0.02, 0.06, size=(n, 1))
    quad = np.stack([c - s * a - s * b, c + s * a - s * b, c + s * a + s * b, c - s * a + s * b], 1)
    quad[:, 3] += rng.normal(scale=0.004, size=(n, 3))  # slightly non-planar
    verts = quad.reshape(-1, 3).astype(F32)
    faces = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    faces[::3, 3] = 0xFFFFFFFF  # every third face is a triangle
    lo, hi = verts.min(0), verts.max(0)
    rays = scenes.incoherent_rays(rng, 20000, lo, hi)
    exp = oracle.closest_brute(verts, faces, rays)
    assert (exp["prim"] != oracle.NO_HIT).mean() > 0.2
    ta, tg = oracle.tree_bound(verts, faces)
    used = verts[faces[faces != 0xFFFFFFFF]]
    assert np.all(ta < used.min(0)) and np.all(tg > used.max(0))
    quad_v3 = verts[faces[faces[:, 3] != 0xFFFFFFFF, 3]]
    assert np.all(ta < quad_v3.min(0)) and np.all(tg > quad_v3.max(0))
    hit = exp["prim"] != oracle.NO_HIT
    r = oracle.retest(verts, faces, rays, exp["prim"])
    assert np.array_equal(r["t"][hit].view(np.uint32), exp["t"][hit].view(np.uint32))
    kd = oracle.KdTree(verts, faces).closest(rays)
    same = kd["prim"] == exp["prim"]
    assert same.mean() > 0.999
    assert np.array_equal(kd["t"][~same].view(np.uint32), exp["t"][~same].view(np.uint32))  # exact ties only
    srays = rays.copy()
    srays[:, 7] = scenes.shadow_tmax(rng, len(rays))
    assert np.array_equal(oracle.shadow_brute(verts, faces, srays), oracle.KdTree(verts, faces).shadow(srays)[0])


def test_quad_tree_bound_includes_fourth_vertex(oracle):
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [-1, 2, 0]], dtype=F32)
    a, g = oracle.tree_bound(verts, np.array([[0, 1, 2, 3]], np.uint32))
    assert a[0] == -F32(1) - F32(0.002) and g[1] == F32(2) + F32(0.002)
    a3, g3 = oracle.tree_bound(verts, np.array([[0, 1, 2, 0xFFFFFFFF]], np.uint32))
    assert a3[0] == -F32(0.001) and g3[1] == F32(1) + F32(0.001)
