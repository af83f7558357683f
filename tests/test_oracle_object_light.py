"""Pins the oracle's ObjectLight (oracle/ora_pt.c, restating
src/light/light_object_light.cc:87-237 and include/sampler/sample_pdf1d.h):
the area table against closed-form triangle areas, the surface sampler
against uniform area sampling, illumSample's pdf against the closed-form solid
angle of a rectangle, and intersect's closest-hit rule, its sidedness, the
visibility mask and the reference's distance-free inverse pdf."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def P(oracle):
    oracle._pt_lib()
    return oracle


def quad(c, e1, e2):
    """Quad c, c+e1, c+e1+e2, c+e2 as two triangles with face normal e1 x e2."""
    c, e1, e2 = (np.asarray(v, np.float32) for v in (c, e1, e2))
    return np.array([c, c + e1, c + e1 + e2, c + e2], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def merge(parts):
    verts, tris, off = [], [], 0
    for v, t in parts:
        verts.append(v)
        tris.append(t + off)
        off += v.shape[0]
    return np.concatenate(verts), np.concatenate(tris).astype(np.uint32)


def light(first, n, **kw):
    return dict(type="objectlight", first_tri=first, n_tris=n, color=(1, 1, 1), power=1.0, **kw)


def grid(n):
    s = (np.arange(n) + 0.5) / n
    return [(a, b) for a in s for b in s]


def test_area_table_is_the_normalised_running_sum(P):
    # unit square (two triangles of 1/2) and a right triangle of legs 2 and 3 (area 3)
    sq = quad((0, 0, 0), (1, 0, 0), (0, 1, 0))
    tri = (np.array([[0, 0, 1], [2, 0, 1], [0, 3, 1]], np.float32), np.array([[0, 1, 2]], np.uint32))
    verts, tris = merge([sq, tri])
    area, cdf = P.ObjectLight(verts, tris, light(0, 3)).setup()
    assert area == pytest.approx(4.0, rel=1e-7)
    assert np.allclose(cdf, [0.125, 0.25, 1.0], rtol=1e-7)
    assert cdf[-1] == 1.0


def test_surface_samples_are_uniform_over_the_area(P):
    # two triangles of areas 1 and 3, far from each other
    t0 = np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0]], np.float32)          # area 1
    t1 = np.array([[5, 0, 0], [8, 0, 0], [5, 2, 0]], np.float32)          # area 3
    verts = np.concatenate([t0, t1])
    tris = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    ol = P.ObjectLight(verts, tris, light(0, 2, double_sided=True))
    recv = np.array([1.0, 1.0, 3.0], np.float32)
    pts = []
    for s1, s2 in grid(64):
        hit, ldir, dist, pdf = ol.sample(recv, s1, s2)
        assert hit == 1
        pts.append(recv + ldir.astype(np.float64) * dist)
    pts = np.array(pts)
    assert np.allclose(pts[:, 2], 0.0, atol=1e-5)
    in1 = pts[:, 0] > 4.0
    assert in1.mean() == pytest.approx(0.75, abs=1.0 / 64)
    # each triangle's samples are uniform over it: their mean is its centroid
    for sel, t in ((~in1, t0), (in1, t1)):
        assert np.allclose(pts[sel].mean(0), t.mean(0), atol=2e-2)
        # and they lie inside it (barycentric coordinates in [0, 1])
        e1, e2 = t[1] - t[0], t[2] - t[0]
        m = np.array([e1[:2], e2[:2]]).T
        bc = np.linalg.solve(m, (pts[sel][:, :2] - t[0][:2]).T)
        assert bc.min() > -1e-5 and bc.sum(0).max() < 1 + 1e-5


@pytest.mark.parametrize("h", [0.5, 1.0, 3.0])
def test_illum_pdf_integrates_to_the_solid_angle(P, h):
    """E[1/pdf] over uniform area samples is Omega / pi for pdf = d^2 pi / (A cos):
    the solid angle of a 2 x 2 square at height h above a point on its axis is
    4 asin(1 / (1 + h^2))."""
    verts, tris = quad((-1, -1, h), (0, 2, 0), (2, 0, 0))  # face normal -z: faces the receiver
    ol = P.ObjectLight(verts, tris, light(0, 2))
    inv = [1.0 / ol.sample((0, 0, 0), s1, s2)[3] for s1, s2 in grid(128)]
    omega = 4.0 * math.asin(1.0 / (1.0 + h * h))
    assert math.pi * float(np.mean(inv)) == pytest.approx(omega, rel=2e-3)


def test_illum_sample_sidedness(P):
    verts, tris = quad((-1, -1, 1), (0, 2, 0), (2, 0, 0))  # normal -z
    above = np.array([0.1, 0.2, 2.0], np.float32)
    single = P.ObjectLight(verts, tris, light(0, 2))
    double = P.ObjectLight(verts, tris, light(0, 2, double_sided=True))
    for s1, s2 in grid(8):
        assert single.sample(above, s1, s2)[0] == 0
        hit, ldir, dist, pdf = double.sample(above, s1, s2)
        assert hit == 1 and ldir[2] < 0
        # mirrored receiver below the light sees the same pdf from the front
        below = np.array([0.1, 0.2, 0.0], np.float32)
        h2, _, d2, p2 = single.sample(below, s1, s2)
        assert h2 == 1 and d2 == pytest.approx(dist, rel=1e-6) and p2 == pytest.approx(pdf, rel=1e-5)


def test_intersect_closest_hit_and_distance_free_ipdf(P):
    near = quad((-1, -1, 1), (0, 2, 0), (2, 0, 0))   # z = 1, normal -z (faces the origin)
    far = quad((-1, -1, 2), (2, 0, 0), (0, 2, 0))    # z = 2, normal +z (back to the origin)
    verts, tris = merge([near, far])
    up = np.array([0, 0, 1], np.float32)
    o = np.zeros(3, np.float32)
    both = P.ObjectLight(verts, tris, light(0, 4))
    hit, ipdf = both.intersect(o, up, 5e-5)
    area = 8.0
    assert hit == 1
    # the reference's ipdf: area * cos / pi with 1 / t^2 taken at t = -1
    assert ipdf == pytest.approx(area / math.pi, rel=1e-6)
    # same ipdf from further away: no distance term
    assert both.intersect(np.array([0, 0, -7], np.float32), up, 5e-5) == (1, ipdf)
    # oblique ray: cos term
    d = np.array([0.3, 0.0, 1.0], np.float32)
    d /= np.linalg.norm(d)
    hit, ip2 = both.intersect(o, d, 5e-5)
    assert hit == 1 and ip2 == pytest.approx(area * d[2] / math.pi, rel=1e-6)
    # from between the quads the closest is the far one, seen from its back: no hit
    mid = np.array([0, 0, 1.5], np.float32)
    assert both.intersect(mid, up, 5e-5)[0] == 0
    assert P.ObjectLight(verts, tris, light(0, 4, double_sided=True)).intersect(mid, up, 5e-5) == (1, ipdf)
    # the closest hit decides even when it faces away: the far quad alone is hit from
    # below only through its back
    assert P.ObjectLight(verts, tris, light(2, 2)).intersect(o, up, 5e-5)[0] == 0
    # invisible triangles are not hit: with the near quad invisible the far one is the closest
    vis = np.array([0, 0, 3, 3], np.uint8)
    assert P.ObjectLight(verts, tris, light(0, 4), vis=vis).intersect(o, up, 5e-5)[0] == 0
    vis = np.array([3, 3, 0, 0], np.uint8)
    assert P.ObjectLight(verts, tris, light(0, 4), vis=vis).intersect(mid, up, 5e-5)[0] == 0
    # miss
    assert both.intersect(o, np.array([1, 0, 0], np.float32), 5e-5)[0] == 0
