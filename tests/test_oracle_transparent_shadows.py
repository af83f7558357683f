"""Pins the oracle's transparent shadows (oracle/ora_pt.c with the kd-tree
port's TransparentShadow query, restating include/accelerator/accelerator.h:113-168
and src/material/material_glass.cc:297-304): a point light seen by a floor
point through fake-shadow glass plates is attenuated by the product of the
plates' kt * filter colour, with kt from the oracle's Fresnel terms (pinned
against textbook optics in test_oracle_materials); more plates than
"shadowDepth", an opaque plate, or glass without fake shadows block it."""
import numpy as np
import pytest

LIGHT = (0.3, 0.2, 2.0)


def quad(c, e1, e2):
    c, e1, e2 = (np.asarray(v, np.float32) for v in (c, e1, e2))
    return np.array([c, c + e1, c + e1 + e2, c + e2], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def scene(plates, transp=True, depth=4):
    """Floor z = 0 (albedo 0.8), camera at z = 0.5 looking straight down at the
    origin, a point light above; plates: list of (z, material dict)."""
    parts = [quad((-2, -2, 0), (4, 0, 0), (0, 4, 0))]
    mats = [dict(color=(0.8, 0.8, 0.8))]
    mat_id = [0, 0]
    for k, (z, m) in enumerate(plates):
        parts.append(quad((-1, -1, z), (0, 2, 0), (2, 0, 0)))
        mats.append(m)
        mat_id += [k + 1, k + 1]
    verts, tris, off = [], [], 0
    for v, t in parts:
        verts.append(v)
        tris.append(t + off)
        off += len(v)
    return dict(verts=np.concatenate(verts), tris=np.concatenate(tris).astype(np.uint32),
                mat_id=np.array(mat_id, np.uint32), materials=mats,
                lights=[dict(type="pointlight", **{"from": LIGHT}, color=(1.0, 0.9, 0.8), power=3.0)],
                camera=dict(from_=(0, 0, 0.5), to=(0, 0, 0), up=(0, 1, 0.5), focal=1.0, resx=9, resy=9),
                settings=dict(path_samples=1, bounces=1, rr_min_bounces=0, aa_samples=1, aa_pixelwidth=1.5,
                              background=None, seed=0, raydepth=0, transpShad=transp, shadowDepth=depth))


@pytest.fixture(scope="module")
def P(oracle):
    oracle._pt_lib()
    oracle._mat_lib()
    return oracle


def centre(P, sc):
    col, _ = P.PathTracer(sc).pixel_sample(4, 4, 0)
    return col[:3].astype(np.float64)


def glass(ior, filt=(0.9, 0.7, 0.5), tf=0.6, fake=True):
    return dict(type="glass", IOR=ior, filter_color=filt, transmit_filter=tf, fake_shadows=fake)


def plate_factor(P, m):
    """kt * filter colour of one plate for the shadow ray from the origin to the light."""
    d = np.asarray(LIGHT, np.float64)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    kr, kt = P.fresnel(d, np.array([0, 0, 1], np.float32), m["IOR"])
    filt = np.array([m["transmit_filter"] * c + (1 - m["transmit_filter"]) for c in m["filter_color"]], np.float64)
    return kt * filt


def test_unoccluded_is_the_point_light_closed_form(P):
    c0 = centre(P, scene([]))
    d2 = float(np.dot(LIGHT, LIGHT))
    cos = LIGHT[2] / np.sqrt(d2)
    expect = 0.8 * np.array([1.0, 0.9, 0.8]) * 3.0 / d2 * cos  # Lambert eval (no 1/pi) * L / d^2 * cos
    assert np.allclose(c0, expect, rtol=1e-5)


@pytest.mark.parametrize("iors", [(1.5,), (1.5, 1.33), (2.4, 1.2, 1.7)])
def test_plates_multiply_their_transmission(P, iors):
    c0 = centre(P, scene([]))
    ms = [glass(ior) for ior in iors]
    c = centre(P, scene([(1.0 + 0.3 * k, m) for k, m in enumerate(ms)]))
    expect = c0.copy()
    for m in ms:
        expect *= plate_factor(P, m)
    assert np.allclose(c, expect, rtol=2e-6)
    assert c.max() < c0.max()


def test_blocking_cases(P):
    two = [(1.0, glass(1.5)), (1.3, glass(1.33))]
    # transparent shadows off: glass is an ordinary occluder
    assert centre(P, scene(two, transp=False)).max() == 0.0
    # more transparent casters than shadowDepth
    assert centre(P, scene(two, depth=1)).max() == 0.0
    assert centre(P, scene(two, depth=2)).max() > 0.0
    # glass without fake shadows is not transparent
    assert centre(P, scene([(1.0, glass(1.5, fake=False))])).max() == 0.0
    # an opaque plate behind a transparent one
    assert centre(P, scene([(1.0, glass(1.5)), (1.3, dict(color=(0.5, 0.5, 0.5)))])).max() == 0.0
