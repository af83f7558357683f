"""Pins the oracle's glossy, mirror, glass and light materials and its
recursive raytrace (oracle/ora_pt.c): Snell's law and the Fresnel terms
against textbook optics, the glossy lobe against Ashikhmin and Shirley's
published BRDF and its own sampler, and whole-integrator closed forms for a
mirror, a glass plate and a glossy floor in an emitting box."""
import math

import numpy as np
import pytest

Z = (0.0, 0.0, 1.0)


@pytest.fixture(scope="module")
def P(oracle):
    oracle._mat_lib()
    return oracle


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def rand_dirs(rng, n, upper=True):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    if upper:
        v[:, 2] = np.abs(v[:, 2]) + 1e-3
        v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32)


# ---------------------------------------------------------------- refraction and Fresnel
def test_refract_obeys_snell(P):
    rng = np.random.default_rng(1)
    n = np.array(Z, np.float32)
    for ior in (1.33, 1.5, 2.4):
        for wi in rand_dirs(rng, 200):
            wt = P.refract(n, wi, ior)  # entering: wi above the surface
            assert wt is not None
            sin_i = float(np.hypot(wi[0], wi[1]))
            sin_t = float(np.hypot(wt[0], wt[1]))
            assert wt[2] < 0 and abs(np.linalg.norm(wt) - 1) < 1e-6
            assert abs(sin_t - sin_i / ior) < 2e-6
            # in the plane of incidence, on the far side of the normal
            assert abs(np.dot(np.cross(n, wi), wt)) < 2e-6
            assert np.dot(wi[:2], wt[:2]) <= 1e-7
            # leaving the medium: wi below the surface, eta = ior; total inner reflection past the critical angle
            wb = np.array([wi[0], wi[1], -wi[2]], np.float32)
            out = P.refract(n, wb, ior)
            if ior * sin_i >= 1.0 + 1e-5:
                assert out is None
            elif ior * sin_i < 1.0 - 1e-5:
                assert out is not None and out[2] > 0
                assert abs(float(np.hypot(out[0], out[1])) - ior * sin_i) < 2e-5


def s_polarised_reflectance(cos_i, eta):
    sin_t = math.sqrt(max(0.0, 1 - cos_i * cos_i)) / eta
    cos_t = math.sqrt(max(0.0, 1 - sin_t * sin_t))
    return ((cos_i - eta * cos_t) / (cos_i + eta * cos_t)) ** 2


def test_fresnel_reference_form(P):
    """Vec3f::fresnel is 0.5 * Rs * (1 + q) with Rs the textbook s-polarised
    reflectance; its q has c(g + c) + 1 in the denominator where the textbook
    unpolarised form (Cook-Torrance) has c(g - c) + 1.  At normal incidence
    that gives R0 / 2 * (1 + eta^2 / (eta + 2)^2), not R0 = ((eta-1)/(eta+1))^2."""
    n = np.array(Z, np.float32)
    for eta in (1.2, 1.33, 1.5, 2.0, 2.4):
        kr, kt = P.fresnel(n, n, eta)
        r0 = ((eta - 1) / (eta + 1)) ** 2
        assert abs(kr - r0 / 2 * (1 + eta ** 2 / (eta + 2) ** 2)) < 2e-7
        assert kr + kt == pytest.approx(1.0, abs=1e-7)
        for deg in (10, 30, 50, 70, 85):
            c = math.cos(math.radians(deg))
            wi = np.array([math.sqrt(1 - c * c), 0.0, c], np.float32)
            kr, kt = P.fresnel(wi, n, eta)
            rs = s_polarised_reflectance(float(wi[2]), eta)
            g = math.sqrt(eta * eta + float(wi[2]) ** 2 - 1)
            a = float(wi[2]) * (g + float(wi[2]))
            assert kr / (0.5 * rs) == pytest.approx(1 + ((a - 1) / (a + 1)) ** 2, rel=2e-5)
            # the incident side does not matter
            assert P.fresnel(-wi, n, eta)[0] == pytest.approx(kr, rel=1e-6)
        # grazing incidence reflects everything
        wi = np.array([1.0, 0.0, 1e-4], np.float32)
        assert P.fresnel(wi, n, eta)[0] > 0.999


# ---------------------------------------------------------------- mirror
def test_mirror(P):
    m = P.Material(type="mirror", color=(0.9, 0.5, 0.25), reflect=0.8)
    assert m.flags == P.F_SPECULAR
    rng = np.random.default_rng(2)
    for wo in rand_dirs(rng, 50):
        r, t = m.specular(Z, wo)
        assert t is None
        assert np.allclose(r["dir"], [-wo[0], -wo[1], wo[2]], atol=1e-6)
        assert np.allclose(r["col"], np.float32(0.8) * np.float32([0.9, 0.5, 0.25]), rtol=1e-7)
        s = m.sample(Z, wo, 0.3, 0.7)
        assert np.allclose(s["wi"], r["dir"], atol=1e-6)
        assert s["w"] == 1.0 and s["sampled"] == P.F_SPECULAR | P.F_REFLECT
        # colour / |cos|: the path throughput after the |cos| weight is the mirror colour
        assert np.allclose(s["col"] * abs(s["wi"][2]), r["col"], rtol=1e-5)
        # seen from below the surface the normal is flipped for getSpecular
        rb, _ = m.specular(Z, -wo)
        assert np.allclose(rb["dir"], [wo[0], wo[1], -wo[2]], atol=1e-6)
    assert np.allclose(P.Material(type="mirror", reflect=3.0).specular(Z, Z)[0]["col"], 1.0)
    assert np.all(m.eval(Z, Z, Z) == 0) and m.pdf(Z, Z, Z) == 0


# ---------------------------------------------------------------- glass
def test_glass_specular_split(P):
    filt, mirror = (0.2, 0.6, 1.0), (1.0, 0.9, 0.8)
    m = P.Material(type="glass", IOR=1.5, filter_color=filt, transmit_filter=0.5, mirror_color=mirror)
    assert m.flags == P.F_SPECULAR | P.F_REFLECT | P.F_TRANSMIT
    fc = 0.5 * np.float32(filt) + 0.5  # transmit_filter blends the filter colour with white
    rng = np.random.default_rng(3)
    for wo in rand_dirs(rng, 40):
        kr, kt = P.fresnel(wo, np.float32(Z), 1.5)
        r, t = m.specular(Z, wo, ray_level=1)
        assert np.allclose(r["dir"], [-wo[0], -wo[1], wo[2]], atol=1e-6)
        assert np.allclose(r["col"], kr * np.float32(mirror), rtol=1e-6)
        assert np.allclose(t["col"], kt * fc, rtol=1e-6)
        assert np.allclose(t["dir"], P.refract(np.float32(Z), wo, 1.5), atol=1e-6)
        # inside the glass the internal reflection is followed only for ray levels below 3
        wi_in = np.array([wo[0] * 0.3, wo[1] * 0.3, -1.0], np.float32)
        wi_in /= np.linalg.norm(wi_in)
        r2, t2 = m.specular(Z, wi_in, ray_level=2)
        r3, t3 = m.specular(Z, wi_in, ray_level=3)
        assert r2 is not None and t2 is not None and t2["dir"][2] > 0
        assert r3 is None and t3 is not None
    # total inner reflection: the whole mirror colour, no transmission
    wi_tir = np.array([0.9, 0.0, -math.sqrt(1 - 0.81)], np.float32)
    r, t = m.specular(Z, wi_tir, ray_level=5)
    assert t is None and np.allclose(r["col"], mirror)
    assert np.allclose(r["dir"], [-0.9, 0.0, -math.sqrt(1 - 0.81)], atol=1e-6)


def test_glass_sample_picks_by_fresnel(P):
    m = P.Material(type="glass", IOR=1.5, mirror_color=(0.5, 0.5, 0.5))
    wo = unit([0.5, 0.2, 0.8]).astype(np.float32)
    kr, kt = P.fresnel(wo, np.float32(Z), 1.5)
    p_kt = np.float32(0.01 + 0.99 * kt)
    below = m.sample(Z, wo, float(np.nextafter(p_kt, np.float32(0))), 0.5)
    above = m.sample(Z, wo, float(p_kt), 0.5)
    assert below["sampled"] == P.F_SPECULAR | P.F_TRANSMIT and below["pdf"] == pytest.approx(float(p_kt))
    assert below["wi"][2] < 0 and np.allclose(below["col"], 1.0) and below["w"] == 1.0
    assert above["sampled"] == P.F_SPECULAR | P.F_REFLECT
    assert above["pdf"] == pytest.approx(0.01 + 0.99 * kr, rel=1e-6) and np.allclose(above["col"], 0.5)
    # the glass never samples for the light-sampling flags (no Specular bit)
    s = m.sample(Z, wo, 0.1, 0.1, P.F_GLOSSY | P.F_DIFFUSE | P.F_DISPERSIVE | P.F_REFLECT | P.F_TRANSMIT)
    assert s["pdf"] == 0.0 and np.all(s["col"] == 0)
    # fake shadows: the transmission is a filter event
    f = P.Material(type="glass", IOR=1.5, fake_shadows=True)
    assert f.flags & P.F_FILTER
    assert f.sample(Z, wo, 0.0, 0.5)["sampled"] == P.F_FILTER | P.F_TRANSMIT


# ---------------------------------------------------------------- glossy
def ashikhmin_shirley_specular(wo, wi, e, rs):
    """Ashikhmin & Shirley, "An anisotropic Phong BRDF model" (2000), eq. 4 with
    n_u = n_v = e and Schlick's Fresnel: (e+1)/(8 pi) (n.h)^e F(k.h) / ((h.k) max(n.k1, n.k2))."""
    h = unit(np.add(wo, wi))
    kh = float(np.dot(wi, h))
    f = rs + (1 - rs) * (1 - kh) ** 5
    return (e + 1) / (8 * math.pi) * h[2] ** e * f / (kh * max(wo[2], wi[2]))


def test_glossy_eval_is_ashikhmin_shirley(P):
    """The reference's divisor is 8 pi ((k.h) max(..) 0.99 + 0.04) instead of
    8 pi (k.h) max(..): the ratio is removed before comparing."""
    rng = np.random.default_rng(4)
    for e, rs, col in ((50.0, 1.0, (1, 1, 1)), (10.0, 0.3, (0.9, 0.6, 0.3)), (200.0, 0.7, (1, 1, 1))):
        m = P.Material(type="glossy", exponent=e, glossy_reflect=rs, color=col, as_diffuse=False)
        for _ in range(300):
            wo, wi = rand_dirs(rng, 2)
            h = unit(np.add(wo, wi))
            if h[2] < 0.3:
                continue
            x = float(np.dot(wi, h)) * max(float(wo[2]), float(wi[2]))
            fudge = x / (x * 0.99 + 0.04)
            got = m.eval(Z, wo, wi)
            exp = ashikhmin_shirley_specular(wo, wi, e, rs) * fudge * np.asarray(col, np.float64)
            assert np.allclose(got, exp, rtol=3e-3, atol=1e-7)
            # reciprocity
            assert np.allclose(m.eval(Z, wi, wo), got, rtol=1e-5, atol=1e-9)
        # eval needs a Diffuse bit in the request; opposite sides give 0
        assert np.all(m.eval(Z, Z, Z, P.F_GLOSSY | P.F_REFLECT) == 0)
        assert np.all(m.eval(Z, Z, (0.0, 0.6, -0.8)) == 0)


def test_glossy_diffuse_parts(P):
    """eval's diffuse part is Lambert without 1/pi: diffuse_reflect (1 - glossy_reflect)
    diffuse_color; sample's is Ashikhmin-Shirley's diffuse term 28/(23 pi) ..."""
    m = P.Material(type="glossy", exponent=30.0, glossy_reflect=0.25, diffuse_reflect=0.8,
                   diffuse_color=(0.5, 1.0, 0.2), color=(0, 0, 0))
    assert m.flags == P.F_DIFFUSE | P.F_REFLECT
    rng = np.random.default_rng(5)
    for wo, wi in zip(rand_dirs(rng, 50), rand_dirs(rng, 50)):
        assert np.allclose(m.eval(Z, wo, wi), 0.8 * 0.75 * np.float32([0.5, 1.0, 0.2]), rtol=1e-6)
    assert abs(0.387507688 - 28 / (23 * math.pi)) < 1e-9
    wo = unit([0.3, 0.1, 0.9]).astype(np.float32)
    for s1 in np.linspace(0.01, 0.99, 25):
        s = m.sample(Z, wo, float(s1), 0.37, P.F_DIFFUSE | P.F_REFLECT | P.F_TRANSMIT)
        wi = s["wi"]
        if wi[2] <= 0:
            continue
        kd = 28 / (23 * math.pi) * 0.8 * 0.75 * (1 - (1 - wi[2] / 2) ** 5) * (1 - (1 - wo[2] / 2) ** 5)
        glossy_part = s["col"][0] - kd * 0.5  # red channel: diffuse colour 0.5, glossy colour 0
        assert abs(glossy_part) < 2e-6
        assert s["col"][1] == pytest.approx(kd, rel=2e-5)


def true_pdf(m, P, wo, wi, e):
    """The reference's pdf with its 0.99 / 0.04 divisor ratio removed."""
    h = unit(np.add(wo, wi))
    c = float(np.dot(wo, h))
    return m.pdf(Z, wo, wi, P.F_GLOSSY | P.F_REFLECT) * (c * 0.99 + 0.04) / c


def sphere_grid(n_t, n_p):
    t = (np.arange(n_t) + 0.5) / n_t * (math.pi / 2)
    p = (np.arange(n_p) + 0.5) / n_p * 2 * math.pi
    T, Pp = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1)
    dA = np.sin(T) * (math.pi / 2 / n_t) * (2 * math.pi / n_p)
    return d.reshape(-1, 3).astype(np.float32), dA.reshape(-1)


def test_glossy_pdf_normalised(P):
    e = 40.0
    m = P.Material(type="glossy", exponent=e, as_diffuse=False)
    dirs, dA = sphere_grid(400, 200)
    total = sum(true_pdf(m, P, np.float32(Z), d, e) * a for d, a in zip(dirs, dA))
    assert total == pytest.approx(1.0, abs=5e-3)


def test_glossy_sampler_matches_pdf(P):
    e = 20.0
    m = P.Material(type="glossy", exponent=e, as_diffuse=False)
    wo = unit([math.sin(0.4), 0.0, math.cos(0.4)]).astype(np.float32)
    rng = np.random.default_rng(6)
    n = 60000
    n_t, n_p = 12, 24
    hist = np.zeros((n_t, n_p))
    for s1, s2 in rng.random((n, 2)):
        s = m.sample(Z, wo, float(s1), float(s2), P.F_GLOSSY | P.F_REFLECT)
        wi = s["wi"]
        if s["pdf"] == 0.0:
            continue
        # the sampled half vector is built with the fast sine and cosine, so it is unit
        # length only to about 1e-3; pdf() renormalises wi + wo, and cos^e amplifies that by e
        assert s["pdf"] == pytest.approx(m.pdf(Z, wo, wi, P.F_GLOSSY | P.F_REFLECT), rel=1.5e-3 * e)
        assert s["w"] == pytest.approx(wi[2] / (s["pdf"] * 0.99 + 0.01), rel=1e-5)
        th = math.acos(min(1.0, float(wi[2])))
        ph = math.atan2(float(wi[1]), float(wi[0])) % (2 * math.pi)
        hist[min(n_t - 1, int(th / (math.pi / 2) * n_t)), min(n_p - 1, int(ph / (2 * math.pi) * n_p))] += 1
    # expected counts from the pdf integrated over each bin
    sub = 6
    exp = np.zeros_like(hist)
    for i in range(n_t):
        for j in range(n_p):
            t = (i + (np.arange(sub) + 0.5) / sub) * (math.pi / 2 / n_t)
            p = (j + (np.arange(sub) + 0.5) / sub) * (2 * math.pi / n_p)
            acc = 0.0
            for tt in t:
                for pp in p:
                    d = np.array([math.sin(tt) * math.cos(pp), math.sin(tt) * math.sin(pp), math.cos(tt)], np.float32)
                    acc += true_pdf(m, P, wo, d, e) * math.sin(tt)
            exp[i, j] = acc * (math.pi / 2 / n_t / sub) * (2 * math.pi / n_p / sub) * n
    big = exp > 400
    assert big.sum() > 10
    assert np.all(np.abs(hist[big] - exp[big]) < 5 * np.sqrt(exp[big]) + 0.03 * exp[big])


# ---------------------------------------------------------------- light material
def test_light_material_emits_one_side(P):
    m = P.Material(type="light", color=(1.0, 0.5, 0.25), power=4.0)
    assert m.flags == P.F_EMIT
    assert np.allclose(m.emit(Z, unit([0.2, 0.1, 0.9])), [4.0, 2.0, 1.0])
    assert np.all(m.emit(Z, unit([0.2, 0.1, -0.9])) == 0)
    d = P.Material(type="light", color=(1.0, 0.5, 0.25), power=4.0, double_sided=True)
    assert np.allclose(d.emit(Z, unit([0.2, 0.1, -0.9])), [4.0, 2.0, 1.0])
    s = m.sample(Z, Z, 0.5, 0.5)
    assert s["w"] == 0.0 and np.all(s["col"] == 0)


# ---------------------------------------------------------------- whole integrator
def quad(c, e1, e2):
    c, e1, e2 = (np.asarray(x, np.float32) for x in (c, e1, e2))
    return np.array([c, c + e1, c + e1 + e2, c + e2], np.float32)


def scene_of(quads, materials, camera, **settings):
    verts = np.concatenate([q for q, _ in quads])
    tris, mat_id = [], []
    for k, (_, mid) in enumerate(quads):
        b = 4 * k
        tris += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        mat_id += [mid, mid]
    st = dict(path_samples=1, bounces=3, rr_min_bounces=0, aa_samples=1, aa_pixelwidth=1.5, background=None,
              seed=0, raydepth=5)
    st.update(settings)
    return dict(verts=verts, tris=np.array(tris, np.uint32), mat_id=np.array(mat_id, np.uint32),
                materials=materials, lights=[], camera=camera, settings=st)


def one_pixel_camera(from_, to, up):
    return dict(from_=from_, to=to, up=up, focal=50.0, resx=1, resy=1)


def test_mirror_shows_emitter(P):
    """Camera -> mirror at 45 degrees -> emitter: the pixel is emit * mirror colour."""
    mirror = quad((-1, -1, 1), (2, 0, 2), (0, 2, 0))       # plane z = x + 2: the downward ray turns to -x
    emitter = quad((-5, -5, -5), (0, 0, 20), (0, 10, 0))  # plane x = -5
    mats = [dict(type="mirror", color=(0.9, 0.8, 0.7), reflect=0.5),
            dict(color=(1.0, 0.5, 0.25), emit=2.0, diffuse_reflect=0.0)]
    sc = scene_of([(mirror, 0), (emitter, 1)], mats, one_pixel_camera((0, 0, 10), (0, 0, 0), (0, 1, 10)))
    col, _ = P.PathTracer(sc).pixel_sample(0, 0, 0)
    expected = np.float32([1.0, 0.5, 0.25]) * np.float32(2.0) * (np.float32([0.9, 0.8, 0.7]) * np.float32(0.5))
    assert np.allclose(col[:3], expected, rtol=1e-6)
    # no recursion allowed: the mirror alone is black
    sc["settings"]["raydepth"] = 0
    col0, _ = P.PathTracer(sc).pixel_sample(0, 0, 0)
    assert np.all(col0[:3] == 0)


def test_glass_plate_transmits_fresnel_squared(P):
    """Normal incidence through a plate (front face z = 1 facing +z, back face
    z = 0.9 facing -z) onto an emitter at z = -1; raydepth 2 follows exactly
    refract -> refract (the internal reflection's child hits the front face from
    inside at level 2, where nothing more is spawned): pixel = kt * kt * E."""
    front = quad((-5, -5, 1), (10, 0, 0), (0, 10, 0))
    back = quad((-5, -5, 0.9), (0, 10, 0), (10, 0, 0))
    emitter = quad((-5, -5, -1), (10, 0, 0), (0, 10, 0))
    mats = [dict(type="glass", IOR=1.5), dict(color=(1.0, 1.0, 1.0), emit=3.0, diffuse_reflect=0.0)]
    sc = scene_of([(front, 0), (back, 0), (emitter, 1)], mats,
                  one_pixel_camera((0, 0, 10), (0, 0, 0), (0, 1, 10)), raydepth=2)
    col, _ = P.PathTracer(sc).pixel_sample(0, 0, 0)
    kr, kt = P.fresnel(np.float32(Z), np.float32(Z), 1.5)
    assert np.allclose(col[:3], np.float32(3.0) * np.float32(kt) * np.float32(kt), rtol=2e-6)
    assert col[0] == pytest.approx(3.0 * (1 - 0.02 * (1 + 2.25 / 12.25)) ** 2, rel=1e-5)
    # raydepth 1: the refracted child reaches the back face but may not recurse further
    sc["settings"]["raydepth"] = 1
    assert np.all(P.PathTracer(sc).pixel_sample(0, 0, 0)[0][:3] == 0)


def test_glossy_branches_average_their_samples(P):
    """A glossy floor inside a box of emitters: the camera sample is the mean of
    the 8 glossy branches, each E * colour * w of its BSDF sample, drawn from
    Halton bases 2 and 3 started at 8 * sample + pixel offset."""
    e_col = np.float32([2.0, 1.0, 0.5])
    floor = quad((-10, -10, 0), (20, 0, 0), (0, 20, 0))
    walls = [quad((-10, -10, 0), (0, 20, 0), (0, 0, 20)), quad((10, -10, 0), (0, 0, 20), (0, 20, 0)),
             quad((-10, -10, 0), (0, 0, 20), (20, 0, 0)), quad((-10, 10, 0), (20, 0, 0), (0, 0, 20)),
             quad((-10, -10, 20), (20, 0, 0), (0, 20, 0))]
    mats = [dict(type="glossy", exponent=8.0, as_diffuse=False, color=(0.9, 0.9, 0.9)),
            dict(color=tuple(e_col), emit=1.0, diffuse_reflect=0.0)]
    cam = one_pixel_camera((0, -3, 3), (0, 0, 0), (0, -3, 4))
    sc = scene_of([(floor, 0)] + [(w, 1) for w in walls], mats, cam)
    pt = P.PathTracer(sc)
    col, _ = pt.pixel_sample(0, 0, 0)
    L = P._pt_lib()
    wo = -P.camera_dir(sc, 0.5, 0.5)
    offset = L.ora_fnv32a(0 * L.ora_fnv32a(0))
    acc = np.zeros(3)
    for k in range(8):
        s1 = L.ora_halton_next(2, offset + k)
        s2 = L.ora_halton_next(3, offset + k)
        s = P.Material(**mats[0]).sample(Z, wo, s1, s2, P.F_GLOSSY | P.F_REFLECT)
        acc += e_col * (s["col"] * s["w"])
    assert np.allclose(col[:3], acc / 8, rtol=1e-4)
    assert col[0] > 0.5


# ---------------------------------------------------------------- shinydiffuse
def test_shinydiffuse_eval_pdf_closed_form(P):
    """Lambert without 1/pi: eval = diffuse_reflect * color on the side of wo,
    pdf = |cos wi| (material_shiny_diffuse.cc:297-347, :441-485)."""
    m = P.Material(type="shinydiffuse", color=(0.9, 0.5, 0.25), diffuse_reflect=0.8)
    assert m.flags & P.F_DIFFUSE
    rng = np.random.default_rng(8)
    for wo, wi in zip(rand_dirs(rng, 100), rand_dirs(rng, 100)):
        assert np.allclose(m.eval(Z, wo, wi), 0.8 * np.float32([0.9, 0.5, 0.25]), rtol=1e-6)
        assert m.pdf(Z, wo, wi) == pytest.approx(abs(float(wi[2])), rel=1e-6)
        # the transmit side gets nothing: no translucency built
        assert np.all(m.eval(Z, wo, wi * np.float32([1, 1, -1])) == 0)


def test_shinydiffuse_sampler_matches_cosine_density(P):
    """sample() draws wi from cosHemisphere (include/sampler/sample.h:43-52):
    cos(theta) = sqrt(s1) exactly, azimuth 2 pi s2 through the fast sine and
    cosine.  A histogram of 60k samples against the cosine density cos/pi
    over (theta, phi) bins; the returned pdf is |cos wi| and the weight
    |cos| / (0.99 pdf + 0.01) (:432); the colour is eval's."""
    m = P.Material(type="shinydiffuse", color=(0.7, 0.7, 0.7))
    wo = unit([0.5, -0.2, 0.8]).astype(np.float32)
    rng = np.random.default_rng(9)
    n = 60000
    n_t, n_p = 10, 16
    hist = np.zeros((n_t, n_p))
    for s1, s2 in rng.random((n, 2)):
        s = m.sample(Z, wo, float(s1), float(s2))
        wi = s["wi"]
        assert wi[2] > 0 and s["sampled"] & P.F_DIFFUSE
        assert float(wi[2]) == pytest.approx(math.sqrt(s1), abs=2e-7)
        assert s["pdf"] == pytest.approx(float(wi[2]), rel=1e-6)
        assert s["w"] == pytest.approx(float(wi[2]) / (s["pdf"] * 0.99 + 0.01), rel=1e-6)
        assert np.allclose(s["col"], 0.7, rtol=1e-6)
        # the fast sine/cosine leave the direction unit only to about 2e-3
        assert abs(float(np.linalg.norm(wi)) - 1.0) < 3e-3
        th = math.acos(min(1.0, float(wi[2])))
        ph = math.atan2(float(wi[1]), float(wi[0])) % (2 * math.pi)
        hist[min(n_t - 1, int(th / (math.pi / 2) * n_t)), min(n_p - 1, int(ph / (2 * math.pi) * n_p))] += 1
    # cos/pi integrated over a bin: (sin^2 t1 - sin^2 t0) / 2 * dphi / pi
    t = np.linspace(0, math.pi / 2, n_t + 1)
    ring = (np.sin(t[1:]) ** 2 - np.sin(t[:-1]) ** 2) / 2 * (2 * math.pi / n_p) / math.pi
    exp = np.repeat(ring[:, None], n_p, axis=1) * n
    assert np.all(np.abs(hist - exp) < 5 * np.sqrt(exp) + 0.02 * exp)
    assert abs(exp.sum() - n) < 1e-6 * n


def test_shinydiffuse_sampler_below_the_surface(P):
    """From below (wo on the back), sample() flips to the side of wo:
    the reference's face-forwarded normal."""
    m = P.Material(type="shinydiffuse", color=(0.7, 0.7, 0.7))
    wo = unit([0.3, 0.1, -0.9]).astype(np.float32)
    for s1, s2 in ((0.1, 0.2), (0.5, 0.7), (0.9, 0.05)):
        s = m.sample(Z, wo, s1, s2)
        assert s["wi"][2] < 0 and float(-s["wi"][2]) == pytest.approx(math.sqrt(s1), abs=2e-7)
