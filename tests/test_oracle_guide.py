"""Pins the path-guiding oracle (oracle/ora_guide.py): BF16 rounding, the hash
encoding against its definition (partition of unity, exact on linear fields,
dense / hashed indexing), the vMF lobe against its normalisation and mean
cosine, Jakob's sampler against the density, and the analytic gradients of
the loss against central finite differences of the float64 network."""
import math

import numpy as np
import pytest

from oracle import ora_guide as G


def test_bf16_rounding():
    one = np.float32(1.0)
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 3e38, 1e-40], np.float32)
    r = G.bf16(vals)
    assert r[0] == one
    assert r[1] == one                       # a tie rounds to even (mantissa ...0)
    assert r[2] == np.float32(1 + 2 ** -7)  # above the tie rounds up
    assert r[3] == np.float32(1 + 2 ** -6)  # a tie from an odd mantissa rounds up to even
    assert r[4] == np.float32(-2.5)
    assert np.isnan(G.bf16(np.array([np.nan], np.float32)))[0]
    # every result is representable: low 16 bits clear
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)


def test_level_resolutions():
    res = G.level_resolutions()
    assert res[0] == 16 and res[-1] in (2047, 2048) and np.all(np.diff(res) > 0)
    ratio = res[1:] / res[:-1]
    assert np.all(np.abs(ratio - 128 ** (1 / 15)) < 0.07)


def test_grid_indexing_dense_then_hashed():
    T = 1 << 19
    res = G.level_resolutions()
    # coarse levels index a dense (res + 1)^3 lattice one to one
    r = int(res[0])
    g = np.arange(r + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    idx = G.grid_index(r, T, X.ravel(), Y.ravel(), Z.ravel())
    assert len(np.unique(idx)) == (r + 1) ** 3 and idx.max() < T
    # fine levels hash: x ^ 2654435761 y ^ 805459861 z mod T
    r = int(res[-1])
    assert (r + 1) ** 3 > T
    one = np.array([1])
    zero = np.array([0])
    assert G.grid_index(r, T, one, zero, zero)[0] == 1
    assert G.grid_index(r, T, zero, one, zero)[0] == 2654435761 % T
    assert G.grid_index(r, T, zero, zero, one)[0] == 805459861 % T
    assert G.grid_index(r, T, one, one, one)[0] == (1 ^ 2654435761 ^ 805459861) % T


def small_params(seed=0, T=1 << 12):
    p = G.init_params(seed, table_size=T, bmin=(-1, -2, 0), bmax=(3, 2, 2))
    return p


def test_encoding_partition_of_unity_and_linear_fields():
    p = small_params()
    rng = np.random.default_rng(1)
    pos = rng.uniform([-1, -2, 0], [3, 2, 2], (500, 3)).astype(np.float32)
    n = np.tile([0, 0, 1], (500, 1)).astype(np.float32)
    p["grid"][:] = 1.0
    x, _ = G.encode(p, pos, n, n)
    assert np.allclose(x[:, :32], 1.0, atol=2e-6)
    # dense levels: lattice values a + b.(i, j, k) interpolate exactly to a + b.(p * res)
    T = p["grid"].shape[1]
    coef = np.array([0.25, -0.5, 0.125])
    for lev in range(G.LEVELS):
        r = int(p["res"][lev])
        if (r + 1) ** 3 > T:
            continue
        g = np.arange(r + 1)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        idx = G.grid_index(r, T, X.ravel(), Y.ravel(), Z.ravel())
        p["grid"][lev, idx, 0] = (0.5 + coef[0] * X.ravel() + coef[1] * Y.ravel() + coef[2] * Z.ravel())
    x, _ = G.encode(p, pos, n, n)
    pn = np.clip((pos - p["bmin"]) / (p["bmax"] - p["bmin"]), 0, 1)
    for lev in range(G.LEVELS):
        r = int(p["res"][lev])
        if (r + 1) ** 3 > T:
            continue
        exp = 0.5 + (pn * r) @ coef
        assert np.allclose(x[:, 2 * lev], exp, atol=2e-5)
    assert np.all(x[:, 38] == 1.0) and np.all(x[:, 39:] == 0.0)


def sphere_quadrature(n=600):
    t = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, P = np.meshgrid(t, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n) * (np.pi / n)).reshape(-1)
    return d, dA


def test_vmf_density_normalised_with_mean_cosine():
    d, dA = sphere_quadrature()
    mu = np.array([[0.48, 0.6, 0.64]])
    for k in (0.05, 0.5, 3.0, 40.0):
        pdf = G.vmf_pdf(np.repeat(mu[:, None, :], 1, 1), np.array([[k]]), d)[:, 0]
        assert (pdf * dA).sum() == pytest.approx(1.0, abs=2e-3)
        mean_cos = (pdf * (d @ mu[0]) * dA).sum()
        assert mean_cos == pytest.approx(1 / math.tanh(k) - 1 / k, abs=3e-3)


def test_vmf_sampler_follows_density():
    rng = np.random.default_rng(2)
    n = 200000
    mu = np.tile(np.array([0.0, 0.6, 0.8]), (n, 1))
    for k in (0.3, 5.0, 60.0):
        w = G.vmf_sample(mu, np.full(n, k), rng.random(n), rng.random(n))
        assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-9)
        c = w @ mu[0]
        # the cosine to mu has density k e^(k (c - 1)) / (1 - e^-2k) on [-1, 1]
        edges = np.linspace(-1, 1, 41)
        hist, _ = np.histogram(c, edges)
        cdf = lambda x: (np.exp(k * (x - 1)) - np.exp(-2 * k)) / (1 - np.exp(-2 * k))
        exp = np.diff(cdf(edges)) * n
        big = exp > 200
        assert np.all(np.abs(hist[big] - exp[big]) < 5 * np.sqrt(exp[big]) + 0.01 * exp[big])
        # azimuth uniform around mu
        t = np.cross(mu[0], [1.0, 0, 0])
        t /= np.linalg.norm(t)
        b = np.cross(mu[0], t)
        az = np.arctan2(w @ b, w @ t)
        h2, _ = np.histogram(az, np.linspace(-np.pi, np.pi, 9))
        assert np.all(np.abs(h2 - n / 8) < 5 * np.sqrt(n / 8))


def f64_net(seed=3):
    p = small_params(seed, T=1 << 10)
    p["grid"] = (np.random.default_rng(seed).normal(0, 0.5, p["grid"].shape)).astype(np.float32)
    p["W"] = [w.astype(np.float64) * 1.5 for w in p["W"]]
    return p


def test_dloss_dout_matches_finite_differences():
    rng = np.random.default_rng(4)
    n = 6
    out = rng.normal(0, 1.0, (n, G.OUT_DIM))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    c = rng.uniform(0.1, 2.0, n)
    g = G.dloss_dout(out, w, c)
    h = 1e-6
    for i in range(n):
        for j in range(5 * G.LOBES):
            o1 = out.copy()
            o1[i, j] += h
            o2 = out.copy()
            o2[i, j] -= h
            num = (G.loss(o1, w, c) - G.loss(o2, w, c)) / (2 * h)
            assert g[i, j] == pytest.approx(num, rel=1e-5, abs=1e-7), (i, j)
    assert np.all(g[:, 5 * G.LOBES:] == 0)


def net_loss(p, pos, n, wo, w, c):
    x, _ = G.encode(p, pos, n, wo, dtype=np.float64)
    hs, out = G.mlp_forward(p, x, bf=False)
    return G.loss(out, w, c)


def test_backward_matches_finite_differences():
    p = f64_net()
    rng = np.random.default_rng(5)
    m = 5
    pos = rng.uniform([-1, -2, 0], [3, 2, 2], (m, 3)).astype(np.float32)
    nrm = rng.normal(size=(m, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(m, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    w = rng.normal(size=(m, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    c = rng.uniform(0.5, 1.5, m)
    x, corners = G.encode(p, pos, nrm, wo, dtype=np.float64)
    hs, out = G.mlp_forward(p, x, bf=False)
    gW, ggrid = G.backward(p, hs, out, corners, w, c, bf=False)
    h = 1e-6
    for k in range(4):
        rows, cols = p["W"][k].shape
        for _ in range(12):
            i, j = rng.integers(rows if k < 3 else 40), rng.integers(cols)
            p["W"][k][i, j] += h
            lp = net_loss(p, pos, nrm, wo, w, c)
            p["W"][k][i, j] -= 2 * h
            lm = net_loss(p, pos, nrm, wo, w, c)
            p["W"][k][i, j] += h
            num = (lp - lm) / (2 * h)
            assert gW[k][i, j] == pytest.approx(num, rel=1e-4, abs=1e-7), (k, i, j)
    # grid entries that the samples touch
    touched = [(lev, int(idx[0]), f) for lev in (0, 7, 15) for idx, _ in corners[lev][:3] for f in (0, 1)]
    grid64 = p["grid"].astype(np.float64)
    for lev, e, f in touched:
        base = p["grid"][lev, e, f]
        hh = 1e-3
        p["grid"] = grid64.copy()
        p["grid"][lev, e, f] = base + hh
        lp = net_loss(p, pos, nrm, wo, w, c)
        p["grid"][lev, e, f] = base - hh
        lm = net_loss(p, pos, nrm, wo, w, c)
        p["grid"] = grid64.copy()
        num = (lp - lm) / (2 * hh)
        assert ggrid[lev, e, f] == pytest.approx(num, rel=2e-3, abs=1e-6), (lev, e, f)


# ---------------------------------------------------------------- Adam
def test_adam_first_steps_closed_form():
    """From zero moments the bias-corrected moments of step 1 are exactly g
    and g^2, so the step is lr g / (|g| + eps); with the same gradient again
    they stay g and g^2.  With g1 then g2 the corrected first moment is
    (b1 (1 - b1) g1 + (1 - b1) g2) / (1 - b1^2) and the second
    (b2 (1 - b2) g1^2 + (1 - b2) g2^2) / (1 - b2^2)."""
    from oracle import ora_guide as G
    rng = np.random.default_rng(0)
    p0 = rng.normal(size=1000).astype(np.float32)
    g1 = rng.normal(size=1000).astype(np.float32)
    g2 = rng.normal(size=1000).astype(np.float32)
    z = np.zeros(1000, np.float32)
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    p1, m1, v1 = G.adam_step(p0, z, z, g1, 1, lr, b1, b2, eps)
    assert np.allclose(p1, p0 - lr * g1 / (np.abs(g1) + eps), rtol=0, atol=2e-8)
    assert np.allclose(m1, 0.1 * g1, rtol=1e-6) and np.allclose(v1, 0.01 * g1 * g1, rtol=1e-5)
    p1b, _, _ = G.adam_step(p1, m1, v1, g1, 2, lr, b1, b2, eps)
    assert np.allclose(p1b, p1 - lr * g1 / (np.abs(g1) + eps), rtol=0, atol=5e-8)
    p2, m2, v2 = G.adam_step(p1, m1, v1, g2, 2, lr, b1, b2, eps)
    mh = (b1 * (1 - b1) * g1.astype(float) + (1 - b1) * g2) / (1 - b1 ** 2)
    vh = (b2 * (1 - b2) * g1.astype(float) ** 2 + (1 - b2) * g2.astype(float) ** 2) / (1 - b2 ** 2)
    assert np.allclose(p2, p1 - lr * mh / (np.sqrt(vh) + eps), rtol=0, atol=4e-7)
    # a zero gradient from zero moments leaves the weight where it is
    assert np.array_equal(G.adam_step(p0, z, z, z, 1)[0], p0)


# ---------------------------------------------------------------- learned selection probability
def test_selection_gradient_is_the_second_moments_derivative():
    """Over a discrete set of directions with integrand F, BSDF density p and
    mixture density q, the second moment of F / p_mix(alpha) under its own
    density is M(theta) = sum F^2 / p_mix.  The expectation of the per-record
    gradient under the drawing density p_old (any alpha_old) equals dM/dtheta
    by central differences; its sign sends alpha towards the BSDF when the
    mixture misses F and towards the mixture when it matches F."""
    from oracle import ora_guide as G
    rng = np.random.default_rng(3)
    n = 50
    F = rng.gamma(0.5, 1.0, n)
    p = rng.dirichlet(np.ones(n))
    for q, expect_sign in ((F / F.sum(), +1), (rng.dirichlet(np.ones(n) * 0.1), None)):
        for a_old, theta in ((0.5, -0.5), (0.9, 1.3), (0.25, -2.0)):
            p_old = a_old * p + (1 - a_old) * q
            c = F / p_old
            grad = float((p_old * G.selection_grad(c, p, q, a_old, theta)).sum())

            def M(th):
                al = G.selection_alpha(th)
                return float((F * F / (al * p + (1 - al) * q)).sum())
            h = 1e-5
            num = (M(theta + h) - M(theta - h)) / (2 * h)
            assert grad == pytest.approx(num, rel=1e-5)
            if expect_sign is not None:
                # q proportional to F is the zero-variance density: lowering alpha lowers M
                assert grad > 0


def test_fallback_moments_estimate_the_second_moments():
    """The fallback's per-record terms, averaged over draws from p_old, are
    unbiased for the second moments sum F^2 / p_mix(alpha) and sum F^2 / p of
    the guided and the BSDF-only one-sample estimators (exact expectation over
    a discrete set of directions, and a Monte Carlo average of 200k draws within
    5 of its standard errors).  The decision follows the exact comparison: a mixture proportional to
    F (zero variance) stays guided, a mixture that misses F falls back, and
    with F proportional to p (uniform incoming radiance, diffuse BSDF) every
    mixture other than p falls back."""
    from oracle import ora_guide as G
    rng = np.random.default_rng(11)
    n = 64
    F = rng.gamma(0.5, 1.0, n)
    p = rng.dirichlet(np.ones(n))
    cases = ((F / F.sum(), False), (rng.dirichlet(np.ones(n) * 0.05), True))
    for q, falls_back in cases:
        for a_old, alpha in ((0.5, 0.5), (0.2, 0.7), (1.0, 0.3)):
            p_old = a_old * p + (1 - a_old) * q
            c = F / p_old
            tg, tb = G.fallback_terms(c, p, q, a_old, alpha)
            pm = alpha * p + (1 - alpha) * q
            exact_g, exact_b = float((F * F / pm).sum()), float((F * F / p).sum())
            assert float((p_old * tg).sum()) == pytest.approx(exact_g, rel=1e-12)
            assert float((p_old * tb).sum()) == pytest.approx(exact_b, rel=1e-12)
            # Monte Carlo: within 5 standard errors (exact variance of the terms)
            m = 200_000
            draws = rng.choice(n, size=m, p=p_old / p_old.sum())
            for t, exact in ((tg, exact_g), (tb, exact_b)):
                se = np.sqrt(max(float((p_old * t * t).sum()) - exact * exact, 0.0) / m)
                assert abs(t[draws].mean() - exact) <= 5 * se + 1e-12 * exact
            assert G.fallback_update(0.0, 0.0, exact_g, exact_b)[3] == falls_back
    # F proportional to p: BSDF sampling has zero variance, any other mixture falls back
    Fp = 3.0 * p
    for q in (rng.dirichlet(np.ones(n)), 0.5 * p + 0.5 * rng.dirichlet(np.ones(n))):
        tg, tb = G.fallback_terms(Fp / p, p, q, 1.0, 0.5)
        assert float((p * tg).sum()) > float((p * tb).sum()) == pytest.approx(9.0)
    # the running update: first batch taken as is, then an exponential average;
    # no fallback before 256 records
    g, b, seen, fb = G.fallback_update(0.0, 0.0, 2.0, 1.0, 0, 100)
    assert (g, b, seen, fb) == (2.0, 1.0, 100, False)
    g, b, seen, fb = G.fallback_update(g, b, 0.5, 1.0, seen, 156)
    assert g == pytest.approx(1.85) and b == pytest.approx(1.0) and seen == 256 and fb
    for _ in range(20):
        g, b, seen, fb = G.fallback_update(g, b, 0.5, 1.0, seen, 100)
    assert not fb
