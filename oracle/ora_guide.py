"""CPU oracle of the path-guiding network (TEST INFRASTRUCTURE ONLY: the
checker of libyafaray_amd/csrc/guide.hip; never imported by the library).

The reference has no path guiding; this restates the network the GPU kernels
implement, from the published methods it follows:
  * multi-resolution hash encoding (Mueller, Evans, Schied, Keller, "Instant
    Neural Graphics Primitives with a Multiresolution Hash Encoding", 2022):
    L levels of T entries and F features, resolution N_l = floor(N_min b^l),
    dense indexing while (N_l + 1)^3 <= T, otherwise the spatial hash
    x ^ 2654435761 y ^ 805459861 z mod T, trilinear interpolation;
  * a ReLU MLP without biases (the input carries a constant 1), BF16 inputs,
    weights and activations with FP32 accumulation;
  * an 8-lobe von Mises-Fisher mixture on the sphere (density
    kappa / (2 pi (1 - e^-2kappa)) e^(kappa (mu.w - 1)); sampling after
    W. Jakob, "Numerically stable sampling of the von Mises-Fisher
    distribution on S^2", 2012), trained by the cross-entropy / KL loss of
    neural importance sampling (Mueller et al., "Neural Importance Sampling",
    2019): loss = -c log q(w) for a sample w with target weight c.

Two precisions: `bf16=True` rounds inputs, weights and activations to BF16
the way the GPU does (round to nearest even) and accumulates in float32;
`bf16=False` runs the same network in float64, for finite-difference checks
of the analytic gradients.
"""
from __future__ import annotations

import numpy as np

LEVELS = 16
FEATURES = 2
IN_DIM = 64          # 32 grid features, normal (3), wo (3), constant 1, zero padding
HIDDEN = 64
OUT_DIM = 48         # 8 lobes x (mu 3, log kappa 1, weight logit 1) = 40, padded to 48
LOBES = 8
KAPPA_LOG_MIN, KAPPA_LOG_MAX = -2.0, 9.0
PRIME_Y, PRIME_Z = 2654435761, 805459861


def bf16(x):
    """Round float32 to bfloat16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    r = ((u + rounding) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32)
    return np.where(nan, a, out)


def level_resolutions(n_min=16, n_max=2048, levels=LEVELS):
    b = np.exp((np.log(n_max) - np.log(n_min)) / (levels - 1))
    return np.floor(n_min * b ** np.arange(levels)).astype(np.int64)


def grid_index(level_res, table_size, x, y, z):
    r1 = level_res + 1
    if r1 ** 3 <= table_size:
        # a corner past res only occurs on the upper face of the cube (weight 0)
        x, y, z = (np.minimum(np.asarray(a), level_res) for a in (x, y, z))
        return (x + y * r1 + z * r1 * r1).astype(np.int64)
    h = (x.astype(np.uint64) ^ (y.astype(np.uint64) * PRIME_Y) ^ (z.astype(np.uint64) * PRIME_Z)) & 0xFFFFFFFF
    return (h & (table_size - 1)).astype(np.int64)


def encode(params, pos, normal, wo, dtype=np.float32):
    """Input vectors (n, 64): the trilinear grid features of every level, then
    normal, wo, 1.  Returns the inputs and what the backward pass needs.  In
    float32 (the GPU's precision) the grid is read as BF16, as the kernels
    gather it."""
    grid = params["grid"]                      # (L, T, F)
    if dtype == np.float32:
        grid = bf16(grid)
    T = grid.shape[1]
    res = params["res"]
    lo, hi = params["bmin"], params["bmax"]
    pos = np.asarray(pos, np.float32)
    p = ((pos - lo) / (hi - lo)).astype(np.float32)
    p = np.clip(p, 0.0, 1.0).astype(np.float32)
    n = pos.shape[0]
    x = np.zeros((n, IN_DIM), dtype=dtype)
    corners = []
    for lev in range(LEVELS):
        s = (p * np.float32(res[lev])).astype(np.float32)
        b = np.floor(s).astype(np.int64)
        f = (s - b.astype(np.float32)).astype(np.float32)
        acc = np.zeros((n, FEATURES), dtype=dtype)
        cl = []
        for c in range(8):
            d = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
            wx = f[:, 0] if d[0] else (np.float32(1) - f[:, 0])
            wy = f[:, 1] if d[1] else (np.float32(1) - f[:, 1])
            wz = f[:, 2] if d[2] else (np.float32(1) - f[:, 2])
            w = ((wx * wy).astype(np.float32) * wz).astype(np.float32)
            idx = grid_index(int(res[lev]), T, b[:, 0] + d[0], b[:, 1] + d[1], b[:, 2] + d[2])
            g = grid[lev][idx].astype(dtype)
            acc = (acc + w.astype(dtype)[:, None] * g).astype(dtype)
            cl.append((idx, w))
        x[:, 2 * lev:2 * lev + 2] = acc
        corners.append(cl)
    x[:, 32:35] = normal
    x[:, 35:38] = wo
    x[:, 38] = 1.0
    return x, corners


def mlp_forward(params, x, bf=True):
    """Activations of the 3 hidden layers and the 48 outputs."""
    W = params["W"]                              # [W1 (64x64), W2, W3, W4 (48x64)]
    rnd = bf16 if bf else (lambda a: a)
    dt = np.float32 if bf else np.float64
    hs = [rnd(x.astype(dt))]
    for k in range(3):
        a = (hs[-1].astype(dt) @ rnd(W[k].astype(dt)).T).astype(dt)
        hs.append(rnd(np.maximum(a, 0)))
    out = (hs[-1] @ rnd(W[3].astype(dt)).T).astype(dt)
    return hs, out


def decode(out):
    """Mixture parameters from the raw outputs: unit mu (8, 3), kappa (8), pi (8)."""
    raw = out[:, :5 * LOBES].reshape(-1, LOBES, 5)
    m = raw[..., 0:3]
    norm = np.sqrt((m * m).sum(-1))
    mu = m / np.maximum(norm, 1e-8)[..., None]
    s = np.clip(raw[..., 3], KAPPA_LOG_MIN, KAPPA_LOG_MAX)
    kappa = np.exp(s)
    a = raw[..., 4]
    a = a - a.max(-1, keepdims=True)
    e = np.exp(a)
    pi = e / e.sum(-1, keepdims=True)
    return mu, kappa, pi, norm


def vmf_pdf(mu, kappa, w):
    """Density of each lobe at direction w (n, 3): (n, 8)."""
    cos = (mu * w[:, None, :]).sum(-1)
    # kappa / (2 pi (1 - e^-2k)) e^(k (cos - 1)), written with expm1 for small kappa
    norm = kappa / (2 * np.pi * (-np.expm1(-2 * kappa)))
    return norm * np.exp(kappa * (cos - 1))


def mixture_pdf(out, w):
    mu, kappa, pi, _ = decode(out)
    return (pi * vmf_pdf(mu, kappa, w)).sum(-1)


def loss(out, w, c):
    """sum over samples of -c log q(w)."""
    return float(-(c * np.log(mixture_pdf(out, w))).sum())


def dloss_dout(out, w, c):
    """Analytic gradient of loss() with respect to the raw outputs (n, 48)."""
    mu, kappa, pi, norm = decode(out)
    v = pi * vmf_pdf(mu, kappa, w)
    q = v.sum(-1, keepdims=True)
    r = v / q                                           # responsibilities
    cos = (mu * w[:, None, :]).sum(-1)
    g = np.zeros_like(out)
    gr = g[:, :5 * LOBES].reshape(-1, LOBES, 5)
    # weight logits: d log q / d a_j = r_j - pi_j
    gr[..., 4] = r - pi
    # log kappa: d log vMF / d kappa = 1/k - 2 e^-2k / (1 - e^-2k) + cos - 1
    e2 = np.exp(-2 * kappa)
    dk = 1 / kappa - 2 * e2 / (-np.expm1(-2 * kappa)) + cos - 1
    raw_s = out[:, :5 * LOBES].reshape(-1, LOBES, 5)[..., 3]
    inside = (raw_s > KAPPA_LOG_MIN) & (raw_s < KAPPA_LOG_MAX)
    gr[..., 3] = np.where(inside, r * kappa * dk, 0.0)
    # mu = m / |m|: d log vMF / d m = kappa (w - (mu.w) mu) / |m|
    proj = w[:, None, :] - cos[..., None] * mu
    gr[..., 0:3] = (r * kappa / np.maximum(norm, 1e-8))[..., None] * proj
    return -(c[:, None] * g)


def backward(params, hs, out, corners, w, c, bf=True):
    """Gradients of loss() for the four weight matrices and the grid."""
    W = params["W"]
    rnd = bf16 if bf else (lambda a: a)
    dt = np.float32 if bf else np.float64
    d = dloss_dout(out.astype(np.float64), w.astype(np.float64), c.astype(np.float64)).astype(dt)
    d = rnd(d)
    gW = [None] * 4
    gW[3] = (d.T @ hs[3]).astype(dt)
    dh = (d @ rnd(W[3].astype(dt))).astype(dt)
    for k in (2, 1, 0):
        dh = rnd(np.where(hs[k + 1] > 0, dh, 0))
        gW[k] = (dh.T @ hs[k]).astype(dt)
        dh = (dh @ rnd(W[k].astype(dt))).astype(dt)
    dx = dh                                             # gradient of the 64 inputs
    ggrid = np.zeros(params["grid"].shape, dtype=np.float64)
    for lev in range(LEVELS):
        for idx, wt in corners[lev]:
            for f in range(FEATURES):
                np.add.at(ggrid[lev, :, f], idx, wt.astype(np.float64) * dx[:, 2 * lev + f])
    return gW, ggrid


def init_params(seed=0, table_size=1 << 19, bmin=(0, 0, 0), bmax=(1, 1, 1)):
    """The GPU's initial weights for a given seed (see guide.hip): grid
    features U(-1e-4, 1e-4), weights U(-s, s) with s = sqrt(6 / fan_in)."""
    rng = np.random.default_rng(seed)
    grid = rng.uniform(-1e-4, 1e-4, (LEVELS, table_size, FEATURES)).astype(np.float32)
    W = []
    for shape in ((HIDDEN, IN_DIM), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (OUT_DIM, HIDDEN)):
        s = np.sqrt(6.0 / shape[1])
        W.append(rng.uniform(-s, s, shape).astype(np.float32))
    W[3][5 * LOBES:] = 0.0
    return dict(grid=grid, W=W, res=level_resolutions(), bmin=np.asarray(bmin, np.float32),
                bmax=np.asarray(bmax, np.float32))


def vmf_sample(mu, kappa, u1, u2):
    """Jakob's stable vMF sampling around mu (one lobe per row)."""
    wc = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    phi = 2 * np.pi * u2
    s = np.sqrt(np.maximum(0, 1 - wc * wc))
    # orthonormal frame around mu (Frisvad-style with the sign trick)
    sign = np.where(mu[:, 2] >= 0, 1.0, -1.0)
    a = -1 / (sign + mu[:, 2])
    b = mu[:, 0] * mu[:, 1] * a
    t = np.stack([1 + sign * mu[:, 0] ** 2 * a, sign * b, -sign * mu[:, 0]], -1)
    bb = np.stack([b, sign + mu[:, 1] ** 2 * a, -mu[:, 1]], -1)
    return (wc[:, None] * mu + s[:, None] * (np.cos(phi)[:, None] * t + np.sin(phi)[:, None] * bb))


def adam_step(p, m, v, g, t, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8):
    """One Adam step (Kingma & Ba, "Adam: A Method for Stochastic
    Optimization", 2015, Algorithm 1) in float32 in the kernels' order:
    m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g^2;
    p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
    Returns the new (p, m, v)."""
    f = np.float32
    p, m, v, g = (np.asarray(a, np.float32) for a in (p, m, v, g))
    b1, b2 = f(beta1), f(beta2)
    c1 = f(1) - f(np.power(np.float64(b1), t))
    c2 = f(1) - f(np.power(np.float64(b2), t))
    m = b1 * m + (f(1) - b1) * g
    v = b2 * v + (f(1) - b2) * g * g
    p = p - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


ALPHA_MIN = 0.2


def selection_alpha(theta):
    """Learned selection probability of a region from its logit."""
    return ALPHA_MIN + (1.0 - ALPHA_MIN) / (1.0 + np.exp(-np.asarray(theta, np.float64)))


FALLBACK_DECAY = 0.9


def fallback_terms(c, p, q, alpha_old, alpha):
    """Per-record terms of the per-region fallback (the path tracer's
    k_fallback_sum): a record drawn from p_old = alpha_old p + (1 - alpha_old) q
    with weight c has integrand F = c p_old; F^2 / (r p_old) has expectation
    integral F^2 / r under the draws, the second moment of the one-sample
    estimator with density r.  Returns the terms for r = the mixture at the
    region's alpha and for r = p (BSDF sampling alone)."""
    c, p, q, a0, al = (np.asarray(v, np.float64) for v in (c, p, q, alpha_old, alpha))
    p_old = a0 * p + (1.0 - a0) * q
    f = c * p_old
    return f * f / ((al * p + (1.0 - al) * q) * p_old), f * f / (p * p_old)


FALLBACK_MIN_RECORDS = 256


def fallback_update(run_guided, run_bsdf, batch_guided, batch_bsdf, seen=0, batch_n=FALLBACK_MIN_RECORDS,
                    decay=FALLBACK_DECAY):
    """Running moments after one mini-batch of batch_n records in the region
    (their means given; k_fallback_update): the batch's means when the region
    has no estimate yet (both 0), else decay * running + (1 - decay) * batch;
    and the records seen.  The region samples the BSDF alone while the guided
    moment is the larger, once it has seen FALLBACK_MIN_RECORDS records.
    Returns (guided, bsdf, seen, falls back)."""
    if run_guided == 0.0 and run_bsdf == 0.0:
        g, b = batch_guided, batch_bsdf
    else:
        g = decay * run_guided + (1.0 - decay) * batch_guided
        b = decay * run_bsdf + (1.0 - decay) * batch_bsdf
    seen = seen + batch_n
    return g, b, seen, seen >= FALLBACK_MIN_RECORDS and g > b


def selection_grad(c, p, q, alpha_old, theta):
    """Per-record gradient of the one-sample MIS estimator's second moment in
    the region's logit theta (the path tracer's k_alpha_grad): a record drawn
    with densities p (BSDF), q (mixture) at alpha_old has weight c = F / p_old;
    d M / d alpha = -F^2 (p - q) / (p_mix(alpha)^2 p_old), times d alpha / d theta."""
    c, p, q, a0 = (np.asarray(v, np.float64) for v in (c, p, q, alpha_old))
    s = 1.0 / (1.0 + np.exp(-np.asarray(theta, np.float64)))
    alpha = ALPHA_MIN + (1.0 - ALPHA_MIN) * s
    p_old = a0 * p + (1.0 - a0) * q
    pm = alpha * p + (1.0 - alpha) * q
    f = c * p_old
    return -f * f * (p - q) / (pm * pm * p_old) * (1.0 - ALPHA_MIN) * s * (1.0 - s)
