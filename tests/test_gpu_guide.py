"""GPU path-guiding network (libyafaray_amd/csrc/guide.hip) against its CPU
oracle (oracle/ora_guide.py): the mixture it infers, the gradients of the
training loss (MLP weights and hash grid), and that Adam on those gradients
learns a position-dependent lobe.  The oracle rounds the grid, inputs,
weights and activations to BF16 like the kernels and accumulates in float32;
the MFMA accumulation order differs, so results agree to BF16 rounding
(relative 5e-3 on densities at the 99th percentile, 5e-3 on gradient norms), not bit
for bit."""
import sys

import numpy as np
import pytest

from oracle import ora_guide as G

pytestmark = pytest.mark.gpu

BMIN, BMAX = (-1.0, -2.0, 0.0), (3.0, 2.0, 2.0)
LOG2 = 14


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def unit(v):
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def case(seed, n):
    rng = np.random.default_rng(seed)
    p = G.init_params(seed, table_size=1 << LOG2, bmin=BMIN, bmax=BMAX)
    p["grid"] = rng.normal(0, 0.5, p["grid"].shape).astype(np.float32)
    p["W"] = [w * np.float32(1.3) for w in p["W"]]
    pos = rng.uniform(BMIN, BMAX, (n, 3)).astype(np.float32)
    nrm = unit(rng.normal(size=(n, 3)))
    wo = unit(rng.normal(size=(n, 3)))
    w = unit(rng.normal(size=(n, 3)))
    c = rng.uniform(0.2, 2.0, n).astype(np.float32)
    return p, pos, nrm, wo, w, c


def gpu_guide(p):
    from libyafaray_amd.guide import GpuGuide, mlp_to_flat
    g = GpuGuide(BMIN, BMAX, log2_table_size=LOG2, max_batch=1 << 16)
    g.set_weights(p["grid"], mlp_to_flat(p["W"]))
    return g


def test_forward_matches_oracle():
    from libyafaray_amd.guide import pack_samples
    p, pos, nrm, wo, w, c = case(0, 3000)
    say("guide forward 3000 queries")
    g = gpu_guide(p)
    mix = g.infer(pack_samples(pos, nrm, wo)).cpu().numpy()
    g.close()
    x, _ = G.encode(p, pos, nrm, wo)
    hs, out = G.mlp_forward(p, x, bf=True)
    mu, kappa, pi, _ = G.decode(out)
    gm = mix.reshape(-1, 8, 6)
    assert np.all(np.isfinite(gm))
    # lobe weights sum to one, directions are unit
    assert np.allclose(gm[..., 4].sum(-1), 1.0, atol=1e-5)
    assert np.allclose(np.linalg.norm(gm[..., 0:3], axis=-1), 1.0, atol=1e-5)
    # mixture density at random directions, and the parameters themselves
    qg = (gm[..., 4] * gm[..., 5] * np.exp(gm[..., 3] * ((gm[..., 0:3] * w[:, None, :]).sum(-1) - 1))).sum(-1)
    qo = G.mixture_pdf(out, w)
    rel = np.abs(qg - qo) / qo
    say(f"  pdf rel err median {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e}")
    assert np.quantile(rel, 0.99) < 5e-3 and np.median(rel) < 1e-4
    assert np.quantile(np.abs(gm[..., 4] - pi), 0.99) < 1e-2
    assert np.quantile(np.abs(np.log(gm[..., 3] / kappa)), 0.99) < 3e-2


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradients_match_oracle():
    from libyafaray_amd.guide import flat_to_mlp, pack_samples
    p, pos, nrm, wo, w, c = case(1, 4000)
    say("guide backward 4000 samples")
    g = gpu_guide(p)
    g.backward(pack_samples(pos, nrm, wo, w, c))
    ggrid, gmlp, gloss = g.gradients()
    g.close()
    x, corners = G.encode(p, pos, nrm, wo)
    hs, out = G.mlp_forward(p, x, bf=True)
    oloss = G.loss(out.astype(np.float64), w.astype(np.float64), c.astype(np.float64))
    gW, ogrid = G.backward(p, hs, out, corners, w, c, bf=True)
    say(f"  loss gpu {gloss:.6f} oracle {oloss:.6f}")
    assert gloss == pytest.approx(oloss, rel=2e-3)
    for k, (a, b) in enumerate(zip(flat_to_mlp(gmlp), gW)):
        e = rel_err(a[:40] if k == 3 else a, b[:40] if k == 3 else b)
        cos = float((a * b).sum() / (np.linalg.norm(a) * np.linalg.norm(b)))
        say(f"  dW{k + 1}: rel err {e:.2e}, cosine {cos:.6f}")
        assert e < 5e-3 and cos > 0.99999
    assert np.all(flat_to_mlp(gmlp)[3][40:] == 0)
    touched = ogrid != 0
    e = rel_err(ggrid[touched], ogrid[touched])
    say(f"  grid: {touched.sum()} entries, rel err {e:.2e}, untouched nonzero {np.count_nonzero(ggrid[~touched])}")
    assert e < 5e-3
    assert np.count_nonzero(ggrid[~touched]) == 0


def test_adam_steps_match_oracle():
    """Two Adam steps (k_adam on the MLP, k_adam_grid on the lattice entries in
    use) on the gradients the GPU formed: the FP32 master weights equal the
    oracle's Adam (Kingma & Ba 2015) applied to those gradients, to float32
    rounding of the bias corrections (the host forms 1 - b^t with powf)."""
    from libyafaray_amd.guide import flat_to_mlp, mlp_to_flat, pack_samples
    p, pos, nrm, wo, w, c = case(2, 3000)
    say("guide two Adam steps")
    g = gpu_guide(p)
    grid0, mlp0 = g.get_weights()
    state = [(grid0.reshape(-1), np.zeros(grid0.size, np.float32), np.zeros(grid0.size, np.float32)),
             (mlp0, np.zeros(mlp0.size, np.float32), np.zeros(mlp0.size, np.float32))]
    for step, (lo, hi) in enumerate(((0, 1500), (1500, 3000)), start=1):
        g.backward(pack_samples(pos[lo:hi], nrm[lo:hi], wo[lo:hi], w[lo:hi], c[lo:hi]))
        ggrid, gmlp, _ = g.gradients()
        g.adam_step()
        grid, mlp = g.get_weights()
        state = [G.adam_step(*state[0], ggrid.reshape(-1), step), G.adam_step(*state[1], gmlp, step)]
        for name, got, exp, prev in (("grid", grid.reshape(-1), state[0][0], None), ("mlp", mlp, state[1][0], None)):
            err = np.abs(got - exp)
            moved = np.abs(exp - (grid0.reshape(-1) if name == "grid" else mlp0)) > 0
            say(f"  step {step} {name}: {moved.sum()} weights moved, max abs err {err.max():.2e}")
            assert moved.sum() > 100
            assert np.all(err <= 2e-6 * np.maximum(1.0, np.abs(exp))), (name, step, err.max())
    g.close()


def test_training_learns_a_position_dependent_lobe():
    """Directions follow a vMF lobe (kappa 20) whose axis turns with the
    position; Adam on the network's own gradients raises the held-out log
    likelihood from near the uniform sphere's towards the lobe's entropy."""
    from libyafaray_amd.guide import GpuGuide, pack_samples
    rng = np.random.default_rng(7)

    def draw(n):
        pos = rng.uniform(BMIN, BMAX, (n, 3)).astype(np.float32)
        t = (pos[:, 0] - BMIN[0]) / (BMAX[0] - BMIN[0])
        axis = unit(np.stack([np.cos(3 * t), np.sin(3 * t), 0.5 + 0 * t], -1))
        w = G.vmf_sample(axis.astype(np.float64), np.full(n, 20.0), rng.random(n), rng.random(n)).astype(np.float32)
        nrm = np.tile(np.float32([0, 0, 1]), (n, 1))
        return pack_samples(pos, nrm, nrm, w, np.ones(n, np.float32)), w

    held, hw = draw(8192)

    def loglik(g):
        mix = g.infer(held).cpu().numpy().reshape(-1, 8, 6)
        q = (mix[..., 4] * mix[..., 5] * np.exp(mix[..., 3] * ((mix[..., 0:3] * hw[:, None, :]).sum(-1) - 1))).sum(-1)
        return float(np.log(q).mean())

    g = GpuGuide(BMIN, BMAX, log2_table_size=LOG2, max_batch=1 << 14, lr=1e-2)
    before = loglik(g)
    for step in range(300):
        batch, _ = draw(1 << 14)
        g.train_step(batch)
    after = loglik(g)
    g.close()
    k = 20.0
    best = np.log(k / (2 * np.pi * (1 - np.exp(-2 * k)))) + k * (1 / np.tanh(k) - 1 / k - 1)  # -entropy of the lobe
    say(f"  held-out log q: before {before:.3f}, after {after:.3f}, lobe optimum {best:.3f}")
    assert before < -1.5
    assert after > before + 1.5 and after > best - 1.0


# ---------------------------------------------------------------- guided path tracing
def _render(sc, spp, guide=None, train_passes=0, seed=0, seeds=(0,), depth=None, learned=False, selection=None):
    """Film images (one per seed) of a plain or guided render; with
    train_passes the guide first trains on that many 1-spp passes (learned:
    with the selection probability learned per region, returned in
    selection[0]); depth: the path vertices after the sub-path start that are
    guided (None: all)."""
    from libyafaray_amd.pt import GpuPathTracer
    s = dict(sc)
    s["settings"] = dict(sc["settings"], aa_samples=spp, seed=seed)
    pt = GpuPathTracer(s)
    pt.set_guide_depth(depth)
    alpha = None
    if guide is not None and train_passes:
        pt.set_guide(guide, bsdf_prob=0.3, train=True, train_steps=8)
        if learned:
            pt.set_guide_selection("learn", lr=0.05)
        pt.clear()
        for k in range(train_passes):
            pt.render(k, 1)
        if learned:
            alpha = pt.guide_selection()
            if selection is not None:
                selection.append(alpha)
    pt.close()
    imgs = []
    for sd in seeds:
        s["settings"] = dict(sc["settings"], aa_samples=spp, seed=sd)
        pt = GpuPathTracer(s)
        pt.set_guide_depth(depth)
        if guide is not None:
            pt.set_guide(guide, bsdf_prob=0.3, train=False)
            if learned:
                pt.set_guide_selection("keep")
        pt.clear()
        pt.render(0, spp)
        rgba, w = pt.film()
        pt.close()
        imgs.append(rgba[..., :3] / w[..., None])
    return imgs


@pytest.mark.parametrize("depth,learned", [(None, False), (1, False), (1, True)])
def test_guided_render_converges_to_the_plain_image(depth, learned):
    """One-sample MIS keeps the estimator unbiased: with an untrained and with a
    trained guide, the image mean of independent guided renders (seeds 1..8,
    256 spp each: the guided directions and the roulette draw from the seed)
    agrees with a 4096-spp plain render's within four standard errors of
    their spread (a wrong MIS weight or a biased roulette is off by per cents),
    and the guided images are finite and within 10x the plain relMSE.  Guiding
    every vertex, and only the camera vertex and the next one; with the
    selection probability learned per region, which also stays within
    [0.2, 1] and moves away from its start."""
    from libyafaray_amd import scenes
    from libyafaray_amd.guide import GpuGuide
    sc = scenes.cornell_path_scene(n_random=2000, res=48, spp=16)
    bmin, bmax = sc["verts"].min(0) - 1e-3, sc["verts"].max(0) + 1e-3
    ref = _render(sc, 4096, seeds=(99,))[0]
    rel = lambda x: float(((x - ref) ** 2 / (ref ** 2 + 1e-2)).mean())
    plain = _render(sc, 256)[0]
    say(f"  plain 256 spp: mean {plain.mean():.5f} (reference {ref.mean():.5f}), relMSE {rel(plain):.2e}")
    for train in (0, 8):
        g = GpuGuide(bmin, bmax, log2_table_size=16, max_batch=1 << 14, seed=3)
        sel = []
        imgs = _render(sc, 256, guide=g, train_passes=train, seeds=tuple(range(1, 9)), depth=depth, learned=learned,
                       selection=sel)
        g.close()
        if learned and train:
            a = sel[0]
            moved = np.abs(a - 0.3) > 0.05
            say(f"  learned selection: min {a.min():.3f} max {a.max():.3f}, {moved.sum()} of {a.size} regions moved")
            assert np.all((a >= 0.2 - 1e-6) & (a <= 1 + 1e-6)) and moved.sum() > 0
        means = np.array([im.mean() for im in imgs], np.float64)
        m, se = means.mean(), means.std(ddof=1) / np.sqrt(len(means))
        errs = [rel(im) for im in imgs]
        say(f"  guided ({train} training passes) 8 x 256 spp: mean {m:.5f} +- {se:.5f}, "
            f"ratio {m / ref.mean():.4f}, relMSE median {np.median(errs):.2e}")
        assert all(np.isfinite(im).all() for im in imgs)
        assert abs(m - ref.mean()) < 4 * se + 1e-3 * ref.mean()
        # measured 1.4-6.9x plain's (the trained guide varies between runs:
        # its training sums gradients with float atomics): with bsdf_prob
        # 0.3 the guided vertices draw 70 % of their directions from the
        # mixture, and their own uniforms where the plain tracer has the
        # reference's low-discrepancy samples
        assert np.median(errs) < 10 * rel(plain)


def _plates(res=24, spp=16, seed=0):
    """Two facing Lambertian plates (z = 0 and z = 1, albedo 0.8) under a
    downward area light plane at z = 0.5 covering them: every vertex sees the
    same radiance from every direction of its hemisphere, so BSDF sampling is
    (up to the reference's 0.01 sample-weight term) the zero-variance density
    and any other guided mixture can only add variance."""
    S, h = 1000.0, 0.5
    verts = np.array([[-S, -S, 0], [S, -S, 0], [S, S, 0], [-S, S, 0],
                      [-S, -S, 1], [S, -S, 1], [S, S, 1], [-S, S, 1]], np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]], np.uint32)
    return dict(verts=verts, tris=tris, mat_id=np.zeros(4, np.uint32), materials=[dict(color=(0.8,) * 3)],
                lights=[dict(type="area", corner=(-S, -S, h), point1=(-S, S, h), point2=(S, -S, h), power=1.0)],
                camera=dict(from_=(0.0, 0.0, 0.9), to=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.9), focal=0.6,
                            resx=res, resy=res),
                settings=dict(path_samples=1, bounces=4, rr_min_bounces=10, aa_samples=spp, aa_pixelwidth=1.5,
                              background=None, seed=seed))


def test_fallback_to_bsdf_sampling_where_guiding_loses():
    """ymx_pt_set_guide_fallback on a scene where BSDF sampling cannot be
    beaten: after training, nearly every region with 256 records or more has a
    running guided second moment above BSDF sampling's and samples the BSDF alone
    (its selection probability reads 1); the renders stay unbiased (mean of 8
    seeds within 4 standard errors of a 4096-spp plain render), and their
    pixel variance across seeds is below that of the same guide without the
    fallback (measured: 1.4e-4 against 0.14; what remains comes from the few
    regions still guided and from the guide's own uniforms in the BSDF branch,
    since the plain tracer's samples here do not depend on the seed)."""
    from libyafaray_amd.guide import GpuGuide
    from libyafaray_amd.pt import GpuPathTracer
    sc = _plates()
    bmin, bmax = np.float32([-2, -2, -0.01]), np.float32([2, 2, 1.01])
    ref = _render(sc, 4096, seeds=(99,))[0]
    plain = np.array(_render(sc, 64, seeds=tuple(range(1, 9))))
    results = {}
    for fallback in (False, True):
        g = GpuGuide(bmin, bmax, log2_table_size=16, max_batch=1 << 14, seed=5)
        s = dict(sc)
        s["settings"] = dict(sc["settings"], aa_samples=1, seed=0)
        pt = GpuPathTracer(s)
        pt.set_guide(g, bsdf_prob=0.3, train=True, train_steps=8)
        if fallback:
            pt.set_guide_fallback("start")
        pt.clear()
        for k in range(12):
            pt.render(k, 1)
        if fallback:
            mg, mb, records = pt.guide_fallback()
            alpha = pt.guide_selection()
            seen = records >= 256
            down = seen & (mg > mb)
            say(f"  fallback: {seen.sum()} regions with 256 records or more, {down.sum()} fallen back; "
                f"median guided / BSDF moment {np.median(mg[seen] / mb[seen]):.3f}")
            assert seen.sum() >= 20
            assert down.sum() >= 0.9 * seen.sum()
            assert np.all(alpha[down] == 1.0) and np.all(alpha[seen & ~down] == np.float32(0.3))
        pt.close()
        imgs = []
        for sd in range(1, 9):
            s["settings"] = dict(sc["settings"], aa_samples=64, seed=sd)
            pt = GpuPathTracer(s)
            pt.set_guide(g, bsdf_prob=0.3, train=False)
            if fallback:
                pt.set_guide_fallback("keep")
            pt.clear()
            pt.render(0, 64)
            rgba, w = pt.film()
            pt.close()
            imgs.append(rgba[..., :3] / w[..., None])
        g.close()
        results[fallback] = np.array(imgs)
    for name, imgs in (("plain", plain), ("guided", results[False]), ("guided + fallback", results[True])):
        means = imgs.mean(axis=(1, 2, 3))
        m, se = means.mean(), means.std(ddof=1) / np.sqrt(len(means))
        var = float(imgs.var(axis=0, ddof=1).mean())
        say(f"  {name}: mean {m:.5f} +- {se:.5f} (reference {ref.mean():.5f}), pixel variance {var:.3e}")
        assert abs(m - ref.mean()) < 4 * se + 1e-3 * ref.mean()
        results[name] = var
    assert results["guided + fallback"] < results["guided"]


def test_trained_guide_beats_bsdf_sampling_on_a_spot_lit_room():
    """What the trained network gains over the method without it: on a box lit
    mostly by a small bright spot of the floor (scenes.spot_room_scene), the
    guide trained on the render's first 32 spp and combined with BSDF
    sampling (probability 0.5) gives a lower relMSE at 64 further spp than
    plain BSDF sampling, against a 8192-spp reference, averaged over 4 seeds
    (measured: 2.0-2.5x; asserted: 1.4x), and stays unbiased in the mean."""
    from libyafaray_amd import scenes
    from libyafaray_amd.guide import GpuGuide
    from libyafaray_amd.pt import GpuPathTracer
    sc = scenes.spot_room_scene(res=64)

    def render(spp, seed, guide=None):
        s = dict(sc)
        s["settings"] = dict(sc["settings"], aa_samples=spp + (32 if guide else 0), seed=seed)
        pt = GpuPathTracer(s)
        first = 0
        if guide is not None:
            pt.set_guide_depth(1)
            pt.set_guide(guide, bsdf_prob=0.5, train=True, train_steps=16)
            pt.clear()
            for k in range(8):
                pt.render(4 * k, 4)
            pt.set_guide(guide, bsdf_prob=0.5, train=False)
            first = 32
        pt.clear()
        pt.render(first, spp)
        rgba, w = pt.film()
        pt.close()
        return rgba[..., :3].astype(np.float64) / w[..., None]

    ref = (render(4096, 91) + render(4096, 92)) / 2
    rel = lambda x: float(((x - ref) ** 2 / (ref ** 2 + 1e-2)).mean())
    plain, guided, means = [], [], []
    bmin, bmax = sc["verts"].min(0) - 1e-3, sc["verts"].max(0) + 1e-3
    for sd in range(1, 5):
        plain.append(rel(render(64, sd)))
        g = GpuGuide(bmin, bmax, log2_table_size=19, n_min=4, n_max=32, max_batch=1 << 16, lr=1e-2, seed=sd)
        img = render(64, sd, guide=g)
        g.close()
        guided.append(rel(img))
        means.append(img.mean())
    gain = np.mean(plain) / np.mean(guided)
    m, se = np.mean(means), np.std(means, ddof=1) / 2
    say(f"  spot room 64 spp: plain relMSE {np.mean(plain):.3e}, guided {np.mean(guided):.3e}, gain {gain:.2f}; "
        f"guided mean {m:.5f} +- {se:.5f} (reference {ref.mean():.5f})")
    assert gain > 1.4
    assert abs(m - ref.mean()) < 4 * se + 2e-3 * ref.mean()
