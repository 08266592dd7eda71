"""GPU parity of the learned directional sampler against its FP64 oracle.

* network forward (BF16 MFMA, FP32 accumulation) against the oracle that
  rounds weights and activations to bfloat16 at the same points; the
  un-rounded FP64 network is farther away, which shows the rounding points
  (and the operand packing) are the ones the oracle models;
* mixture decoding and von Mises-Fisher sampling from the same outputs and
  uniforms: same lobe, direction within 2e-5, density within 1e-4 relative
  plus the FP32 cosine error times the largest concentration;
* loss gradient of a minibatch (forward, output gradient, backward, weight
  gradients) and one Adam step;
* the guided estimator is unbiased: after online training its image agrees
  with the plain path tracer's on 4x4-pixel cells within 3 sigma and in the
  image mean within 3.5 standard errors;
* training lowers the loss.
"""
import numpy as np
import pytest

import orc
from lmx import guide, render, scenes

pytestmark = pytest.mark.gpu


def _params(seed):
    r = np.random.default_rng(seed)
    P = np.zeros(guide.NPARAM)
    sizes = [(169, 64), (64, 64), (64,), (64, 64), (64,), (64, 64), (64,), (64, 41), (41,)]
    o = 0
    for s in sizes:
        k = int(np.prod(s))
        P[o:o + k] = r.normal(size=k) * (np.sqrt(2.0 / s[0]) if len(s) == 2 else 0.1)
        o += k
    return P.astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def cornell():
    return render.Scene(scenes.cornell_box(32, 32))


def test_forward_matches_bf16_oracle(gpu, cornell):
    import torch
    g = guide.Guide(cornell)
    P = _params(1)
    g.params = P.astype(np.float32)
    x = np.random.default_rng(2).random((3000, 7)).astype(np.float32)
    yg = g.eval(torch.from_numpy(x).to(gpu)).cpu().numpy().astype(np.float64)
    yb = orc.guide_forward(P, x.astype(np.float64), bf16=True)
    yx = orc.guide_forward(P, x.astype(np.float64), bf16=False)
    scale = np.abs(yb).max()
    err_b = np.abs(yg - yb)
    assert err_b.max() < 0.03 * scale, err_b.max() / scale
    assert err_b.mean() < 2e-3 * scale, err_b.mean() / scale
    assert err_b.mean() < 0.5 * np.abs(yg - yx).mean()


def test_mixture_sampling_matches_oracle(gpu):
    import torch
    r = np.random.default_rng(3)
    n = 2000
    y = r.normal(size=(n, 41))
    y[:, 24:32] = r.uniform(-3, 8, (n, 8))
    u = r.random((n, 3))
    out = guide.Guide.mix_sample(torch.from_numpy(y.astype(np.float32)).to(gpu),
                                 torch.from_numpy(u.astype(np.float32)).to(gpu)).cpu().numpy()
    y32, u32 = y.astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64)
    same = 0
    for i in range(n):
        d, k = orc.mix_sample(y32[i], *u32[i])
        p, a = orc.mix_pdf(y32[i], out[i, :3].astype(np.float64))
        assert abs(out[i, 4] - a) < 1e-6
        if int(out[i, 5]) != k:
            continue                   # u on a lobe boundary within FP32 rounding
        same += 1
        assert np.abs(out[i, :3] - d).max() < 2e-5, (i, out[i, :3], d)
        # exp(kappa (mu.w - 1)) turns the FP32 rounding of the cosine
        # (a few 1e-8) into a relative error of kappa times that
        kmax = np.exp(np.clip(y32[i, 24:32], -4, 9)).max()
        assert abs(out[i, 3] - p) <= (1e-4 + 4e-7 * kmax) * p + 1e-6, (i, out[i, 3], p, kmax)
    assert same > 0.995 * n


def _records(n, seed):
    r = np.random.default_rng(seed)
    x = r.random((n, 7))
    d = r.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    pb = r.uniform(0.05, 1.0, n)
    w = r.uniform(0.0, 2.0, n)
    rec = guide.Guide.records(x, d, pb, w)
    x, d, pb, w = (rec[:, 0:7].astype(np.float64), rec[:, 8:11].astype(np.float64),
                   rec[:, 7].astype(np.float64), rec[:, 15].astype(np.float64))
    return rec, x, d, pb, w


def test_loss_gradient_matches_oracle(gpu, cornell):
    import torch
    g = guide.Guide(cornell)
    P = _params(4)
    g.params = P.astype(np.float32)
    rec, x, d, pb, w = _records(512, 5)
    gg, loss = g.grad(torch.from_numpy(rec).to(gpu))
    gg = gg.cpu().numpy().astype(np.float64)
    L, go = orc.guide_loss(P, x, d, pb, w, bf16=True)
    assert abs(loss - L) < 2e-3 * abs(L) + 1e-4, (loss, L)
    # per block: relative L2 error and direction
    blocks = {"W1": (0, 10816), "W2": (10816, 14912), "b2": (14912, 14976), "W3": (14976, 19072),
              "b3": (19072, 19136), "W4": (19136, 23232), "b4": (23232, 23296), "W5": (23296, 25920),
              "b5": (25920, 25961)}
    for name, (a, b) in blocks.items():
        ga, gb = gg[a:b], go[a:b]
        rel = np.linalg.norm(ga - gb) / max(np.linalg.norm(gb), 1e-12)
        assert rel < 0.03, (name, rel)
    cos = gg @ go / (np.linalg.norm(gg) * np.linalg.norm(go))
    assert cos > 0.9995, cos


def test_adam_step_matches_oracle(gpu, cornell):
    import torch
    g = guide.Guide(cornell, lr=1e-3)
    P = _params(6)
    g.params = P.astype(np.float32)
    rec, *_ = _records(256, 7)
    t = torch.from_numpy(rec).to(gpu)
    grad1, _ = g.grad(t)                       # gradient at P, no step
    grad1 = grad1.cpu().numpy().astype(np.float64)
    g.grad(t, adam=True)                       # one step from P
    P1 = g.params.astype(np.float64)
    Po, m, v = P.copy(), np.zeros_like(P), np.zeros_like(P)
    orc.guide_adam(Po, m, v, grad1, 1, lr=1e-3)
    np.testing.assert_allclose(P1, Po, rtol=1e-5, atol=2e-7)


def _batches(fn, k):
    return np.stack([fn(s).cpu().numpy().astype(np.float64) for s in range(k)])


def test_guided_render_unbiased_and_training_helps(gpu, cornell):
    g = guide.Guide(cornell, steps_per_pass=16, batch=1 << 14)
    npix = 32 * 32
    losses = []
    for k in range(12):
        _, _, gs = g.render(npix * 64, max_num_vertices=9, seed=1000 + k, train=True)
        losses.append(gs["loss"])
        assert gs["records"] > 0
    assert np.mean(losses[-4:]) < np.mean(losses[:4]), losses
    # One-sample MIS keeps the estimator unbiased for any network, but its
    # weights f cos / q reach 1 / alpha_min times the BSDF's per bounce, so
    # per-batch means of a trained network are strongly skewed (a cell's
    # batch means have skewness near 3 at 512 spp) and a t statistic over
    # few batches leans negative: compare 4x4-pixel cells and the image mean,
    # each against its standard error, over 128 guided batches.
    kg, kp = 128, 32
    G = _batches(lambda s: g.render(npix * 512, max_num_vertices=9, seed=300 + s)[0], kg)
    P = _batches(lambda s: cornell.render_pt(npix * 8192, max_num_vertices=9, seed=200 + s)[0], kp)
    cg = G.reshape(kg, 8, 4, 8, 4, 3).mean(axis=(2, 4))
    cp = P.reshape(kp, 8, 4, 8, 4, 3).mean(axis=(2, 4))
    se = np.sqrt(cg.var(0, ddof=1) / kg + cp.var(0, ddof=1) / kp) + 1e-12
    z = np.abs(cg.mean(0) - cp.mean(0)) / se
    lit = (cg.mean(0) + cp.mean(0)) > 1e-3
    assert lit.mean() > 0.5
    assert (z[lit] <= 3.0).mean() > 0.95, (z[lit] > 3).mean()
    mg, mp = G.mean(axis=(1, 2, 3)), P.mean(axis=(1, 2, 3))
    se = np.sqrt(mg.var(ddof=1) / kg + mp.var(ddof=1) / kp)
    assert abs(mg.mean() - mp.mean()) < 3.5 * se, (mg.mean(), mp.mean(), se)
