"""The oracle of the learned directional sampler, pinned by closed forms.

* von Mises-Fisher lobe: the density integrates to one over the sphere and
  its mean cosine is coth(kappa) - 1/kappa; sampled cosines follow the
  closed-form CDF (e^{kappa (c-1)} - e^{-2 kappa}) / (1 - e^{-2 kappa}) and
  the azimuth is uniform about the mean -- a wrong normalisation, a sign or a
  swapped frame axis fails these.
* mixture: the density of the decoded outputs integrates to one; sampling
  by lobe reproduces it.
* loss gradient: central finite differences of the FP64 loss in every layer
  and in each output group (means, concentrations, weights, BSDF fraction).
* Adam: one step against the closed form.
"""
import numpy as np
import pytest

import orc


def _sphere_grid(n_theta=800, n_phi=1600):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1)
    dA = np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)
    return w.reshape(-1, 3), dA.ravel()


@pytest.mark.parametrize("kappa", [0.02, 1.0, 10.0, 200.0])
def test_vmf_normalised_and_mean_cosine(kappa):
    mu = np.array([0.3, -0.5, 0.8]); mu /= np.linalg.norm(mu)
    w, dA = _sphere_grid()
    c = w @ mu
    # vectorised restatement of the density only to integrate it; the value
    # at a few points is checked against the oracle itself
    pdf = kappa / (2 * np.pi * -np.expm1(-2 * kappa)) * np.exp(kappa * (c - 1))
    for i in (0, len(w) // 3, int(np.argmax(c))):
        assert abs(orc.vmf_pdf(mu, kappa, w[i]) - pdf[i]) <= 1e-12 * max(1.0, pdf[i])
    assert abs((pdf * dA).sum() - 1.0) < 2e-3
    mean_cos = (pdf * c * dA).sum()
    assert abs(mean_cos - (1 / np.tanh(kappa) - 1 / kappa)) < 2e-3


@pytest.mark.parametrize("kappa", [0.5, 30.0, 2000.0])
def test_vmf_sampling_follows_closed_form(kappa):
    mu = np.array([-0.2, 0.9, 0.1]); mu /= np.linalg.norm(mu)
    rng = np.random.default_rng(4)
    u = rng.random((20000, 2))
    d = np.array([orc.vmf_sample(mu, kappa, a, b) for a, b in u])
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    c = np.sort(d @ mu)
    cdf = (np.exp(kappa * (c - 1)) - np.exp(-2 * kappa)) / (-np.expm1(-2 * kappa))
    ks = np.abs(cdf - (np.arange(len(c)) + 0.5) / len(c)).max()
    assert ks < 0.015, ks
    # azimuth uniform around mu
    t = np.cross(mu, [1.0, 0.0, 0.0]); t /= np.linalg.norm(t)
    b = np.cross(mu, t)
    perp = d - np.outer(d @ mu, mu)
    ok = np.linalg.norm(perp, axis=1) > 1e-9
    az = np.sort(np.mod(np.arctan2(perp[ok] @ b, perp[ok] @ t), 2 * np.pi) / (2 * np.pi))
    assert np.abs(az - (np.arange(len(az)) + 0.5) / len(az)).max() < 0.015


def _random_outputs(seed):
    r = np.random.default_rng(seed)
    y = r.normal(size=41)
    y[24:32] = r.uniform(-1, 6, 8)         # log kappa
    return y


def test_mixture_normalised_and_sampled():
    y = _random_outputs(3)
    w, dA = _sphere_grid(400, 800)
    p = np.array([orc.mix_pdf(y, wi)[0] for wi in w[::7]])
    assert abs((p * dA[::7]).sum() * 7 - 1.0) < 2e-2
    # lobe frequencies follow the softmax weights
    e = np.exp(y[32:40] - y[32:40].max()); pi = e / e.sum()
    rng = np.random.default_rng(9)
    lobes = np.array([orc.mix_sample(y, *rng.random(3))[1] for _ in range(20000)])
    freq = np.bincount(lobes, minlength=8) / len(lobes)
    assert np.abs(freq - pi).max() < 0.015
    _, alpha = orc.mix_pdf(y, [0, 0, 1.0])
    assert abs(alpha - (0.5 + 0.4 / (1 + np.exp(-y[40])))) < 1e-12


def _batch(seed, n):
    r = np.random.default_rng(seed)
    x = r.random((n, 7))
    d = r.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    pb = r.uniform(0.05, 1.0, n)
    w = r.uniform(0.0, 2.0, n)
    return x, d, pb, w


def _params(seed):
    r = np.random.default_rng(seed)
    P = np.zeros(orc.GUIDE_NPARAM)
    # He-style scales per layer so that every unit is active on some records
    sizes = [(169, 64), (64, 64), (64,), (64, 64), (64,), (64, 64), (64,), (64, 41), (41,)]
    o = 0
    for s in sizes:
        k = int(np.prod(s))
        P[o:o + k] = r.normal(size=k) * (np.sqrt(2.0 / s[0]) if len(s) == 2 else 0.1)
        o += k
    assert o == orc.GUIDE_NPARAM
    return P


def test_loss_gradient_matches_finite_differences():
    P = _params(5)
    x, d, pb, w = _batch(6, 6)
    L, g = orc.guide_loss(P, x, d, pb, w)
    assert np.isfinite(L)
    rng = np.random.default_rng(7)
    # indices in every block: W1 (incl. bias row), W2, b2, W3, b3, W4, b4, W5
    # per output group, b5 per output group
    blocks = [(0, 10816), (10816, 14912), (14912, 14976), (14976, 19072), (19072, 19136),
              (19136, 23232), (23232, 23296)]
    idx = [int(rng.integers(a, b)) for a, b in blocks for _ in range(6)]
    idx += [10752 + int(rng.integers(0, 64)) for _ in range(3)]          # W1 bias row
    for col in (0, 5, 24, 27, 32, 38, 40):                             # W5 / b5 by output group
        idx += [23296 + int(rng.integers(0, 64)) * 41 + col, 25920 + col]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h; Pm[i] -= h
        fd = (orc.guide_loss(Pp, x, d, pb, w, grad=False) - orc.guide_loss(Pm, x, d, pb, w, grad=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
    assert np.count_nonzero(g) > 0.5 * g.size


def test_adam_first_steps_closed_form():
    rng = np.random.default_rng(1)
    n = 1000
    p0 = rng.normal(size=n)
    g1, g2 = rng.normal(size=n), rng.normal(size=n)
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    orc.guide_adam(p, m, v, g1, 1, lr=1e-3)
    np.testing.assert_allclose(p, p0 - 1e-3 * g1 / (np.abs(g1) + 1e-8), rtol=1e-12, atol=1e-15)
    orc.guide_adam(p, m, v, g2, 2, lr=1e-3)
    m2 = 0.9 * 0.1 * g1 + 0.1 * g2
    v2 = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    exp = p0 - 1e-3 * g1 / (np.abs(g1) + 1e-8) - 1e-3 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    np.testing.assert_allclose(p, exp, rtol=1e-10, atol=1e-14)


def test_bf16_rounding_is_nearest_even():
    assert orc.bf16_round(1.0) == 1.0
    assert orc.bf16_round(1.0 + 2 ** -8) == 1.0                 # tie -> even (1.0)
    assert orc.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6   # tie -> even (odd below)
    assert orc.bf16_round(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
    assert orc.bf16_round(-3.0) == -3.0
