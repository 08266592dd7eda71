"""The oracle's path sampler, pinned by published vectors and closed forms.

* Philox4x32-10: the Random123 known-answer vectors.
* Lambertian and normalized Phong: the sampling pdf integrates to one, the
  sampled distribution follows it (cos a has CDF x^(n+1) about the mirror
  direction), and the directional albedo at normal incidence is exactly R
  (what the (n+2)/(2 pi) normalisation means) -- a dropped or shifted
  normalisation term fails these.
* renderer::pt inside an emitting, reflecting closed sphere: radiance
  Le * sum_{j < m-1} rho^j (m = max_num_vertices), Le / (1 - rho) unbounded,
  with and without Russian roulette.
"""
import numpy as np
import pytest

import orc
from lmx import scenes


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert list(orc.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _sphere_dirs(n):
    """quadrature points and weights on the unit sphere (midpoint in cos, phi)"""
    k = n
    ct = (np.arange(k) + 0.5) / k * 2 - 1
    ph = (np.arange(2 * k) + 0.5) / (2 * k) * 2 * np.pi
    C, P = np.meshgrid(ct, ph, indexing="ij")
    S = np.sqrt(1 - C ** 2)
    d = np.stack([S * np.cos(P), S * np.sin(P), C], axis=-1).reshape(-1, 3)
    w = (2.0 / k) * (2 * np.pi / (2 * k))
    return d, w


@pytest.mark.parametrize("mat", [
    dict(type=scenes.MAT_DIFFUSE, R=(0.7, 0.5, 0.3), exponent=0.0, Le=(0, 0, 0)),
    dict(type=scenes.MAT_PHONG, R=(0.7, 0.5, 0.3), exponent=12.0, Le=(0, 0, 0)),
    dict(type=scenes.MAT_PHONG, R=(0.7, 0.5, 0.3), exponent=150.0, Le=(0, 0, 0)),
])
def test_bsdf_pdf_normalised_and_albedo_at_normal_incidence(mat):
    n = np.array([0, 0, 1], np.float32)
    wi = np.array([0, 0, 1], np.float32)
    # at normal incidence both lobes are symmetric about n: integrate the
    # solid-angle pdf over cos(theta) with 2 pi for the azimuth
    k = 20000
    tot = 0.0
    for c in (np.arange(k) + 0.5) / k:
        x = np.array([np.sqrt(1 - c * c), 0.0, c], np.float32)
        pdf, f = orc.bsdf_eval(mat, n, wi, x)
        tot += pdf * c * (2 * np.pi / k)           # projected -> solid angle
    assert abs(tot - 1.0) < 2e-3, tot
    # and over the whole hemisphere on a coarse 2-D grid for the diffuse lobe
    if mat["type"] == scenes.MAT_DIFFUSE:
        d, w = _sphere_dirs(200)
        tot2 = sum(orc.bsdf_eval(mat, n, wi, x)[0] * x[2] * w for x in d[d[:, 2] > 0])
        assert abs(tot2 - 1.0) < 5e-3, tot2
    rng = np.random.default_rng(3)
    acc = np.zeros(3)
    m = 20000
    for u in rng.random((m, 2), dtype=np.float32):
        ok, wo, pdf, f = orc.bsdf_query(mat, n, wi, u)
        assert ok
        if pdf > 0:
            acc += f / pdf
    np.testing.assert_allclose(acc / m, mat["R"], rtol=1e-2)


def test_phong_lobe_follows_its_pdf():
    mat = dict(type=scenes.MAT_PHONG, R=(1, 1, 1), exponent=30.0, Le=(0, 0, 0))
    n = np.array([0, 0, 1], np.float32)
    wi = np.array([0.6, 0.0, 0.8], np.float32)       # mirror direction (-0.6, 0, 0.8)
    r = np.array([-0.6, 0.0, 0.8])
    rng = np.random.default_rng(5)
    cos_a = []
    for u in rng.random((20000, 2), dtype=np.float32):
        ok, wo, pdf, f = orc.bsdf_query(mat, n, wi, u)
        assert ok
        cos_a.append(float(np.dot(wo, r)))
    cos_a = np.sort(np.array(cos_a))
    # CDF of cos a is x^(n+1): Kolmogorov-Smirnov distance
    emp = np.arange(1, cos_a.size + 1) / cos_a.size
    ks = np.max(np.abs(emp - np.clip(cos_a, 0, 1) ** 31.0))
    assert ks < 0.015, ks


@pytest.mark.parametrize("maxv,rr_depth", [(4, 1), (7, 3), (-1, 2)])
def test_pt_furnace_closed_form(maxv, rr_depth):
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, width=32, height=32)
    out = orc.Scene(sc).pt_samples(0, 32 * 32 * 256, max_num_vertices=maxv, rr_depth=rr_depth, seed=5)
    expect = Le / (1 - rho) if maxv == -1 else Le * sum(rho ** j for j in range(maxv - 1))
    assert abs(out[:, 2:5].mean() - expect) < 0.01 * expect


def test_uniform_table_and_philox_agree():
    """A table holding the generator's own values reproduces the generated
    samples exactly (the slot layout is the one both use)."""
    sc = scenes.cornell_box(16, 16)
    s = orc.Scene(sc)
    n, maxv = 4096, 9
    w = orc.table_width(maxv)
    lib = orc._rlib()
    tab = np.array([[lib.orc_uniform(42, i, k) for k in range(w)] for i in range(n)], np.float32)
    a = s.pt_samples(0, n, max_num_vertices=maxv, seed=42)
    b = s.pt_samples(0, n, max_num_vertices=maxv, table=tab)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("rr_depth", [1, 3])
def test_pt_furnace_throughput_roulette(rr_depth):
    """The throughput roulette (survive with min(1, max channel)) is unbiased:
    the furnace closed form holds with it too."""
    rho, Le, maxv = 0.5, 1.0, 7
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, width=32, height=32)
    out = orc.Scene(sc).pt_samples(0, 32 * 32 * 256, max_num_vertices=maxv, rr_depth=rr_depth, seed=6, rr_mode=1)
    expect = Le * sum(rho ** j for j in range(maxv - 1))
    assert abs(out[:, 2:5].mean() - expect) < 0.01 * expect


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("maxv,rr_depth", [(4, 1), (7, 3)])
def test_nee_furnace_closed_form(mode, maxv, rr_depth):
    """renderer::ptdirect (1) and renderer::ptmis (2) inside the emitting,
    reflecting sphere: every path length k carries Le rho^(k-2), found by
    light sampling (and, for ptmis, by BSDF hits weighted against it)."""
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, width=32, height=32)
    out = orc.Scene(sc).nee_samples(0, 32 * 32 * 128, mode, max_num_vertices=maxv, rr_depth=rr_depth, seed=8)
    expect = Le * sum(rho ** j for j in range(maxv - 1))
    img = out[:, 2:5].mean() + out[:, 10:13].mean()
    assert abs(img - expect) < 0.01 * expect, (img, expect)


@pytest.mark.parametrize("mode", [1, 2])
def test_nee_cornell_matches_pt(mode):
    """Next-event estimation and plain path tracing estimate the same image."""
    sc = scenes.cornell_box(32, 32)
    o = orc.Scene(sc)
    n = 32 * 32 * 256
    a = o.nee_samples(0, n, mode, max_num_vertices=9, seed=3)
    ma = (a[:, 2:5].astype(np.float64) + a[:, 10:13]).mean(axis=1)
    b = o.pt_samples(0, 32 * 32 * 4096, max_num_vertices=9, seed=4)[:, 2:5].astype(np.float64).mean(axis=1)
    se = np.hypot(ma.std() / np.sqrt(len(ma)), b.std() / np.sqrt(len(b)))
    assert abs(ma.mean() - b.mean()) < 3.5 * se, (ma.mean(), b.mean(), se)
    # and light sampling is far less noisy than BSDF sampling here
    assert ma.std() < 0.5 * b.std()


@pytest.mark.parametrize("roughness", [0.05, 0.3, 0.8])
def test_ggx_distribution_normalised(roughness):
    """bsdf::cooktorrance's GGX D integrates to one against cos(theta_h):
    D cos = r^2 / (pi cos^3 (r^2 + tan^2)^2), integrated over theta_h in
    closed form (substituting t = tan^2)."""
    mat = dict(type=scenes.MAT_GGX, R=(1, 1, 1), exponent=0.0, Le=(0, 0, 0), roughness=roughness,
               eta=(1.0, 1.0, 1.0), k=(0.0, 0.0, 0.0))
    n = np.array([0, 0, 1], np.float32)
    # at normal incidence H = normalize(wi + wo); pdf_proj * cos_o = D cos_h / (4 wo.H)
    wi = np.array([0, 0, 1], np.float32)
    k = 40000
    tot = 0.0
    for c in (np.arange(k) + 0.5) / k:
        x = np.array([np.sqrt(1 - c * c), 0.0, c], np.float32)
        pdf, f = orc.bsdf_eval(mat, n, wi, x)
        tot += pdf * c * (2 * np.pi / k)
    # every reflected direction of a normal-incidence ray stays above the
    # surface for H within 45 degrees; the pdf integrates to the probability
    # mass of D cos over those H, in closed form 1 / (1 + r^2)
    assert abs(tot - 1.0 / (1.0 + roughness ** 2)) < 3e-3, (tot, 1.0 / (1.0 + roughness ** 2))


@pytest.mark.parametrize("roughness", [0.1, 0.5])
def test_ggx_sampling_matches_pdf_and_is_reciprocal(roughness):
    """E[1 / p_sa] over the sampler's draws is the solid angle of the
    directions it can reach (2 pi: the whole hemisphere), which fails if the
    half-vector Jacobian 1 / (4 wo.H) is off; f is symmetric in wi, wo; and
    E[f cos / p] is the directional albedo, at most R."""
    mat = dict(type=scenes.MAT_GGX, R=(0.9, 0.9, 0.9), exponent=0.0, Le=(0, 0, 0), roughness=roughness)
    n = np.array([0, 0, 1], np.float32)
    wi = np.array([0.5, 0.1, 0.86], np.float32); wi /= np.linalg.norm(wi)
    rng = np.random.default_rng(11)
    m = 40000
    inv, alb = 0.0, np.zeros(3)
    for u in rng.random((m, 2), dtype=np.float32):
        ok, wo, pdf, f = orc.bsdf_query(mat, n, wi, u)
        if ok and pdf > 0:
            inv += 1.0 / (pdf * wo[2])
            alb += f / pdf
            if rng.random() < 0.02:
                _, f2 = orc.bsdf_eval(mat, n, wo, wi)
                np.testing.assert_allclose(f2, f, rtol=2e-4)
    assert abs(inv / m - 2 * np.pi) < 0.05 * 2 * np.pi, inv / m
    assert np.all(alb / m <= 0.9 * 1.001) and np.all(alb / m > 0.5)
