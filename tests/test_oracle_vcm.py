"""The oracle's renderer::vcm (modes vcm and bdpm) pinned by what the
mathematics fixes, on the CPU:

* the furnace closed form: inside an emitting sphere of albedo rho every
  path length k (in vertices) carries Le rho^(k-2); mode vcm sums its
  connections and merges per length, mode bdpm its merges alone.  Merging is
  consistent, not unbiased: the bias scales with the radius and the
  tessellation, so the bound is 4 standard errors plus 1.5%;
* the vanishing-radius limit: the merge pdfs carry pi r^2, so at r -> 0
  every merge weight term vanishes and mode vcm's connection records are the
  bdpt-mode records of the same uniforms (orc_bdpt_samples, itself pinned by
  its furnace and by the path tracer);
* the weights: a merge's power-heuristic weight counts every merge strategy
  N times, itself included, so N w <= 1 for every merge;
* images: on the Cornell box mode vcm agrees with the bdpt-mode image, and
  mode bdpm with its paths of 3 or more vertices (a merge needs a photon
  vertex past the light and an eye vertex past the pinhole), within noise
  plus the merge bias.
"""
import numpy as np
import pytest

import orc
from lmx import scenes


def _length_totals(rec, m, n, ks):
    out = {}
    for k in ks:
        tot = np.zeros(n)
        for t in range(1, k + 1):
            for merge in (0, 1):
                s = k - t + merge
                if s > m:
                    continue
                tot += rec[:, 2 * (t * (m + 1) + s) + merge, 2:5].astype(np.float64).mean(axis=1)
        out[k] = (tot.mean(), tot.std() / np.sqrt(n))
    return out


@pytest.mark.parametrize("mode", [orc.VCM_VCM, orc.VCM_BDPM])
def test_vcm_furnace_closed_form(mode):
    rho, Le, m = 0.5, 1.0, 4
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=3, width=16, height=16)
    n = 16 * 16 * 32
    rec = orc.Scene(sc).vcm_samples(0, n, radius=0.02, num_photons=100000, mode=mode, max_num_vertices=m,
                                    seed=3, pseed=4)
    tot = _length_totals(rec, m, n, (2, 3, 4))
    for k, (est, se) in tot.items():
        if mode == orc.VCM_BDPM and k == 2:
            assert est == 0.0                  # length 2 has no merge (the pinhole is never merged)
            continue
        want = Le * rho ** (k - 2)
        assert abs(est - want) < 4 * se + 0.015 * want, (mode, k, est, se)


def test_vcm_vanishing_radius_is_bdpt():
    sc = scenes.cornell_box(16, 16)
    n, m = 16 * 16 * 4, 6
    tab = np.random.default_rng(5).random((n, 4096), dtype=np.float32)
    o = orc.Scene(sc)
    v = o.vcm_samples(0, n, radius=1e-7, num_photons=2000, mode=orc.VCM_VCM, max_num_vertices=m, table=tab)
    b = o.bdpt_samples(0, n, max_num_vertices=m, table=tab)
    conn = v[:, 0::2, :]                       # merge == 0 records, index t (m+1) + s
    np.testing.assert_array_equal(conn[:, :, 5], b[:, :, 5])
    assert (b[:, :, 5] > 0).sum() > 5000
    np.testing.assert_allclose(conn[:, :, 6], b[:, :, 6], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(conn[:, :, 2:5], b[:, :, 2:5], rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("mode", [orc.VCM_VCM, orc.VCM_BDPM])
def test_vcm_merge_weights_bounded_by_one_over_n(mode):
    sc = scenes.cornell_box(16, 16)
    n, m, nph = 16 * 16 * 4, 6, 5000
    rec = orc.Scene(sc).vcm_samples(0, n, radius=0.05, num_photons=nph, mode=mode, max_num_vertices=m, seed=9,
                                    pseed=10)
    mg = rec[:, 1::2, :]
    cnt, wsum = mg[:, :, 5], mg[:, :, 6]
    has = cnt > 0
    assert has.sum() > 500
    assert np.all(wsum[has] * nph <= cnt[has] * (1 + 1e-5))
    if mode == orc.VCM_BDPM:
        # no connections in bdpm mode
        assert not np.any(rec[:, 0::2, 5] > 0)


def test_vcm_and_bdpm_images_match_bdpt():
    sc = scenes.cornell_box(16, 16)
    o = orc.Scene(sc)
    n, m = 16 * 16 * 256, 6
    img = {}
    for name, fn in (("bdpt", lambda sd: o.bdpt_samples(0, n, max_num_vertices=m, seed=sd)),
                     ("vcm", lambda sd: o.vcm_samples(0, n, radius=0.03, num_photons=20000, mode=orc.VCM_VCM,
                                                      max_num_vertices=m, seed=sd, pseed=sd + 1000)),
                     ("bdpm", lambda sd: o.vcm_samples(0, n, radius=0.03, num_photons=20000, mode=orc.VCM_BDPM,
                                                       max_num_vertices=m, seed=sd, pseed=sd + 1000))):
        means, longer = [], []
        for sd in range(4):
            rec = fn(100 + sd).astype(np.float64)
            means.append(rec[:, :, 2:5].sum(axis=1).mean())
            if name == "bdpt":
                # the same image without length 2 (the camera seeing the light), which no merge makes
                t = np.arange(m + 1)
                k2 = [t_ * (m + 1) + (2 - t_) for t_ in t if 1 <= t_ <= 2]
                longer.append(means[-1] - rec[:, k2, 2:5].sum(axis=1).mean())
        img[name] = (np.mean(means), np.std(means, ddof=1) / 2.0)
        if longer:
            img["bdpt3"] = (np.mean(longer), np.std(longer, ddof=1) / 2.0)
    for name, ref in (("vcm", "bdpt"), ("bdpm", "bdpt3")):
        (mv, sv), (mb, sb) = img[name], img[ref]
        assert abs(mv - mb) < 4 * np.hypot(sv, sb) + 0.03 * mb, (name, mv, mb, sv, sb)
