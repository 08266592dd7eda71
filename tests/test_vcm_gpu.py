"""GPU parity of renderer::vcm with vertex merging (modes vcm and bdpm)
against the oracle's literal restatement (one pass, brute-force weights):

* per (sample, t, s, connect / merge) from identical eye and photon uniform
  tables: the same connections and merges contribute (up to shadow rays and
  merge-radius tests that an ulp of a vertex position decides), connection
  C and w and merge sums of C and of weights within 1e-4 relative for at
  least 99.99 % of the records (counts equal) with bounded worst cases
  (Lambertian, Phong and Cook-Torrance boxes);
* the furnace closed form of both modes per path length;
* a multi-pass render with the radius schedule agrees with the bdpt-mode
  image (both estimate the same image; merging's bias shrinks with r).
"""
import numpy as np
import pytest

import orc
from lmx import render, scenes

pytestmark = pytest.mark.gpu


def _records(g, c, merge, floor=1e-9):
    g, c = g[:, merge::2, :], c[:, merge::2, :]
    fl_g = (g[:, :, 5] > 0) & (np.abs(g[:, :, 2:5]).max(axis=2) > floor)
    fl_c = (c[:, :, 5] > 0) & (np.abs(c[:, :, 2:5]).max(axis=2) > floor)
    one = int((fl_g != fl_c).sum())
    both = fl_g & fl_c
    cg, cc = g[both][:, 2:5].astype(np.float64), c[both][:, 2:5].astype(np.float64)
    rel = (np.abs(cg - cc) / np.maximum(np.abs(cc), floor)).max(axis=1)
    wg, wc = g[both][:, 6].astype(np.float64), c[both][:, 6].astype(np.float64)
    relw = np.abs(wg - wc) / np.maximum(np.abs(wc), 1e-30)
    cnt = g[both][:, 5] != c[both][:, 5]
    # a record whose photon count differs (a photon within an ulp of the
    # radius) is counted, not compared
    return one, int(both.sum()), np.maximum(rel, relw)[~cnt], int(cnt.sum())


def _run(sc, n, m, nph, r, mode, tab, ptab, gpu):
    import torch
    _, _, so = render.Scene(sc).render_vcm(n, mode=mode, num_iteration_pass=1, num_photon_trace_samples=nph,
                                           initial_radius=r, max_num_vertices=m,
                                           table=torch.from_numpy(tab).to(gpu),
                                           photon_table=torch.from_numpy(ptab).to(gpu), sample_out=True,
                                           batch=2048)
    c = orc.Scene(sc).vcm_samples(0, n, radius=r, num_photons=nph, mode=mode, max_num_vertices=m, table=tab,
                                  ptable=ptab)
    return so.cpu().numpy(), c


@pytest.mark.parametrize("mode", [render.VCM_MODE_VCM, render.VCM_MODE_BDPM])
@pytest.mark.parametrize("boxes", [None, "phong", "ggx"])
def test_vcm_per_record_parity_uniform_tables(gpu, mode, boxes):
    sc = scenes.cornell_box(16, 16, phong_boxes=30.0 if boxes == "phong" else None,
                            ggx_boxes=0.3 if boxes == "ggx" else None)
    n, m, nph, r = 16 * 16 * 16, 6, 20000, 0.04
    w = render.vcm_bdpt_table_width(m)
    rng = np.random.default_rng(21)
    tab = rng.random((n, w), dtype=np.float32)
    ptab = rng.random((nph, w), dtype=np.float32)
    g, c = _run(sc, n, m, nph, r, mode, tab, ptab, gpu)
    # merges: the same (sample, t, s) receive merges, with the same count of
    # photons, contribution sum and weight sum
    one, both, err, cnt = _records(g, c, 1)
    assert both > 2000, both
    assert one <= 2e-3 * both, f"{one} of {both} merge records on one side only"
    assert cnt <= 2e-3 * both, f"{cnt} of {both} merge records with another photon count"
    print(f"merge records {both}: 99.99 % quantile {np.quantile(err, 0.9999):.2e}, worst {err.max():.2e}")
    assert (err <= 1e-4).mean() >= 0.9999, f"{(err > 1e-4).sum()} of {both} merge records beyond 1e-4"
    assert err.max() <= 1e-3, f"worst merge record error {err.max():.3e}"
    if mode == render.VCM_MODE_BDPM:
        assert not np.any(g[:, 0::2, 5] > 0)
        return
    one, both, err, _ = _records(g, c, 0)
    assert both > 5000
    assert one <= 5e-4 * both, f"{one} of {both} connections on one side only"
    print(f"connections {both}: 99.99 % quantile {np.quantile(err, 0.9999):.2e}, worst {err.max():.2e}")
    assert (err <= 1e-4).mean() >= 0.9999, f"{(err > 1e-4).sum()} of {both} connections beyond 1e-4"
    assert err.max() <= 5e-4, f"worst connection error {err.max():.3e}"


@pytest.mark.parametrize("mode", [render.VCM_MODE_VCM, render.VCM_MODE_BDPM])
def test_vcm_furnace_closed_form(gpu, mode):
    rho, Le, m = 0.5, 1.0, 4
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=3, width=16, height=16)
    n = 16 * 16 * 256
    _, _, so = render.Scene(sc).render_vcm(n, mode=mode, num_iteration_pass=1, num_photon_trace_samples=200000,
                                           initial_radius=0.02, max_num_vertices=m, seed=5, sample_out=True)
    rec = so.cpu().numpy().astype(np.float64)
    for k in (2, 3, 4):
        tot = np.zeros(n)
        for t in range(1, k + 1):
            for merge in (0, 1):
                s = k - t + merge
                if s <= m:
                    tot += rec[:, 2 * (t * (m + 1) + s) + merge, 2:5].mean(axis=1)
        est, se = tot.mean(), tot.std() / np.sqrt(n)
        if mode == render.VCM_MODE_BDPM and k == 2:
            assert est == 0.0
            continue
        want = Le * rho ** (k - 2)
        assert abs(est - want) < 4 * se + 0.015 * want, (k, est, se)


def test_vcm_passes_match_bdpt_image(gpu):
    sc = scenes.cornell_box(32, 32, phong_boxes=30.0)
    rs = render.Scene(sc)
    n = 32 * 32 * 64
    fv, st, _ = rs.render_vcm(n, num_iteration_pass=8, num_photon_trace_samples=100000, initial_radius=0.05,
                              max_num_vertices=8, seed=3)
    assert st["passes"] == 8
    imgs = [rs.render_vcm_bdpt(n * 8, max_num_vertices=8, seed=10 + k)[0].cpu().numpy() for k in range(4)]
    mb = np.mean([im.mean() for im in imgs])
    sb = np.std([im.mean() for im in imgs], ddof=1)
    mv = fv.cpu().numpy().mean()
    assert abs(mv - mb) < 4 * sb + 0.02 * mb, (mv, mb, sb)


def test_vcm_argument_errors(gpu):
    import torch
    from lmx._lib import LmxError
    sc = scenes.cornell_box(8, 8)
    rs = render.Scene(sc)
    n, m = 64, 5
    tab = torch.zeros((n, render.vcm_bdpt_table_width(m)), dtype=torch.float32, device=gpu)
    with pytest.raises(LmxError, match="num_iteration_pass 1"):
        rs.render_vcm(n, num_iteration_pass=2, max_num_vertices=m, table=tab)
    with pytest.raises(LmxError, match="unknown mode"):
        rs.render_vcm(n, mode=1, num_iteration_pass=1, max_num_vertices=m)
    with pytest.raises(LmxError, match="num_photon_trace_samples"):
        rs.render_vcm(n, num_iteration_pass=1, num_photon_trace_samples=0, max_num_vertices=m)
    with pytest.raises(LmxError, match="initial_radius"):
        rs.render_vcm(n, num_iteration_pass=1, initial_radius=0.0, max_num_vertices=m)
    with pytest.raises(LmxError, match="max_num_vertices"):
        rs.render_vcm(n, num_iteration_pass=1, max_num_vertices=-1)
    # a light-free scene renders black
    dark = scenes.cornell_box(8, 8)
    dark["lights"] = []
    f, _, _ = render.Scene(dark).render_vcm(n, num_iteration_pass=2, num_photon_trace_samples=100,
                                            max_num_vertices=m)
    assert float(f.abs().max()) == 0.0
