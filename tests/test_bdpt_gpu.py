"""GPU parity of the batched bidirectional sampler (renderer::vcm, bdpt mode)
against the oracle's literal restatement of the reference.

* per (sample, s, t) from identical uniform tables: the same strategies
  contribute (up to shadow rays through an edge), C = f w / p within 1e-4
  relative and the MIS weight w within 1e-4 (the GPU evaluates w in O(n)
  from stored per-vertex pdfs; the oracle re-evaluates every strategy's path
  pdf as the reference does);
* the furnace closed form: inside an emitting, reflecting sphere every path
  length k contributes Le rho^(k-2), summed over its strategies;
* converged images: BDPT and the plain path tracer agree per pixel within
  3 sigma (both are unbiased estimators of the same image).
"""
import numpy as np
import pytest

import orc
from lmx import render, scenes

pytestmark = pytest.mark.gpu


def _connections(g, c, floor=1e-9):
    """Per-connection agreement of two (n, (m+1)^2, 8) record sets.

    Returns (present in exactly one, present in both, relative error of C
    per common connection, absolute error of w per common connection).
    Contributions below `floor` count as absent: the reference form f w / p
    multiplies ~2n FP32 factors and underflows long, dim paths to zero where
    the throughput form used on the GPU keeps a denormal-sized value."""
    fl_g = (g[:, :, 5] > 0) & (np.abs(g[:, :, 2:5]).max(axis=2) > floor)
    fl_c = (c[:, :, 5] > 0) & (np.abs(c[:, :, 2:5]).max(axis=2) > floor)
    one = int((fl_g != fl_c).sum())
    both = fl_g & fl_c
    cg, cc = g[both][:, 2:5].astype(np.float64), c[both][:, 2:5].astype(np.float64)
    rel = (np.abs(cg - cc) / np.maximum(np.abs(cc), floor)).max(axis=1)
    dw = np.abs(g[both][:, 6] - c[both][:, 6])
    return one, int(both.sum()), rel, dw


def _check(g, c, worst=5e-4, edge_frac=5e-4, min_both=5000):
    """Same strategies contribute, except for shadow rays that graze an edge
    (a visibility decided by the last ulp of a vertex position, at most
    `edge_frac` of the connections); contributions and MIS weights agree
    within 1e-4 for at least 99.99 % of the common connections and within
    `worst` for all of them.  The vertices, BSDFs and contributions are
    computed with the same FP32 operations on both sides (shared sine,
    cosine and power); what differs is the MIS weight, an O(1) ratio chain in
    FP32 on the GPU against the oracle's re-evaluation of every strategy's
    path pdf, whose rounding differences grow along long glossy chains."""
    one, both, rel, dw = _connections(g, c)
    assert both > min_both
    assert one <= edge_frac * both, f"{one} of {both} connections contribute on one side only"
    err = np.maximum(rel, dw)
    print(f"connections {both}: 99.99 % quantile {np.quantile(err, 0.9999):.2e}, worst {err.max():.2e}")
    assert (err <= 1e-4).mean() >= 0.9999, f"{(err > 1e-4).sum()} of {both} beyond 1e-4"
    assert err.max() <= worst, f"worst error {err.max():.3e}"


@pytest.mark.parametrize("raster_mode", [render.RASTER_STRATIFIED, render.RASTER_RANDOM])
@pytest.mark.parametrize("phong", [None, 50.0])
def test_bdpt_per_strategy_parity_uniform_table(gpu, raster_mode, phong):
    import torch
    sc = scenes.cornell_box(32, 32, phong_boxes=phong)
    n, m = 32 * 32 * 8, 8
    tab = np.random.default_rng(11).random((n, render.vcm_bdpt_table_width(m)), dtype=np.float32)
    rs = render.Scene(sc)
    film, stats, so = rs.render_vcm_bdpt(n, max_num_vertices=m, raster_mode=raster_mode,
                                         table=torch.from_numpy(tab).to(gpu), sample_out=True, batch=4096)
    g = so.cpu().numpy()
    c = orc.Scene(sc).bdpt_samples(0, n, max_num_vertices=m, raster_mode=raster_mode, table=tab)
    _check(g, c)
    # raster positions of light-tracing (t = 1) contributions
    t1 = (c[:, m + 1:2 * (m + 1), 5] > 0) & (g[:, m + 1:2 * (m + 1), 5] > 0)
    np.testing.assert_allclose(g[:, m + 1:2 * (m + 1), 0:2][t1], c[:, m + 1:2 * (m + 1), 0:2][t1], atol=1e-4)


def test_bdpt_per_strategy_parity_philox(gpu):
    sc = scenes.sphere_grid(48, 48, grid=3, subdiv=3, seed=3)
    n, m = 48 * 48 * 4, 10
    _, _, so = render.Scene(sc).render_vcm_bdpt(n, max_num_vertices=m, seed=7, sample_out=True, batch=3000)
    c = orc.Scene(sc).bdpt_samples(0, n, max_num_vertices=m, seed=7)
    _check(so.cpu().numpy(), c, min_both=1000)


def test_bdpt_connection_rounds_parity(gpu):
    """A closed box at max_num_vertices 12: nearly every sample holds all 67
    of its connections, more than the 24 shadow-queue slots per sample of
    one connection round, so each batch needs the graph's two rounds and
    more from the host; every connection must still be made exactly once."""
    import torch
    sc = scenes.cornell_box(32, 32)
    n, m = 32 * 32 * 2, 12
    tab = np.random.default_rng(23).random((n, render.vcm_bdpt_table_width(m)), dtype=np.float32)
    _, st, so = render.Scene(sc).render_vcm_bdpt(n, max_num_vertices=m, table=torch.from_numpy(tab).to(gpu),
                                                  sample_out=True, batch=1024)
    c = orc.Scene(sc).bdpt_samples(0, n, max_num_vertices=m, table=tab)
    _check(so.cpu().numpy(), c, min_both=20000)


@pytest.mark.parametrize("maxv", [3, 5, 12])
def test_bdpt_furnace_closed_form(gpu, maxv):
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=2, width=16, height=16)
    n = 16 * 16 * 1024
    _, _, so = render.Scene(sc).render_vcm_bdpt(n, max_num_vertices=maxv, seed=3, sample_out=True)
    rec = so.cpu().numpy().astype(np.float64)
    m = maxv
    for k in range(2, maxv + 1):               # path length in vertices
        tot = 0.0
        for t in range(1, k + 1):
            s = k - t
            tot += rec[:, t * (m + 1) + s, 2:5].sum() / 3.0
        est = tot / n                          # W H / N film scale, averaged over the W H pixels
        assert abs(est - Le * rho ** (k - 2)) < 0.02 * rho ** (k - 2), (k, est)


def test_bdpt_and_pt_images_agree_within_3_sigma(gpu):
    sc = scenes.cornell_box(32, 32)
    rs = render.Scene(sc)
    # per-pixel means and standard errors from 16 independent batches each;
    # the path tracer reaches the small light rarely, so its batches hold
    # 16x more samples than the bidirectional ones (4096 vs 256 spp), enough
    # for batch means close to normal
    def batches(fn, k):
        imgs = [fn(seed).cpu().numpy().astype(np.float64) for seed in range(k)]
        a = np.stack(imgs)
        return a.mean(axis=0), a.std(axis=0, ddof=1) / np.sqrt(k)
    mb, sb = batches(lambda sd: rs.render_vcm_bdpt(32 * 32 * 256, max_num_vertices=9, seed=100 + sd)[0], 16)
    mp, sp = batches(lambda sd: rs.render_pt(32 * 32 * 4096, max_num_vertices=9, seed=200 + sd)[0], 16)
    z = np.abs(mb - mp) / np.sqrt(sb ** 2 + sp ** 2 + 1e-12)
    lit = (mb + mp) > 1e-3
    assert lit.mean() > 0.5
    assert (z[lit] <= 3.0).mean() > 0.97, (z[lit] > 3).mean()
    # and the image means agree to 1%
    assert abs(mb.mean() - mp.mean()) < 0.01 * mp.mean()


def test_bdpt_per_strategy_parity_ggx(gpu):
    import torch
    sc = scenes.cornell_box(32, 32, ggx_boxes=0.3)
    n, m = 32 * 32 * 8, 8
    tab = np.random.default_rng(12).random((n, render.vcm_bdpt_table_width(m)), dtype=np.float32)
    _, _, so = render.Scene(sc).render_vcm_bdpt(n, max_num_vertices=m, table=torch.from_numpy(tab).to(gpu),
                                                sample_out=True, batch=4096)
    c = orc.Scene(sc).bdpt_samples(0, n, max_num_vertices=m, table=tab)
    _check(so.cpu().numpy(), c)
