"""GPU parity of renderer::bdpt (roulette subpaths, direct light vertices,
power-heuristic or simple weights) against the oracle's literal restatement
of renderer_bdpt.cpp.

* per (sample, s, t, direct) from identical uniform tables and from the
  Philox generator: the same strategies contribute (up to shadow rays through
  an edge), C = C* w / SelectionPDF within 1e-4 relative and w within 1e-4
  (the GPU keeps the roulette compensation in the subpath throughput and
  evaluates w in O(1) from per-vertex pdf chains, the direct strategies as
  s = 1 connections of a fresh light point; the oracle rebuilds every
  strategy's path and re-evaluates all path pdfs as the reference does);
* bounded and unbounded (max_num_vertices = -1) renders;
* the furnace closed form: Le rho^(k-2) per length, Le / (1 - rho) unbounded;
* converged images agree with the plain path tracer within 3 sigma.
"""
import numpy as np
import pytest

import orc
from lmx import render, scenes
from test_bdpt_gpu import _check

pytestmark = pytest.mark.gpu


def _run(sc, n, m, gpu, seed=None, table_seed=None, batch=0, **kw):
    import torch
    rs = render.Scene(sc)
    tab = None
    if table_seed is not None:
        tab = np.random.default_rng(table_seed).random((n, render.bdpt_table_width(m)), dtype=np.float32)
    _, _, so = rs.render_bdpt(n, max_num_vertices=m, seed=seed or 42, sample_out=True, batch=batch,
                              table=None if tab is None else torch.from_numpy(tab).to(gpu), **kw)
    g = so.cpu().numpy()
    c = orc.Scene(sc).bdptref_samples(0, n, max_num_vertices=m, seed=seed or 42, table=tab, **kw)
    return g, c


@pytest.mark.parametrize("raster_mode", [render.RASTER_STRATIFIED, render.RASTER_RANDOM])
@pytest.mark.parametrize("phong", [None, 50.0])
def test_renderer_bdpt_per_strategy_parity_uniform_table(gpu, raster_mode, phong):
    sc = scenes.cornell_box(32, 32, phong_boxes=phong)
    g, c = _run(sc, 32 * 32 * 8, 8, gpu, table_seed=13, raster_mode=raster_mode, batch=4096)
    _check(g, c, min_both=3000)
    # every kind of strategy is exercised: direct (s = 0 with a light point),
    # light hits, connections, and t = 1 (light subpath to the camera)
    M = 8
    direct = c[:, 1::2, 5].sum()
    assert direct > 1000 and c[:, [2 * (t * (M + 1)) for t in range(2, M + 1)], 5].sum() > 10
    assert c[:, [2 * ((M + 1) + s) for s in range(1, M)], 5].sum() > 100


@pytest.mark.parametrize("mis", [render.MIS_POWER, render.MIS_SIMPLE])
def test_renderer_bdpt_per_strategy_parity_philox(gpu, mis):
    sc = scenes.cornell_box(32, 32, phong_boxes=20.0)
    g, c = _run(sc, 32 * 32 * 8, 9, gpu, seed=77, mis=mis, batch=3000)
    _check(g, c, min_both=3000)


def test_renderer_bdpt_unbounded_parity(gpu):
    """max_num_vertices = -1: records of 2 x 33^2 strategies per sample."""
    sc = scenes.cornell_box(16, 16)
    g, c = _run(sc, 16 * 16 * 8, -1, gpu, seed=5)
    _check(g, c, min_both=500)
    assert (c[:, :, 5] > 0).sum(axis=1).max() > 10


def test_renderer_bdpt_min_num_vertices(gpu):
    sc = scenes.cornell_box(16, 16)
    g, c = _run(sc, 16 * 16 * 8, 7, gpu, seed=9, min_num_vertices=4)
    _check(g, c, min_both=300)
    M = 7
    for t in range(M + 1):
        for s in range(M + 1 - t):
            if s + t < 4:
                assert not (g[:, 2 * (t * (M + 1) + s):2 * (t * (M + 1) + s) + 2, 5] > 0).any()


@pytest.mark.parametrize("maxv", [5, -1])
@pytest.mark.parametrize("mis", [render.MIS_POWER, render.MIS_SIMPLE])
def test_renderer_bdpt_furnace_closed_form(gpu, maxv, mis):
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=2, width=16, height=16)
    rs = render.Scene(sc)
    imgs = np.stack([rs.render_bdpt(16 * 16 * 512, max_num_vertices=maxv, seed=40 + k, mis=mis)[0]
                     .cpu().numpy().astype(np.float64).mean() for k in range(16)])
    expect = Le / (1 - rho) if maxv == -1 else Le * sum(rho ** j for j in range(maxv - 1))
    se = imgs.std(ddof=1) / np.sqrt(len(imgs))
    assert abs(imgs.mean() - expect) < max(4 * se, 1e-3 * expect), (imgs.mean(), expect, se)


def test_renderer_bdpt_and_pt_images_agree_within_3_sigma(gpu):
    sc = scenes.cornell_box(32, 32)
    rs = render.Scene(sc)

    def batches(fn, k):
        a = np.stack([fn(seed).cpu().numpy().astype(np.float64) for seed in range(k)])
        return a.mean(axis=0), a.std(axis=0, ddof=1) / np.sqrt(k)
    mb, sb = batches(lambda sd: rs.render_bdpt(32 * 32 * 256, max_num_vertices=9, seed=300 + sd)[0], 16)
    mp, sp = batches(lambda sd: rs.render_pt(32 * 32 * 4096, max_num_vertices=9, seed=500 + sd)[0], 16)
    z = np.abs(mb - mp) / np.sqrt(sb ** 2 + sp ** 2 + 1e-12)
    lit = (mb + mp) > 1e-3
    assert lit.mean() > 0.5
    assert (z[lit] <= 3.0).mean() > 0.97, (z[lit] > 3).mean()
    assert abs(mb.mean() - mp.mean()) < 0.01 * mp.mean()


def test_renderer_bdpt_per_strategy_parity_ggx(gpu):
    """bsdf::cooktorrance boxes: connections evaluate the GGX lobe from both
    sides and the direct strategies at glossy eye vertices."""
    sc = scenes.cornell_box(32, 32, ggx_boxes=0.3)
    g, c = _run(sc, 32 * 32 * 8, 8, gpu, table_seed=17, batch=4096)
    _check(g, c, min_both=3000)
