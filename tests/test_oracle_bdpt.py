"""The oracle's bidirectional sampler (renderer::vcm, bdpt mode), pinned by
closed forms and by its own unbiased parts.

* Inside an emitting, reflecting closed sphere every path of k vertices
  carries Le rho^(k-2).  Each (s, t) strategy alone, unweighted (C / w), is an
  unbiased estimator of that value, so a dropped geometry term, a pdf of the
  wrong side or a misplaced emission factor in any one strategy fails; the
  MIS-weighted sum over the strategies of one length hits it only if the
  weights of every path sum to one.
* In the Cornell box the unweighted strategies of one length agree with each
  other, and the whole image agrees with the oracle's plain path tracer.
"""
import numpy as np
import pytest

import orc
from lmx import scenes


def _strategy(rec, m, s, t):
    """Per-sample weighted C, unweighted C / w (zero where absent)."""
    r = rec[:, t * (m + 1) + s].astype(np.float64)
    on = r[:, 5] > 0
    c = r[:, 2:5].mean(axis=1)
    un = np.where(on, c / np.where(on, r[:, 6], 1.0), 0.0)
    return c, un


@pytest.mark.parametrize("maxv", [4, 5])
def test_bdpt_furnace_every_strategy_unbiased(maxv):
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=2, width=16, height=16)
    n = 16 * 16 * 256
    rec = orc.Scene(sc).bdpt_samples(0, n, max_num_vertices=maxv, seed=9)
    m = maxv
    for k in range(2, maxv + 1):
        expect = Le * rho ** (k - 2)
        tot = np.zeros(n)
        for t in range(1, k + 1):
            s = k - t
            c, un = _strategy(rec, m, s, t)
            tot += c
            se = un.std() / np.sqrt(n)
            assert abs(un.mean() - expect) < max(4 * se, 1e-3 * expect), (k, s, t, un.mean(), se)
        assert abs(tot.mean() - expect) < 0.01 * expect, (k, tot.mean())


def test_bdpt_cornell_strategies_agree_and_match_pt():
    sc = scenes.cornell_box(32, 32)
    o = orc.Scene(sc)
    m, n = 9, 32 * 32 * 128
    rec = o.bdpt_samples(0, n, max_num_vertices=m, seed=100)
    # strategies of one length: unweighted estimates agree within 4 sigma
    for k in (3, 4):
        vals = []
        for t in range(1, k + 1):
            _, un = _strategy(rec, m, k - t, t)
            vals.append((un.mean(), un.std() / np.sqrt(n)))
        best = min(vals, key=lambda v: v[1])
        for mu, se in vals:
            assert abs(mu - best[0]) < 4 * np.hypot(se, best[1]), (k, vals)
    b = rec[:, :, 2:5].astype(np.float64).sum(axis=1).mean(axis=1)
    del rec
    npt = 32 * 32 * 2048
    p = o.pt_samples(0, npt, max_num_vertices=m, seed=200)[:, 2:5].astype(np.float64).mean(axis=1)
    se = np.hypot(b.std() / np.sqrt(n), p.std() / np.sqrt(npt))
    assert abs(b.mean() - p.mean()) < 3 * se, (b.mean(), p.mean(), se)


# ---------------------------------------------------------------------------
# renderer::bdpt (roulette subpaths, direct light vertices, power or simple
# weights): the same closed form per strategy, with the direct strategies
# (s = 0 with a light point sampled from eye vertex t - 2) among them.  Each
# strategy alone, unweighted and divided by its selection probability, is
# unbiased, so the roulette compensation (1 / SelectionPDF) and the direct
# strategy's pdf conversion are pinned per strategy; the weighted sum per
# length pins that the power-heuristic (or simple) weights of every path sum
# to one over all strategies, the direct ones included.

def _ref_strategy(rec, M, s, t, d):
    r = rec[:, 2 * (t * (M + 1) + s) + d].astype(np.float64)
    on = r[:, 5] > 0
    c = r[:, 2:5].mean(axis=1)
    un = np.where(on, c / np.where(on, r[:, 6], 1.0), 0.0)
    return c, un, on


@pytest.mark.parametrize("mis", [0, 1])
def test_renderer_bdpt_furnace_every_strategy_unbiased(mis):
    rho, Le, maxv = 0.5, 1.0, 5
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=2, width=16, height=16)
    n = 16 * 16 * 512
    rec = orc.Scene(sc).bdptref_samples(0, n, max_num_vertices=maxv, seed=21, mis=mis)
    M = maxv
    for k in range(2, maxv + 1):
        expect = Le * rho ** (k - 2)
        tot = np.zeros(n)
        seen = 0
        for s in range(0, k + 1):
            for d in (0, 1):
                t = k - s
                if s > 0 and t > 0 and d:
                    continue
                c, un, on = _ref_strategy(rec, M, s, t, d)
                tot += c
                if t == 0:
                    assert not on.any()                      # pinhole: t = 0 never contributes
                    continue
                seen += 1
                se = un.std() / np.sqrt(n)
                assert abs(un.mean() - expect) < max(4 * se, 1e-3 * expect), (k, s, t, d, un.mean(), se)
        assert seen == k + 1                                 # s = 0..k-1 and the direct strategy
        se = tot.std() / np.sqrt(n)
        assert abs(tot.mean() - expect) < max(4 * se, 2e-3 * expect), (k, tot.mean(), se)


@pytest.mark.parametrize("mis", [0, 1])
def test_renderer_bdpt_unbounded_furnace_closed_form(mis):
    """max_num_vertices = -1: every length, the roulette alone ends the
    subpaths; the image is Le / (1 - rho)."""
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=2, width=16, height=16)
    n = 16 * 16 * 512
    rec = orc.Scene(sc).bdptref_samples(0, n, max_num_vertices=-1, seed=5, mis=mis, nthreads=8)
    tot = rec[:, :, 2:5].astype(np.float64).sum(axis=1).mean(axis=1)
    del rec
    se = tot.std() / np.sqrt(n)
    assert abs(tot.mean() - Le / (1 - rho)) < 4 * se, (tot.mean(), se)


def test_renderer_bdpt_min_num_vertices_only_when_bounded():
    sc = scenes.cornell_box(16, 16)
    o = orc.Scene(sc)
    n = 16 * 16 * 8
    rec = o.bdptref_samples(0, n, max_num_vertices=6, min_num_vertices=4, seed=3)
    M = 6
    for t in range(M + 1):
        for s in range(M + 1 - t):
            for d in (0, 1):
                on = rec[:, 2 * (t * (M + 1) + s) + d, 5] > 0
                if s + t < 4:
                    assert not on.any(), (s, t, d)
    # unbounded: min_num_vertices is ignored (renderer_bdpt.cpp:887)
    a = o.bdptref_samples(0, n, max_num_vertices=-1, min_num_vertices=4, seed=3)
    b = o.bdptref_samples(0, n, max_num_vertices=-1, min_num_vertices=0, seed=3)
    np.testing.assert_array_equal(a, b)


def test_renderer_bdpt_cornell_matches_pt():
    sc = scenes.cornell_box(32, 32)
    o = orc.Scene(sc)
    m, n = 9, 32 * 32 * 128
    rec = o.bdptref_samples(0, n, max_num_vertices=m, seed=100)
    b = rec[:, :, 2:5].astype(np.float64).sum(axis=1).mean(axis=1)
    del rec
    npt = 32 * 32 * 2048
    p = o.pt_samples(0, npt, max_num_vertices=m, seed=200)[:, 2:5].astype(np.float64).mean(axis=1)
    se = np.hypot(b.std() / np.sqrt(n), p.std() / np.sqrt(npt))
    assert abs(b.mean() - p.mean()) < 3 * se, (b.mean(), p.mean(), se)
