"""The oracle's renderer::ppm and renderer::sppm pinned by what the mathematics
fixes, on the CPU:

* furnace (emitting sphere of albedo rho, Le): every measurement point sits
  on the first vertex with its emission Le exactly, and the progressive
  estimate plus it gives the closed form sum_{k<=M} Le rho^(k-2) within 4
  standard errors plus 1.5 %;
* Cornell box: the image (no emitter is a measurement point there, so paths
  of 2 vertices are missing) equals the bidirectional oracle's paths of 3 ..
  M vertices within 4 standard errors plus 2 %;
* the progressive update: radii never grow, shrink where photons were
  found, and N counts alpha of the gathered photons per pass.
"""
import numpy as np

import orc
from lmx import scenes


def test_ppm_furnace_closed_form():
    rho, Le, M = 0.5, 1.0, 5
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=5, width=16, height=16)
    n = 16 * 16 * 16
    rec = orc.Scene(sc).ppm_points(n, 8, 100000, M, initial_radius=0.05, alpha=0.7, seed=3,
                                   pseed=4).astype(np.float64)
    assert (rec[:, 13] == 1).all()
    np.testing.assert_allclose(rec[:, 5:8], Le, rtol=1e-6)
    C = rec[:, 2:5].mean(axis=1)
    want = sum(Le * rho ** (k - 2) for k in range(2, M + 1))
    assert abs(C.mean() - want) < 4 * C.std() / np.sqrt(n) + 0.015 * want, (C.mean(), want)
    # radii shrink from the initial one; N = alpha M summed over the passes
    assert (rec[:, 12] <= 0.05 * (1 + 1e-6)).all() and (rec[:, 12] < 0.05).mean() > 0.99
    assert (rec[:, 11] > 0).mean() > 0.99


def test_ppm_cornell_matches_bidirectional_lengths():
    sc = scenes.cornell_box(16, 16)
    o = orc.Scene(sc)
    n, m = 16 * 16 * 64, 6
    rec = o.ppm_points(n, 16, 100000, m, initial_radius=0.03, seed=3, pseed=4).astype(np.float64)
    b = o.bdpt_samples(0, 4 * n, max_num_vertices=m, seed=7).astype(np.float64)
    L = sum(b[:, t * (m + 1) + (k - t), 2:5].sum(axis=1) for k in range(3, m + 1) for t in range(1, k + 1))
    C = rec[:, 2:5].sum(axis=1)
    se = np.hypot(C.std() / np.sqrt(n), L.std() / np.sqrt(4 * n))
    assert abs(C.mean() - L.mean()) < 4 * se + 0.02 * L.mean(), (C.mean(), L.mean(), se)
    assert (rec[:, 5:8] == 0).all()                      # no emitter is diffuse or glossy here


def test_sppm_furnace_closed_form():
    """renderer::sppm: a point per pixel re-traced every pass; the film is
    the closed form (emission averaged over the passes: Le exactly)."""
    rho, Le, M, P = 0.5, 1.0, 5, 16
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=5, width=32, height=32)
    rec = orc.Scene(sc).sppm_pixels(P, 100000, M, initial_radius=0.05, seed=3, pseed=4).astype(np.float64)
    assert (rec[:, 11] == P).all()                       # every pass made a point in every pixel
    np.testing.assert_allclose(rec[:, 3:6] / P, Le, rtol=1e-5)
    C = rec[:, 0:3].mean(axis=1)
    want = sum(Le * rho ** (k - 2) for k in range(2, M + 1))
    assert abs(C.mean() - want) < 4 * C.std() / np.sqrt(len(C)) + 0.015 * want, (C.mean(), want)


def test_sppm_cornell_matches_bidirectional_lengths():
    sc = scenes.cornell_box(32, 32)
    o = orc.Scene(sc)
    m = 6
    rec = o.sppm_pixels(32, 100000, m, initial_radius=0.03, seed=3, pseed=4).astype(np.float64)
    n = 32 * 32 * 256
    b = o.bdpt_samples(0, n, max_num_vertices=m, seed=7).astype(np.float64)
    L = sum(b[:, t * (m + 1) + (k - t), 2:5].sum(axis=1) for k in range(3, m + 1) for t in range(1, k + 1))
    C = rec[:, 0:3].sum(axis=1)
    se = np.hypot(C.std() / np.sqrt(len(C)), L.std() / np.sqrt(n))
    assert abs(C.mean() - L.mean()) < 4 * se + 0.02 * L.mean(), (C.mean(), L.mean(), se)
