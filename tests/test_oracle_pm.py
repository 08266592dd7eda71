"""The oracle's renderer::pm pinned by what the mathematics fixes, on the CPU:

* furnace (emitting sphere of albedo rho, Le): the emitter hits of the eye
  path are exact -- Le at the first vertex, rho Le at the second with final
  gather -- and the density estimate carries the rest of the closed form
  sum_{k<=M} Le rho^(k-2) (Le / (1 - rho) unbounded).  The kernel estimate on
  a faceted sphere is consistent, not unbiased, so the bound is 4 standard
  errors plus 1.5 %;
* Cornell box: the density estimate per maximum path length equals the sum
  of the bidirectional oracle's paths of 4 .. M vertices (the lengths a
  final-gather photon completes), and the emitter hits equal the path
  tracer's paths of 2 and 3 vertices.
"""
import numpy as np
import pytest

import orc
from lmx import scenes


@pytest.mark.parametrize("fg,M", [(1, 5), (0, 5), (1, -1)])
def test_pm_furnace_closed_form(fg, M):
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=5, width=16, height=16)
    n = 16 * 16 * 64
    rec = orc.Scene(sc).pm_samples(0, n, radius=0.02, num_photons=200000, max_num_vertices=M, finalgather=fg,
                                   seed=3, pseed=4).astype(np.float64)
    em, de = rec[:, 2:5].mean(axis=1), rec[:, 6:9].mean(axis=1)
    # every eye path sees the emitter at vertex 1 (and 2 with final gather),
    # except where the sampled direction is tangent to the facet (a disk
    # sample on the unit circle: cos = 0 in float, so f = 0 at vertex 1)
    exact = np.abs(em - Le * (1.0 + rho * fg)) <= 1e-5 * Le * (1.0 + rho * fg)
    assert exact.mean() >= 0.999 and np.all(np.abs(em[~exact] - Le) <= 1e-5 * Le), em[~exact]
    kmax = M if M > 0 else 200
    total = sum(Le * rho ** (k - 2) for k in range(2, kmax + 1))
    want = total - Le * (1.0 + rho * fg)
    assert abs(de.mean() - want) < 4 * de.std() / np.sqrt(n) + 0.015 * want, (de.mean(), want)
    assert (rec[:, 10] == 2 + fg).all()                 # the gather vertex


def test_pm_cornell_lengths_match_bidirectional_and_pt():
    sc = scenes.cornell_box(16, 16)
    o = orc.Scene(sc)
    n, m = 16 * 16 * 256, 7
    b = o.bdpt_samples(0, n, max_num_vertices=m, seed=7).astype(np.float64)
    per_len = {k: sum(b[:, t * (m + 1) + (k - t), 2:5].sum(axis=1) for t in range(1, k + 1)) for k in range(2, m + 1)}
    for M in (5, 7):
        rec = o.pm_samples(0, n, radius=0.01, num_photons=400000, max_num_vertices=M, seed=3,
                           pseed=4).astype(np.float64)
        de = rec[:, 6:9].sum(axis=1)
        ref = sum(per_len[k] for k in range(4, M + 1))
        se = np.hypot(de.std(), ref.std()) / np.sqrt(n)
        assert abs(de.mean() - ref.mean()) < 4 * se + 0.03 * ref.mean(), (M, de.mean(), ref.mean(), se)
    # emitter hits: paths of 2 and 3 vertices, against the path tracer (the
    # light covers few pixels: many samples)
    n2 = 16 * 16 * 4096
    rec = o.pm_samples(0, n2, radius=0.01, num_photons=100, max_num_vertices=3, seed=11, pseed=4)
    pt = o.pt_samples(0, n2, max_num_vertices=3, seed=12)
    a, c = rec[:, 2:5].astype(np.float64).sum(axis=1), pt[:, 2:5].astype(np.float64).sum(axis=1)
    assert abs(a.mean() - c.mean()) < 4 * np.hypot(a.std(), c.std()) / np.sqrt(n2), (a.mean(), c.mean())
