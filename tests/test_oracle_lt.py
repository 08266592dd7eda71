"""The oracle's renderer::ltdirect pinned by what the mathematics fixes, on
the CPU:

* furnace (emitting sphere of albedo rho, Le): the connection of light
  vertex k to the camera carries the path length k + 2, so its mean is
  Le rho^k, zero past max_num_vertices, and the image the closed form;
* Cornell box: the image equals the bidirectional oracle's sum over every
  path length (both are unbiased estimators of the same image).
"""
import numpy as np

import orc
from lmx import scenes


def test_ltdirect_furnace_per_vertex():
    rho, Le, M = 0.5, 1.0, 5
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=3, width=16, height=16)
    n = 1 << 18
    rec = orc.Scene(sc).ltdirect_samples(0, n, max_num_vertices=M, seed=3).astype(np.float64)
    for k in range(M):
        c = rec[:, k, 2:5].mean(axis=1)
        want = Le * rho ** k if k + 2 <= M else 0.0
        assert abs(c.mean() - want) <= 4 * c.std() / np.sqrt(n) + 1e-3 * want, (k, c.mean(), want)
    assert not rec[:, M - 1, 5].any()


def test_ltdirect_cornell_matches_bidirectional():
    sc = scenes.cornell_box(16, 16)
    o = orc.Scene(sc)
    m, n = 6, 1 << 18
    rec = o.ltdirect_samples(0, n, max_num_vertices=m, seed=3).astype(np.float64)
    b = o.bdpt_samples(0, 16 * 16 * 256, max_num_vertices=m, seed=7).astype(np.float64)
    L = sum(b[:, t * (m + 1) + (k - t), 2:5].sum(axis=1) for k in range(2, m + 1) for t in range(1, k + 1))
    C = rec[:, :, 2:5].sum(axis=(1, 2))
    se = np.hypot(C.std() / np.sqrt(n), L.std() / np.sqrt(len(L)))
    assert abs(C.mean() - L.mean()) < 4 * se + 0.01 * L.mean(), (C.mean(), L.mean(), se)
