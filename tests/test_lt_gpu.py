"""GPU parity of renderer::ltdirect against the oracle's literal
restatement:

* per (sample, light vertex) from identical uniform tables: the same camera
  connections contribute (up to shadow rays through an edge), at the same
  raster positions, with C within 1e-4 relative, bounded and unbounded,
  Lambertian, Phong and Cook-Torrance boxes;
* the furnace per light vertex from Philox streams;
* the image against the bidirectional sampler's.
"""
import numpy as np
import pytest

import orc
from lmx import render, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,boxes", [(6, None), (-1, None), (6, "phong"), (6, "ggx")])
def test_ltdirect_per_connection_parity_uniform_table(gpu, M, boxes):
    import torch
    sc = scenes.cornell_box(16, 16, phong_boxes=30.0 if boxes == "phong" else None,
                            ggx_boxes=0.3 if boxes == "ggx" else None)
    n = 1 << 15
    w = render.table_width(32 if M == -1 else M, nee=True)
    tab = np.random.default_rng(51).random((n, w), dtype=np.float32)
    _, _, so = render.Scene(sc).render_ltdirect(n, max_num_vertices=M, table=torch.from_numpy(tab).to(gpu),
                                                sample_out=True, batch=8192)
    g = so.cpu().numpy().astype(np.float64)
    c = orc.Scene(sc).ltdirect_samples(0, n, max_num_vertices=M, table=tab).astype(np.float64)
    on_g, on_c = g[:, :, 5] > 0, c[:, :, 5] > 0
    both = on_g & on_c
    assert both.sum() > 5000
    assert (on_g != on_c).sum() <= 5e-4 * both.sum() + 2
    np.testing.assert_allclose(g[both][:, 0:2], c[both][:, 0:2], atol=1e-4)
    rel = (np.abs(g[both][:, 2:5] - c[both][:, 2:5]) / np.maximum(np.abs(c[both][:, 2:5]), 1e-12)).max(axis=1)
    assert (rel <= 1e-4).mean() >= 0.9999, (rel > 1e-4).sum()
    assert rel.max() <= 1e-3, rel.max()


def test_ltdirect_furnace_per_vertex(gpu):
    rho, Le, M = 0.5, 1.0, 5
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, subdiv=3, width=16, height=16)
    n = 1 << 20
    _, _, so = render.Scene(sc).render_ltdirect(n, max_num_vertices=M, seed=5, sample_out=True)
    rec = so.cpu().numpy().astype(np.float64)
    for k in range(M):
        c = rec[:, k, 2:5].mean(axis=1)
        want = Le * rho ** k if k + 2 <= M else 0.0
        assert abs(c.mean() - want) <= 4 * c.std() / np.sqrt(n) + 1e-3 * want, (k, c.mean(), want)


def test_ltdirect_image_matches_bdpt(gpu):
    sc = scenes.cornell_box(32, 32)
    rs = render.Scene(sc)
    lt = [rs.render_ltdirect(1 << 22, max_num_vertices=7, seed=20 + k)[0].cpu().numpy().mean() for k in range(4)]
    bd = [rs.render_vcm_bdpt(32 * 32 * 1024, max_num_vertices=7, seed=30 + k)[0].cpu().numpy().mean()
          for k in range(4)]
    se = np.hypot(np.std(lt, ddof=1), np.std(bd, ddof=1)) / 2.0
    assert abs(np.mean(lt) - np.mean(bd)) < 4 * se + 0.01 * np.mean(bd), (np.mean(lt), np.mean(bd), se)


def test_ltdirect_argument_errors(gpu):
    import torch
    from lmx._lib import LmxError
    rs = render.Scene(scenes.cornell_box(8, 8))
    with pytest.raises(LmxError, match="max_num_vertices"):
        rs.render_ltdirect(64, max_num_vertices=1)
    narrow = torch.zeros((64, 9), dtype=torch.float32, device=gpu)
    with pytest.raises(LmxError, match="narrower"):
        rs.render_ltdirect(64, max_num_vertices=4, table=narrow)
