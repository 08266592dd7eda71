"""GPU parity of the wavefront renderer::pt against the CPU oracle.

* per-sample contributions from identical uniform tables: the same splat
  raster position, the same number of splats and C within 1e-4 relative for
  at least 99.99 % of the samples, and no sample beyond 1e-3 among those
  with the same splats (kernels and oracle share the FP32 sine, cosine and
  power of lmx_fmath.h and both round every operation as written, so the
  records agree bit for bit in practice);
* the same with Philox-generated uniforms (the generator itself is pinned
  in the CPU tests by the published Random123 known-answer vectors);
* a closed form: radiance inside an emitting, reflecting closed sphere;
* converged images: GPU and oracle renders of the Cornell box agree per
  pixel within 3 sigma.
"""
import numpy as np
import pytest

import orc
from lmx import render, scenes

pytestmark = pytest.mark.gpu


def _table(n, width, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, width), dtype=np.float32)


def _compare(gpu_s, cpu_s, rtol=1e-4, worst=1e-3):
    """per-sample records: raster x, y, C rgb, #splats.  Returns the mask of
    samples that agree within rtol; asserts the worst case among samples
    with the same splats."""
    g = gpu_s[:, :6]
    c = cpu_s
    same = (g[:, 5] == c[:, 5]) & np.all(g[:, 0:2] == c[:, 0:2], axis=1)
    rel = (np.abs(g[:, 2:5] - c[:, 2:5]) / np.maximum(np.abs(c[:, 2:5]), 1e-3)).max(axis=1)
    if same.any():
        assert rel[same].max() <= worst, f"worst relative error {rel[same].max():.3e}"
    return same & (rel <= rtol)


@pytest.mark.parametrize("raster_mode", [render.RASTER_STRATIFIED, render.RASTER_RANDOM])
def test_pt_per_sample_parity_uniform_table(gpu, raster_mode):
    import torch
    sc = scenes.cornell_box(64, 64)
    n = 64 * 64 * 16
    maxv = 9
    tab = _table(n, render.table_width(maxv), 7)
    rs = render.Scene(sc)
    film, stats, so = rs.render_pt(n, max_num_vertices=maxv, raster_mode=raster_mode,
                                   table=torch.from_numpy(tab).to(gpu), sample_out=True)
    g = so.cpu().numpy()
    c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=maxv, raster_mode=raster_mode, table=tab)
    ok = _compare(g, c)
    assert (c[:, 5] > 0).sum() > 200                     # paths that reach the light
    assert ok.mean() >= 0.9999, f"{(~ok).sum()} of {n} samples differ"
    # the film is the sum of the per-sample splats, rescaled by W H / N
    f_cpu = orc.Scene(sc).film(c) * (64 * 64 / n)
    np.testing.assert_allclose(film.cpu().numpy(), f_cpu, rtol=2e-3, atol=2e-3)


@pytest.mark.parametrize("exponent", [20.0, 400.0])
def test_pt_per_sample_parity_phong(gpu, exponent):
    """Cornell box with normalized-Phong boxes, and the displaced-sphere mesh
    with smooth vertex normals (shading-normal correction in play)."""
    sc = scenes.cornell_box(64, 64, phong_boxes=exponent)
    n = 64 * 64 * 16
    _, _, so = render.Scene(sc).render_pt(n, max_num_vertices=9, seed=77, sample_out=True)
    c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=9, seed=77)
    ok = _compare(so.cpu().numpy(), c)
    assert ok.mean() >= 0.9999, f"{(~ok).sum()} of {n} samples differ"


def test_pt_per_sample_parity_sphere_grid_small(gpu):
    sc = scenes.sphere_grid(96, 96, grid=3, subdiv=3, seed=3)
    n = 96 * 96 * 8
    _, _, so = render.Scene(sc).render_pt(n, max_num_vertices=17, rr_depth=3, seed=5, sample_out=True)
    c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=17, rr_depth=3, seed=5)
    ok = _compare(so.cpu().numpy(), c)
    assert (c[:, 5] > 0).sum() > 100
    assert ok.mean() >= 0.9999, f"{(~ok).sum()} of {n} samples differ"


@pytest.mark.parametrize("rr_mode", [render.RR_FIXED, render.RR_THROUGHPUT])
def test_pt_per_sample_parity_philox(gpu, rr_mode):
    sc = scenes.cornell_box(64, 64)
    n = 64 * 64 * 16
    rs = render.Scene(sc)
    _, _, so = rs.render_pt(n, max_num_vertices=9, seed=1234, sample_out=True, rr_mode=rr_mode, rr_depth=2)
    c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=9, seed=1234, rr_mode=rr_mode, rr_depth=2)
    ok = _compare(so.cpu().numpy(), c)
    assert ok.mean() >= 0.9999, f"{(~ok).sum()} of {n} samples differ"


@pytest.mark.parametrize("maxv,rr_depth", [(4, 1), (7, 3), (-1, 2)])
def test_pt_furnace_closed_form(gpu, maxv, rr_depth):
    rho, Le = 0.5, 1.0
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, width=32, height=32)
    spp = 512
    n = 32 * 32 * spp
    rs = render.Scene(sc)
    film, stats, _ = rs.render_pt(n, max_num_vertices=maxv, rr_depth=rr_depth, seed=5)
    img = film.cpu().numpy()
    if maxv == -1:
        expect = Le / (1 - rho)
    else:
        expect = Le * sum(rho ** j for j in range(maxv - 1))
    # the pixel mean over 1024 pixels x 512 spp is within 1% of the closed form
    assert abs(img.mean() - expect) < 0.01 * expect, (img.mean(), expect)
    assert stats["rays"] > n


def test_pt_cornell_image_within_3_sigma(gpu):
    sc = scenes.cornell_box(32, 32)
    spp = 1024
    n = 32 * 32 * spp
    rs = render.Scene(sc)
    _, _, so = rs.render_pt(n, max_num_vertices=9, seed=99, sample_out=True)
    g = so.cpu().numpy()[:, 2:5].reshape(spp, 32 * 32, 3)       # stratified: sample i -> pixel i mod WH
    c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=9, seed=4242)[:, 2:5].reshape(spp, 32 * 32, 3)
    mg, mc = g.mean(axis=0), c.mean(axis=0)
    sg, sc_ = g.var(axis=0, ddof=1) / spp, c.var(axis=0, ddof=1) / spp
    z = np.abs(mg - mc) / np.sqrt(sg + sc_ + 1e-12)
    lit = (mg + mc) > 0
    assert lit.mean() > 0.5
    assert (z[lit] <= 3.0).mean() > 0.99, (z[lit] > 3).mean()


def _compare_nee(g, c, rtol=1e-4):
    """(n, 16) records of renderer::ptdirect / ptmis: the pixel part (0-5)
    and the camera-vertex connection (8-13) must agree like _compare's."""
    ok = _compare(g[:, 0:8], c[:, 0:8], rtol)
    cam_g, cam_c = g[:, 13] > 0, c[:, 13] > 0
    same = cam_g == cam_c
    close = np.all(np.abs(g[:, 10:13] - c[:, 10:13]) <= rtol * np.maximum(np.abs(c[:, 10:13]), 1e-6), axis=1)
    ras = np.all(~cam_c[:, None] | (np.abs(g[:, 8:10] - c[:, 8:10]) <= 1e-5), axis=1)
    return ok & same & close & ras


@pytest.mark.parametrize("entry,mode", [("ptdirect", 1), ("ptmis", 2)])
@pytest.mark.parametrize("raster_mode", [render.RASTER_STRATIFIED, render.RASTER_RANDOM])
def test_nee_per_sample_parity_uniform_table(gpu, entry, mode, raster_mode):
    import torch
    sc = scenes.cornell_box(32, 32, phong_boxes=20.0)
    n, m = 32 * 32 * 16, 9
    tab = np.random.default_rng(21).random((n, render.table_width(m, nee=True)), dtype=np.float32)
    rs = render.Scene(sc)
    _, _, so = getattr(rs, "render_" + entry)(n, max_num_vertices=m, raster_mode=raster_mode,
                                              table=torch.from_numpy(tab).to(gpu), sample_out=True)
    c = orc.Scene(sc).nee_samples(0, n, mode, max_num_vertices=m, raster_mode=raster_mode, table=tab)
    assert (c[:, 5] > 0).sum() > 1000 and (c[:, 13] > 0).sum() > 10
    ok = _compare_nee(so.cpu().numpy(), c)
    assert ok.mean() > 0.998, f"{(~ok).sum()} of {n} samples differ"


@pytest.mark.parametrize("entry,mode", [("ptdirect", 1), ("ptmis", 2)])
def test_nee_per_sample_parity_philox(gpu, entry, mode):
    sc = scenes.cornell_box(64, 64)
    n = 64 * 64 * 8
    _, _, so = getattr(render.Scene(sc), "render_" + entry)(n, max_num_vertices=9, seed=77, sample_out=True,
                                                            rr_depth=2)
    c = orc.Scene(sc).nee_samples(0, n, mode, max_num_vertices=9, seed=77, rr_depth=2)
    ok = _compare_nee(so.cpu().numpy(), c)
    assert ok.mean() > 0.998, f"{(~ok).sum()} of {n} samples differ"


@pytest.mark.parametrize("entry", ["ptdirect", "ptmis"])
def test_nee_furnace_closed_form(gpu, entry):
    rho, Le, maxv = 0.5, 1.0, 6
    sc = scenes.furnace_sphere(albedo=rho, Le=Le, width=32, height=32)
    film, _, _ = getattr(render.Scene(sc), "render_" + entry)(32 * 32 * 256, max_num_vertices=maxv, rr_depth=2,
                                                              seed=9)
    expect = Le * sum(rho ** j for j in range(maxv - 1))
    assert abs(float(film.mean()) - expect) < 0.01 * expect


def test_nee_images_agree_with_pt(gpu):
    sc = scenes.cornell_box(32, 32)
    rs = render.Scene(sc)
    def batches(fn, k):
        a = np.stack([fn(s).cpu().numpy().astype(np.float64) for s in range(k)])
        return a.mean(0), a.std(0, ddof=1) / np.sqrt(k)
    mp, sp = batches(lambda s: rs.render_pt(32 * 32 * 4096, max_num_vertices=9, seed=50 + s)[0], 16)
    for entry in ("ptdirect", "ptmis"):
        mn, sn = batches(lambda s: getattr(rs, "render_" + entry)(32 * 32 * 256, max_num_vertices=9, seed=90 + s)[0], 16)
        z = np.abs(mn - mp) / np.sqrt(sn ** 2 + sp ** 2 + 1e-12)
        lit = (mn + mp) > 1e-3
        assert (z[lit] <= 3.0).mean() > 0.97, (entry, (z[lit] > 3).mean())
        assert abs(mn.mean() - mp.mean()) < 0.01 * mp.mean(), entry


@pytest.mark.parametrize("entry,mode", [("pt", 0), ("ptmis", 2)])
def test_ggx_per_sample_parity(gpu, entry, mode):
    """bsdf::cooktorrance (GGX) boxes: GPU and oracle per sample."""
    import torch
    sc = scenes.cornell_box(32, 32, ggx_boxes=0.3)
    n, m = 32 * 32 * 16, 9
    rs = render.Scene(sc)
    tab = np.random.default_rng(31).random((n, render.table_width(m, nee=mode > 0)), dtype=np.float32)
    _, _, so = getattr(rs, "render_" + entry)(n, max_num_vertices=m, table=torch.from_numpy(tab).to(gpu),
                                              sample_out=True)
    if mode == 0:
        c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=m, table=tab)
        ok = _compare(so.cpu().numpy(), c)
    else:
        c = orc.Scene(sc).nee_samples(0, n, mode, max_num_vertices=m, table=tab)
        ok = _compare_nee(so.cpu().numpy(), c)
    assert ok.mean() >= 0.9999, f"{(~ok).sum()} of {n} samples differ"
