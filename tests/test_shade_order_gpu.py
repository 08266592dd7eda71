"""Material-binned shading (shade_order = LMX_SHADE_BY_MATERIAL): the traced
paths are regrouped by the BSDF of the surface they hit before shading.  It
changes which lane shades which path, never what a path computes: every
per-sample record equals the queue-order render's bit for bit, and the films
agree up to the order of their float additions.  Scenes with two and three
materials in one launch (Lambertian with Phong or Cook-Torrance boxes, the
sphere grid's 70/30 Lambertian/Phong mix), with and without next-event
estimation.
"""
import numpy as np
import pytest

from lmx import render, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["phong", "ggx", "grid", "ptmis"])
def test_binned_shading_equals_queue_order(gpu, case):
    if case == "grid":
        sc = scenes.sphere_grid(96, 96, grid=3, subdiv=3, seed=3)
    else:
        sc = scenes.cornell_box(64, 64, phong_boxes=50.0 if case != "ggx" else None,
                                ggx_boxes=0.3 if case == "ggx" else None)
    rs = render.Scene(sc)
    w, h = sc["camera"]["width"], sc["camera"]["height"]
    n = w * h * 16
    fn = rs.render_ptmis if case == "ptmis" else rs.render_pt
    kw = dict(max_num_vertices=12, rr_depth=3, seed=31, sample_out=True, queue_capacity=1 << 16)
    fa, sa, ra = fn(n, shade_order=render.SHADE_QUEUE, **kw)
    fb, sb, rb = fn(n, shade_order=render.SHADE_BY_MATERIAL, **kw)
    ra, rb = ra.cpu().numpy(), rb.cpu().numpy()
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert sa["rays"] == sb["rays"]
    a, b = fa.cpu().numpy().astype(np.float64), fb.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-7 * max(a.max(), 1e-30))
    assert (ra[:, 5] > 0).sum() > 100


def test_unknown_shade_order_is_refused(gpu):
    from lmx._lib import LmxError
    rs = render.Scene(scenes.cornell_box(8, 8))
    with pytest.raises(LmxError, match="shade_order"):
        rs.render_pt(64, shade_order=7)
