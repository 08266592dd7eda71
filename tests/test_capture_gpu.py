"""The measurement hook that prices the samplers' traversal: ray_out copies
the rays a render traces so that the counting walkers can count their node
visits.  Every traced ray is captured exactly once (the count equals the
render's ray statistics), the path tracer's first rays are its camera rays,
and the bidirectional sampler's shadow rays are the ones the occlusion
kernel answers (re-tracing them gives the render's visibility count).
"""
import numpy as np
import pytest

from lmx import render, scenes

pytestmark = pytest.mark.gpu


def test_pt_capture_counts_every_extension_ray(gpu):
    import torch
    sc = scenes.cornell_box(64, 64)
    rs = render.Scene(sc)
    n = 64 * 64 * 4
    buf = torch.full((1 << 16, 8), float("nan"), dtype=torch.float32, device=gpu)
    _, st, _ = rs.render_pt(n, max_num_vertices=9, rr_depth=3, seed=3, ray_out=buf)
    h = buf.cpu().numpy()
    k = int((~np.isnan(h[:, 0])).sum())
    assert k == st["rays"] and st["rays"] > n
    assert np.isnan(h[k:, 0]).all()
    cam = np.asarray(sc["camera"]["eye"], np.float32)
    np.testing.assert_allclose(h[:n, 0:3], np.broadcast_to(cam, (n, 3)), atol=1e-6)
    # a capacity smaller than the render keeps the first rays only
    small = torch.full((1000, 8), float("nan"), dtype=torch.float32, device=gpu)
    rs.render_pt(n, max_num_vertices=9, rr_depth=3, seed=3, ray_out=small)
    np.testing.assert_array_equal(small.cpu().numpy(), h[:1000])


def test_bidirectional_capture_splits_extension_and_shadow_rays(gpu):
    import torch
    sc = scenes.cornell_box(32, 32)
    rs = render.Scene(sc)
    n = 32 * 32 * 2
    cap = 1 << 18
    buf = torch.full((cap, 8), float("nan"), dtype=torch.float32, device=gpu)
    _, st, _ = rs.render_vcm_bdpt(n, max_num_vertices=6, seed=5, ray_out=buf)
    h = buf.cpu().numpy()
    ext, sh = h[:cap // 2], h[cap // 2:]
    ne, ns = int((~np.isnan(ext[:, 0])).sum()), int((~np.isnan(sh[:, 0])).sum())
    assert ne > n and ns > n
    # slots of connections that carry nothing hold a ray with tmax < tmin,
    # which the statistics do not count as traced
    live = int((sh[:ns, 7] >= sh[:ns, 3]).sum())
    assert 0 < live <= ns
    assert ne + live == st["rays"]
