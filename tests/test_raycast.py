"""renderer::raycast: the oracle pinned by a closed form on the CPU, the GPU
kernel against the oracle per pixel.

A quad facing the camera along its axis: the ray through pixel centre
(rx, ry) of a pinhole with vertical field of view fov and aspect a has
direction ((2 rx - 1) a tan(fov/2), (2 ry - 1) tan(fov/2), -1) normalised,
so the image is R |d_z| inside the quad and black outside."""
import numpy as np
import pytest

import orc
from lmx import render, scenes


def _quad_scene(w=24, h=16, R=(0.2, 0.4, 0.6), half=0.6):
    v = np.array([[-half, -half, -2.0], [half, -half, -2.0], [half, half, -2.0], [-half, half, -2.0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    mats = [dict(type=scenes.MAT_DIFFUSE, R=R, exponent=0.0, Le=(0, 0, 0))]
    return dict(verts=v, tris=f, normals=None, tri_mat=np.zeros(2, np.uint32), materials=mats, lights=[],
                camera=dict(eye=(0.0, 0.0, 0.0), center=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov_deg=45.0,
                            width=w, height=h))


def test_raycast_oracle_closed_form():
    sc = _quad_scene()
    img = orc.Scene(sc).raycast().astype(np.float64)
    h, w = img.shape[:2]
    t = np.tan(np.radians(45.0) / 2)
    y, x = np.mgrid[0:h, 0:w]
    dx = (2 * (x + 0.5) / w - 1) * (w / h) * t
    dy = (2 * (y + 0.5) / h - 1) * t
    dz = 1.0 / np.sqrt(dx * dx + dy * dy + 1.0)
    inside = (np.abs(2.0 * dx) < 0.6) & (np.abs(2.0 * dy) < 0.6)          # the hit point at z = -2
    assert inside.sum() > 20 and (~inside).sum() > 20
    want = dz[..., None] * np.array([0.2, 0.4, 0.6])
    np.testing.assert_allclose(img[inside], want[inside], rtol=1e-5)
    assert (img[~inside] == 0).all()


@pytest.mark.gpu
def test_raycast_gpu_matches_oracle(gpu):
    for sc in (scenes.cornell_box(64, 48), _quad_scene(), scenes.cornell_box(32, 32, ggx_boxes=0.3)):
        g, st = render.Scene(sc).render_raycast()
        c = orc.Scene(sc).raycast()
        g = g.cpu().numpy()
        assert st["rays"] == g.shape[0] * g.shape[1]
        close = np.isclose(g, c, rtol=1e-5, atol=1e-7).all(axis=2)
        assert close.mean() > 0.999, (~close).sum()           # a ray through an edge may pick the other triangle
