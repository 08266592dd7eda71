"""lmx_accel_intersect_noalloc: the closest-hit query with no device
allocation, its launch scratch in one static device array per GPU.  Through
the C ABI it must return exactly the hits of lmx_accel_intersect (the same
kernel over the same BVH), on two calls in a row, the second with fewer rays
than the first (what the first call left in the shared scratch must not
leak into the second), in both triangle tests; and it must refuse an accel
that is not resident on the device instead of allocating one.
"""
import numpy as np
import pytest

from lmx import accel, scenes
from lmx._lib import LmxError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("isect", [accel.ISECT_WATERTIGHT, accel.ISECT_TRIACCEL])
def test_noalloc_matches_intersect_on_two_calls(gpu, isect):
    import torch
    v, f = scenes.displaced_icosphere(5, seed=4)
    coh, inc = scenes.traversal_rays(v, res=192, n_incoherent=90000)
    rays = np.concatenate([coh, inc])
    a = accel.Accel(v, f, isect=isect)
    a.upload()
    for n in (rays.shape[0], 7777):          # the second call is smaller
        print(f"noalloc: isect {isect}, {n} rays", flush=True)
        d = torch.from_numpy(rays[-n:].copy()).to(gpu)
        want = a.intersect(d)
        got = torch.full((n, 4), float("nan"), dtype=torch.float32, device=gpu)
        a.intersect_noalloc(d, got)
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(w.view(np.uint32), g.view(np.uint32))
        assert (accel.hit_ids(g) != accel.MISS).mean() > 0.3
    assert a.info()["stack_overflow"] == 0


def test_noalloc_refuses_an_accel_not_on_the_device(gpu):
    import torch
    v, f = scenes.displaced_icosphere(2, seed=1)
    a = accel.Accel(v, f)
    rays = torch.zeros((16, 8), dtype=torch.float32, device=gpu)
    hits = torch.zeros((16, 4), dtype=torch.float32, device=gpu)
    with pytest.raises(LmxError, match="lmx_accel_upload"):
        a.intersect_noalloc(rays, hits)
    a.upload()
    rays[:, 5] = 1.0                     # +y directions from the origin
    rays[:, 7] = 3.0e38
    a.intersect_noalloc(rays, hits)
    torch.cuda.synchronize()
    assert (accel.hit_ids(hits.cpu().numpy()) != accel.MISS).all()
