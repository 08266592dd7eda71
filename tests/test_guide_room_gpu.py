"""Learned path guiding on the occluded-light room (the guiding workload:
12 quads plus 96 displaced spheres, 1,966,104 triangles, 40 % normalized
Phong e = 200, the only emitter behind a wall with a 0.1-wide gap), at 256^2
with the same generator, the same MLP (168-64x4-41, BF16, 8 vMF lobes + BSDF
mix) and the workload's Adam learning rate 1e-3, trained online in 16 passes
of 16 Adam steps on the records of 16 spp each -- the training data of the
1024^2 workload's first 16 spp (16 passes of 2^20 samples) at a quarter of
its resolution:

* unbiasedness of one-sample MIS on this scene: the guided image agrees with
  the plain path tracer's per 8x8-pixel cell within 3 sigma (batch means) and
  in the image mean within 3.5 standard errors;
* what the trained network gains: at equal samples per pixel, against a
  high-spp plain reference, the trained guide's MSE is at least 1.5x lower
  than plain path tracing's and lower than an untrained network's (the same
  mixture with random weights), which isolates the learning from the MIS
  mixture itself.
"""
import numpy as np
import pytest

from lmx import guide, render, scenes

pytestmark = pytest.mark.gpu

W = H = 256
KW = dict(max_num_vertices=17, rr_depth=3, rr_prob=0.5, rr_mode=render.RR_THROUGHPUT)


@pytest.fixture(scope="module")
def room():
    sc = scenes.occluded_room(W, H)
    rs = render.Scene(sc)
    rs.render_pt(W * H * 4, seed=1, **KW)                 # capacity autotune and warm-up
    return rs


@pytest.fixture(scope="module")
def reference(room):
    import torch
    ref = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda:0")
    spp, chunk = 8192, 256
    for k in range(spp // chunk):
        f, _, _ = room.render_pt(W * H * chunk, seed=5_000_000 + k, **KW)
        ref += f.double()
    return (ref / (spp // chunk)).cpu().numpy()


def _trained(room, seed, passes=16):
    g = guide.Guide(room, lr=1e-3, seed=seed)
    for k in range(passes):
        g.render(W * H * 16, seed=100 * seed + k, train=True, **KW)
    return g


def test_room_guided_unbiased_per_cell(gpu, room):
    g = _trained(room, 1)
    kb = 32
    G = np.stack([g.render(W * H * 32, seed=7000 + s, **KW)[0].cpu().numpy().astype(np.float64) for s in range(kb)])
    P = np.stack([room.render_pt(W * H * 64, seed=8000 + s, **KW)[0].cpu().numpy().astype(np.float64)
                  for s in range(kb)])
    cg = G.reshape(kb, H // 8, 8, W // 8, 8, 3).mean(axis=(2, 4))
    cp = P.reshape(kb, H // 8, 8, W // 8, 8, 3).mean(axis=(2, 4))
    se = np.sqrt(cg.var(0, ddof=1) / kb + cp.var(0, ddof=1) / kb) + 1e-12
    z = np.abs(cg.mean(0) - cp.mean(0)) / se
    lit = (cg.mean(0) + cp.mean(0)) > 1e-4
    print(f"room cells: {lit.sum()} lit, fraction within 3 sigma {(z[lit] <= 3.0).mean():.4f}")
    assert lit.mean() > 0.5
    assert (z[lit] <= 3.0).mean() > 0.95, (z[lit] > 3).mean()
    mg, mp = G.mean(axis=(1, 2, 3)), P.mean(axis=(1, 2, 3))
    se_m = np.sqrt(mg.var(ddof=1) / kb + mp.var(ddof=1) / kb)
    print(f"room image mean: guided {mg.mean():.6f} plain {mp.mean():.6f} se {se_m:.2e}")
    assert abs(mg.mean() - mp.mean()) < 3.5 * se_m, (mg.mean(), mp.mean(), se_m)


def test_room_trained_network_gain_equal_spp(gpu, room, reference):
    spp = 64
    trained = _trained(room, 2)
    untrained = guide.Guide(room, lr=1e-3, seed=3)
    mse = {"plain": [], "trained": [], "untrained": []}
    for s in range(3):
        fp = room.render_pt(W * H * spp, seed=9000 + s, **KW)[0].cpu().numpy().astype(np.float64)
        ft = trained.render(W * H * spp, seed=9100 + s, **KW)[0].cpu().numpy().astype(np.float64)
        fu = untrained.render(W * H * spp, seed=9200 + s, **KW)[0].cpu().numpy().astype(np.float64)
        mse["plain"].append(((fp - reference) ** 2).mean())
        mse["trained"].append(((ft - reference) ** 2).mean())
        mse["untrained"].append(((fu - reference) ** 2).mean())
    m = {k: float(np.mean(v)) for k, v in mse.items()}
    print(f"room MSE at {spp} spp: plain {m['plain']:.3e} trained {m['trained']:.3e} untrained {m['untrained']:.3e}; "
          f"plain/trained {m['plain'] / m['trained']:.2f}, untrained/trained {m['untrained'] / m['trained']:.2f}")
    assert m["plain"] / m["trained"] >= 1.5
    assert m["untrained"] / m["trained"] >= 1.2
