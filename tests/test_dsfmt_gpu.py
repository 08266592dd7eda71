"""The reference's own random stream driving a GPU render.

The reference's renderers draw every uniform from Random::Next(), a
dSFMT-19937 stream per thread (src/liblightmetrica/random.cpp:37-105,
seeded as scheduler.cpp:157-164 does from the initial Random of seed 42).
Here that stream -- the oracle's restatement, pinned bit for bit against the
vendored dSFMT.c's output -- fills the uniform table of a Cornell-box
renderer::pt render (Next() rounded to float as the reference's Float = float
build does, row after row in the canonical slot layout), and the HIP sampler
and the oracle's restatement of renderer_pt.cpp agree per sample: the same
splat positions and splat counts, C within 1e-4 relative.
"""
import numpy as np
import pytest

import orc
from lmx import render, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("maxv,rr_depth", [(9, 3), (5, 1)])
def test_pt_from_dsfmt_stream_per_sample(gpu, maxv, rr_depth):
    import torch
    sc = scenes.cornell_box(64, 64)
    n = 64 * 64 * 16
    w = render.table_width(maxv)
    tab = orc.dsfmt_stream(42, n * w).astype(np.float32).reshape(n, w)
    _, _, so = render.Scene(sc).render_pt(n, max_num_vertices=maxv, rr_depth=rr_depth,
                                          table=torch.from_numpy(tab).to(gpu), sample_out=True)
    g = so.cpu().numpy()[:, :6]
    c = orc.Scene(sc).pt_samples(0, n, max_num_vertices=maxv, rr_depth=rr_depth, table=tab)
    assert (c[:, 5] > 0).sum() > 200
    same = (g[:, 5] == c[:, 5]) & np.all(g[:, 0:2] == c[:, 0:2], axis=1)
    rel = (np.abs(g[:, 2:5] - c[:, 2:5]) / np.maximum(np.abs(c[:, 2:5]), 1e-3)).max(axis=1)
    ok = same & (rel <= 1e-4)
    assert ok.mean() >= 0.9999, f"{(~ok).sum()} of {n} samples differ"
    assert rel[same].max() <= 1e-3
