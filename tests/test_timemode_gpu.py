"""The scheduler's time mode (render_time; src/liblightmetrica/scheduler.cpp:
57,108,191-199): samples start until the deadline, every started sample
finishes, and the film is rescaled by W H / (samples taken).

* the call lasts at least render_time and not much longer (the samples in
  flight drain);
* the image equals a sample-count render of exactly the samples taken, with
  the same seed (the samples are the same Philox streams), up to the order of
  the film's float additions -- which pins the rescale by the samples taken;
* the same for the bidirectional sampler, whose time mode runs whole batches;
* time mode refuses uniform tables and per-sample output.
"""
import time

import numpy as np
import pytest
import torch

from lmx import render, scenes
from lmx._lib import LmxError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cornell():
    rs = render.Scene(scenes.cornell_box(64, 64))
    rs.render_pt(64 * 64 * 64, max_num_vertices=9, seed=1)             # capacity autotune
    return rs


def test_pt_time_mode_equals_sample_count_render(gpu, cornell):
    T = 0.25
    t0 = time.perf_counter()
    ft, st, _ = cornell.render_pt(1, max_num_vertices=9, rr_depth=3, seed=17, render_time=T)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n = st["samples"]
    assert wall >= T and wall < T + 1.0, wall
    assert n > 64 * 64 * 16, n
    fc, sc_, _ = cornell.render_pt(n, max_num_vertices=9, rr_depth=3, seed=17)
    assert sc_["samples"] == n
    a, b = ft.cpu().numpy().astype(np.float64), fc.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * b.max())


def test_bidirectional_time_mode_whole_batches(gpu, cornell):
    T = 0.2
    B = 1 << 16
    t0 = time.perf_counter()
    ft, st, _ = cornell.render_vcm_bdpt(1, max_num_vertices=6, seed=23, batch=B, render_time=T)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n = st["samples"]
    assert wall >= T and wall < T + 1.0, wall
    assert n % B == 0 and n >= B, n
    fc, _, _ = cornell.render_vcm_bdpt(n, max_num_vertices=6, seed=23, batch=B)
    a, b = ft.cpu().numpy().astype(np.float64), fc.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * b.max())
    fb, sb, _ = cornell.render_bdpt(1, max_num_vertices=6, seed=29, batch=B, render_time=T)
    assert sb["samples"] % B == 0 and sb["samples"] >= B
    assert np.isfinite(fb.cpu().numpy()).all()


def test_time_mode_argument_errors(gpu, cornell):
    tab = torch.zeros((64, render.table_width(9)), dtype=torch.float32, device=gpu)
    with pytest.raises(LmxError, match="render_time"):
        cornell.render_pt(64, max_num_vertices=9, table=tab, render_time=0.1)
    with pytest.raises(LmxError, match="render_time"):
        cornell.render_pt(64, max_num_vertices=9, sample_out=True, render_time=0.1)
    with pytest.raises(LmxError, match="render_time"):
        cornell.render_vcm_bdpt(64, max_num_vertices=6, sample_out=True, render_time=0.1)
