"""Argument checks of the C ABI: every entry point rejects bad arguments with
its error code and a message naming the call, and leaves the scene usable
(the reference reports bool false plus a log line and carries on,
component.h:224-248)."""
import ctypes

import numpy as np
import pytest
import torch

from lmx import render, scenes
from lmx._lib import LmxError, lib, lmx_render_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rs(gpu):
    return render.Scene(scenes.cornell_box(16, 16))


def _call(entry, rs, p, film=None):
    f = film if film is not None else torch.zeros((16, 16, 3), device="cuda")
    return getattr(lib, entry)(rs._h, ctypes.byref(p), ctypes.c_void_p(f.data_ptr()), None, None)


def _params(**kw):
    p = lmx_render_params()
    p.max_num_vertices, p.num_samples, p.rr_depth, p.rr_prob, p.raster_mode = 6, 1024, 1, 0.5, 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("entry", ["lmx_render_pt", "lmx_render_ptdirect", "lmx_render_ptmis", "lmx_render_bdpt",
                                   "lmx_render_vcm_bdpt"])
def test_render_argument_errors(rs, entry):
    assert _call(entry, rs, _params(num_samples=0)) == -1
    assert entry in lib.lmx_last_error().decode()
    assert _call(entry, rs, _params(raster_mode=7)) == -1
    tab = torch.zeros((1024, 4), device="cuda")
    p = _params(uniform_table=ctypes.c_void_p(tab.data_ptr()), table_width=4)
    assert _call(entry, rs, p) == -1 and "table" in lib.lmx_last_error().decode()
    assert getattr(lib, entry)(None, ctypes.byref(_params()), None, None, None) == -1
    # the scene still renders afterwards
    assert _call(entry, rs, _params()) == 0


@pytest.mark.parametrize("entry", ["lmx_render_bdpt", "lmx_render_vcm_bdpt"])
def test_failed_batch_allocation_leaves_the_scene_usable(rs, entry):
    """A batch whose buffers cannot all be allocated (2^27 samples of 10
    vertices: 172 GB of vertex records and 310 GB of shadow-queue records) is
    an error, and the next render on the same scene allocates its own buffers
    again instead of taking the failed ones for allocated."""
    big = _params(max_num_vertices=10, num_samples=1 << 27, queue_capacity=1 << 27)
    assert _call(entry, rs, big) == -3 and "bdpt" in lib.lmx_last_error().decode()
    film = torch.zeros((16, 16, 3), device="cuda")
    assert _call(entry, rs, _params(max_num_vertices=10, num_samples=16 * 16 * 64), film) == 0
    torch.cuda.synchronize()
    f = film.cpu().numpy()
    assert np.isfinite(f).all() and f.mean() > 0.0


def test_bdpt_specific_errors(rs):
    assert _call("lmx_render_bdpt", rs, _params(mis=5)) == -1 and "mis" in lib.lmx_last_error().decode()
    assert _call("lmx_render_bdpt", rs, _params(max_num_vertices=40)) == -1
    assert _call("lmx_render_bdpt", rs, _params(max_num_vertices=-1)) == 0          # unbounded is allowed
    assert _call("lmx_render_vcm_bdpt", rs, _params(max_num_vertices=-1)) == -1     # not for renderer::vcm
    assert _call("lmx_render_vcm_bdpt", rs, _params(max_num_vertices=1)) == -1


def test_film_save_and_load_errors(tmp_path):
    f = torch.zeros((4, 4, 3), device="cuda")
    assert lib.lmx_film_save(ctypes.c_void_p(f.data_ptr()), 4, 4, 9, str(tmp_path / "x").encode(), None) == -1
    assert lib.lmx_film_save(ctypes.c_void_p(f.data_ptr()), 0, 4, 0, str(tmp_path / "x").encode(), None) == -1
    assert lib.lmx_film_save(ctypes.c_void_p(f.data_ptr()), 4, 4, 0,
                             str(tmp_path / "no" / "such" / "dir").encode(), None) == -5   # LMX_EIO
    w, h = ctypes.c_int32(), ctypes.c_int32()
    assert lib.lmx_image_load(str(tmp_path / "missing.hdr").encode(), None, ctypes.byref(w), ctypes.byref(h)) == -5
    (tmp_path / "a.txt").write_bytes(b"x")
    assert lib.lmx_image_load(str(tmp_path / "a.txt").encode(), None, ctypes.byref(w), ctypes.byref(h)) == -1


def test_scene_without_lights_renders_black(gpu):
    sc = scenes.cornell_box(16, 16)
    sc["lights"] = []
    for m in sc["materials"]:
        m["Le"] = (0.0, 0.0, 0.0)
    rs = render.Scene(sc)
    for fn in (rs.render_pt, rs.render_ptmis, rs.render_bdpt, rs.render_vcm_bdpt):
        film = fn(16 * 16 * 4, max_num_vertices=6)[0]
        assert float(film.abs().max()) == 0.0


def test_guide_errors(rs):
    from lmx import guide
    with pytest.raises(LmxError):
        guide.Guide(rs, lr=-1.0)
    g = guide.Guide(rs)
    assert lib.lmx_guide_eval(g._h, None, 4, None, None) == -1
    x = torch.rand((0, 7), device="cuda")
    assert g.eval(x).shape == (0, 41)
