"""The learned directional sampler (path guiding) over liblmx.

`Guide(scene)` owns the MLP of one `render.Scene`: `render(...)` draws guided
paths (one-sample MIS between the BSDF and the network's von Mises-Fisher
mixture), optionally recording the sampled directions and training on them
afterwards (`train=True`).  `eval`, `grad` and `mix_sample` expose the
network, the loss gradient and the mixture sampler on device arrays for
tests.  Parameter layout and outputs: include/lmx.h.
"""
import ctypes

import numpy as np

from ._lib import check, lib, lmx_guide_opts, lmx_guide_stats, lmx_render_params, lmx_render_stats
from .render import RASTER_STRATIFIED, RR_FIXED

NPARAM = 25961
NOUT = 41


def default_opts():
    o = lmx_guide_opts()
    lib.lmx_guide_default_opts(ctypes.byref(o))
    return o


class Guide:
    def __init__(self, scene, lr=None, batch=None, steps_per_pass=None, seed=None, wclamp=None):
        o = default_opts()
        if wclamp is not None:
            o.wclamp = wclamp
        if lr is not None:
            o.lr = lr
        if batch is not None:
            o.batch = batch
        if steps_per_pass is not None:
            o.steps_per_pass = steps_per_pass
        if seed is not None:
            o.seed = seed
        self.scene = scene
        h = ctypes.c_void_p()
        check(lib.lmx_guide_create(scene._h, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._free = lib.lmx_guide_free

    def close(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        a = np.zeros(NPARAM, np.float32)
        check(lib.lmx_guide_params(self._h, ctypes.c_void_p(a.ctypes.data), 0))
        return a

    @params.setter
    def params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == (NPARAM,)
        check(lib.lmx_guide_params(self._h, ctypes.c_void_p(a.ctypes.data), 1))

    def eval(self, x):
        """x: torch (n, 7) float32 on the GPU -> (n, 41)."""
        import torch
        x = x.contiguous()
        y = torch.empty((x.shape[0], NOUT), dtype=torch.float32, device=x.device)
        st = torch.cuda.current_stream(x.device)
        check(lib.lmx_guide_eval(self._h, ctypes.c_void_p(x.data_ptr()), x.shape[0], ctypes.c_void_p(y.data_ptr()),
                                 ctypes.c_void_p(st.cuda_stream)))
        torch.cuda.synchronize(x.device)
        return y

    @staticmethod
    def records(x, dirs, pb, w):
        """Pack host arrays into the (n, 20) float32 record layout."""
        n = x.shape[0]
        r = np.zeros((n, 20), np.float32)
        r[:, 0:7] = x
        r[:, 7] = pb
        r[:, 8:11] = dirs
        r[:, 15] = w
        return r

    def grad(self, recs, adam=False):
        """recs: torch (n, 20) float32 on the GPU -> (gradient (NPARAM,) torch, mean loss)."""
        import torch
        recs = recs.contiguous()
        g = torch.empty(NPARAM, dtype=torch.float32, device=recs.device)
        loss = ctypes.c_float()
        st = torch.cuda.current_stream(recs.device)
        check(lib.lmx_guide_grad(self._h, ctypes.c_void_p(recs.data_ptr()), recs.shape[0],
                                 ctypes.c_void_p(g.data_ptr()), ctypes.byref(loss), int(adam),
                                 ctypes.c_void_p(st.cuda_stream)))
        return g, loss.value

    @staticmethod
    def mix_sample(y, u):
        """y: (n, 41), u: (n, 3) torch float32 on the GPU -> (n, 8)."""
        import torch
        y, u = y.contiguous(), u.contiguous()
        out = torch.empty((y.shape[0], 8), dtype=torch.float32, device=y.device)
        st = torch.cuda.current_stream(y.device)
        check(lib.lmx_guide_mix_sample(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(u.data_ptr()), y.shape[0],
                                       ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st.cuda_stream)))
        torch.cuda.synchronize(y.device)
        return out

    def render(self, num_samples, max_num_vertices=9, min_num_vertices=0, rr_depth=1, rr_prob=0.5, seed=42,
               raster_mode=RASTER_STRATIFIED, train=False, queue_capacity=0, table=None, stream=None,
               rr_mode=RR_FIXED):
        """Guided renderer::pt.  Returns (film (H, W, 3) torch, render stats,
        guide stats)."""
        import torch
        dev = torch.device("cuda")
        sc = self.scene
        film = torch.empty((sc.height, sc.width, 3), dtype=torch.float32, device=dev)
        p = lmx_render_params()
        p.max_num_vertices = max_num_vertices
        p.min_num_vertices = min_num_vertices
        p.rr_depth = rr_depth
        p.rr_prob = rr_prob
        p.seed = seed
        p.num_samples = num_samples
        p.raster_mode = raster_mode
        p.queue_capacity = queue_capacity
        p.rr_mode = rr_mode
        if table is not None:
            assert table.is_cuda and table.dtype == torch.float32 and table.is_contiguous()
            assert table.shape[0] == num_samples
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        st, gs = lmx_render_stats(), lmx_guide_stats()
        if stream is None:
            stream = sc._render_stream()
        check(lib.lmx_render_guided(sc._h, self._h, ctypes.byref(p), int(train), ctypes.c_void_p(film.data_ptr()),
                                    ctypes.c_void_p(stream.cuda_stream), ctypes.byref(gs), ctypes.byref(st)))
        stats = {"rays": st.rays, "iterations": st.iterations, "queue_capacity": st.queue_capacity, "ms": st.ms}
        gstats = {"records": gs.records, "records_nonzero": gs.records_nonzero, "steps": gs.steps, "loss": gs.loss, "render_ms": gs.render_ms,
                  "train_ms": gs.train_ms}
        return film, stats, gstats
