"""Renderers over liblmx's wavefront pipeline.

`Scene` uploads a renderer scene (the dicts of lmx.scenes: world-space mesh,
materials, light primitives, pinhole camera); `Scene.render_pt` is the
reference's renderer::pt (src/liblightmetrica/renderer/renderer_pt.cpp:64-242)
with its parameters: max_num_vertices / min_num_vertices as in the reference
(vertices, camera included), Russian roulette with continuation probability
rr_prob after the hit of segment j once j + 1 >= rr_depth (rr_depth = 1,
rr_prob = 0.5 is the reference's fixed rule, :207-215).
"""
import ctypes

import numpy as np

from ._lib import (check, lib, lmx_camera, lmx_material, lmx_render_params, lmx_render_stats,
                   lmx_scene_desc, lmx_pm_params, lmx_ppm_params, lmx_vcm_params)
from .scenes import camera_frame

RASTER_RANDOM = 0
RASTER_STRATIFIED = 1
RR_FIXED = 0
RR_THROUGHPUT = 1
MIS_POWER = 0
MIS_SIMPLE = 1
SHADE_QUEUE = 0
SHADE_BY_MATERIAL = 1
# renderer::vcm modes with merging (renderer_vcm.cpp:563-570)
VCM_MODE_VCM = 0
VCM_MODE_BDPM = 2
# bsdf::cooktorrance defaults (bsdf_cooktorrance.cpp:60-63)
GGX_DEFAULT = dict(roughness=0.1, eta=(0.14, 0.129, 0.1585), k=(4.58625, 3.348125, 2.329375))
SLOT_SEG0 = 8


def table_width(max_num_vertices, nee=False):
    """Uniform-table width of renderer::pt (nee: ptdirect / ptmis) for this
    path-length bound."""
    return SLOT_SEG0 + (8 if nee else 4) * (max_num_vertices - 1)


def vcm_bdpt_table_width(max_num_vertices):
    """Uniform-table width of renderer::vcm's bdpt mode (eye + light block)."""
    return 2 * table_width(max_num_vertices)


def pm_table_widths(max_num_vertices, finalgather=1):
    """Uniform-table widths of renderer::pm: (eye, photon); -1 = 32 vertices."""
    m = 32 if max_num_vertices == -1 else max_num_vertices
    return table_width(min(m, 2 + finalgather)), table_width(m, nee=True)


def bdpt_table_width(max_num_vertices):
    """Uniform-table width of renderer::bdpt (eye + light block; -1 = 32)."""
    m = 32 if max_num_vertices == -1 else max_num_vertices
    return 2 * table_width(m, nee=True)


class Scene:
    def _render_stream(self):
        """The stream a render runs on when the caller names none: one
        non-default stream per scene and device, made once (a new torch
        stream costs ~130 us at its first use), ordered after the work
        already queued on torch's current stream.  Renders return only once
        their stream is idle, so the next render may reuse it."""
        import torch
        idx = torch.cuda.current_device()
        s = self._streams.get(idx)
        if s is None:
            s = self._streams[idx] = torch.cuda.Stream(idx)
        s.wait_stream(torch.cuda.current_stream(idx))
        return s

    def __init__(self, sc):
        self._keep = []
        self._streams = {}

        def keep(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return ctypes.c_void_p(a.ctypes.data)
        mats = (lmx_material * len(sc["materials"]))()
        for i, m in enumerate(sc["materials"]):
            mats[i].type = m["type"]
            mats[i].R[:] = list(m["R"])
            mats[i].exponent = m["exponent"]
            mats[i].Le[:] = list(m["Le"])
            mats[i].roughness = m.get("roughness", GGX_DEFAULT["roughness"])
            mats[i].eta[:] = list(m.get("eta", GGX_DEFAULT["eta"]))
            mats[i].k[:] = list(m.get("k", GGX_DEFAULT["k"]))
        self._keep.append(mats)
        d = lmx_scene_desc()
        d.verts = keep(sc["verts"], np.float32)
        d.nvert = sc["verts"].shape[0]
        d.tris = keep(sc["tris"], np.uint32)
        d.ntri = sc["tris"].shape[0]
        d.normals = keep(sc["normals"], np.float32) if sc.get("normals") is not None else None
        d.tri_mat = keep(sc["tri_mat"], np.uint32)
        d.mats = ctypes.cast(mats, ctypes.c_void_p)
        d.nmat = len(sc["materials"])
        lp = np.array(sc["lights"], np.uint32).reshape(-1, 2)
        d.light_prims = keep(lp, np.uint32)
        d.nlight = lp.shape[0]
        cam = sc["camera"]
        # eye / center / up, or the sensor's frame (position and the three
        # columns of its transform) as a scene file gives it
        pos, vx, vy, vz = cam["frame"] if "frame" in cam else camera_frame(cam)
        c = lmx_camera()
        c.pos[:] = list(pos); c.vx[:] = list(vx); c.vy[:] = list(vy); c.vz[:] = list(vz)
        c.fov_deg = cam["fov_deg"]
        c.width = cam["width"]
        c.height = cam["height"]
        d.camera = c
        self.width, self.height = cam["width"], cam["height"]
        h = ctypes.c_void_p()
        check(lib.lmx_scene_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._free = lib.lmx_scene_free
        self._keep = []           # the library copied everything

    def close(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _params(self, num_samples, max_num_vertices, min_num_vertices, rr_depth, rr_prob, seed, raster_mode,
                queue_capacity):
        p = lmx_render_params()
        p.max_num_vertices = max_num_vertices
        p.min_num_vertices = min_num_vertices
        p.rr_depth = rr_depth
        p.rr_prob = rr_prob
        p.seed = seed
        p.num_samples = num_samples
        p.raster_mode = raster_mode
        p.queue_capacity = queue_capacity
        return p

    def render_vcm_bdpt(self, num_samples, max_num_vertices=10, min_num_vertices=0, seed=42,
                        raster_mode=RASTER_STRATIFIED, batch=0, table=None, sample_out=False, stream=None,
                        render_time=0.0, ray_out=None):
        """renderer::vcm in bdpt mode.  Returns (film (H, W, 3), stats,
        per-sample records (n, (m+1)^2, 8) or None).  render_time > 0: the
        scheduler's time mode (num_samples is then ignored; stats["samples"]
        says how many were taken)."""
        return self._bidir("lmx_render_vcm_bdpt", (max_num_vertices + 1) ** 2, num_samples, max_num_vertices,
                           min_num_vertices, seed, raster_mode, batch, table, sample_out, stream, MIS_POWER,
                           render_time, ray_out)

    def render_bdpt(self, num_samples, max_num_vertices=-1, min_num_vertices=0, seed=42,
                    raster_mode=RASTER_STRATIFIED, batch=0, table=None, sample_out=False, stream=None,
                    mis=MIS_POWER, render_time=0.0, ray_out=None):
        """renderer::bdpt (roulette subpaths, direct light vertices, power or
        simple weights).  Returns (film (H, W, 3), stats, per-sample records
        (n, 2 (M+1)^2, 8) at 2 (t (M+1) + s) + direct, or None); M =
        max_num_vertices, 32 when unbounded (-1)."""
        m = 32 if max_num_vertices == -1 else max_num_vertices
        return self._bidir("lmx_render_bdpt", 2 * (m + 1) ** 2, num_samples, max_num_vertices,
                           min_num_vertices, seed, raster_mode, batch, table, sample_out, stream, mis, render_time,
                           ray_out)

    def render_vcm(self, num_samples, mode=VCM_MODE_VCM, num_iteration_pass=100, num_photon_trace_samples=10000,
                   initial_radius=0.1, alpha=0.7, max_num_vertices=10, min_num_vertices=0, seed=42,
                   photon_seed=None, raster_mode=RASTER_STRATIFIED, batch=0, table=None, photon_table=None,
                   sample_out=False, stream=None):
        """renderer::vcm in mode vcm or bdpm (defaults: renderer_vcm.cpp:
        558-577); num_samples = num_eye_trace_samples per pass.  Returns (film
        (H, W, 3), the mean of the passes' images; stats; per-sample records
        (n, 2 (m+1)^2, 8) at 2 (t (m+1) + s) + merge, or None)."""
        import torch
        dev = torch.device("cuda")
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        p = self._params(num_samples, max_num_vertices, min_num_vertices, 1, 0.5, seed, raster_mode, batch)
        v = lmx_vcm_params()
        v.mode = mode
        v.num_iteration_pass = num_iteration_pass
        v.num_photon_trace_samples = num_photon_trace_samples
        v.initial_radius = initial_radius
        v.alpha = alpha
        v.photon_seed = (seed + 1) if photon_seed is None else photon_seed
        so = None
        for t, n in ((table, num_samples), (photon_table, num_photon_trace_samples)):
            if t is not None:
                assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape[0] == n
        if table is not None:
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        if photon_table is not None:
            v.photon_table = ctypes.c_void_p(photon_table.data_ptr())
            v.photon_table_width = photon_table.shape[1]
        if sample_out:
            so = torch.empty((num_samples, 2 * (max_num_vertices + 1) ** 2, 8), dtype=torch.float32, device=dev)
            p.sample_out = ctypes.c_void_p(so.data_ptr())
        st = lmx_render_stats()
        if stream is None:
            stream = self._render_stream()
        check(lib.lmx_render_vcm(self._h, ctypes.byref(p), ctypes.byref(v), ctypes.c_void_p(film.data_ptr()),
                                 ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st)))
        stats = {"rays": st.rays, "passes": st.iterations, "batch": st.queue_capacity, "ms": st.ms,
                 "dropped": st.dropped}
        return film, stats, so

    def render_pm(self, num_samples, num_photon_trace_samples=100000, finalgather=1, radius=0.01,
                  max_num_vertices=-1, seed=42, photon_seed=None, raster_mode=RASTER_STRATIFIED, batch=0,
                  table=None, photon_table=None, sample_out=False, stream=None):
        """renderer::pm (defaults: renderer_pm.cpp:61-70).  Returns (film (H, W,
        3), stats, per-sample records (n, 16) or None: raster x, y, emission C
        rgb, emitter vertices, density C rgb, photons gathered, vertex count at
        the gather vertex)."""
        import torch
        dev = torch.device("cuda")
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        p = self._params(num_samples, max_num_vertices, 0, 1, 0.5, seed, raster_mode, batch)
        v = lmx_pm_params()
        v.num_photon_trace_samples = num_photon_trace_samples
        v.finalgather = finalgather
        v.radius = radius
        v.photon_seed = (seed + 1) if photon_seed is None else photon_seed
        for t, n in ((table, num_samples), (photon_table, num_photon_trace_samples)):
            if t is not None:
                assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape[0] == n
        if table is not None:
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        if photon_table is not None:
            v.photon_table = ctypes.c_void_p(photon_table.data_ptr())
            v.photon_table_width = photon_table.shape[1]
        so = None
        if sample_out:
            so = torch.zeros((num_samples, 16), dtype=torch.float32, device=dev)
            p.sample_out = ctypes.c_void_p(so.data_ptr())
        st = lmx_render_stats()
        if stream is None:
            stream = self._render_stream()
        check(lib.lmx_render_pm(self._h, ctypes.byref(p), ctypes.byref(v), ctypes.c_void_p(film.data_ptr()),
                                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st)))
        stats = {"rays": st.rays, "batches": st.iterations, "batch": st.queue_capacity, "ms": st.ms}
        return film, stats, so

    def render_ppm(self, num_samples, max_num_vertices, num_iteration_pass=1000, num_photon_trace_samples=100,
                   initial_radius=0.1, alpha=0.7, seed=42, photon_seed=None, raster_mode=RASTER_STRATIFIED,
                   table=None, photon_table=None, sample_out=False, stream=None):
        """renderer::ppm (defaults: renderer_ppm.cpp:76-89; max_num_vertices
        has none there either); num_samples = measurement points.  Returns
        (film (H, W, 3), stats, per-point records (n, 16) or None: raster x,
        y, C rgb, emission rgb, tau rgb, N, radius, made-a-point flag)."""
        import torch
        dev = torch.device("cuda")
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        p = self._params(num_samples, max_num_vertices, 0, 1, 0.5, seed, raster_mode, 0)
        v = lmx_ppm_params()
        v.num_iteration_pass = num_iteration_pass
        v.num_photon_trace_samples = num_photon_trace_samples
        v.initial_radius = initial_radius
        v.alpha = alpha
        v.photon_seed = (seed + 1) if photon_seed is None else photon_seed
        for t, n in ((table, num_samples), (photon_table, num_iteration_pass * num_photon_trace_samples)):
            if t is not None:
                assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape[0] == n
        if table is not None:
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        if photon_table is not None:
            v.photon_table = ctypes.c_void_p(photon_table.data_ptr())
            v.photon_table_width = photon_table.shape[1]
        so = None
        if sample_out:
            so = torch.zeros((num_samples, 16), dtype=torch.float32, device=dev)
            p.sample_out = ctypes.c_void_p(so.data_ptr())
        st = lmx_render_stats()
        if stream is None:
            stream = self._render_stream()
        check(lib.lmx_render_ppm(self._h, ctypes.byref(p), ctypes.byref(v), ctypes.c_void_p(film.data_ptr()),
                                 ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st)))
        stats = {"rays": st.rays, "passes": st.iterations, "ms": st.ms}
        return film, stats, so

    def render_sppm(self, max_num_vertices, num_iteration_pass=1000, num_photon_trace_samples=100,
                    initial_radius=0.1, alpha=0.7, seed=42, photon_seed=None, table=None, photon_table=None,
                    sample_out=False, stream=None):
        """renderer::sppm (defaults: renderer_sppm.cpp:81-96): a point per
        pixel.  Returns (film (H, W, 3), stats, per-pixel records (W H, 16) or
        None: film C rgb, emission sum rgb, tau rgb, N, radius, passes with a
        point)."""
        import torch
        dev = torch.device("cuda")
        npix = self.width * self.height
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        p = self._params(npix, max_num_vertices, 0, 1, 0.5, seed, RASTER_STRATIFIED, 0)
        v = lmx_ppm_params()
        v.num_iteration_pass = num_iteration_pass
        v.num_photon_trace_samples = num_photon_trace_samples
        v.initial_radius = initial_radius
        v.alpha = alpha
        v.photon_seed = (seed + 1) if photon_seed is None else photon_seed
        for t, n in ((table, num_iteration_pass * npix), (photon_table, num_iteration_pass * num_photon_trace_samples)):
            if t is not None:
                assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape[0] == n
        if table is not None:
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        if photon_table is not None:
            v.photon_table = ctypes.c_void_p(photon_table.data_ptr())
            v.photon_table_width = photon_table.shape[1]
        so = None
        if sample_out:
            so = torch.zeros((npix, 16), dtype=torch.float32, device=dev)
            p.sample_out = ctypes.c_void_p(so.data_ptr())
        st = lmx_render_stats()
        if stream is None:
            stream = self._render_stream()
        check(lib.lmx_render_sppm(self._h, ctypes.byref(p), ctypes.byref(v), ctypes.c_void_p(film.data_ptr()),
                                  ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st)))
        stats = {"rays": st.rays, "passes": st.iterations, "ms": st.ms}
        return film, stats, so

    def render_ltdirect(self, num_samples, max_num_vertices=-1, seed=42, batch=0, table=None, sample_out=False,
                        stream=None):
        """renderer::ltdirect (renderer_ltdirect.cpp:63-68): light subpaths
        connected to the pinhole at every vertex.  Returns (film (H, W, 3),
        stats, per-sample records (n, M, 8) or None; M = max_num_vertices, 32
        when unbounded)."""
        m = 32 if max_num_vertices == -1 else max_num_vertices
        return self._bidir("lmx_render_ltdirect", m, num_samples, max_num_vertices, 0, seed, RASTER_RANDOM, batch,
                           table, sample_out, stream, MIS_POWER)

    def render_raycast(self, stream=None):
        """renderer::raycast: (H, W, 3) film of |sn . d| x reflectance at the
        pixel centres, and stats."""
        import torch
        dev = torch.device("cuda")
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        st = lmx_render_stats()
        if stream is None:
            stream = self._render_stream()
        check(lib.lmx_render_raycast(self._h, ctypes.c_void_p(film.data_ptr()), ctypes.c_void_p(stream.cuda_stream),
                                     ctypes.byref(st)))
        return film, {"rays": st.rays, "ms": st.ms}

    def _bidir(self, entry, nrec, num_samples, max_num_vertices, min_num_vertices, seed, raster_mode, batch,
               table, sample_out, stream, mis, render_time=0.0, ray_out=None):
        import torch
        dev = torch.device("cuda")
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        p = self._params(num_samples, max_num_vertices, min_num_vertices, 1, 0.5, seed, raster_mode, batch)
        p.mis = mis
        p.render_time = render_time
        if ray_out is not None:
            # torch float32 (k, 8): subpath rays in the first half, shadow rays in the second
            assert ray_out.is_cuda and ray_out.dtype == torch.float32 and ray_out.is_contiguous()
            p.ray_out = ctypes.c_void_p(ray_out.data_ptr())
            p.ray_out_cap = ray_out.shape[0]
        so = None
        if table is not None:
            assert table.is_cuda and table.dtype == torch.float32 and table.is_contiguous()
            assert table.shape[0] == num_samples
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        if sample_out:
            so = torch.zeros((num_samples, nrec, 8), dtype=torch.float32, device=dev)
            p.sample_out = ctypes.c_void_p(so.data_ptr())
        st = lmx_render_stats()
        if stream is None:
            stream = self._render_stream()
        check(getattr(lib, entry)(self._h, ctypes.byref(p), ctypes.c_void_p(film.data_ptr()),
                                  ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st)))
        stats = {"rays": st.rays, "batches": st.iterations, "batch": st.queue_capacity, "ms": st.ms,
                 "samples": st.samples}
        return film, stats, so

    def render_ptdirect(self, num_samples, **kw):
        """renderer::ptdirect: renderer::pt with light sampling at every
        vertex (records (n, 16) with sample_out)."""
        return self.render_pt(num_samples, _entry="lmx_render_ptdirect", **kw)

    def render_ptmis(self, num_samples, **kw):
        """renderer::ptmis: light and BSDF sampling, balance heuristic."""
        return self.render_pt(num_samples, _entry="lmx_render_ptmis", **kw)

    def render_pt(self, num_samples, max_num_vertices=9, min_num_vertices=0, rr_depth=1, rr_prob=0.5,
                  seed=42, raster_mode=RASTER_STRATIFIED, queue_capacity=0, table=None, sample_out=False,
                  stream=None, rr_mode=RR_FIXED, render_time=0.0, shade_order=0, ray_out=None,
                  _entry="lmx_render_pt"):
        """Returns (film as a torch tensor (H, W, 3) on the GPU, stats dict,
        per-sample records (n, 8) or None).  `table`: torch float32 tensor
        (num_samples, width) on the GPU, or None for Philox.  render_time > 0:
        the scheduler's time mode (num_samples ignored; stats["samples"] is the
        number taken)."""
        import torch
        dev = torch.device("cuda")
        film = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=dev)
        so = None
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
        p.render_time = render_time
        p.shade_order = shade_order
        if ray_out is not None:
            # torch float32 (k, 8) on the GPU: the first k extension rays traced
            assert ray_out.is_cuda and ray_out.dtype == torch.float32 and ray_out.is_contiguous()
            assert ray_out.dim() == 2 and ray_out.shape[1] == 8
            p.ray_out = ctypes.c_void_p(ray_out.data_ptr())
            p.ray_out_cap = ray_out.shape[0]
        if table is not None:
            assert table.is_cuda and table.dtype == torch.float32 and table.is_contiguous()
            assert table.shape[0] == num_samples
            p.uniform_table = ctypes.c_void_p(table.data_ptr())
            p.table_width = table.shape[1]
        if sample_out:
            so = torch.zeros((num_samples, 8 if _entry == "lmx_render_pt" else 16), dtype=torch.float32, device=dev)
            p.sample_out = ctypes.c_void_p(so.data_ptr())
        st = lmx_render_stats()
        if stream is None:
            # a non-legacy stream: the renderer replays hipGraphs on it
            stream = self._render_stream()
        check(getattr(lib, _entry)(self._h, ctypes.byref(p), ctypes.c_void_p(film.data_ptr()),
                                   ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st)))
        stats = {"rays": st.rays, "iterations": st.iterations, "queue_capacity": st.queue_capacity, "ms": st.ms,
                 "samples": st.samples}
        return film, stats, so
