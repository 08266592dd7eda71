"""ctypes binding of the CPU oracle (liborc.so) and of the reference build
(_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LMX_ORC_LIB names another build of the same sources (bench.py's
# -march=native build for its CPU baseline)
ORC_PATH = os.environ.get("LMX_ORC_LIB") or os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

TEST_TRIACCEL = 0
TEST_WATERTIGHT = 1
TIE_LAST_WINS = 0
TIE_MIN_ID = 1
MISS = 0xFFFFFFFF

_orc = None
_ref = None


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            raise ImportError(f"{ORC_PATH} missing: run `make -C oracle`")
        so = ctypes.CDLL(ORC_PATH)
        vp, u32, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
        so.orc_triaccel_load.argtypes = [vp, vp, vp, vp]
        so.orc_triaccel_intersect.argtypes = [vp, vp, vp, f32, f32, vp, vp, vp]
        so.orc_watertight_intersect.argtypes = [vp, vp, vp, vp, vp, f32, f32, vp, vp, vp]
        so.orc_bound_intersect.argtypes = [vp, vp, vp, vp, f32, f32]
        so.orc_closest_hit_naive.argtypes = [vp, vp, u32, vp, u32, i32, i32, vp, i32]
        so.orc_closest_hit_naive.restype = None
        so.orc_qbvh_build.argtypes = [vp, vp, u32]
        so.orc_qbvh_build.restype = vp
        so.orc_qbvh_free.argtypes = [vp]
        so.orc_qbvh_free.restype = None
        so.orc_qbvh_num_nodes.argtypes = [vp]
        so.orc_qbvh_num_nodes.restype = u32
        so.orc_qbvh_intersect.argtypes = [vp, vp, u32, vp, i32]
        so.orc_qbvh_intersect.restype = None
        so.orc_bvh_build.argtypes = [vp, vp, u32]
        so.orc_bvh_build.restype = vp
        so.orc_bvh_free.argtypes = [vp]
        so.orc_bvh_free.restype = None
        so.orc_bvh_num_nodes.argtypes = [vp]
        so.orc_bvh_num_nodes.restype = u32
        so.orc_bvh_intersect.argtypes = [vp, vp, u32, vp, i32]
        so.orc_bvh_intersect.restype = None
        for name in ("orc_sample_concentric_disk", "orc_sample_cosine_hemisphere",
                     "orc_sample_uniform_sphere", "orc_sample_uniform_triangle"):
            getattr(so, name).argtypes = [vp, vp]
            getattr(so, name).restype = None
        _orc = so
    return _orc


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        so = ctypes.CDLL(REF_PATH)
        vp, f32, i32, u32 = ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_uint32
        so.ref_triaccel_load.argtypes = [vp, vp, vp, vp]
        so.ref_triaccel_intersect.argtypes = [vp, vp, vp, f32, f32, vp, vp, vp]
        so.ref_bound_intersect.argtypes = [vp, vp, vp, vp, f32, f32]
        for name in ("ref_sample_concentric_disk", "ref_sample_cosine_hemisphere",
                     "ref_sample_uniform_sphere", "ref_sample_uniform_triangle"):
            getattr(so, name).argtypes = [vp, vp]
            getattr(so, name).restype = None
        so.ref_dist_sample_reuse.argtypes = [vp, i32, f32, vp]
        so.ref_dsfmt_stream.argtypes = [u32, i32, vp, vp]
        so.ref_dsfmt_stream.restype = None
        _ref = so
    return _ref


FM_SIN, FM_COS, FM_TAN, FM_EXP, FM_LOG, FM_POW = range(6)


def fmath(kind, x, y=None):
    """The host build of lmx_fmath.h (orc_fmath.c) over float32 arrays."""
    so = lib()
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float32)
    out = np.empty_like(x)
    so.orc_fmath_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    so.orc_fmath_eval.restype = None
    so.orc_fmath_eval(kind, _p(x), _p(y), x.size, _p(out))
    return out


def dsfmt_stream(seed, n, uint32=False):
    """n values of the reference's Random seeded with `seed`: Next() as
    float64 in [0, 1), or NextUInt() words (orc_dsfmt.c)."""
    so = lib()
    so.orc_dsfmt_stream.argtypes = [ctypes.c_uint32, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    so.orc_dsfmt_stream.restype = None
    if uint32:
        out = np.zeros(n, np.uint32)
        so.orc_dsfmt_stream(seed, n, None, _p(out))
    else:
        out = np.zeros(n, np.float64)
        so.orc_dsfmt_stream(seed, n, _p(out), None)
    return out


def closest_hit_naive(verts, tris, rays, test=TEST_WATERTIGHT, tie=TIE_MIN_ID, nthreads=8):
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    h = np.zeros((r.shape[0], 4), dtype=np.float32)
    lib().orc_closest_hit_naive(_p(v), _p(t), t.shape[0], _p(r), r.shape[0], test, tie, _p(h), nthreads)
    return h


class QBVH:
    """The reference's 4-wide QBVH (accel_qbvh.cpp) restated in C."""

    def __init__(self, verts, tris):
        self._v = np.ascontiguousarray(verts, dtype=np.float32)
        self._t = np.ascontiguousarray(tris, dtype=np.uint32)
        self._h = lib().orc_qbvh_build(_p(self._v), _p(self._t), self._t.shape[0])
        self._free = lib().orc_qbvh_free

    def intersect(self, rays, nthreads=8):
        r = np.ascontiguousarray(rays, dtype=np.float32)
        h = np.zeros((r.shape[0], 4), dtype=np.float32)
        lib().orc_qbvh_intersect(self._h, _p(r), r.shape[0], _p(h), nthreads)
        return h

    def num_nodes(self):
        return lib().orc_qbvh_num_nodes(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None


class BVH:
    """The reference's binned-SAH binary BVH (accel_bvh_sahbin.cpp) restated in C."""

    def __init__(self, verts, tris):
        self._v = np.ascontiguousarray(verts, dtype=np.float32)
        self._t = np.ascontiguousarray(tris, dtype=np.uint32)
        self._h = lib().orc_bvh_build(_p(self._v), _p(self._t), self._t.shape[0])
        self._free = lib().orc_bvh_free

    def intersect(self, rays, nthreads=8):
        r = np.ascontiguousarray(rays, dtype=np.float32)
        h = np.zeros((r.shape[0], 4), dtype=np.float32)
        lib().orc_bvh_intersect(self._h, _p(r), r.shape[0], _p(h), nthreads)
        return h

    def num_nodes(self):
        return lib().orc_bvh_num_nodes(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None


def hit_ids(h):
    return np.ascontiguousarray(h, dtype=np.float32).view(np.uint32)[:, 3]


# --------------------------------------------------------------------------
# path samplers (orc_render.c)

PT_RASTER = 1
PT_SEG0 = 8
RASTER_RANDOM = 0
RASTER_STRATIFIED = 1
PHILOX_TAG = 0x6C6D7870


class orc_material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("R", ctypes.c_float * 3), ("exponent", ctypes.c_float),
                ("Le", ctypes.c_float * 3), ("roughness", ctypes.c_float), ("eta", ctypes.c_float * 3),
                ("k", ctypes.c_float * 3), ("reserved", ctypes.c_float)]


GGX_DEFAULT = dict(roughness=0.1, eta=(0.14, 0.129, 0.1585), k=(4.58625, 3.348125, 2.329375))


def _material(mat):
    m = orc_material()
    m.type = mat["type"]
    m.R[:] = list(mat["R"])
    m.exponent = mat.get("exponent", 0.0)
    m.Le[:] = list(mat.get("Le", (0, 0, 0)))
    m.roughness = mat.get("roughness", GGX_DEFAULT["roughness"])
    m.eta[:] = list(mat.get("eta", GGX_DEFAULT["eta"]))
    m.k[:] = list(mat.get("k", GGX_DEFAULT["k"]))
    return m


class orc_scene_desc(ctypes.Structure):
    _fields_ = [("verts", ctypes.c_void_p), ("nvert", ctypes.c_uint32),
                ("tris", ctypes.c_void_p), ("ntri", ctypes.c_uint32),
                ("normals", ctypes.c_void_p), ("tri_mat", ctypes.c_void_p),
                ("mats", ctypes.c_void_p), ("nmat", ctypes.c_uint32),
                ("light_prims", ctypes.c_void_p), ("nlight", ctypes.c_uint32),
                ("cam_pos", ctypes.c_float * 3), ("cam_vx", ctypes.c_float * 3),
                ("cam_vy", ctypes.c_float * 3), ("cam_vz", ctypes.c_float * 3),
                ("fov_deg", ctypes.c_float), ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class orc_pt_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("min_num_vertices", ctypes.c_int32),
                ("rr_depth", ctypes.c_int32), ("rr_prob", ctypes.c_float),
                ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32), ("raster_mode", ctypes.c_int32),
                ("rr_mode", ctypes.c_int32)]


class orc_bdpt_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("min_num_vertices", ctypes.c_int32),
                ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32), ("raster_mode", ctypes.c_int32)]


class orc_bdptref_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("min_num_vertices", ctypes.c_int32),
                ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32), ("raster_mode", ctypes.c_int32),
                ("mis", ctypes.c_int32)]


class orc_vcm_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("min_num_vertices", ctypes.c_int32),
                ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32),
                ("pseed_lo", ctypes.c_uint32), ("pseed_hi", ctypes.c_uint32), ("raster_mode", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("num_photons", ctypes.c_uint32), ("radius", ctypes.c_float)]


class orc_pm_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32),
                ("pseed_lo", ctypes.c_uint32), ("pseed_hi", ctypes.c_uint32), ("raster_mode", ctypes.c_int32),
                ("finalgather", ctypes.c_int32), ("num_photons", ctypes.c_uint32), ("radius", ctypes.c_float)]


class orc_ppm_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32),
                ("pseed_lo", ctypes.c_uint32), ("pseed_hi", ctypes.c_uint32), ("raster_mode", ctypes.c_int32),
                ("num_passes", ctypes.c_uint32), ("num_photons", ctypes.c_uint32),
                ("initial_radius", ctypes.c_float), ("alpha", ctypes.c_float)]


class orc_lt_params(ctypes.Structure):
    _fields_ = [("max_num_vertices", ctypes.c_int32), ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32)]


VCM_VCM = 0
VCM_BDPM = 2


def _render_sigs(so):
    vp = ctypes.c_void_p
    so.orc_philox4x32_10.argtypes = [vp, vp, vp]
    so.orc_philox4x32_10.restype = None
    so.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
    so.orc_uniform.restype = ctypes.c_float
    so.orc_scene_new.argtypes = [ctypes.POINTER(orc_scene_desc)]
    so.orc_scene_new.restype = vp
    so.orc_scene_free.argtypes = [vp]
    so.orc_scene_free.restype = None
    so.orc_pt_samples.argtypes = [vp, ctypes.POINTER(orc_pt_params), vp, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_pt_samples.restype = None
    so.orc_film_splat.argtypes = [vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, vp]
    so.orc_film_splat.restype = None
    so.orc_bdpt_samples.argtypes = [vp, ctypes.POINTER(orc_bdpt_params), vp, ctypes.c_uint32, ctypes.c_uint64,
                                    ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_bdpt_samples.restype = None
    so.orc_bdptref_samples.argtypes = [vp, ctypes.POINTER(orc_bdptref_params), vp, ctypes.c_uint32, ctypes.c_uint64,
                                       ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_bdptref_samples.restype = None
    so.orc_vcm_samples.argtypes = [vp, ctypes.POINTER(orc_vcm_params), vp, ctypes.c_uint32, vp, ctypes.c_uint32,
                                   ctypes.c_uint64, ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_vcm_samples.restype = None
    so.orc_pm_samples.argtypes = [vp, ctypes.POINTER(orc_pm_params), vp, ctypes.c_uint32, vp, ctypes.c_uint32,
                                  ctypes.c_uint64, ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_pm_samples.restype = None
    so.orc_ppm_points.argtypes = [vp, ctypes.POINTER(orc_ppm_params), vp, ctypes.c_uint32, vp, ctypes.c_uint32,
                                  ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_ppm_points.restype = None
    so.orc_sppm_pixels.argtypes = [vp, ctypes.POINTER(orc_ppm_params), vp, ctypes.c_uint32, vp, ctypes.c_uint32,
                                   vp, ctypes.c_int]
    so.orc_sppm_pixels.restype = None
    so.orc_ltdirect_samples.argtypes = [vp, ctypes.POINTER(orc_lt_params), vp, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_ltdirect_samples.restype = None
    so.orc_raycast.argtypes = [vp, vp]
    so.orc_raycast.restype = None
    so.orc_film_rgbe.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    so.orc_film_rgbe.restype = None
    so.orc_film_png.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    so.orc_film_png.restype = None
    so.orc_ptnee_samples.argtypes = [vp, ctypes.POINTER(orc_pt_params), ctypes.c_int, vp, ctypes.c_uint32,
                                     ctypes.c_uint64, ctypes.c_uint64, vp, ctypes.c_int]
    so.orc_ptnee_samples.restype = None
    so.orc_bsdf_query.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    so.orc_bsdf_eval.argtypes = [vp, vp, vp, vp, vp, vp]
    so.orc_bsdf_eval.restype = None


def _rlib():
    so = lib()
    if not hasattr(so, "_render_sigs"):
        _render_sigs(so)
        so._render_sigs = True
    return so


def bsdf_query(mat, n, wi, u):
    """(ok, wo, pdf_projected, f) of one oracle BSDF sample on a flat frame."""
    m = _material(mat)
    n = np.ascontiguousarray(n, np.float32); wi = np.ascontiguousarray(wi, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    wo = np.zeros(3, np.float32); pdf = np.zeros(1, np.float32); f = np.zeros(3, np.float32)
    ok = _rlib().orc_bsdf_query(ctypes.byref(m), _p(n), _p(wi), _p(u), _p(wo), _p(pdf), _p(f))
    return ok, wo, float(pdf[0]), f


def bsdf_eval(mat, n, wi, wo):
    m = _material(mat)
    n = np.ascontiguousarray(n, np.float32); wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    pdf = np.zeros(1, np.float32); f = np.zeros(3, np.float32)
    _rlib().orc_bsdf_eval(ctypes.byref(m), _p(n), _p(wi), _p(wo), _p(pdf), _p(f))
    return float(pdf[0]), f


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    so = lib()
    if not hasattr(so, "_render_sigs"):
        _render_sigs(so)
        so._render_sigs = True
    so.orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def table_width(max_num_vertices):
    return PT_SEG0 + 4 * (max_num_vertices - 1)


class Scene:
    """The oracle's view of a renderer scene dict (lmx.scenes)."""

    def __init__(self, sc):
        so = lib()
        if not hasattr(so, "_render_sigs"):
            _render_sigs(so)
            so._render_sigs = True
        self._keep = []

        def keep(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return ctypes.c_void_p(a.ctypes.data)
        mats = (orc_material * len(sc["materials"]))()
        for i, m in enumerate(sc["materials"]):
            mats[i] = _material(m)
        self._keep.append(mats)
        d = orc_scene_desc()
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
        import sys
        sys.path.insert(0, os.path.join(HERE, "..", "lightmetrica-v2_amd"))
        from lmx.scenes import camera_frame
        pos, vx, vy, vz = camera_frame(cam)
        d.cam_pos[:] = list(pos); d.cam_vx[:] = list(vx); d.cam_vy[:] = list(vy); d.cam_vz[:] = list(vz)
        d.fov_deg = cam["fov_deg"]
        d.width = cam["width"]
        d.height = cam["height"]
        self.desc = d
        self.width, self.height = cam["width"], cam["height"]
        self._h = so.orc_scene_new(ctypes.byref(d))
        self._free = so.orc_scene_free

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    def pt_samples(self, first, n, max_num_vertices=9, min_num_vertices=0, rr_depth=1, rr_prob=0.5,
                   seed=42, raster_mode=RASTER_STRATIFIED, table=None, nthreads=8, rr_mode=0):
        """(n, 6) float32: raster x, y, C rgb, number of splats."""
        p = orc_pt_params(max_num_vertices, min_num_vertices, rr_depth, rr_prob, seed & 0xFFFFFFFF,
                          seed >> 32, raster_mode, rr_mode)
        out = np.zeros((n, 6), np.float32)
        tw = 0
        tp = None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw = table.shape[1]
            tp = ctypes.c_void_p(table.ctypes.data)
        lib().orc_pt_samples(self._h, ctypes.byref(p), tp, tw, first, n, _p(out), nthreads)
        return out

    def nee_samples(self, first, n, mode, max_num_vertices=9, min_num_vertices=0, rr_depth=1, rr_prob=0.5,
                    seed=42, raster_mode=RASTER_STRATIFIED, table=None, nthreads=8, rr_mode=0):
        """renderer::ptdirect (mode 1) / renderer::ptmis (mode 2): (n, 16)
        float32 -- raster x, y, C rgb, count, 0, 0, then the camera-vertex
        connection's raster x, y, C rgb, count, 0, 0."""
        p = orc_pt_params(max_num_vertices, min_num_vertices, rr_depth, rr_prob, seed & 0xFFFFFFFF,
                          seed >> 32, raster_mode, rr_mode)
        out = np.zeros((n, 16), np.float32)
        tw, tp = 0, None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw = table.shape[1]
            tp = ctypes.c_void_p(table.ctypes.data)
        lib().orc_ptnee_samples(self._h, ctypes.byref(p), int(mode), tp, tw, first, n, _p(out), nthreads)
        return out

    def bdpt_samples(self, first, n, max_num_vertices=10, min_num_vertices=0, seed=42,
                     raster_mode=RASTER_STRATIFIED, table=None, nthreads=8):
        """(n, (m+1)^2, 8) float32 per (t (m+1) + s): raster x, y, C rgb
        (f w / p, before the W H / N film scale), contributed flag, w, 0."""
        m = max_num_vertices
        p = orc_bdpt_params(m, min_num_vertices, seed & 0xFFFFFFFF, seed >> 32, raster_mode)
        out = np.zeros((n, (m + 1) * (m + 1), 8), np.float32)
        tw, tp = 0, None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw = table.shape[1]
            tp = ctypes.c_void_p(table.ctypes.data)
        lib().orc_bdpt_samples(self._h, ctypes.byref(p), tp, tw, first, n, _p(out), nthreads)
        return out

    def bdptref_samples(self, first, n, max_num_vertices=-1, min_num_vertices=0, seed=42,
                        raster_mode=RASTER_STRATIFIED, mis=0, table=None, nthreads=8):
        """renderer::bdpt: (n, 2 (M+1)^2, 8) float32 per 2 (t (M+1) + s) + direct
        (M = max_num_vertices, 32 when unbounded): raster x, y, C rgb
        (C* w / SelectionPDF, before the W H / N film scale), contributed
        flag, w, 0.  mis: 0 power heuristic, 1 simple."""
        M = 32 if max_num_vertices == -1 else max_num_vertices
        p = orc_bdptref_params(max_num_vertices, min_num_vertices, seed & 0xFFFFFFFF, seed >> 32, raster_mode, mis)
        out = np.zeros((n, 2 * (M + 1) * (M + 1), 8), np.float32)
        tw, tp = 0, None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw = table.shape[1]
            tp = ctypes.c_void_p(table.ctypes.data)
        lib().orc_bdptref_samples(self._h, ctypes.byref(p), tp, tw, first, n, _p(out), nthreads)
        return out

    def vcm_samples(self, first, n, radius, num_photons, mode=VCM_VCM, max_num_vertices=10, min_num_vertices=0,
                    seed=42, pseed=43, raster_mode=RASTER_STRATIFIED, table=None, ptable=None, nthreads=8):
        """renderer::vcm (vcm or bdpm) for one pass: (n, 2 (m+1)^2, 8) float32 per
        2 (t (m+1) + s) + merge: raster x, y, C rgb (merges summed over
        photon vertices), count, sum of w, 0."""
        m = max_num_vertices
        p = orc_vcm_params(m, min_num_vertices, seed & 0xFFFFFFFF, seed >> 32, pseed & 0xFFFFFFFF, pseed >> 32,
                           raster_mode, mode, num_photons, radius)
        out = np.zeros((n, 2 * (m + 1) * (m + 1), 8), np.float32)
        tw = ptw = 0
        tp = pp = None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw, tp = table.shape[1], ctypes.c_void_p(table.ctypes.data)
        if ptable is not None:
            ptable = np.ascontiguousarray(ptable, np.float32)
            ptw, pp = ptable.shape[1], ctypes.c_void_p(ptable.ctypes.data)
        lib().orc_vcm_samples(self._h, ctypes.byref(p), tp, tw, pp, ptw, first, n, _p(out), nthreads)
        return out

    def pm_samples(self, first, n, radius, num_photons, max_num_vertices=-1, finalgather=1, seed=42, pseed=43,
                   raster_mode=RASTER_STRATIFIED, table=None, ptable=None, nthreads=8):
        """renderer::pm: (n, 16) float32 per eye sample: raster x, y, emission
        C rgb, emission count, density C rgb, photons gathered, vertex count
        at the gather vertex, 0 x 5."""
        p = orc_pm_params(max_num_vertices, seed & 0xFFFFFFFF, seed >> 32, pseed & 0xFFFFFFFF, pseed >> 32,
                          raster_mode, finalgather, num_photons, radius)
        out = np.zeros((n, 16), np.float32)
        tw = ptw = 0
        tp = pp = None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw, tp = table.shape[1], ctypes.c_void_p(table.ctypes.data)
        if ptable is not None:
            ptable = np.ascontiguousarray(ptable, np.float32)
            ptw, pp = ptable.shape[1], ctypes.c_void_p(ptable.ctypes.data)
        lib().orc_pm_samples(self._h, ctypes.byref(p), tp, tw, pp, ptw, first, n, _p(out), nthreads)
        return out

    def ppm_points(self, n, num_passes, num_photons, max_num_vertices, initial_radius=0.1, alpha=0.7, seed=42,
                   pseed=43, raster_mode=RASTER_STRATIFIED, table=None, ptable=None, nthreads=8):
        """renderer::ppm: (n, 16) float32 per measurement point: raster x, y,
        C rgb, emission rgb, tau rgb, N, radius, made-a-point flag, 0, 0."""
        p = orc_ppm_params(max_num_vertices, seed & 0xFFFFFFFF, seed >> 32, pseed & 0xFFFFFFFF, pseed >> 32,
                           raster_mode, num_passes, num_photons, initial_radius, alpha)
        out = np.zeros((n, 16), np.float32)
        tw = ptw = 0
        tp = pp = None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw, tp = table.shape[1], ctypes.c_void_p(table.ctypes.data)
        if ptable is not None:
            ptable = np.ascontiguousarray(ptable, np.float32)
            ptw, pp = ptable.shape[1], ctypes.c_void_p(ptable.ctypes.data)
        lib().orc_ppm_points(self._h, ctypes.byref(p), tp, tw, pp, ptw, n, _p(out), nthreads)
        return out

    def sppm_pixels(self, num_passes, num_photons, max_num_vertices, initial_radius=0.1, alpha=0.7, seed=42,
                    pseed=43, table=None, ptable=None, nthreads=8):
        """renderer::sppm: (H W, 16) float32 per pixel: film C rgb, emission
        sum rgb, tau rgb, N, radius, passes with a point, 0 x 4."""
        p = orc_ppm_params(max_num_vertices, seed & 0xFFFFFFFF, seed >> 32, pseed & 0xFFFFFFFF, pseed >> 32,
                           RASTER_STRATIFIED, num_passes, num_photons, initial_radius, alpha)
        out = np.zeros((self.height * self.width, 16), np.float32)
        tw = ptw = 0
        tp = pp = None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw, tp = table.shape[1], ctypes.c_void_p(table.ctypes.data)
        if ptable is not None:
            ptable = np.ascontiguousarray(ptable, np.float32)
            ptw, pp = ptable.shape[1], ctypes.c_void_p(ptable.ctypes.data)
        lib().orc_sppm_pixels(self._h, ctypes.byref(p), tp, tw, pp, ptw, _p(out), nthreads)
        return out

    def ltdirect_samples(self, first, n, max_num_vertices=-1, seed=42, table=None, nthreads=8):
        """renderer::ltdirect: (n, M, 8) float32, M = max_num_vertices (32
        unbounded); record k: raster x, y, C rgb, contributed flag, 0, 0."""
        m = 32 if max_num_vertices == -1 else max_num_vertices
        p = orc_lt_params(max_num_vertices, seed & 0xFFFFFFFF, seed >> 32)
        out = np.zeros((n, m, 8), np.float32)
        tw, tp = 0, None
        if table is not None:
            table = np.ascontiguousarray(table, np.float32)
            tw, tp = table.shape[1], ctypes.c_void_p(table.ctypes.data)
        lib().orc_ltdirect_samples(self._h, ctypes.byref(p), tp, tw, first, n, _p(out), nthreads)
        return out

    def raycast(self):
        """renderer::raycast: (H, W, 3) float32, row 0 = bottom."""
        out = np.zeros((self.height, self.width, 3), np.float32)
        lib().orc_raycast(self._h, _p(out))
        return out

    def film(self, samples):
        f = np.zeros((self.height, self.width, 3), np.float64)
        s = np.ascontiguousarray(samples, np.float32)
        lib().orc_film_splat(_p(s), s.shape[0], self.width, self.height, _p(f))
        return f


# ---------------------------------------------------------------------------
# learned directional sampler (orc_guide.c)

GUIDE_NPARAM = 25961
GUIDE_NOUT = 41
_guide_ready = False


def _glib():
    global _guide_ready
    so = lib()
    if not _guide_ready:
        vp, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
        so.orc_guide_features.argtypes = [vp, vp]
        so.orc_guide_features.restype = None
        so.orc_guide_forward_batch.argtypes = [vp, vp, u64, vp, i32]
        so.orc_guide_forward_batch.restype = None
        so.orc_guide_vmf_pdf.argtypes = [vp, f64, vp]
        so.orc_guide_vmf_pdf.restype = f64
        so.orc_guide_vmf_sample.argtypes = [vp, f64, f64, f64, vp]
        so.orc_guide_vmf_sample.restype = None
        so.orc_guide_mix_pdf.argtypes = [vp, vp, vp]
        so.orc_guide_mix_pdf.restype = f64
        so.orc_guide_mix_sample.argtypes = [vp, f64, f64, f64, vp]
        so.orc_guide_mix_sample.restype = i32
        so.orc_guide_loss.argtypes = [vp, u64, vp, vp, vp, vp, vp, i32]
        so.orc_guide_loss.restype = f64
        so.orc_guide_adam.argtypes = [vp, vp, vp, vp, u64, i32, f64, f64, f64, f64]
        so.orc_guide_adam.restype = None
        so.orc_bf16_round.argtypes = [f64]
        so.orc_bf16_round.restype = f64
        _guide_ready = True
    return so


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def guide_features(x):
    x = _d(x)
    f = np.zeros(168)
    _glib().orc_guide_features(_p(x), _p(f))
    return f


def guide_forward(params, x, bf16=False):
    """(n, 7) inputs -> (n, 41) outputs."""
    params, x = _d(params), _d(np.atleast_2d(x))
    y = np.zeros((x.shape[0], GUIDE_NOUT))
    _glib().orc_guide_forward_batch(_p(params), _p(x), x.shape[0], _p(y), int(bf16))
    return y


def vmf_pdf(mu, kappa, w):
    mu, w = _d(mu), _d(w)
    return _glib().orc_guide_vmf_pdf(_p(mu), float(kappa), _p(w))


def vmf_sample(mu, kappa, u1, u2):
    mu = _d(mu)
    o = np.zeros(3)
    _glib().orc_guide_vmf_sample(_p(mu), float(kappa), float(u1), float(u2), _p(o))
    return o


def mix_pdf(y, w):
    """(p_vmf(w), alpha) of the mixture encoded by outputs y."""
    y, w = _d(y), _d(w)
    a = np.zeros(1)
    p = _glib().orc_guide_mix_pdf(_p(y), _p(w), _p(a))
    return p, float(a[0])


def mix_sample(y, u, u1, u2):
    y = _d(y)
    o = np.zeros(3)
    k = _glib().orc_guide_mix_sample(_p(y), float(u), float(u1), float(u2), _p(o))
    return o, k


def guide_loss(params, x, dirs, pb, w, grad=True, bf16=False):
    params, x, dirs, pb, w = _d(params), _d(x), _d(dirs), _d(pb), _d(w)
    g = np.zeros(GUIDE_NPARAM) if grad else None
    L = _glib().orc_guide_loss(_p(params), x.shape[0], _p(x), _p(dirs), _p(pb), _p(w),
                               _p(g) if grad else None, int(bf16))
    return (L, g) if grad else L


def guide_adam(params, m, v, grad, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """In place on float64 arrays params, m, v."""
    _glib().orc_guide_adam(_p(params), _p(m), _p(v), _p(_d(grad)), params.shape[0], int(step), lr, beta1, beta2, eps)


def bf16_round(x):
    return _glib().orc_bf16_round(float(x))


def film_rgbe(film):
    """Film_HDR::Save's RGBE bytes of an (H, W, 3) film (row 0 = bottom), in
    file row order (top row first): (H, W, 4) uint8."""
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    lib().orc_film_rgbe(_p(f), w, h, _p(out))
    return out


def film_png(film):
    """Film_HDR::Save's PNG tone-mapped bytes, file row order: (H, W, 3)."""
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    out = np.zeros((h, w, 3), np.uint8)
    lib().orc_film_png(_p(f), w, h, _p(out))
    return out
