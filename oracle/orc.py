"""ctypes binding of the CPU oracle (liborc.so) and of the reference build
(_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "liborc.so")
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
