"""Accel: the batched closest-hit component over liblmx's compressed wide BVH.

Mirrors the reference's Accel3 interface (include/lightmetrica/accel3.h:43-68):
`build` plays Accel::Build (accel.h:79) over world-space triangles and
`intersect` is the batched form of Accel3::Intersect; each ray carries its own
[tmin, tmax] (Scene3::Intersect passes [1e-4, FLT_MAX], scene3.cpp:461).

Rays are float32 arrays of shape (n, 8): o.xyz, tmin, d.xyz, tmax.
Hits are float32 arrays of shape (n, 4): t, b0, b1, id (id bit pattern as
uint32; MISS = 0xffffffff).
"""
import ctypes

import numpy as np

from ._lib import lib, check, lmx_build_opts, lmx_accel_info

ISECT_TRIACCEL = 0
ISECT_WATERTIGHT = 1
MISS = 0xFFFFFFFF
FLT_MAX = np.float32(np.finfo(np.float32).max)


# struct WNode (csrc/wbvh.h), 128 bytes
WNODE_DTYPE = np.dtype([
    ("o", np.float32, 3), ("e", np.uint8, 3), ("imask", np.uint8),
    ("child_base", np.uint32), ("tri_base", np.uint32), ("tvalid", np.uint32), ("reserved", np.uint32),
    ("q", np.uint8, (3, 8, 2))])
assert WNODE_DTYPE.itemsize == 80
WNODE_QBIAS = 151     # grid step of a node axis = 2^(e - WNODE_QBIAS)


def _c(a):
    return ctypes.c_void_p(a.ctypes.data)


class Accel:
    """Owns one lmx_accel handle."""

    def __init__(self, verts, tris, isect=ISECT_WATERTIGHT, nthreads=0):
        verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        self.nvert = verts.shape[0]
        self.ntri = tris.shape[0]
        self.isect = isect
        opts = lmx_build_opts(isect, nthreads)
        h = ctypes.c_void_p()
        check(lib.lmx_accel_build(_c(verts), self.nvert, _c(tris), self.ntri, ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self._free = lib.lmx_accel_free     # kept for use during interpreter shutdown

    def close(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self):
        i = lmx_accel_info()
        check(lib.lmx_accel_info_get(self._h, ctypes.byref(i)))
        return {"num_nodes": i.num_nodes, "num_tris": i.num_tris, "max_depth": i.max_depth,
                "isect": i.isect, "node_bytes": i.node_bytes, "tri_bytes": i.tri_bytes,
                "stack_overflow": i.stack_overflow}

    def export(self):
        """(nodes, tris) as structured numpy arrays (layouts of csrc/wbvh.h)."""
        info = self.info()
        nodes = np.zeros(info["num_nodes"], dtype=WNODE_DTYPE)
        tris = np.zeros((info["num_tris"], 12), dtype=np.float32)
        check(lib.lmx_accel_export(self._h, _c(nodes), _c(tris)))
        return nodes, tris

    def upload(self):
        check(lib.lmx_accel_upload(self._h))

    def intersect(self, rays, hits=None, stream=None):
        """GPU closest hit.  `rays`: torch float32 tensor (n, 8) on the GPU.
        Returns (and fills) a torch float32 tensor (n, 4).  Asynchronous on
        `stream` (default: torch's current stream)."""
        import torch
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous()
        assert rays.dim() == 2 and rays.shape[1] == 8
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        assert hits.is_cuda and hits.dtype == torch.float32 and hits.is_contiguous() and hits.shape == (n, 4)
        if stream is None:
            stream = torch.cuda.current_stream(rays.device)
        check(lib.lmx_accel_intersect(self._h, ctypes.c_void_p(rays.data_ptr()), n,
                                      ctypes.c_void_p(hits.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
        return hits

    def intersect_noalloc(self, rays, hits, stream=None):
        """lmx_accel_intersect_noalloc: the same query with no device
        allocation; the accel must have been uploaded and `hits` (n, 4) is
        the caller's."""
        import torch
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous()
        assert rays.dim() == 2 and rays.shape[1] == 8
        n = rays.shape[0]
        assert hits.is_cuda and hits.dtype == torch.float32 and hits.is_contiguous() and hits.shape == (n, 4)
        if stream is None:
            stream = torch.cuda.current_stream(rays.device)
        check(lib.lmx_accel_intersect_noalloc(self._h, ctypes.c_void_p(rays.data_ptr()), n,
                                              ctypes.c_void_p(hits.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
        return hits

    def occluded(self, rays, hits=None, stream=None):
        """GPU any-hit query: returns (and fills) a torch float32 tensor
        (n, 4) whose id column is MISS exactly for unoccluded rays."""
        import torch
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous()
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        if stream is None:
            stream = torch.cuda.current_stream(rays.device)
        check(lib.lmx_accel_occluded(self._h, ctypes.c_void_p(rays.data_ptr()), n,
                                     ctypes.c_void_p(hits.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
        return hits

    def intersect_host(self, rays, nthreads=0):
        """Host walker over the same BVH (numpy in, numpy out)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), dtype=np.float32)
        check(lib.lmx_accel_intersect_host(self._h, _c(rays), rays.shape[0], _c(hits), nthreads))
        return hits

    def count(self, rays, nthreads=0, anyhit=False):
        """(node visits, triangle tests) the traversal kernel performs (the
        occlusion kernel with anyhit=True)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        nv = ctypes.c_uint64()
        nt = ctypes.c_uint64()
        fn = lib.lmx_accel_count_occluded if anyhit else lib.lmx_accel_count
        check(fn(self._h, _c(rays), rays.shape[0], ctypes.byref(nv), ctypes.byref(nt), nthreads))
        return nv.value, nt.value


def hit_ids(hits):
    """uint32 view of the id column of a (n, 4) float32 hit array."""
    h = np.ascontiguousarray(hits, dtype=np.float32)
    return h.view(np.uint32)[:, 3]
