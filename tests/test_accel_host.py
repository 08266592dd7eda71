"""liblmx on the CPU: the C ABI, the wide-BVH builder and the host walker.

* every symbol include/lmx.h declares is exported by liblmx.so;
* the host walker over the compressed BVH equals the oracle's brute force
  bit for bit (hit index, t, barycentrics) under both triangle tests;
* structural invariants of the compressed BVH: every triangle sits in
  exactly one leaf, quantised child boxes contain their triangles and
  subtrees, leaves hold 1..3 triangles;
* error paths return codes with messages (no exception crosses the ABI).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import orc
from lmx import accel, scenes
from lmx._lib import LIB_PATH, LmxError, SIGNATURES, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lmx.h")).read()
    return sorted(set(re.findall(r"\b(lmx_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lmx_\w+)", out))
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    assert set(decl) <= exported, set(decl) - exported
    assert set(decl) == set(SIGNATURES), "ctypes signatures out of sync with include/lmx.h"
    for name in decl:
        assert getattr(lib, name) is not None


MESHES = {
    "quads": (np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, -1], [1, 0, -1], [1, 1, -1],
                        [0, 1, -1]], np.float32), np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)),
    "one": (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32)),
}


def mesh(name):
    if name in MESHES:
        return MESHES[name]
    if name == "sphere4":
        return scenes.displaced_icosphere(4, seed=3)
    if name == "sphere6":
        return scenes.displaced_icosphere(6, seed=4)
    if name == "degenerate":
        v, f = scenes.displaced_icosphere(3, seed=5)
        f = f.copy()
        f[::7, 2] = f[::7, 1]        # collapse every 7th triangle to a segment
        return v, f
    if name == "coincident":
        # many copies of one triangle: all centroids coincide
        v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
        return v, np.tile(np.array([[0, 1, 2]], np.uint32), (40, 1))
    raise KeyError(name)


def rays_for(v, n=6000, seed=1):
    _, inc = scenes.traversal_rays(v, res=4, n_incoherent=n, seed=seed)
    lo, hi = v.min(axis=0), v.max(axis=0)
    c = 0.5 * (lo + hi)
    eye = c + np.array([0.3, 0.2, 2.5], np.float32) * max(1.0, float((hi - lo).max()))
    cam = scenes.pinhole_rays(40, 40, eye, c, (0, 1, 0), 50.0)
    return np.concatenate([inc, cam])


@pytest.mark.parametrize("name", ["quads", "one", "sphere4", "sphere6", "degenerate", "coincident"])
@pytest.mark.parametrize("mode", [accel.ISECT_TRIACCEL, accel.ISECT_WATERTIGHT])
def test_host_walker_bitexact_vs_bruteforce(name, mode):
    v, f = mesh(name)
    rays = rays_for(v)
    a = accel.Accel(v, f, isect=mode)
    hh = a.intersect_host(rays)
    ho = orc.closest_hit_naive(v, f, rays, test=mode, tie=orc.TIE_MIN_ID)
    np.testing.assert_array_equal(hh.view(np.uint32), ho.view(np.uint32))


@pytest.mark.parametrize("scale,shift", [(1e-12, 0.0), (1e12, 0.0), (1.0, 1e6), (1e-6, 1.0)])
def test_host_walker_bitexact_at_extreme_scales(scale, shift):
    """The byte grid's power-of-two steps and the binary16 subnormals of the
    box test at scene sizes from 1e-12 to 1e12 (beyond them the watertight
    test's own FP32 products leave range) and far from the origin: the
    walker still finds exactly the brute force's hits (ray origins and tmin
    scaled with the scene)."""
    v, f = mesh("sphere4")
    v = (v.astype(np.float64) * scale + shift).astype(np.float32)
    rays = rays_for(mesh("sphere4")[0], n=3000).copy()
    rays[:, 0:3] = (rays[:, 0:3].astype(np.float64) * scale + shift).astype(np.float32)
    rays[:, 3] = np.float32(rays[0, 3] * scale)
    a = accel.Accel(v, f)
    hh = a.intersect_host(rays)
    ho = orc.closest_hit_naive(v, f, rays, test=orc.TEST_WATERTIGHT, tie=orc.TIE_MIN_ID)
    assert (orc.hit_ids(ho) != orc.MISS).sum() > 500
    np.testing.assert_array_equal(hh.view(np.uint32), ho.view(np.uint32))


def test_counting_walker_matches_host_walk_cost():
    v, f = mesh("sphere6")
    rays = rays_for(v, n=3000)
    a = accel.Accel(v, f)
    nv1, nt1 = a.count(rays, nthreads=1)
    nv8, nt8 = a.count(rays, nthreads=8)
    assert (nv1, nt1) == (nv8, nt8)
    assert nv1 >= rays.shape[0]           # at least the root per ray
    # splitting the batch changes nothing (counts are per ray)
    nva, nta = a.count(rays[:1000])
    nvb, ntb = a.count(rays[1000:])
    assert (nva + nvb, nta + ntb) == (nv1, nt1)


def test_occlusion_counting_walker():
    """The any-hit count (lmx_accel_count_occluded): per ray never more node
    visits than the closest-hit walk, the same for rays that hit nothing,
    and exactly one visit more than the nodes opened before the first hit
    is impossible to undercut: at least the root per ray."""
    v, f = mesh("sphere6")
    rays = rays_for(v, n=3000)
    a = accel.Accel(v, f)
    hits = a.intersect_host(rays)
    miss = accel.hit_ids(hits) == accel.MISS
    assert 0 < miss.sum() < rays.shape[0]
    nvc, ntc = a.count(rays)
    nva, nta = a.count(rays, anyhit=True)
    assert nva <= nvc and nta <= ntc and nva >= rays.shape[0]
    assert a.count(rays[miss]) == a.count(rays[miss], anyhit=True)
    # per ray: the any-hit walk stops early only on rays that hit
    for i in np.nonzero(~miss)[0][:200]:
        c1, _ = a.count(rays[i:i + 1])
        c2, _ = a.count(rays[i:i + 1], anyhit=True)
        assert 1 <= c2 <= c1


def decode_boxes(node):
    o = node["o"].astype(np.float64)
    sc = np.ldexp(1.0, node["e"].astype(np.int64) - accel.WNODE_QBIAS)
    lo = o[:, None] + node["q"][:, :, 0].astype(np.float64) * sc[:, None]
    hi = o[:, None] + node["q"][:, :, 1].astype(np.float64) * sc[:, None]
    return lo, hi      # (3, 8) each


@pytest.mark.parametrize("name", ["sphere6", "degenerate", "coincident", "one"])
def test_bvh_structure_invariants(name):
    v, f = mesh(name)
    a = accel.Accel(v, f)
    nodes, tris = a.export()
    info = a.info()
    assert info["num_tris"] == f.shape[0]
    ids = tris.view(np.uint32)[:, 3]
    assert np.array_equal(np.sort(ids), np.arange(f.shape[0], dtype=np.uint32))   # each triangle once
    # watertight records keep the vertices exactly
    np.testing.assert_array_equal(tris[:, 0:3], v[f[ids, 0]])
    np.testing.assert_array_equal(tris[:, 4:7], v[f[ids, 1]])
    np.testing.assert_array_equal(tris[:, 8:11], v[f[ids, 2]])
    # geometry bounds of every subtree, children before parents (a node's
    # children always have larger indices in the breadth-first layout)
    sub_lo = np.full((len(nodes), 3), np.inf)
    sub_hi = np.full((len(nodes), 3), -np.inf)
    seen_children = []
    for ni in range(len(nodes) - 1, -1, -1):
        node = nodes[ni]
        lo, hi = decode_boxes(node)
        rank = 0
        for s in range(8):
            internal = (node["imask"] >> s) & 1
            tv = int(node["tvalid"])
            c = bin((tv >> (3 * s)) & 7).count("1")                 # triangles of leaf slot s
            off = bin(tv & ((1 << (3 * s)) - 1)).count("1")        # triangles of the leaf slots before it
            if internal:
                child = node["child_base"] + rank
                rank += 1
                assert child > ni
                seen_children.append(child)
                glo, ghi = sub_lo[child], sub_hi[child]
            elif c:
                assert (tv >> (3 * s)) & 7 in (1, 3, 7)                 # the first c bits of the group
                p = tris[node["tri_base"] + off: node["tri_base"] + off + c].reshape(-1, 3, 4)[:, :, 0:3]
                p = p.reshape(-1, 3).astype(np.float64)
                glo, ghi = p.min(axis=0), p.max(axis=0)
            else:
                continue
            # the quantised slot box contains all geometry below the slot
            assert (glo >= lo[:, s]).all() and (ghi <= hi[:, s]).all(), (ni, s)
            sub_lo[ni] = np.minimum(sub_lo[ni], glo)
            sub_hi[ni] = np.maximum(sub_hi[ni], ghi)
    # every node but the root is the child of exactly one node
    assert sorted(seen_children) == list(range(1, len(nodes)))


def test_build_errors_are_reported():
    v, f = mesh("one")
    with pytest.raises(LmxError, match="out of range"):
        accel.Accel(v, np.array([[0, 1, 3]], np.uint32))
    with pytest.raises(LmxError, match="empty mesh"):
        accel.Accel(v, np.zeros((0, 3), np.uint32))
    with pytest.raises(LmxError, match="non-finite"):
        accel.Accel(np.array([[0, 0, np.nan], [1, 0, 0], [0, 1, 0]], np.float32), f)
    with pytest.raises(LmxError, match="unknown triangle test"):
        accel.Accel(v, f, isect=7)
    h = ctypes.c_void_p()
    assert lib.lmx_accel_build(None, 0, None, 0, None, ctypes.byref(h)) == -1
    assert "NULL" in lib.lmx_last_error().decode()
    assert lib.lmx_accel_intersect(None, None, 0, None, None) == -1


def test_intersect_without_gpu_reports_hip_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    v, f = mesh("one")
    a = accel.Accel(v, f)
    buf = (ctypes.c_float * 8)()
    rc = lib.lmx_accel_intersect(a._h, ctypes.cast(buf, ctypes.c_void_p), 1, ctypes.cast(buf, ctypes.c_void_p), None)
    assert rc == -3
    assert "hip" in lib.lmx_last_error().decode().lower()
