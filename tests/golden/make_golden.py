"""Writes the golden fixtures under tests/golden/ from the reference's own
code: the header-only hot-path pieces (include/lightmetrica/triaccel.h,
bound.h, sampler.h, dist.h) and the vendored dSFMT, compiled from
/root/reference by `make -C oracle ref` into oracle/_ref/libref.so.

Run in the development container (where the reference is present):
    make -C oracle ref && python tests/golden/make_golden.py
The fixtures are inputs and the reference's outputs only.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import orc  # noqa: E402


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def triangles_and_rays(rng, n):
    tri = rng.normal(size=(n, 3, 3)).astype(np.float32)
    # axis-aligned, sliver and degenerate cases
    tri[0] = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    tri[1] = [[0, 0, 0], [0, 1, 0], [0, 0, 1]]
    tri[2] = [[0, 0, 0], [1, 1, 1], [2, 2, 2]]          # collinear: k = 3
    tri[3] = [[1, 1, 1], [1, 1, 1], [1, 1, 1]]          # point
    tri[4] = [[0, 0, 0], [1, 0, 0], [1, 1e-7, 0]]       # sliver
    o = rng.normal(size=(n, 3)).astype(np.float32) * 3
    c = tri.mean(axis=1)
    d = (c - o + rng.normal(size=(n, 3)) * 0.3).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[5] = [0, 0, 1]
    d[6] = [1, 0, 0]
    return tri, o.astype(np.float32), d.astype(np.float32)


def main():
    ref = orc.ref()
    rng = np.random.default_rng(20251202)
    out = {}

    # TriAccel load + intersect (triaccel.h:50-151)
    n = 2048
    tri, o, d = triangles_and_rays(rng, n)
    rec = np.zeros((n, 12), dtype=np.uint32)
    load_rc = np.zeros(n, dtype=np.int32)
    hit = np.zeros(n, dtype=np.int32)
    tuv = np.zeros((n, 3), dtype=np.float32)
    for i in range(n):
        r = np.zeros(12, dtype=np.uint32)
        load_rc[i] = ref.ref_triaccel_load(_p(tri[i, 0]), _p(tri[i, 1]), _p(tri[i, 2]), _p(r))
        rec[i] = r
        u = np.zeros(1, np.float32); v = np.zeros(1, np.float32); t = np.zeros(1, np.float32)
        hit[i] = ref.ref_triaccel_intersect(_p(r), _p(o[i]), _p(d[i]), ctypes.c_float(0.0),
                                            ctypes.c_float(np.finfo(np.float32).max), _p(u), _p(v), _p(t))
        tuv[i] = (t[0], u[0], v[0])
    out.update(triaccel_tri=tri, triaccel_o=o, triaccel_d=d, triaccel_rec=rec, triaccel_load_rc=load_rc,
               triaccel_hit=hit, triaccel_tuv=tuv)

    # Bound::Intersect (bound.h:67-89), including zero direction components
    m = 2048
    lo = rng.normal(size=(m, 3)).astype(np.float32)
    hi = (lo + rng.random((m, 3)) * 2).astype(np.float32)
    bo = (rng.normal(size=(m, 3)) * 3).astype(np.float32)
    bd = rng.normal(size=(m, 3)).astype(np.float32)
    bd[rng.random((m, 3)) < 0.1] = 0.0
    bres = np.zeros(m, dtype=np.int32)
    for i in range(m):
        bres[i] = ref.ref_bound_intersect(_p(lo[i]), _p(hi[i]), _p(bo[i]), _p(bd[i]), ctypes.c_float(1e-4),
                                          ctypes.c_float(np.finfo(np.float32).max))
    out.update(bound_lo=lo, bound_hi=hi, bound_o=bo, bound_d=bd, bound_hit=bres)

    # sampler.h:44-101
    k = 2048
    u = rng.random((k, 2)).astype(np.float32)
    u[0] = [0.5, 0.5]
    u[1] = [0.0, 0.0]
    u[2] = [1.0 - 2 ** -24, 0.5]
    disk = np.zeros((k, 2), np.float32); hemi = np.zeros((k, 3), np.float32)
    sph = np.zeros((k, 3), np.float32); trib = np.zeros((k, 2), np.float32)
    for i in range(k):
        ref.ref_sample_concentric_disk(_p(u[i]), _p(disk[i]))
        ref.ref_sample_cosine_hemisphere(_p(u[i]), _p(hemi[i]))
        ref.ref_sample_uniform_sphere(_p(u[i]), _p(sph[i]))
        ref.ref_sample_uniform_triangle(_p(u[i]), _p(trib[i]))
    out.update(sample_u=u, sample_disk=disk, sample_hemi=hemi, sample_sphere=sph, sample_tri=trib)

    # Distribution1D::SampleReuse (dist.h:63-82)
    w = rng.random(7).astype(np.float32) + 0.1
    du = rng.random(512).astype(np.float32)
    didx = np.zeros(512, np.int32); du2 = np.zeros(512, np.float32)
    for i in range(512):
        x = np.zeros(1, np.float32)
        didx[i] = ref.ref_dist_sample_reuse(_p(w), 7, ctypes.c_float(du[i]), _p(x))
        du2[i] = x[0]
    out.update(dist_w=w, dist_u=du, dist_idx=didx, dist_u2=du2)

    # dSFMT-19937 streams as Random::Next / NextUInt (random.cpp:65-90)
    for seed in (1, 42):
        nx = np.zeros(2048, np.float64); nu = np.zeros(2048, np.uint32)
        ref.ref_dsfmt_stream(seed, 2048, _p(nx), _p(nu))
        out[f"dsfmt_next_{seed}"] = nx
        out[f"dsfmt_uint_{seed}"] = nu

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
