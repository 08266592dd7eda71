"""GPU parity of the wide-BVH traversal kernel against the CPU oracle.

Bar: hit triangle index, t and barycentrics bit-exact against the oracle's
brute force under the same triangle test and the canonical tie rule
(smallest id among equal t), for both the reference-compatible TriAccel test
and the watertight test.  At the benchmark's full size (1,310,720 triangles,
2048^2 coherent + 2048^2 incoherent rays) the kernel is compared bit-exactly
with the host walker over the same BVH on every ray and with the oracle's
brute force on a seeded sample of rays (watertight mode), and in TriAccel
mode with the oracle's restatement of the reference's QBVH on every ray.
"""
import numpy as np
import pytest

import orc
from lmx import accel, scenes

pytestmark = pytest.mark.gpu


def _gpu_hits(a, rays, dev):
    import torch
    r = torch.from_numpy(rays).to(dev)
    h = a.intersect(r)
    torch.cuda.synchronize()
    return h.cpu().numpy()


@pytest.mark.parametrize("mode", [accel.ISECT_TRIACCEL, accel.ISECT_WATERTIGHT])
def test_small_mesh_bitexact_vs_bruteforce(gpu, mode):
    v, f = scenes.displaced_icosphere(4, seed=3)
    coh, inc = scenes.traversal_rays(v, res=64, n_incoherent=8192)
    rays = np.concatenate([coh, inc])
    a = accel.Accel(v, f, isect=mode)
    hg = _gpu_hits(a, rays, gpu)
    ho = orc.closest_hit_naive(v, f, rays, test=mode, tie=orc.TIE_MIN_ID)
    assert (orc.hit_ids(ho) != orc.MISS).sum() > 1000
    np.testing.assert_array_equal(hg.view(np.uint32), ho.view(np.uint32))


@pytest.mark.parametrize("n", [0, 1, 5, 63, 64, 65, 1000])
def test_ragged_ray_counts(gpu, n):
    v, f = scenes.displaced_icosphere(3, seed=5)
    _, inc = scenes.traversal_rays(v, res=4, n_incoherent=max(n, 1), seed=9)
    rays = inc[:n]
    a = accel.Accel(v, f)
    if n == 0:
        import torch
        h = a.intersect(torch.zeros((0, 8), dtype=torch.float32, device=gpu))
        assert h.shape == (0, 4)
        return
    hg = _gpu_hits(a, rays, gpu)
    ho = orc.closest_hit_naive(v, f, rays, test=orc.TEST_WATERTIGHT, tie=orc.TIE_MIN_ID)
    np.testing.assert_array_equal(hg.view(np.uint32), ho.view(np.uint32))


@pytest.mark.parametrize("scale,shift", [(1e-12, 0.0), (1e12, 0.0), (1.0, 1e6), (1e-6, 1.0)])
def test_extreme_scales_bitexact_vs_bruteforce(gpu, scale, shift):
    """Scene sizes from 1e-12 to 1e12 and far from the origin (ray origins
    and tmin scaled with the scene): the kernel's byte-grid box test, with
    its binary16 subnormal operands, still finds exactly the brute force's
    hits."""
    v, f = scenes.displaced_icosphere(4, seed=3)
    coh, inc = scenes.traversal_rays(v, res=48, n_incoherent=4096)
    rays = np.concatenate([coh, inc]).copy()
    v = (v.astype(np.float64) * scale + shift).astype(np.float32)
    rays[:, 0:3] = (rays[:, 0:3].astype(np.float64) * scale + shift).astype(np.float32)
    rays[:, 3] = np.float32(rays[0, 3] * scale)
    a = accel.Accel(v, f)
    hg = _gpu_hits(a, rays, gpu)
    ho = orc.closest_hit_naive(v, f, rays, test=orc.TEST_WATERTIGHT, tie=orc.TIE_MIN_ID)
    assert (orc.hit_ids(ho) != orc.MISS).sum() > 1000
    np.testing.assert_array_equal(hg.view(np.uint32), ho.view(np.uint32))


@pytest.mark.parametrize("n", [327679, 327680, 327681, 400003])
def test_fixed_and_claimed_chunks(gpu, n):
    """Ray counts around 5,120 waves x 64 rays, where the waves' fixed first
    chunks stop covering the batch and the rest is claimed through the
    region cursors: every ray is traced exactly once (bit-exact against the
    host walker over the same BVH)."""
    v, f = scenes.displaced_icosphere(4, seed=6)
    _, rays = scenes.traversal_rays(v, res=4, n_incoherent=n, seed=10)
    a = accel.Accel(v, f)
    hg = _gpu_hits(a, rays, gpu)
    hh = a.intersect_host(rays)
    assert (accel.hit_ids(hh) != accel.MISS).sum() > n // 10
    np.testing.assert_array_equal(hg.view(np.uint32), hh.view(np.uint32))


def test_full_size_traversal_benchmark(gpu):
    v, f = scenes.traversal_scene()
    assert f.shape[0] == 1310720
    coh, inc = scenes.traversal_rays(v)
    rays = np.concatenate([coh, inc])
    assert rays.shape[0] == 8388608
    a = accel.Accel(v, f)
    hg = _gpu_hits(a, rays, gpu)
    hh = a.intersect_host(rays)
    np.testing.assert_array_equal(hg.view(np.uint32), hh.view(np.uint32))
    # seeded sample against brute force over all 1.3M triangles
    idx = np.random.default_rng(0).choice(rays.shape[0], 512, replace=False)
    ho = orc.closest_hit_naive(v, f, rays[idx], test=orc.TEST_WATERTIGHT, tie=orc.TIE_MIN_ID, nthreads=8)
    np.testing.assert_array_equal(hg[idx].view(np.uint32), ho.view(np.uint32))
    # every coherent ray aimed at the sphere's silhouette interior hits it
    ids = accel.hit_ids(hg)
    assert (ids[: coh.shape[0]] != accel.MISS).mean() > 0.2
    assert (ids[coh.shape[0]:] != accel.MISS).mean() > 0.3


def test_full_size_triaccel_vs_reference_qbvh_every_ray(gpu):
    """The benchmark's full size in the reference's TriAccel mode
    (triaccel.h:93-151) against the oracle's restatement of the reference's
    own accel, the 4-wide QBVH (accel_qbvh.cpp:392-481, a later equal-t hit
    replaces an earlier one), on all 8,388,608 rays: the same triangle on
    every ray except exact t ties (two triangles hit at the same t, where the
    kernel keeps the smallest id and the QBVH its traversal order), and where
    the triangle agrees t and barycentrics bit for bit (the same FP32
    operations in the same order).  Ties are checked with the FP64 solve:
    both triangles are hit at that t."""
    from test_oracle import _fp64_solve
    v, f = scenes.traversal_scene()
    coh, inc = scenes.traversal_rays(v)
    rays = np.concatenate([coh, inc])
    a = accel.Accel(v, f, isect=accel.ISECT_TRIACCEL)
    hg = _gpu_hits(a, rays, gpu)
    ho = orc.QBVH(v, f).intersect(rays, nthreads=16)
    ig, io = accel.hit_ids(hg), orc.hit_ids(ho)
    same = ig == io
    assert (ig != accel.MISS).mean() > 0.5
    np.testing.assert_array_equal(hg[same].view(np.uint32), ho[same].view(np.uint32))
    bad = np.nonzero(~same)[0]
    print(f"TriAccel mode, {rays.shape[0]} rays: {len(bad)} with another triangle")
    assert len(bad) <= 1e-5 * rays.shape[0]
    for k in bad:
        assert ig[k] != accel.MISS and io[k] != accel.MISS, (k, ig[k], io[k])
        assert hg[k, 0] == ho[k, 0], (k, hg[k, 0], ho[k, 0])          # an exact t tie
        for tri in (ig[k], io[k]):
            t, b0, b1 = _fp64_solve(v, f, rays[k:k + 1], np.array([tri]))
            assert abs(t[0] - hg[k, 0]) <= 1e-5 * abs(t[0])
            assert min(b0[0], b1[0], 1.0 - b0[0] - b1[0]) > -1e-5
    assert a.info()["stack_overflow"] == 0


@pytest.mark.parametrize("simple2", [False, True])
@pytest.mark.parametrize("mode", [accel.ISECT_TRIACCEL, accel.ISECT_WATERTIGHT])
def test_accel3_analytic_quads_on_gpu(gpu, simple2, mode):
    """src/lightmetrica-test/test_accel3.cpp:272-345 through the HIP kernel:
    p, gn and uv within EpsLarge = 1e-3 of the analytic answer."""
    from test_oracle import (SIMPLE_F, SIMPLE_UV, SIMPLE_V, SIMPLE2_F, SIMPLE2_UV, SIMPLE2_V,
                             hit_geometry, quad_rays)
    v, f, uv = (SIMPLE2_V, SIMPLE2_F, SIMPLE2_UV) if simple2 else (SIMPLE_V, SIMPLE_F, SIMPLE_UV)
    rays, expect = quad_rays(simple2)
    a = accel.Accel(v, f, isect=mode)
    hg = _gpu_hits(a, rays, gpu)
    assert (accel.hit_ids(hg) != accel.MISS).all()
    p, gn, tuv = hit_geometry(v, f, uv, rays, hg)
    np.testing.assert_allclose(p, expect, atol=1e-3)
    n_exp = np.array([0, 0, 1], np.float32) if not simple2 else np.array([1, 0, 1], np.float32) / np.sqrt(2)
    np.testing.assert_allclose(gn, np.broadcast_to(n_exp, gn.shape), atol=1e-3)
    np.testing.assert_allclose(tuv, expect[:, 0:2], atol=1e-3)
    ho = orc.closest_hit_naive(v, f, rays, test=mode, tie=orc.TIE_MIN_ID)
    np.testing.assert_array_equal(hg.view(np.uint32), ho.view(np.uint32))


@pytest.mark.parametrize("tmax_scale", [0.5, 1.0, 2.0])
def test_occlusion_matches_closest_hit(gpu, tmax_scale):
    """Any-hit traversal: a ray is occluded on [tmin, tmax] exactly when the
    closest hit lies in that range, and the reported triangle really is hit
    there.  Mixed lengths exercise lanes that finish early and are refilled."""
    import torch
    v, f = scenes.displaced_icosphere(5, seed=13)
    coh, inc = scenes.traversal_rays(v, res=96, n_incoherent=40000, seed=21)
    rays = np.concatenate([coh, inc]).copy()
    hc = orc.closest_hit_naive(v, f, rays, test=orc.TEST_WATERTIGHT, tie=orc.TIE_MIN_ID)
    ids = orc.hit_ids(hc)
    rng = np.random.default_rng(4)
    # finite tmax around the closest hit distance (and beyond for misses)
    t_ref = np.where(ids != orc.MISS, hc[:, 0], 3.0)
    rays[:, 7] = (t_ref * tmax_scale * rng.uniform(0.7, 1.3, rays.shape[0])).astype(np.float32)
    a = accel.Accel(v, f)
    hg = a.occluded(torch.from_numpy(rays).to(gpu))
    torch.cuda.synchronize()
    hg = hg.cpu().numpy()
    occ_expect = (ids != orc.MISS) & (hc[:, 0] <= rays[:, 7])
    occ = accel.hit_ids(hg) != accel.MISS
    np.testing.assert_array_equal(occ, occ_expect)
    # the reported triangle is hit by that ray within range
    k = np.nonzero(occ)[0][:2000]
    for i in k[:200]:
        tri = int(accel.hit_ids(hg)[i])
        one = orc.closest_hit_naive(v, f[tri:tri + 1], rays[i:i + 1], test=orc.TEST_WATERTIGHT)
        assert orc.hit_ids(one)[0] == 0
