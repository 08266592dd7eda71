"""Benchmark: closest-hit traversal throughput (Mrays/s) on one MI355X.

Workload: the BVH traversal microbenchmark -- a 1,310,720-triangle icosphere
(8 subdivisions) displaced radially by seeded 3-octave value noise
(amplitude 0.1), and 8,388,608 rays: 2048^2 coherent pinhole primary rays
plus 2048^2 incoherent rays (origins uniform in the mesh bounds, directions
uniform on the sphere); closest hit, t in [1e-4, FLT_MAX].  One step = one
lmx_accel_intersect launch over all 8,388,608 rays, inputs resident in HBM.

Prints one JSON line.  `roofline` prices the traversal kernel's algorithmic
bytes -- 32 B ray read + 16 B hit write per ray, plus 80 B per wide-node
visit and 48 B per triangle test as counted by the host walker over the same
BVH in the same order -- against 8 TB/s HBM, with the kernel's duration
measured by HIP events on the stream it runs on.  `cpu_baseline` is the
oracle's restatement of the reference's 4-wide SSE QBVH
(accel_qbvh.cpp) with its TriAccel test, on host threads, over a fixed sample
of the same rays.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lightmetrica-v2_amd"))

HBM_PEAK_GBS = 8000.0
VALU_CLOCK_HZ = 2.4e9        # MI355X peak engine clock
NODE_BYTES = 80
TRI_BYTES = 48
RAY_BYTES = 32 + 16


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def median(xs):
    xs = sorted(xs)
    k = len(xs)
    return xs[k // 2] if k % 2 else 0.5 * (xs[k // 2 - 1] + xs[k // 2])


# One GPU's job on the MI355X box gets a 16-core share of the host, whatever
# the machine's core count; the CPU baselines use every core of that share.
HOST_SHARE_CORES = 16


def host_cores(requested=0):
    """(threads for the CPU baselines, what the host reports)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    threads = requested or min(affinity, HOST_SHARE_CORES)
    return threads, {"threads": threads, "host_nproc": nproc, "affinity": affinity,
                     "share": HOST_SHARE_CORES, "cpu": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_native():
    """Builds the oracle for this host's CPU (-march=native, as the
    reference's own build, cmake/LMPlatformOptions.cmake:95) into a temporary
    directory and points the oracle's loader at it; the results are those of
    the portable build (no contraction, no fast-math), only the instruction
    selection changes.  Returns the flags used, or None if the build failed
    (the portable build is then timed)."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="lmx_orc_native_"), "liborc.so")
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native", "NATIVE_OUT=" + out],
                       capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(out):
        log(f"native oracle build failed, timing the portable build: {r.stderr[-300:]}")
        return None
    os.environ["LMX_ORC_LIB"] = out
    return "-O3 -march=native -ffp-contract=off"


def traversal_profile():
    """Figures of the committed rocprofv3 counter summary of the traversal
    kernel (profiles/traversal_traffic.txt: `key: value` lines): HBM-side
    bytes per launch from FETCH_SIZE / WRITE_SIZE, the L2 hit rate, VALU
    instructions per launch, the busy fractions of the texture address and
    data units, the L1 hit rate and the profiled kernel time."""
    path = os.path.join(ROOT, "profiles", "traversal_traffic.txt")
    out = {}
    if os.path.exists(path):
        with open(path) as fh:
            for line in fh:
                if not line.startswith("#") and ":" in line:
                    k, v = line.split(":", 1)
                    try:
                        out[k.strip()] = float(v)
                    except ValueError:
                        pass
    return out


def cpu_baseline(verts, tris, rays, threads, reps=3):
    """Reference-algorithm QBVH on the host (oracle restatement): the median
    of `reps` passes over the rays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    t0 = time.perf_counter()
    q = orc.QBVH(verts, tris)
    build_s = time.perf_counter() - t0
    q.intersect(rays[:65536], nthreads=threads)     # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        q.intersect(rays, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    dt = median(ts)
    return rays.shape[0] / dt / 1e6, build_s, dt


def pt_cornell_section(dev, cpu, threads):
    """renderer::pt on the Cornell box (5 wall quads, 2 boxes rotated by
    +-17 degrees, 0.25 x 0.25 emitter of radiance (17, 12, 4); 256x256,
    16 spp, paths of at most 8 segments, roulette 0.5 from depth 3, seed 42),
    with the oracle's per-sample renderer::pt over the whole frame beside it
    (renderer_pt.cpp:64-242)."""
    import torch
    from lmx import render, scenes
    sc = scenes.cornell_box(256, 256)
    rs = render.Scene(sc)
    n = 256 * 256 * 16
    kw = dict(max_num_vertices=9, rr_depth=3, rr_prob=0.5, seed=42)
    rs.render_pt(n, **kw)                                   # capacity autotune + warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        film, st, _ = rs.render_pt(n, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = median(times)
    out = {
        "workload": f"pt_cornell: Cornell box ({sc['tris'].shape[0]} tris incl. the emitter), 256x256, 16 spp, "
                    "max 8 segments, RR 0.5 from depth 3, seed 42",
        "msamples_per_s": round(n / dt / 1e6, 2),
        "mrays_per_s": round(st["rays"] / dt / 1e6, 1),
        "rays_per_sample": round(st["rays"] / n, 4),
        "ms_per_frame": round(dt * 1e3, 3),
        "queue_capacity": st["queue_capacity"],
        "image_mean": round(float(film.mean()), 5),
        "cpu_baseline": None,
    }
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        osc = orc.Scene(sc)
        osc.pt_samples(0, 4096, nthreads=threads, **kw)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            c = osc.pt_samples(0, n, nthreads=threads, **kw)
            ts.append(time.perf_counter() - t0)
        cdt = median(ts)
        img = osc.film(c) * (256 * 256 / n)
        out["cpu_baseline"] = {
            "msamples_per_s": round(n / cdt / 1e6, 3),
            "mrays_per_s_est": round(n / cdt * st["rays"] / n / 1e6, 2),
            "cores": threads, "kind": "port",
            "sample": f"the whole frame ({n} samples), median of 3 runs of the oracle's renderer::pt; "
                      "Mrays/s with the GPU's rays per sample",
            "image_mean": round(float(img.mean()), 5)}
        out["gpu_over_cpu_samples"] = round((n / dt) / (n / cdt), 1)
    return out


# Algorithmic bytes of one path segment of the wavefront path tracer beyond
# its traversal: the shading kernel reads the queued ray (32 B), throughput
# (16 B), path record (16 B) and hit (16 B), the triangle's vertex indices,
# vertices and vertex normals (12 + 36 + 36 B), its material index and
# material (4 + 64 B); every traced ray was written once to a queue (ray,
# throughput, path record: 64 B) by the generation or the shading kernel.
PT_SEG_STATE_BYTES = 80 + 152 + 64


def pt_roofline(sc, rs, kw, n, dt, rays):
    """B_seg roofline of the wavefront path tracer: per extension ray the
    traversal's algorithmic bytes (48 B ray and hit + 80 B per node visit +
    48 B per triangle test, the visits counted by the host walker over the
    same BVH for every ray of a captured 1-spp frame) plus the segment's
    queue and shading bytes, times the rays of the timed frame, over its
    time, against 8 TB/s."""
    import torch
    from lmx import accel
    npix = sc["camera"]["width"] * sc["camera"]["height"]
    cap = 1 << 22
    buf = torch.empty((cap, 8), dtype=torch.float32, device="cuda:0")
    _, st, _ = rs.render_pt(npix, ray_out=buf, **kw)
    k = min(cap, st["rays"])
    captured = buf[:k].cpu().numpy()
    a = accel.Accel(sc["verts"], sc["tris"])
    nv, nt = a.count(captured)
    a.close()
    live = captured[:, 7] > 0.0                 # camera paths that end at once carry tmax < 0
    b_ray = RAY_BYTES + (NODE_BYTES * nv + TRI_BYTES * nt) / k
    b_seg = b_ray + PT_SEG_STATE_BYTES
    achieved = b_seg * rays / dt / 1e9
    return {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "achieved": round(achieved, 1),
            "frac": round(achieved / HBM_PEAK_GBS, 4), "b_seg_bytes": round(b_seg, 1),
            "b_ray_bytes": round(b_ray, 1), "state_bytes_per_segment": PT_SEG_STATE_BYTES,
            "node_visits_per_ray": round(nv / k, 3), "tri_tests_per_ray": round(nt / k, 3),
            "rays_counted": int(k), "live_fraction": round(float(live.mean()), 4),
            "method": "visits counted by the host walker on every extension ray of a captured 1-spp frame "
                      "(lmx_render_params.ray_out); segments = the timed frame's extension rays"}


# Algorithmic bytes of the batched bidirectional sampler beyond traversal.
# A subpath extension: the queued ray, throughput, path record and hit
# (32 + 16 + 16 + 16 B), the hit triangle's shading data (152 B, as for the
# path tracer), the new vertex record (64 B) and the next queue entry (64 B).
# A connection: its four vertex records (4 x 64 B) and two materials
# (2 x 64 B), the shadow-queue record it writes and the splat kernel reads
# back (ray 32 B, contribution, aux and path records 3 x 16 B, read back
# with the hit: 16 + 16 + 16 B) and its film splat (12 B).
BD_EXT_STATE_BYTES = 80 + 152 + 64 + 64
BD_CONN_STATE_BYTES = 4 * 64 + 2 * 64 + 32 + 48 + 48 + 12


def bd_roofline(sc, rs, render, n, dt, label):
    """B_seg roofline of a bidirectional render: every subpath extension ray
    and every shadow ray of a captured random-raster render of 2^17 samples
    is counted by the host walkers (closest hit and occlusion) over the same
    BVH; per-sample ray counts and bytes scale to the timed frame."""
    import numpy as np
    import torch
    from lmx import accel
    m = 1 << 17
    cap = 1 << 25
    buf = torch.full((cap, 8), float("nan"), dtype=torch.float32, device="cuda:0")
    _, st, _ = render(m, buf)
    half = cap // 2
    host = buf.cpu().numpy()
    ext, sh = host[:half], host[half:]
    ne = int((~np.isnan(ext[:, 0])).sum())               # captured records fill each half from its start
    ns = int((~np.isnan(sh[:, 0])).sum())
    ext, sh = ext[:ne], sh[:ns]
    a = accel.Accel(sc["verts"], sc["tris"])
    nve, nte = a.count(ext)
    nvs, nts = a.count(sh, anyhit=True)
    a.close()
    b_ext = RAY_BYTES + (NODE_BYTES * nve + TRI_BYTES * nte) / max(ne, 1) + BD_EXT_STATE_BYTES
    b_sh = RAY_BYTES + (NODE_BYTES * nvs + TRI_BYTES * nts) / max(ns, 1) + BD_CONN_STATE_BYTES
    per_sample = (ne * b_ext + ns * b_sh) / m
    achieved = per_sample * n / dt / 1e9
    return {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "achieved": round(achieved, 1),
            "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_sample": round(per_sample, 1),
            "ext_rays_per_sample": round(ne / m, 3), "shadow_rays_per_sample": round(ns / m, 3),
            "b_seg_ext_bytes": round(b_ext, 1), "b_seg_shadow_bytes": round(b_sh, 1),
            "ext_node_visits_per_ray": round(nve / max(ne, 1), 3), "shadow_node_visits_per_ray": round(nvs / max(ns, 1), 3),
            "captured_rays_match_stats": bool(ne + int((sh[:, 7] >= sh[:, 3]).sum()) == st["rays"]),
            "method": f"{label}: every ray of a captured render of {m} random-raster samples counted by the host "
                      "walkers; bytes per sample times the timed frame's samples"}


def pt_section(dev, cpu, threads):
    """Wavefront renderer::pt on the sphere-grid scene (5,242,886 triangles,
    1024x1024, 64 spp, max 17 vertices = depth 16, roulette from depth 3):
    samples/s and Mrays/s (extension rays), beside the oracle's per-sample
    renderer::pt restatement on host threads over a bounded prefix."""
    import torch
    from lmx import render, scenes
    log("pt: generating the sphere-grid scene")
    sc = scenes.sphere_grid()
    rs = render.Scene(sc)
    n = 1024 * 1024 * 64
    kw = dict(max_num_vertices=17, rr_depth=3, rr_prob=0.5, seed=42)
    # autotune + warm-up: the capacity autotuner probes on this render's
    # first 2^24 samples, enough for its steady state at 2^23 paths in flight
    rs.render_pt(1024 * 1024 * 16, **kw)
    torch.cuda.synchronize()
    times, rays = [], 0
    for _ in range(5):
        t0 = time.perf_counter()
        film, st, _ = rs.render_pt(n, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        rays = st["rays"]
    dt = median(times)
    out = {
        "workload": "pt: 16x16 displaced icospheres (5,242,886 tris), 1024x1024, 64 spp, depth 16, RR from depth 3",
        "samples_per_s": round(n / dt, 1),
        "msamples_per_s": round(n / dt / 1e6, 2),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "rays_per_sample": round(rays / n, 4),
        "s_per_frame": round(dt, 4),
        "iterations": st["iterations"],
        "queue_capacity": st["queue_capacity"],
        "image_mean": round(float(film.mean()), 5),
        "cpu_baseline": None,
    }
    out["roofline"] = pt_roofline(sc, rs, kw, n, dt, rays)
    # the same frame with the traced paths regrouped by material before
    # shading (70 % Lambertian / 30 % Phong), timed the same way
    rs.render_pt(1024 * 1024 * 4, shade_order=render.SHADE_BY_MATERIAL, **kw)
    torch.cuda.synchronize()
    tm = []
    for _ in range(5):
        t0 = time.perf_counter()
        rs.render_pt(n, shade_order=render.SHADE_BY_MATERIAL, **kw)
        torch.cuda.synchronize()
        tm.append(time.perf_counter() - t0)
    out["shade_by_material"] = {"msamples_per_s": round(n / median(tm) / 1e6, 2),
                                "queue_order_msamples_per_s": out["msamples_per_s"],
                                "speedup": round(dt / median(tm), 4)}
    # renderer::ptdirect / ptmis on the same frame: next-event estimation
    # adds one shadow ray per vertex
    for entry in ("ptdirect", "ptmis"):
        fn = getattr(rs, "render_" + entry)
        fn(1024 * 1024 * 4, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st2, _ = fn(n, **kw)
        torch.cuda.synchronize()
        dt2 = time.perf_counter() - t0
        out[entry] = {"msamples_per_s": round(n / dt2 / 1e6, 2), "mrays_per_s": round(st2["rays"] / dt2 / 1e6, 1),
                      "rays_per_sample": round(st2["rays"] / n, 4)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        log("pt: cpu baseline (oracle renderer::pt)")
        t0 = time.perf_counter()
        osc = orc.Scene(sc)
        build = time.perf_counter() - t0
        m = 1 << 19
        osc.pt_samples(0, 4096, nthreads=threads, **kw)
        t0 = time.perf_counter()
        osc.pt_samples(0, m, nthreads=threads, **kw)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "msamples_per_s": round(m / cdt / 1e6, 4),
            "mrays_per_s_est": round(m / cdt * rays / n / 1e6, 3),
            "cores": threads, "kind": "port",
            "sample": f"first {m} samples of the same frame; oracle QBVH build {build:.1f}s not timed; "
                      "Mrays/s estimated with the GPU's rays per sample"}
        out["gpu_over_cpu_samples"] = round((n / dt) / (m / cdt), 1)
    return out


def bdpt_section(dev, cpu, threads):
    """Batched renderer::vcm (bdpt mode) on the sphere-grid scene (5,242,886
    triangles, 1024x1024, 16 spp, subpaths of up to 10 vertices, every (s, t)
    strategy with power-heuristic MIS): samples/s and Mrays/s (subpath
    extension rays plus shadow rays), beside the oracle's per-sample
    restatement on host threads over a bounded prefix; renderer::bdpt on the
    same frame (at most 10 vertices, and unbounded) and renderer::ltdirect
    (16 light subpaths per pixel) beside it."""
    import torch
    from lmx import render, scenes
    log("bdpt: generating the sphere-grid scene")
    sc = scenes.sphere_grid()
    rs = render.Scene(sc)
    n = 1024 * 1024 * 16
    kw = dict(max_num_vertices=10, seed=42)
    rs.render_vcm_bdpt(1 << 22, **kw)                     # autotune + warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        film, st, _ = rs.render_vcm_bdpt(n, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = median(times)
    rays = st["rays"]
    from lmx.render import RASTER_RANDOM
    roof = bd_roofline(sc, rs, lambda m, buf: rs.render_vcm_bdpt(m, raster_mode=RASTER_RANDOM, ray_out=buf, **kw),
                       n, dt, "renderer::vcm bdpt")
    out = {
        "roofline": roof,
        "workload": "bdpt: 16x16 displaced icospheres (5,242,886 tris), 1024x1024, 16 spp, "
                    "subpaths of up to 10 vertices, all (s,t) strategies, power heuristic",
        "samples_per_s": round(n / dt, 1),
        "msamples_per_s": round(n / dt / 1e6, 2),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "rays_per_sample": round(rays / n, 4),
        "s_per_frame": round(dt, 4),
        "batches": st["batches"],
        "batch": st["batch"],
        "image_mean": round(float(film.mean()), 5),
        "cpu_baseline": None,
    }
    # renderer::bdpt on the same frame: roulette subpaths with direct light
    # vertices, power heuristic; bounded at 10 vertices and unbounded
    for key, m in (("renderer_bdpt", 10), ("renderer_bdpt_unbounded", -1)):
        kb = dict(max_num_vertices=m, seed=42)
        rs.render_bdpt(1 << 22, **kb)                    # autotune + warm-up
        torch.cuda.synchronize()
        tb = []
        for _ in range(3):
            t0 = time.perf_counter()
            fb, sb, _ = rs.render_bdpt(n, **kb)
            torch.cuda.synchronize()
            tb.append(time.perf_counter() - t0)
        db = median(tb)
        out[key] = {"max_num_vertices": m, "msamples_per_s": round(n / db / 1e6, 2),
                    "mrays_per_s": round(sb["rays"] / db / 1e6, 1), "rays_per_sample": round(sb["rays"] / n, 4),
                    "s_per_frame": round(db, 4), "batch": sb["batch"], "image_mean": round(float(fb.mean()), 5),
                    "roofline": bd_roofline(sc, rs, lambda k, buf: rs.render_bdpt(k, raster_mode=RASTER_RANDOM,
                                                                                   ray_out=buf, **kb),
                                            n, db, "renderer::bdpt")}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        log("bdpt: cpu baseline (oracle renderer::vcm bdpt)")
        t0 = time.perf_counter()
        osc = orc.Scene(sc)
        build = time.perf_counter() - t0
        osc.bdpt_samples(0, 1024, nthreads=threads, **kw)
        done, chunk = 0, 8192
        t0 = time.perf_counter()
        while done < (1 << 18) and time.perf_counter() - t0 < 10.0:
            osc.bdpt_samples(done, chunk, nthreads=threads, **kw)
            done += chunk
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "msamples_per_s": round(done / cdt / 1e6, 5),
            "cores": threads, "kind": "port",
            "sample": f"first {done} samples of the same frame (about 10 s); oracle QBVH build "
                      f"{build:.1f}s not timed"}
        out["gpu_over_cpu_samples"] = round((n / dt) / (done / cdt), 1)
        log("bdpt: cpu baseline (oracle renderer::bdpt)")
        osc.bdptref_samples(0, 1024, max_num_vertices=10, nthreads=threads)
        done, t0 = 0, time.perf_counter()
        while done < (1 << 18) and time.perf_counter() - t0 < 5.0:
            osc.bdptref_samples(done, chunk, max_num_vertices=10, nthreads=threads)
            done += chunk
        rdt = time.perf_counter() - t0
        out["renderer_bdpt"]["cpu_baseline"] = {
            "msamples_per_s": round(done / rdt / 1e6, 5), "cores": threads, "kind": "port",
            "sample": f"first {done} samples of the same frame (about 5 s)"}
    # renderer::ltdirect on the same frame: 16 light subpaths per pixel, at
    # most 10 vertices, every vertex connected to the camera
    kl = dict(max_num_vertices=10, seed=42)
    rs.render_ltdirect(1 << 22, **kl)                                        # warm-up
    torch.cuda.synchronize()
    tl = []
    for _ in range(3):
        t0 = time.perf_counter()
        fl, sl, _ = rs.render_ltdirect(n, **kl)
        torch.cuda.synchronize()
        tl.append(time.perf_counter() - t0)
    dl = median(tl)
    out["ltdirect"] = {"max_num_vertices": 10, "msamples_per_s": round(n / dl / 1e6, 2),
                       "mrays_per_s": round(sl["rays"] / dl / 1e6, 1), "rays_per_sample": round(sl["rays"] / n, 4),
                       "s_per_frame": round(dl, 4), "image_mean": round(float(fl.mean()), 5), "cpu_baseline": None}
    if cpu:
        log("bdpt: cpu baseline (oracle renderer::ltdirect)")
        osc.ltdirect_samples(0, 1024, nthreads=threads, **kl)
        done, t0 = 0, time.perf_counter()
        while done < (1 << 21) and time.perf_counter() - t0 < 5.0:
            osc.ltdirect_samples(done, 1 << 16, nthreads=threads, **kl)
            done += 1 << 16
        ldt = time.perf_counter() - t0
        out["ltdirect"]["cpu_baseline"] = {
            "msamples_per_s": round(done / ldt / 1e6, 5), "cores": threads, "kind": "port",
            "sample": f"first {done} light subpaths of the same frame (about 5 s)"}
        out["ltdirect"]["gpu_over_cpu_samples"] = round((n / dl) / (done / ldt), 1)
    return out


def vcm_section(dev, cpu, threads):
    """renderer::vcm with vertex merging on the sphere-grid frame of the
    bidirectional section (1024x1024, subpaths of up to 10 vertices): 8 passes
    of 2 spp of eye samples and 2^20 photon subpaths each, initial radius 0.05
    (about 3.5 pixels at the grid's distance), alpha 0.7; mode vcm
    (connections and merges) and mode bdpm (merges alone).  Eye samples/s of
    the whole render (photon passes, grids, eye subpaths, connections, shadow
    rays and merges), beside the oracle's restatement of one pass on host
    threads over a bounded prefix of eye samples, its photon pass timed
    whole.  renderer::pm on the same frame beside it (16 spp, 2^22 photon
    subpaths, radius 0.05, final gather), renderer::ppm (a point per
    pixel, 32 passes of 2^20 photon subpaths) and renderer::sppm (the same,
    points re-traced every pass)."""
    import torch
    from lmx import render, scenes
    log("vcm: generating the sphere-grid scene")
    sc = scenes.sphere_grid()
    rs = render.Scene(sc)
    npix = 1024 * 1024
    passes, ne, nph, r0 = 8, 2 * npix, 1 << 20, 0.05
    kw = dict(num_iteration_pass=passes, num_photon_trace_samples=nph, initial_radius=r0, alpha=0.7,
              max_num_vertices=10, seed=42)
    out = {"workload": "vcm: 16x16 displaced icospheres (5,242,886 tris), 1024x1024, 8 passes x 2 spp, "
                       "2^20 photon subpaths per pass, radius 0.05 (alpha 0.7), subpaths of up to 10 vertices"}
    for key, mode in (("vcm", render.VCM_MODE_VCM), ("bdpm", render.VCM_MODE_BDPM)):
        rs.render_vcm(npix // 4, mode=mode, **dict(kw, num_iteration_pass=1))        # warm-up
        torch.cuda.synchronize()
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            film, st, _ = rs.render_vcm(ne, mode=mode, **kw)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        dt = median(times)
        out[key] = {"msamples_per_s": round(passes * ne / dt / 1e6, 2),
                    "mrays_per_s": round(st["rays"] / dt / 1e6, 1),
                    "rays_per_sample": round(st["rays"] / (passes * ne), 4),
                    "ms_per_pass": round(dt / passes * 1e3, 2), "s_per_frame": round(dt, 4),
                    "image_mean": round(float(film.mean()), 5), "cpu_baseline": None}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        log("vcm: cpu baseline (oracle renderer::vcm, one pass)")
        osc = orc.Scene(sc)
        m = dict(radius=r0, num_photons=nph, mode=orc.VCM_VCM, max_num_vertices=10, seed=42, pseed=43)
        tph = []
        for _ in range(2):                                       # photon pass + grid + one eye sample
            t0 = time.perf_counter()
            osc.vcm_samples(0, 1, nthreads=threads, **m)
            tph.append(time.perf_counter() - t0)
        tph = median(tph)
        chunk = 65536
        t0 = time.perf_counter()
        osc.vcm_samples(0, chunk, nthreads=threads, **m)
        teye = max(time.perf_counter() - t0 - tph, 1e-6) / chunk
        rate = ne / (tph + ne * teye)
        out["vcm"]["cpu_baseline"] = {
            "msamples_per_s": round(rate / 1e6, 5), "cores": threads, "kind": "port",
            "sample": f"one pass on {threads} threads: photon pass of 2^20 subpaths with its grid ({tph:.2f}s) and "
                      f"{chunk} eye samples ({teye * 1e6:.1f} us each), scaled to the pass's {ne} eye samples"}
        out["vcm"]["gpu_over_cpu_samples"] = round(out["vcm"]["msamples_per_s"] * 1e6 / rate, 1)
    # renderer::pm on the same frame: 16 spp, 2^22 photon subpaths, radius
    # 0.05, final gather, at most 10 vertices
    npm, nphm = 16 * npix, 1 << 22
    kp = dict(num_photon_trace_samples=nphm, finalgather=1, radius=0.05, max_num_vertices=10, seed=42)
    rs.render_pm(npix, **kp)                                                   # warm-up
    torch.cuda.synchronize()
    tp = []
    for _ in range(2):
        t0 = time.perf_counter()
        fpm, spm, _ = rs.render_pm(npm, **kp)
        torch.cuda.synchronize()
        tp.append(time.perf_counter() - t0)
    dp = median(tp)
    out["pm"] = {"workload": "pm: same frame, 16 spp, 2^22 photon subpaths (roulette), radius 0.05, final gather, "
                             "at most 10 vertices",
                 "msamples_per_s": round(npm / dp / 1e6, 2), "mrays_per_s": round(spm["rays"] / dp / 1e6, 1),
                 "s_per_frame": round(dp, 4), "image_mean": round(float(fpm.mean()), 5), "cpu_baseline": None}
    if cpu:
        log("vcm: cpu baseline (oracle renderer::pm)")
        mp = dict(radius=0.05, num_photons=nphm, max_num_vertices=10, finalgather=1, seed=42, pseed=43)
        t0 = time.perf_counter()
        osc.pm_samples(0, 1, nthreads=threads, **mp)
        tph = time.perf_counter() - t0
        chunk = 1 << 20
        t0 = time.perf_counter()
        osc.pm_samples(0, chunk, nthreads=threads, **mp)
        teye = max(time.perf_counter() - t0 - tph, 1e-6) / chunk
        rate = npm / (tph + npm * teye)
        out["pm"]["cpu_baseline"] = {
            "msamples_per_s": round(rate / 1e6, 5), "cores": threads, "kind": "port",
            "sample": f"photon pass of 2^22 subpaths with its grid ({tph:.2f}s) and {chunk} eye samples "
                      f"({teye * 1e6:.2f} us each) on {threads} threads, scaled to the frame's {npm} samples"}
        out["pm"]["gpu_over_cpu_samples"] = round(npm / dp / rate, 1)
    # renderer::ppm on the same frame: one measurement point per pixel, 32
    # passes of 2^20 photon subpaths, initial radius 0.05, alpha 0.7
    passes, nphp = 32, 1 << 20
    kq = dict(num_iteration_pass=passes, num_photon_trace_samples=nphp, initial_radius=0.05, alpha=0.7, seed=42)
    rs.render_ppm(npix, 10, **dict(kq, num_iteration_pass=2))                 # warm-up
    torch.cuda.synchronize()
    tq = []
    for _ in range(2):
        t0 = time.perf_counter()
        fpp, spp_, _ = rs.render_ppm(npix, 10, **kq)
        torch.cuda.synchronize()
        tq.append(time.perf_counter() - t0)
    dq = median(tq)
    out["ppm"] = {"workload": "ppm: same frame, one measurement point per pixel (1,048,576), 32 passes of 2^20 "
                              "photon subpaths, initial radius 0.05, alpha 0.7, at most 10 vertices",
                  "ms_per_pass": round(dq / passes * 1e3, 3), "mphotons_per_s": round(passes * nphp / dq / 1e6, 1),
                  "s_per_frame": round(dq, 4), "image_mean": round(float(fpp.mean()), 5), "cpu_baseline": None}
    if cpu:
        log("vcm: cpu baseline (oracle renderer::ppm)")
        mq = dict(initial_radius=0.05, alpha=0.7, seed=42, pseed=43)
        t0 = time.perf_counter()
        osc.ppm_points(npix, 1, nphp, 10, nthreads=threads, **mq)
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        osc.ppm_points(npix, 3, nphp, 10, nthreads=threads, **mq)
        t3 = time.perf_counter() - t0
        cpass = max(t3 - t1, 1e-6) / 2.0
        out["ppm"]["cpu_baseline"] = {
            "ms_per_pass": round(cpass * 1e3, 1), "mphotons_per_s": round(nphp / cpass / 1e6, 3), "cores": threads,
            "kind": "port", "sample": f"the oracle at 3 passes minus 1 pass of the same workload on {threads} "
                                      f"threads ({t1:.1f}s and {t3:.1f}s)"}
        out["ppm"]["gpu_over_cpu_passes"] = round(cpass * passes / dq, 1)
    # renderer::sppm on the same frame: a point per pixel re-traced every
    # pass, 32 passes of 2^20 photon subpaths
    ks = dict(num_iteration_pass=passes, num_photon_trace_samples=nphp, initial_radius=0.05, alpha=0.7, seed=42)
    rs.render_sppm(10, **dict(ks, num_iteration_pass=2))                       # warm-up
    torch.cuda.synchronize()
    tsp = []
    for _ in range(2):
        t0 = time.perf_counter()
        fsp, _, _ = rs.render_sppm(10, **ks)
        torch.cuda.synchronize()
        tsp.append(time.perf_counter() - t0)
    dsp = median(tsp)
    out["sppm"] = {"workload": "sppm: same frame, a point per pixel re-traced every pass, 32 passes of 2^20 "
                               "photon subpaths, initial radius 0.05, alpha 0.7, at most 10 vertices",
                   "ms_per_pass": round(dsp / passes * 1e3, 3),
                   "mphotons_per_s": round(passes * nphp / dsp / 1e6, 1), "s_per_frame": round(dsp, 4),
                   "image_mean": round(float(fsp.mean()), 5), "cpu_baseline": None}
    if cpu:
        log("vcm: cpu baseline (oracle renderer::sppm)")
        t0 = time.perf_counter()
        osc.sppm_pixels(1, nphp, 10, nthreads=threads, **mq)
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        osc.sppm_pixels(3, nphp, 10, nthreads=threads, **mq)
        t3 = time.perf_counter() - t0
        cpass = max(t3 - t1, 1e-6) / 2.0
        out["sppm"]["cpu_baseline"] = {
            "ms_per_pass": round(cpass * 1e3, 1), "mphotons_per_s": round(nphp / cpass / 1e6, 3), "cores": threads,
            "kind": "port", "sample": f"the oracle at 3 passes minus 1 pass of the same workload on {threads} "
                                      f"threads ({t1:.1f}s and {t3:.1f}s)"}
        out["sppm"]["gpu_over_cpu_passes"] = round(cpass * passes / dsp, 1)
    return out


def image_section(dev, cpu):
    """lmx_film_save of a 1024x1024 film in each format (GPU encode, one
    device-to-host copy, host run-length / container write to a temporary
    directory): wall ms per save, median of 5; beside it the oracle's
    per-pixel RGBE and PNG transforms on one host thread."""
    import tempfile
    import numpy as np
    import torch
    from lmx import image
    rng = np.random.default_rng(1)
    f = (rng.random((1024, 1024, 3)) * np.exp(rng.normal(0, 2, (1024, 1024, 1)))).astype(np.float32)
    t = torch.from_numpy(f).to(dev)
    out = {"workload": "film output: 1024x1024 RGB float film -> .hdr (RLE RGBE), .exr (float), .png, .pfm"}
    with tempfile.TemporaryDirectory() as d:
        for name, ty in (("radiancehdr", image.RADIANCEHDR), ("openexr", image.OPENEXR), ("png", image.PNG),
                         ("pfm", image.PFM)):
            image.save(t, os.path.join(d, "w"), ty)
            ts = []
            for k in range(5):
                t0 = time.perf_counter()
                fn = image.save(t, os.path.join(d, f"f{k}"), ty)
                ts.append(time.perf_counter() - t0)
            out[name] = {"ms": round(1e3 * sorted(ts)[2], 2), "bytes": os.path.getsize(fn)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        t0 = time.perf_counter()
        orc.film_rgbe(f)
        orc.film_png(f)
        out["cpu_baseline"] = {"ms_rgbe_plus_png_transforms": round(1e3 * (time.perf_counter() - t0), 2),
                               "cores": 1, "kind": "port",
                               "sample": "the oracle's per-pixel RGBE and PNG transforms of the same film"}
    return out


MLP_FLOP_FWD = 2 * (168 * 64 + 3 * 64 * 64 + 64 * 41)     # 51,328 algorithmic FLOP per query
MFMA_BF16_PEAK = 2.5e15


def guide_section(dev, ref_spp=16384, spp=256, train_passes=8, steps=32, batch=1 << 15, lr=1e-3, cache=None,
                  trials_n=3, mfma=True, train_spp=2):
    """Learned path guiding on the occluded-room scene (12 quads + 96
    displaced spheres = 1,966,104 tris, 40% Phong e=200, emitter behind a wall
    with a 0.1-wide gap), 1024x1024: train online on the first 16 spp (8 passes
    of 2 spp, 32 Adam steps after each), render 256 spp guided; plain PT gets the
    same wall-clock time; both MSEs against a ref_spp plain reference.  Both
    samplers use the throughput roulette from depth 3 and at most 17
    vertices.  `trials_n` independent trials (fresh network, other seeds):
    the record order, and so the minibatches, differ between runs; the trial
    with the median equal-time ratio is reported with every trial's ratios.
    Also: MFMA utilisation of the MLP (inference and training kernels, timed
    with HIP events on their stream)."""
    import numpy as np
    import torch
    from lmx import guide, render, scenes
    kw = dict(max_num_vertices=17, rr_depth=3, rr_prob=0.5, rr_mode=render.RR_THROUGHPUT)
    if cache and "ref" in cache:
        sc, rs, ref, t_ref = cache["sc"], cache["rs"], cache["ref"], cache["t_ref"]
        W, H = sc["camera"]["width"], sc["camera"]["height"]
        npix = W * H
    else:
        log("guide: generating the occluded-room scene")
        sc = scenes.occluded_room()
        rs = render.Scene(sc)
        W, H = sc["camera"]["width"], sc["camera"]["height"]
        npix = W * H
        rs.render_pt(npix * 4, seed=1, **kw)               # autotune + warm-up
        log(f"guide: plain reference, {ref_spp} spp")
        t0 = time.perf_counter()
        ref = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
        chunk, done = 64, 0
        while done < ref_spp:
            c = min(chunk, ref_spp - done)
            f, _, _ = rs.render_pt(npix * c, seed=1_000_000 + done, **kw)
            ref += f.double() * c
            done += c
        ref /= done
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t0
        if cache is not None:
            cache.update(sc=sc, rs=rs, ref=ref, t_ref=t_ref)
    # plain sampling rate (for the equal-time budget)
    rs.render_pt(npix * 16, seed=3, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rs.render_pt(npix * 64, seed=4, **kw)
    torch.cuda.synchronize()
    rate = npix * 64 / (time.perf_counter() - t0)
    # guided: online training then the render, one wall-clock budget; three
    # independent trials (the record order, and so the minibatches, differ
    # between runs), the median ratio reported
    g = guide.Guide(rs, lr=lr, steps_per_pass=steps, batch=batch)
    g.render(npix // 16, seed=5, **kw)                    # kernels and buffers warm
    fq, _, _ = rs.render_pt(npix * spp, seed=998, **kw)
    mse_p_same = float(((fq.double() - ref) ** 2).mean())
    trials = []
    for tr in range(trials_n):
        g = guide.Guide(rs, lr=lr, steps_per_pass=steps, batch=batch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        recs = nz = 0
        for k in range(train_passes):
            _, _, gs = g.render(npix * train_spp, seed=100 + 1000 * tr + k, train=True, **kw)
            recs += gs["records"]
            nz += gs["records_nonzero"]
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        t1 = time.perf_counter()
        fg, stg, _ = g.render(npix * spp, seed=777 + tr, **kw)
        torch.cuda.synchronize()
        t_render = time.perf_counter() - t1
        t_guided = t_train + t_render
        n_plain = max(1, int(t_guided * rate / npix))
        t0 = time.perf_counter()
        fp, stp, _ = rs.render_pt(npix * n_plain, seed=999 + tr, **kw)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        mse_g = float(((fg.double() - ref) ** 2).mean())
        mse_p = float(((fp.double() - ref) ** 2).mean())
        trials.append({"mse_ratio_equal_time": round(mse_p / mse_g, 3), "mse_ratio_equal_spp": round(mse_p_same / mse_g, 3),
                       "mse_guided": mse_g, "mse_plain_equal_time": mse_p, "guided_s": round(t_guided, 3),
                       "train_s": round(t_train, 3), "render_s": round(t_render, 3), "plain_spp_equal_time": n_plain,
                       "plain_s": round(t_plain, 3), "guided_msamples_per_s": round(npix * spp / t_render / 1e6, 1),
                       "train_records": recs, "train_records_nonzero": nz})
    med = sorted(trials, key=lambda t: t["mse_ratio_equal_time"])[len(trials) // 2]
    if not mfma:
        return {"lr": lr, **med, "trials_equal_time": [t["mse_ratio_equal_time"] for t in trials],
                "trials_equal_spp": [t["mse_ratio_equal_spp"] for t in trials]}
    # MFMA utilisation: inference over 2^22 queries, training over 2^16 records
    x = torch.rand((1 << 22, 7), device=dev)
    g.eval(x)
    st = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(5):
        g.eval(x)
    e1.record(st)
    torch.cuda.synchronize()
    t_eval = e0.elapsed_time(e1) / 5 / 1e3
    r = np.random.default_rng(1)
    xs = r.random((1 << 16, 7)); d = r.normal(size=(1 << 16, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    recs_t = torch.from_numpy(guide.Guide.records(xs, d, r.uniform(0.05, 1.0, 1 << 16), r.uniform(0, 2, 1 << 16))).to(dev)
    g.grad(recs_t)
    e0.record(st)
    for _ in range(5):
        g.grad(recs_t)
    e1.record(st)
    torch.cuda.synchronize()
    t_grad = e0.elapsed_time(e1) / 5 / 1e3
    inf_tflops = (1 << 22) * MLP_FLOP_FWD / t_eval / 1e12
    trn_tflops = (1 << 16) * 3 * MLP_FLOP_FWD / t_grad / 1e12
    return {
        "workload": "guide: occluded room (12 quads + 96 displaced spheres, 1,966,104 tris, 40% Phong e=200, "
                    "emitter behind a 0.1-wide gap), 1024x1024, MLP 168-64x4-41 bf16 -> 8 vMF lobes + BSDF mix, "
                    "online training on the first 16 spp, 256 spp guided vs plain PT at equal time",
        "metric": "equal-time MSE ratio, guided vs plain GPU PT (median of the trials)",
        **med,
        "mse_plain_256spp": mse_p_same,
        "plain_msamples_per_s": round(rate / 1e6, 1),
        "trials_equal_time": [t["mse_ratio_equal_time"] for t in trials],
        "trials_equal_spp": [t["mse_ratio_equal_spp"] for t in trials],
        "reference": {"spp": ref_spp, "s": round(t_ref, 1), "mean": round(float(ref.mean()), 6)},
        "training": {"passes": train_passes, "spp_per_pass": train_spp, "steps_per_pass": steps, "batch": batch,
                     "lr": lr},
        "mfma": {"peak_tflops_bf16": MFMA_BF16_PEAK / 1e12,
                 "inference_tflops": round(inf_tflops, 1), "inference_frac": round(inf_tflops * 1e12 / MFMA_BF16_PEAK, 4),
                 "inference_mqueries_per_s": round((1 << 22) / t_eval / 1e6, 1),
                 "training_tflops": round(trn_tflops, 2), "training_frac": round(trn_tflops * 1e12 / MFMA_BF16_PEAK, 5),
                 "flop_per_query": MLP_FLOP_FWD},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-pt", action="store_true", help="skip the wavefront path-tracer section")
    ap.add_argument("--no-bdpt", action="store_true", help="skip the bidirectional section")
    ap.add_argument("--no-guide", action="store_true", help="skip the path-guiding section")
    ap.add_argument("--no-image", action="store_true", help="skip the film-output section")
    ap.add_argument("--no-vcm", action="store_true", help="skip the vertex-merging section")
    ap.add_argument("--guide-ref-spp", type=int, default=16384,
                    help="spp of the plain reference image of the path-guiding section")
    ap.add_argument("--only", choices=["pt", "pt_cornell", "bdpt", "vcm", "guide", "image"], default=None,
                    help="run one renderer section alone and print its JSON (for profiling runs)")
    ap.add_argument("--no-per-set", action="store_true",
                    help="skip the separate coherent/incoherent timings (for profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (steps times); used to "
                         "calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE against known bytes")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "launches": args.steps}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(rays shard across GPUs with no exchange; this benchmark runs on one GPU)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from lmx import accel, scenes

    dev = torch.device("cuda:0")
    threads, host = host_cores(args.cpu_threads)
    native = None if args.no_cpu_baseline else oracle_native()
    host["oracle_build"] = native or "-O3 -march=x86-64-v3 -ffp-contract=off (portable build)"
    if args.only:
        if args.only == "guide":
            cache = {}
            out = guide_section(dev, args.guide_ref_spp, lr=1e-3, cache=cache)
            out["alt_lr"] = guide_section(dev, args.guide_ref_spp, lr=5e-3, cache=cache, mfma=False)
        elif args.only == "image":
            out = image_section(dev, not args.no_cpu_baseline)
        elif args.only == "vcm":
            out = vcm_section(dev, not args.no_cpu_baseline, threads)
        elif args.only == "pt_cornell":
            out = pt_cornell_section(dev, not args.no_cpu_baseline, threads)
        else:
            fn = pt_section if args.only == "pt" else bdpt_section
            out = fn(dev, not args.no_cpu_baseline, threads)
        print(json.dumps({args.only: out}), flush=True)
        return
    log("generating scene")
    verts, tris = scenes.traversal_scene()
    coh, inc = scenes.traversal_rays(verts)
    rays = np.concatenate([coh, inc])
    nray = rays.shape[0]
    log(f"{tris.shape[0]} triangles, {nray} rays; building wide BVH")
    t0 = time.perf_counter()
    a = accel.Accel(verts, tris, isect=accel.ISECT_WATERTIGHT)
    build_s = time.perf_counter() - t0
    info = a.info()
    log(f"built in {build_s:.2f}s: {info}")
    a.upload()

    d_rays = torch.from_numpy(rays).to(dev)
    d_hits = torch.empty((nray, 4), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    for _ in range(args.warmup):
        a.intersect(d_rays, d_hits, stream)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        a.intersect(d_rays, d_hits, stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    ms_per_step = wall / args.steps * 1e3
    mrays = nray * args.steps / wall / 1e6
    log(f"{ms_per_step:.3f} ms/step, kernel {avg_kernel_s * 1e3:.3f} ms, {mrays:.1f} Mrays/s")

    # per-set throughput (coherent / incoherent halves), same kernel
    per_set = {}
    for name, sl in () if args.no_per_set else (("coherent", slice(0, coh.shape[0])), ("incoherent", slice(coh.shape[0], nray))):
        sub = d_rays[sl].contiguous()
        subh = torch.empty((sub.shape[0], 4), dtype=torch.float32, device=dev)
        a.intersect(sub, subh, stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        reps = 5
        for _ in range(reps):
            a.intersect(sub, subh, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        per_set[name] = round(sub.shape[0] * reps / (e0.elapsed_time(e1) / 1e3) / 1e6, 1)

    # algorithmic bytes from the counting walker (same BVH, same order)
    log("counting node visits / triangle tests on the host")
    nodes_v, tris_t = a.count(rays)
    alg_bytes = RAY_BYTES * nray + NODE_BYTES * nodes_v + TRI_BYTES * tris_t
    achieved = alg_bytes / avg_kernel_s / 1e9
    prof = traversal_profile()
    traffic = prof.get("traffic_bytes_per_launch")
    l2 = prof.get("l2_hit_rate")
    valu = prof.get("valu_insts_per_launch")

    hits_np = d_hits.cpu().numpy()
    hit_frac = float((accel.hit_ids(hits_np) != accel.MISS).mean())

    out = {
        "metric": "Mrays/s closest-hit",
        "value": round(mrays, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "bvh_traversal: displaced icosphere 1,310,720 tris; 4,194,304 coherent + 4,194,304 incoherent rays",
            "triangles": int(tris.shape[0]),
            "rays": int(nray),
            "isect": "watertight",
            "bvh": "8-wide, child boxes on a 255-step grid (bytes), 80 B nodes",
            "bvh_nodes": info["num_nodes"],
            "bvh_depth": info["max_depth"],
            "bvh_build_s": round(build_s, 3),
            "hit_fraction": round(hit_frac, 4),
            "mrays_per_set": per_set,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this benchmark's command, scaled by a "
                              "copy kernel's counts (profiles/traversal_traffic.txt, the committed profile run; "
                              "the L2 hit rate and VALU count come from the same runs); not measured in this run",
            "dram_gbs": None if traffic is None else round(traffic / avg_kernel_s / 1e9, 1),
            "dram_frac": None if traffic is None else round(traffic / avg_kernel_s / 1e9 / HBM_PEAK_GBS, 4),
            "l2_hit_rate": l2,
            "valu_insts_per_launch": valu,
            "valu_issue_frac": None if valu is None else round(valu * 4 / 1024 / (avg_kernel_s * VALU_CLOCK_HZ), 4),
            "ta_busy_frac": prof.get("ta_busy_frac"),
            "td_busy_frac": prof.get("td_busy_frac"),
            "td_l1_stall_frac": prof.get("td_l1_stall_frac"),
            "l1_hit_rate": prof.get("l1_hit_rate"),
            "limiter": "the vector-memory pipeline, then VALU issue: in the profiled launch the texture "
                       "address (TA) and data (TD) units are busy most cycles, the data unit waiting on the "
                       "L1 for much of it (L1 misses to L2), while VALU issue (wave instructions x 4 cycles "
                       "on 1,024 SIMDs against the launch's cycles at 2.4 GHz) comes next; the nodes and "
                       "triangles (75 MB) are served mostly by L2 and the Infinity Cache, so fabric traffic "
                       "is a fraction of the algorithmic bytes, whose rate exceeds nothing physical",
            "above_hbm_peak_by_cache_reuse": bool(achieved > HBM_PEAK_GBS),
            "alg_bytes_per_launch": int(alg_bytes),
            "node_visits_per_ray": round(nodes_v / nray, 3),
            "tri_tests_per_ray": round(tris_t / nray, 3),
            "kernel_ms": round(avg_kernel_s * 1e3, 4),
        },
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle QBVH, {threads} threads, {nray} rays x 3")
        v, bsec, tsec = cpu_baseline(verts, tris, rays, threads)
        out["cpu_baseline"] = {
            "value": round(v, 2),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"all {nray} rays of both sets, same mesh, median of 3 passes of {tsec:.2f}s "
                      f"({threads} threads, {threads * 3 * tsec:.0f} core-seconds); QBVH build {bsec:.2f}s not timed",
            "host": host,
        }
        out["gpu_over_cpu"] = round(mrays / v, 1)
    if not args.no_pt:
        out["pt_cornell"] = pt_cornell_section(dev, not args.no_cpu_baseline, threads)
        out["pt"] = pt_section(dev, not args.no_cpu_baseline, threads)
    if not args.no_bdpt:
        out["bdpt"] = bdpt_section(dev, not args.no_cpu_baseline, threads)
    if not args.no_vcm:
        out["vcm"] = vcm_section(dev, not args.no_cpu_baseline, threads)
    if not args.no_guide:
        cache = {}
        out["guide"] = guide_section(dev, args.guide_ref_spp, lr=1e-3, cache=cache)
        # the same trials at a larger learning rate, reported beside the
        # workload's 1e-3
        out["guide"]["alt_lr"] = guide_section(dev, args.guide_ref_spp, lr=5e-3, cache=cache, mfma=False)
    if not args.no_image:
        out["image"] = image_section(dev, not args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
