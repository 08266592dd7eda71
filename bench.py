"""Benchmark: closest-hit traversal throughput (Mrays/s) on one MI355X.

Workload: the BVH traversal microbenchmark -- a 1,310,720-triangle icosphere
(8 subdivisions) displaced radially by seeded 3-octave value noise
(amplitude 0.1), and 8,388,608 rays: 2048^2 coherent pinhole primary rays
plus 2048^2 incoherent rays (origins uniform in the mesh bounds, directions
uniform on the sphere); closest hit, t in [1e-4, FLT_MAX].  One step = one
lmx_accel_intersect launch over all 8,388,608 rays, inputs resident in HBM.

Prints one JSON line.  `roofline` prices the traversal kernel's algorithmic
bytes -- 32 B ray read + 16 B hit write per ray, plus 128 B per wide-node
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
NODE_BYTES = 128
TRI_BYTES = 48
RAY_BYTES = 32 + 16


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per traversal launch measured with rocprofv3 PMC counters,
    if a summary with a `traffic_bytes_per_launch:` line is committed."""
    path = os.path.join(ROOT, "profiles", "traversal_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_baseline(verts, tris, rays, threads):
    """Reference-algorithm QBVH on the host (oracle restatement), timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    t0 = time.perf_counter()
    q = orc.QBVH(verts, tris)
    build_s = time.perf_counter() - t0
    q.intersect(rays[:4096], nthreads=threads)      # warm-up
    t0 = time.perf_counter()
    q.intersect(rays, nthreads=threads)
    dt = time.perf_counter() - t0
    return rays.shape[0] / dt / 1e6, build_s, dt


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
    rs.render_pt(1024 * 1024 * 4, **kw)               # autotune + warm-up
    torch.cuda.synchronize()
    times, rays = [], 0
    for _ in range(3):
        t0 = time.perf_counter()
        film, st, _ = rs.render_pt(n, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        rays = st["rays"]
    dt = min(times)
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
    restatement on host threads over a bounded prefix."""
    import torch
    from lmx import render, scenes
    log("bdpt: generating the sphere-grid scene")
    sc = scenes.sphere_grid()
    rs = render.Scene(sc)
    n = 1024 * 1024 * 16
    kw = dict(max_num_vertices=10, seed=42)
    rs.render_bdpt(1 << 22, **kw)                     # autotune + warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        film, st, _ = rs.render_bdpt(n, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = min(times)
    rays = st["rays"]
    out = {
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
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-pt", action="store_true", help="skip the wavefront path-tracer section")
    ap.add_argument("--no-bdpt", action="store_true", help="skip the bidirectional section")
    ap.add_argument("--only", choices=["pt", "bdpt"], default=None,
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
              "(rays shard across GPUs with no exchange, but this build runs on one GPU only)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from lmx import accel, scenes

    dev = torch.device("cuda:0")
    if args.only:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        fn = pt_section if args.only == "pt" else bdpt_section
        print(json.dumps({args.only: fn(dev, not args.no_cpu_baseline, threads)}), flush=True)
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
    traffic = traffic_from_profiles()

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
            "bvh": "8-wide, 16-bit quantised, 128 B nodes",
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
            "alg_bytes_per_launch": int(alg_bytes),
            "node_visits_per_ray": round(nodes_v / nray, 3),
            "tri_tests_per_ray": round(tris_t / nray, 3),
            "kernel_ms": round(avg_kernel_s * 1e3, 4),
        },
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        sample = np.concatenate([coh[::4], inc[::4]])
        log(f"cpu baseline: oracle QBVH, {threads} threads, {sample.shape[0]} rays")
        v, bsec, tsec = cpu_baseline(verts, tris, sample, threads)
        out["cpu_baseline"] = {
            "value": round(v, 2),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"every 4th ray of both sets ({sample.shape[0]} rays), same mesh; "
                      f"QBVH build {bsec:.2f}s (not timed), trace {tsec:.2f}s",
        }
    if not args.no_pt:
        out["pt"] = pt_section(dev, not args.no_cpu_baseline, args.cpu_threads or min(os.cpu_count() or 1, 16))
    if not args.no_bdpt:
        out["bdpt"] = bdpt_section(dev, not args.no_cpu_baseline, args.cpu_threads or min(os.cpu_count() or 1, 16))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
