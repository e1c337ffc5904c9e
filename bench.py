"""Benchmark of the trace step: Mrays/s on the 1080p ~2M-triangle animated
scene (1,024 objects + 64 alpha fences + 32 mirrors + ground, 4 blue-noise
soft-shadow rays per hit, up to 3 traces per path, a TLAS refit every frame).

A step is one frame: TLAS refit with that frame's instance transforms
(already resident in HBM) + the full trace (raygen, closest-hit traversal,
shading, shadow occlusion, resolve). Two frames are in flight, each with its
own TLAS copy, stream and output images, as the reference's frame loop keeps
them (src/main.rs:600-668); --frames-in-flight 1 traces one after another.
Rays counted = primary + bounce + shadow rays actually traced, from the
device counters of a replay of the same frames. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ray-tracing-gallery_amd"))

# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues per process (4 by
# default). This process holds two frames in flight (a stream and a bounce
# stream each) and servers with up to four batches in flight (again two
# streams each); when busy streams share a queue they serialise. Measured on
# MI355X, served mode: 4 queues 396, 8 queues 446-466, 16 queues 516
# requests/s; the frame rate is the same with 8 and 16. Raised to 16 before
# the HIP runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s at 1080p 2M-tri scene; p50/p99 request latency, served vs plain"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(c, kind):
    """Bytes the launches of a traversal kernel on the frame's own stream must
    move per the path's data layout: the ray (32 B) and its result (24 B hit /
    4 B flag), 128 B per 4-wide node record visited, 48 B per triangle record
    tested, 64 B instance record + 48 B world-to-object transform per BLAS
    entered. Shadow rays are walked in pairs of one path's rays, and the
    occlusion counters count a pair's record fetch once, so these are the
    records the walks load. The work of the bounce stream (the side_*
    counters) is left out: it is timed apart."""
    if kind == "closest":
        side = c["side_node_visits"] > 0
        rays = c["primary_rays"] + (0 if side else c["bounce_rays"])
        return (rays * (32 + 24) + (c["node_visits"] - c["side_node_visits"]) * 128 +
                (c["triangle_tests"] - c["side_triangle_tests"]) * 48 +
                (c["instance_visits"] - c["side_instance_visits"]) * 112)
    rays = c["shadow_rays"] - c["side_shadow_rays"]
    return (rays * (32 + 4) + (c["occ_node_visits"] - c["side_occ_node_visits"]) * 128 +
            (c["occ_triangle_tests"] - c["side_occ_triangle_tests"]) * 48 +
            (c["occ_instance_visits"] - c["side_occ_instance_visits"]) * 112)


def shade_bytes(c):
    """Bytes the frame stream's shading launch (trace level 0) must move per
    the path's data layout: per ray its queue entry (36 B) and hit record
    (24 B); per miss its colour (16 B); per textured hit the instance record
    (64 B), ModelInfo (32 B) and GeometryInfo (24 B), the indices (12 B), three
    vertices' positions, normals and uvs (96 B), one texel each of the
    base-colour and metallic-roughness textures (16 B each, RGBA32F) and the
    two shading records (32 B), and per shadow ray its queue entry (36 B) and
    two blue-noise texels (4 B each); per mirror hit the instance record,
    tables and indices (132 B), three normals (36 B) and the continuation
    ray's queue entry (36 B). Level 0 is told apart from the bounce stream's
    levels by the side_* counters."""
    side = c["side_node_visits"] > 0 or c["side_textured_hits"] > 0 or c["side_bounce_rays"] > 0
    rays = c["primary_rays"] if side else c["primary_rays"] + c["bounce_rays"]
    tex = c["textured_hits"] - c["side_textured_hits"]
    shadow = c["shadow_rays"] - c["side_shadow_rays"]
    cont = c["bounce_rays"] - c["side_bounce_rays"] if side else c["bounce_rays"]
    misses = max(0, rays - tex - cont)
    return (rays * (36 + 24) + misses * 16 + tex * (64 + 32 + 24 + 12 + 96 + 2 * 16 + 32) +
            shadow * (36 + 2 * 4) + cont * (132 + 36 + 36))


def measured(kernel):
    """What the committed rocprofv3 passes over this benchmark measured for a
    launch of `kernel` (profiles/hbm_traffic_trace_kernels.json): HBM bytes
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), L2 hit
    rate (TCC_HIT / (TCC_HIT + TCC_MISS)) and L2 requests (TCC_REQ), or None
    for what is missing."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_trace_kernels.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None, None, None
    v = d.get("per_launch_bytes", {}).get(kernel)
    traffic = int(v["fetch"] + v["write"]) if v else None
    return traffic, d.get("l2_hit_rate", {}).get(kernel), d.get("l2_requests_per_launch", {}).get(kernel)


def kernel_roofline(name, bytes_per_launch, ms_per_launch, launches):
    """Algorithmic rate of a kernel's launches against the HBM peak, beside
    the measured DRAM traffic, L2 hit rate and L2 request rate of the same
    launches from the committed profile."""
    achieved = bytes_per_launch / (ms_per_launch / 1e3) / 1e9 if ms_per_launch > 0 else 0.0
    traffic, l2_hit, l2_req = measured(name)
    out = {"kernel": name, "launches": int(launches), "avg_launch_ms": round(ms_per_launch, 4),
           "algorithmic_bytes_per_launch": int(bytes_per_launch), "achieved": round(achieved, 1),
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "dram_gb_s": round(traffic / (ms_per_launch / 1e3) / 1e9, 1) if traffic and ms_per_launch > 0 else None,
           "l2_hit_rate": l2_hit,
           "l2_requests_per_launch": l2_req,
           "l2_g_requests_per_s": round(l2_req / (ms_per_launch / 1e3) / 1e9, 2) if l2_req and ms_per_launch > 0
           else None}
    out["dram_frac"] = round(out["dram_gb_s"] / HBM_PEAK_GBS, 4) if out["dram_gb_s"] is not None else None
    return out


def cpu_baseline(scene, threads, min_seconds=10.0):
    """The CPU oracle tracing frame 0 of the same scene on the host cores,
    repeated until about `min_seconds` of wall time (the first trace is the
    one checked against the GPU). Returns the host scene, the first trace,
    Mrays/s, the seconds spent and the number of traces."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_host as oh
    hs = oh.HostScene(scene)
    first, rays, traces = None, 0, 0
    t0 = time.perf_counter()
    while traces == 0 or time.perf_counter() - t0 < min_seconds:
        rgba, rgb, hits, cnt = hs.trace_image(threads=threads)
        rays += cnt.primary_rays + cnt.bounce_rays + cnt.shadow_rays
        traces += 1
        if first is None:
            first = (rgba, rgb, hits)
    dt = time.perf_counter() - t0
    return hs, first, rays / dt / 1e6, dt, traces


# Served mode: continuous batching with up to 32 requests and 1 M samples per
# batch, 5 batches in flight (the setting with the best throughput and p99
# in a sweep on MI355X with 16 hardware queues; larger batches delay small
# requests). The plain path is no server at all: one request at a time, a
# refit of its scene's TLAS copy and one rtg_trace_image.
SERVED = {"max_batch": 32, "max_inflight": 5, "max_samples": 1 << 20}
# Poisson arrival rates of the sweep, as fractions of the plain path's
# saturation throughput; the plain path's p99 at PLAIN_OPERATING of its
# saturation is the latency that "throughput at equal p99" holds
SWEEP = (0.25, 0.5, 0.7, 0.85, 1.0, 1.15, 1.3, 1.5, 1.75, 2.0)
PLAIN_OPERATING = 0.7


def wait_until(t0, at):
    while True:
        left = at - (time.perf_counter() - t0)
        if left <= 0:
            return
        if left > 1e-3:
            time.sleep(left - 5e-4)


def plain_replay(ctx, items, arrivals, stream):
    """The plain single-request GPU path: each request, at its arrival, refits
    its scene's TLAS copy with the request's transforms and traces one image
    (rtg_trace_image) on `stream`; the host waits for the GPU before it takes
    the next request. items: (tlas copy, instance buffer, instances, job).
    Returns the latencies (s), the makespan (s) and the GPU-busy seconds
    (HIP events around each request's work)."""
    import ctypes as C
    import numpy as np
    import torch
    from rtg.lib import check, load
    L = load()
    sp = C.c_void_p(stream.cuda_stream)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in items]
    done = []
    t0 = time.perf_counter()
    for (tl, buf, n, job), at, (e0, e1) in zip(items, arrivals, ev):
        wait_until(t0, at)
        e0.record(stream)
        check(L.rtg_tlas_refit(ctx.ptr, tl, buf.data_ptr(), n, sp), "rtg_tlas_refit")
        check(L.rtg_trace_image(ctx.ptr, C.byref(job), sp), "rtg_trace_image")
        e1.record(stream)
        stream.synchronize()
        done.append(time.perf_counter() - t0)
    lat = np.asarray(done) - np.asarray(arrivals)
    busy = sum(a.elapsed_time(b) for a, b in ev) / 1e3
    return lat, done[-1] - float(arrivals[0]), busy


def served_replay(srv, reqs, arrivals):
    comp = srv.replay(reqs, arrivals)
    assert all(c.status == 0 for c in comp), [c.status for c in comp if c.status != 0][:4]
    import numpy as np
    lat = np.array([c.done_s - c.submit_s for c in comp])
    span = max(c.done_s for c in comp) - min(c.submit_s for c in comp)
    return lat, span, max(c.batch_size for c in comp)


def serving_sweep(ctx, resident, stream_reqs, seed, label):
    """Plain vs served over one request set: saturation throughput (all
    requests at t = 0), then p50/p99 latency and throughput under seeded
    Poisson arrivals at the SWEEP rates; the served throughput at the plain
    path's p99 (at PLAIN_OPERATING of its saturation); the plain path's GPU
    idle fraction at saturation, which bounds what any batching of the same
    requests can gain by filling idle time."""
    import numpy as np
    import torch
    from rtg import scenes
    from rtg.serve import Server, make_request

    items, reqs, keep = [], [], []
    clones = [ds.clone_tlas() for ds in resident]
    for i, q in enumerate(stream_reqs):
        ds = resident[q["scene"]]
        w, h = q["width"], q["height"]
        buf = ds.instance_buffer(q["transforms"])
        out = torch.empty(w * h, dtype=torch.int32, device="cuda:0")
        r = make_request(q["scene"], ds, w, h, q["spp"], q["shadow_rays"], 3, frame=q["frame"], instances=buf,
                         outputs=(out, None, None), tag=i)
        job = ds.job((out, None, None), tlas=clones[q["scene"]], width=w, height=h, uniforms=r.uniforms,
                     opts=r.opts)
        reqs.append(r)
        items.append((clones[q["scene"]], buf, ds.n_instances, job))
        keep.append((buf, out))
    n = len(reqs)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    srv = Server(ctx, budget_bytes=48 << 30, **SERVED)
    for ds in resident:
        srv.add_scene(ds)
    # warm both paths (arenas sized, code loaded)
    plain_replay(ctx, items[:8], np.zeros(8), stream)
    served_replay(srv, reqs[:8], np.zeros(8))
    lat, span_p, busy = plain_replay(ctx, items, np.zeros(n), stream)
    lat, span_s, maxb = served_replay(srv, reqs, np.zeros(n))
    sat_plain, sat_served = n / span_p, n / span_s
    log(f"{label} saturation: plain {sat_plain:.1f} rps (GPU idle {1 - busy / span_p:.3f}), served "
        f"{sat_served:.1f} rps")
    sweep = []
    for k, frac in enumerate(SWEEP):
        rate = frac * sat_plain
        arrivals = np.cumsum(-np.log(1.0 - scenes.Rng(seed + 100 + k).uniform(n)) / rate)
        arrivals -= arrivals[0]
        lp, sp_, _ = plain_replay(ctx, items, arrivals, stream)
        ls, ss, mb = served_replay(srv, reqs, arrivals)
        pt = {"rate_rps": round(rate, 2), "fraction_of_plain_saturation": frac}
        for name, l, span in (("plain", lp, sp_), ("served", ls, ss)):
            pt[name] = {"p50_ms": round(float(np.percentile(l, 50)) * 1e3, 2),
                        "p99_ms": round(float(np.percentile(l, 99)) * 1e3, 2),
                        "throughput_rps": round(n / span, 2)}
        pt["served"]["max_batch"] = mb
        sweep.append(pt)
        log(f"{label} rate {rate:.1f}: plain p50/p99 {pt['plain']['p50_ms']}/{pt['plain']['p99_ms']} ms, "
            f"served {pt['served']['p50_ms']}/{pt['served']['p99_ms']} ms")
    srv.close()
    op = next(pt for pt in sweep if pt["fraction_of_plain_saturation"] == PLAIN_OPERATING)
    p99_ref = op["plain"]["p99_ms"]
    ok = [pt for pt in sweep if pt["served"]["p99_ms"] <= p99_ref]
    best = max(ok, key=lambda pt: pt["served"]["throughput_rps"]) if ok else None
    equal = {"p99_ms": p99_ref, "plain_throughput_rps": op["plain"]["throughput_rps"],
             "served_throughput_rps": best["served"]["throughput_rps"] if best else None,
             "served_rate_rps": best["rate_rps"] if best else None,
             "ratio": round(best["served"]["throughput_rps"] / op["plain"]["throughput_rps"], 3) if best else None}
    idle = 1.0 - busy / span_p
    for c in clones:
        from rtg.lib import load
        load().rtg_tlas_destroy(ctx.ptr, c)
    return {"requests": n,
            "saturation": {"plain_rps": round(sat_plain, 2), "served_rps": round(sat_served, 2),
                           "served_vs_plain": round(sat_served / sat_plain, 3), "served_max_batch": maxb},
            "plain_gpu_idle_fraction": round(idle, 4),
            "plain_idle_bound": round(1.0 / max(1e-9, 1.0 - idle), 3),
            "sweep": sweep, "throughput_at_equal_p99": equal}


def serving_benchmark(ctx, n_scenes=32, n_requests=160, seed=0x5EED0004, c2_requests=48):
    """Served stream (C5): n_scenes small-tori scenes resident in HBM;
    requests of 256x256, 1280x720 or 1920x1080 with 1-8 samples per pixel,
    each with a random instance-transform update; and the same comparison on
    the 1080p ~2M-triangle animated scene (C2), each request a 1080p frame
    with its own transforms. Plain = no server (see plain_replay); served =
    continuous batching (SERVED). Each reports saturation throughput, a
    Poisson-rate sweep of p50/p99, the served throughput at the plain path's
    p99 and the plain path's GPU idle fraction."""
    import numpy as np
    import torch
    from rtg import scenes
    from rtg.device import DeviceScene

    t0 = time.perf_counter()
    scene_list, stream = scenes.served_stream(n_scenes=n_scenes, n_requests=n_requests, seed=seed)
    resident = [DeviceScene(ctx, s) for s in scene_list]
    for q in stream:
        q["shadow_rays"] = 1
    samples = sum(q["width"] * q["height"] * q["spp"] for q in stream)
    log(f"serving setup {time.perf_counter() - t0:.1f}s: {n_scenes} scenes, {n_requests} requests, "
        f"{samples / 1e6:.1f} M samples")
    c5 = serving_sweep(ctx, resident, stream, seed, "C5")
    # exact ray count of the request set (counters of plain traces)
    ctx.set_counters(True)
    rays = 0
    import ctypes as C
    from rtg.lib import check, load
    for q in stream:
        ds = resident[q["scene"]]
        clone = ds.clone_tlas()
        ds.refit(ds.instance_buffer(q["transforms"]), tlas=clone)
        out = torch.empty(q["width"] * q["height"], dtype=torch.int32, device="cuda:0")
        from rtg.serve import make_request
        r = make_request(q["scene"], ds, q["width"], q["height"], q["spp"], 1, 3, frame=q["frame"])
        j = ds.job((out, None, None), tlas=clone, width=q["width"], height=q["height"], uniforms=r.uniforms,
                   opts=r.opts)
        check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
        c = ctx.counters()
        rays += c.primary_rays + c.bounce_rays + c.shadow_rays
        load().rtg_tlas_destroy(ctx.ptr, clone)
    ctx.set_counters(False)
    c5.update({"scenes": n_scenes, "samples": samples, "rays": int(rays),
               "resolutions": ["256x256", "1280x720", "1920x1080"], "spp": "1-8", "shadow_rays": 1,
               "served_config": SERVED})
    c5["saturation"]["plain_mrays_s"] = round(rays * c5["saturation"]["plain_rps"] / n_requests / 1e6, 1)
    c5["saturation"]["served_mrays_s"] = round(rays * c5["saturation"]["served_rps"] / n_requests / 1e6, 1)
    for d in resident:
        d.destroy()
    del resident
    torch.cuda.empty_cache()
    # C2: the 1080p animated scene, one request per frame
    t0 = time.perf_counter()
    big = scenes.large_scene()
    dbig = DeviceScene(ctx, big)
    reqs2 = [{"scene": 0, "width": big.width, "height": big.height, "spp": 1, "shadow_rays": big.shadow_rays,
              "frame": f, "transforms": big.transforms_for_frame(f)} for f in range(c2_requests)]
    log(f"C2 serving setup {time.perf_counter() - t0:.1f}s: {c2_requests} 1080p frames")
    c2 = serving_sweep(ctx, [dbig], reqs2, seed + 7, "C2")
    c2.update({"scene": "1080p animated scene, 1,121 instances, 2,176,034 instanced triangles", "spp": 1,
               "shadow_rays": big.shadow_rays, "served_config": SERVED})
    dbig.destroy()
    return {"c5": c5, "c2": c2,
            "served_vs_plain_throughput": c5["saturation"]["served_vs_plain"],
            "throughput_at_equal_p99_ratio": c5["throughput_at_equal_p99"]["ratio"]}


def batched_benchmark(ctx, steps, seed=0x5EED0003):
    """Batched launch: 64 independent scenes (triangle counts log-uniform in
    [1k, 500k], resolutions from {128^2, 512^2, 1280x720, 1920x1080}, 1-4
    spp, 4 soft-shadow rays per hit) traced by one graph launch with the
    autotuner's launch setting. Also times the same jobs through the plain
    single-image path one after another, and as one eager batch."""
    import numpy as np
    import torch
    from rtg import abi, scenes
    from rtg.batch import Batch, autotune, trace_batch
    from rtg.device import DeviceScene
    import ctypes as C
    from rtg.lib import check, load

    t0 = time.perf_counter()
    ws = scenes.batch_workload(seed=seed)
    t_gen = time.perf_counter() - t0
    dss = [DeviceScene(ctx, s) for s in ws]
    outs = [ds.make_outputs(hits=False) for ds in dss]
    jobs = [ds.job(o, frame_index=0) for ds, o in zip(dss, outs)]
    torch.cuda.synchronize()
    samples = sum(s.width * s.height * s.spp for s in ws)
    tris = sum(s.instanced_triangles for s in ws)
    log(f"batched setup {time.perf_counter() - t0:.1f}s (generation {t_gen:.1f}s): {len(ws)} scenes, "
        f"{tris} instanced triangles, {samples / 1e6:.1f} M samples")

    def wall(fn, reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            a = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        return float(np.median(ts))

    def plain():
        for j in jobs:
            check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")

    default = ctx.tuning()
    t_plain = wall(plain, 1)
    t_eager = wall(lambda: trace_batch(ctx, jobs), 2)
    a = time.perf_counter()
    best, tuned_ms = autotune(ctx, jobs, repeats=3)
    t_tune = time.perf_counter() - a
    log(f"autotune {t_tune:.1f}s: {best.as_dict()} {tuned_ms:.1f} ms")
    b = Batch(ctx, jobs)
    b.launch()
    torch.cuda.synchronize()
    ctx.set_timing(True)
    a = time.perf_counter()
    for _ in range(steps):
        b.launch()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - a) / steps
    g_ms, g_n = ctx.kernel_time(abi.RTG_KERNEL_BATCH_GRAPH)
    ctx.set_timing(False)
    b.close()
    # exact ray count: a counting replay of the same batch
    ctx.set_counters(True)
    bc = Batch(ctx, jobs)
    bc.launch()
    c = ctx.counters()
    bc.close()
    ctx.set_counters(False)
    rays = c.primary_rays + c.bounce_rays + c.shadow_rays
    ctx.set_tuning(default)
    for d in dss:
        d.destroy()
    del outs
    torch.cuda.empty_cache()
    return {"scenes": len(ws), "instanced_triangles": int(tris), "samples": int(samples), "rays": int(rays),
            "tri_range": [1000, 500000], "resolutions": ["128x128", "512x512", "1280x720", "1920x1080"],
            "spp": "1-4", "shadow_rays": 4, "seed": seed,
            "tuning": best.as_dict(), "autotune_s": round(t_tune, 1),
            "graph": {"ms_per_batch": round(g_ms / max(1, g_n), 3), "wall_ms_per_batch": round(t_graph * 1e3, 3),
                      "mrays_s": round(rays / (g_ms / max(1, g_n) / 1e3) / 1e6, 1), "launches": int(g_n)},
            "eager_batch": {"wall_ms": round(t_eager * 1e3, 3), "mrays_s": round(rays / t_eager / 1e6, 1)},
            "plain_one_by_one": {"wall_ms": round(t_plain * 1e3, 3), "mrays_s": round(rays / t_plain / 1e6, 1)},
            "graph_vs_plain": round(t_plain / t_graph, 3)}


def blas_build_times(ctx, ds, scene, reps=2):
    """Wall time of rtg_blas_build (synchronous) over the scene's first mesh,
    already resident on the device, as the plain LBVH and with
    PREFER_FAST_TRACE (five treelet passes); the best of `reps` builds each."""
    import ctypes as C
    import torch
    from rtg import abi
    from rtg.lib import check, load
    L = load()
    m = scene.meshes[0]
    pos = ds.mesh_bufs[0][0]
    idx = ds.mesh_bufs[0][3]
    geoms = (abi.Geometry * len(m.geometries))()
    for g, (ix, op) in enumerate(m.geometries):
        geoms[g].vertex_data = pos.data_ptr()
        geoms[g].vertex_stride = 12
        geoms[g].max_vertex = max(0, len(m.positions) - 1)
        geoms[g].index_data = idx[g].data_ptr()
        geoms[g].index_count = ix.size
        geoms[g].flags = abi.RTG_GEOMETRY_OPAQUE if op else 0
    out = {"triangles": int(sum(ix.shape[0] for ix, _ in m.geometries))}
    for name, flags in (("plain", 0), ("fast_trace", abi.RTG_BUILD_PREFER_FAST_TRACE)):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            h = C.c_uint64()
            t = time.perf_counter()
            check(L.rtg_blas_build(ctx.ptr, C.cast(geoms, C.c_void_p), len(m.geometries), flags, C.byref(h), None),
                  "rtg_blas_build")
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) * 1e3
            L.rtg_blas_destroy(ctx.ptr, h.value)
            best = ms if best is None else min(best, ms)
        out[name] = round(best, 2)
    return out


def incoherent_benchmark(ctx, n_rays=16 << 20, sample=65536):
    """Incoherent-ray stress: one BLAS of 1,000,000 random triangles
    (lognormal edge lengths, median 0.01) in the unit cube; 16M rays with
    origins uniform in the cube and directions uniform on the sphere, tmin 0,
    tmax +inf; closest-hit and occlusion queries. Reports Mrays/s, node
    visits per ray and algorithmic GB/s, and checks a ray sample bit for bit
    against the oracle (also timed there, on the host cores)."""
    import numpy as np
    import torch
    from rtg import abi, scenes
    from rtg.device import DeviceScene, hits_to_numpy, rays_to_device

    t0 = time.perf_counter()
    scene = scenes.incoherent_scene()
    rays = scenes.incoherent_rays(n_rays)
    ds = DeviceScene(ctx, scene)
    rd = rays_to_device(rays)
    hits = torch.empty(n_rays * 6, dtype=torch.int32, device="cuda:0")
    occ = torch.empty(n_rays, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    log(f"incoherent setup {time.perf_counter() - t0:.1f}s")
    out = {"triangles": scene.instanced_triangles, "rays": n_rays, "seed": scene.seed}
    for name, mode, buf, kind in (("closest", abi.RTG_QUERY_CLOSEST, hits, abi.RTG_KERNEL_TRACE_CLOSEST),
                                  ("occlusion", abi.RTG_QUERY_OCCLUSION, occ, abi.RTG_KERNEL_TRACE_OCCLUSION)):
        ds.trace_rays(rd, n_rays, mode, out=buf)
        torch.cuda.synchronize()
        ctx.set_timing(True)
        for _ in range(3):
            ds.trace_rays(rd, n_rays, mode, out=buf)
        torch.cuda.synchronize()
        ms, n = ctx.kernel_time(kind)
        ctx.set_timing(False)
        ms /= max(1, n)
        ctx.set_counters(True)
        ds.trace_rays(rd, n_rays, mode, out=buf)
        c = ctx.counters()
        ctx.set_counters(False)
        if name == "closest":
            nb = n_rays * (32 + 24) + c.node_visits * 128 + c.triangle_tests * 48 + c.instance_visits * 112
            nodes, trisv = c.node_visits, c.triangle_tests
        else:
            nb = n_rays * (32 + 4) + c.occ_node_visits * 128 + c.occ_triangle_tests * 48 + c.occ_instance_visits * 112
            nodes, trisv = c.occ_node_visits, c.occ_triangle_tests
        traffic, l2_hit, l2_req = measured(f"k_trace_rays_{name}")
        out[name] = {"ms": round(ms, 3), "mrays_s": round(n_rays / (ms / 1e3) / 1e6, 1),
                     "nodes_per_ray": round(nodes / n_rays, 2), "tris_per_ray": round(trisv / n_rays, 2),
                     "algorithmic_gb_s": round(nb / (ms / 1e3) / 1e9, 1),
                     "measured_traffic_gb_s": round(traffic / (ms / 1e3) / 1e9, 1) if traffic else None,
                     "l2_hit_rate": l2_hit,
                     "l2_g_requests_per_s": round(l2_req / (ms / 1e3) / 1e9, 2) if l2_req else None}
    torch.cuda.synchronize()
    # bit-exact check of a ray sample against the oracle (same BVH, same order)
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_host as oh
    idx = np.linspace(0, n_rays - 1, sample).astype(np.int64)
    hs = oh.HostScene(scene)
    threads = len(os.sched_getaffinity(0))
    a = time.perf_counter()
    o_hits, _ = hs.trace_rays(rays[idx], abi.RTG_QUERY_CLOSEST, threads=threads)
    o_occ, _ = hs.trace_rays(rays[idx], abi.RTG_QUERY_OCCLUSION, threads=threads)
    cpu_s = time.perf_counter() - a
    g_hits = hits_to_numpy(hits)[idx]
    g_occ = occ.cpu().numpy()[idx]
    out["parity_sample"] = {"rays": sample,
                            "hits_bit_exact": bool(np.array_equal(g_hits.view(np.uint32), o_hits.view(np.uint32))),
                            "occlusion_exact": bool(np.array_equal(g_occ.astype(np.uint32),
                                                                   np.asarray(o_occ).astype(np.uint32)))}
    out["cpu_oracle_mrays_s"] = round(2 * sample / cpu_s / 1e6, 3)
    out["cpu_cores"] = threads
    out["blas_build_ms"] = blas_build_times(ctx, ds, scene)
    ds.destroy()
    del hits, occ, rd
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames-in-flight", type=int, default=2,
                    help="frames traced concurrently on their own streams and TLAS copies, as the reference's "
                         "frame loop keeps them (src/main.rs:600-668); 1 = one frame after another")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity check")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream benchmark")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene graph-batch benchmark")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-incoherent", action="store_true", help="skip the 16M incoherent-ray benchmark")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one elementwise kernel that reads and writes BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU: the trace step has no cross-GPU exchange; "
                                   "run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    from rtg import abi, scenes
    from rtg.device import DeviceScene, hits_to_numpy
    from rtg.frames import FrameDriver
    from rtg.lib import Context

    torch.cuda.init()  # torch's device state first, then the library's context

    if args.copy_calibration:
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)  # reads n*4 bytes, writes n*4 bytes
        torch.cuda.synchronize()
        del src, dst
        log(f"copy calibration: {n * 4} bytes read and written")

    t_setup = time.perf_counter()
    scene = scenes.large_scene(width=args.width, height=args.height)
    ctx = Context(0)
    ds = DeviceScene(ctx, scene, "cuda:0")
    W, K = args.warmup, args.steps
    # each frame's instance transforms (the reference's scene update,
    # src/scene.rs:167-204), computed ahead; writing them into the slot's
    # mapped instance array is part of the timed frame
    xfs = [np.asarray(scene.transforms_for_frame(f), dtype=np.float32) for f in range(W + K)]
    # frames in flight (src/main.rs:600-668): slot f % FIF has a TLAS copy,
    # a pinned instance array, a frame state, a stream and output images
    FIF = max(1, args.frames_in_flight)
    fd = FrameDriver(ds, FIF)
    slot_stream = [torch.cuda.Stream() for _ in range(FIF)]
    slot_out = [ds.make_outputs(hits=False) for _ in range(FIF)]
    slot_job = [ds.job(slot_out[k], frame_index=0) for k in range(FIF)]
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t_setup:.1f}s: {len(scene.instances)} instances, "
        f"{scene.instanced_triangles} instanced triangles, {args.width}x{args.height}, {FIF} frames in flight")

    def frame(f):
        # wait for the slot, write this frame's transforms into its mapped
        # array, then upload, refit and trace on the slot's stream
        k = f % FIF
        inst, _ = fd.begin(f)
        inst["transform"] = xfs[f]
        slot_job[k].uniforms.frame_index = f
        fd.submit(f, slot_job[k], stream=slot_stream[k])

    for f in range(W):
        frame(f)
    torch.cuda.synchronize()
    log("warmup done")
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(W, W + K):
        frame(f)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ktime = {name: ctx.kernel_time(k) for k, name in enumerate(abi.RTG_KERNEL_NAMES)}
    ctx.set_timing(False)
    log(f"timed {K} frames in {t1 - t0:.3f}s")

    # exact ray and traversal counts of the same frames
    ctx.set_counters(True)
    tot = None
    for f in range(W, W + K):
        frame(f)
        c = ctx.counters().as_dict()
        tot = c if tot is None else {k: tot[k] + c[k] for k in tot}
    ctx.set_counters(False)
    overflow = bool(ctx.status() & abi.RTG_STATUS_STACK_OVERFLOW)
    rays = tot["primary_rays"] + tot["bounce_rays"] + tot["shadow_rays"]
    elapsed = t1 - t0
    value = rays / elapsed / 1e6

    # rooflines of the frame stream's traversal and shading launches: the
    # algorithmic bytes of their launches (the counters of the same frames)
    # over their HIP-event time on that stream, beside what the committed
    # rocprofv3 passes measured for the same launches. The dominant kernel,
    # whose figures are the line's `roofline`, is the one that moves the most
    # algorithmic bytes: the level-0 shadow rays.
    per_kernel = {}
    for kind, name, nbytes in (("trace_closest", "k_trace_closest", algorithmic_bytes(tot, "closest")),
                               ("trace_occlusion", "k_trace_occlusion", algorithmic_bytes(tot, "occlusion")),
                               ("shade", "k_shade", shade_bytes(tot))):
        ms, launches = ktime[kind]
        per_kernel[name] = kernel_roofline(name, nbytes / max(1, launches), ms / max(1, launches), launches)
    dom = max(("k_trace_closest", "k_trace_occlusion"), key=lambda k: per_kernel[k]["algorithmic_bytes_per_launch"])
    roofline = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", **per_kernel[dom],
                "reading": "achieved/frac: algorithmic bytes (node, triangle and instance records per the data "
                           "layout) over launch time; the records are served mostly from L1 and L2, so the "
                           "measured DRAM rate (dram_gb_s, dram_frac) is far lower: the traversal is bound by "
                           "the latency of its dependent loads, not by HBM bandwidth",
                "kernels": per_kernel}
    kernel_ms = {k: round(v[0] / max(1, K), 4) for k, v in ktime.items()}

    result = {"metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": K, "warmup": W,
              "ms_per_step": round(elapsed / K * 1e3, 3), "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32", "data": "synthetic",
              "config": {"workload": "1080p animated scene: 1024 objects (tori 48x24, level-4 icospheres, "
                                     "lognormal triangle soups) + 64 alpha-clipped fences + 32 mirrors + ground; "
                                     "4 blue-noise soft-shadow rays per hit, 2 deg cone; up to 3 traces per path; "
                                     "per frame the host writes the instance transforms into a mapped array, "
                                     "uploads them and refits the TLAS",
                         "width": args.width, "height": args.height, "instances": len(scene.instances),
                         "instanced_triangles": scene.instanced_triangles, "spp": 1,
                         "rays_per_frame": int(rays / K), "frames_in_flight": FIF, "seed": scene.seed},
              "roofline": roofline, "kernel_ms_per_frame": kernel_ms,
              "traversal": {"nodes_per_closest_ray": round(tot["node_visits"] / max(1, tot["primary_rays"] +
                                                                                    tot["bounce_rays"]), 2),
                            # node records loaded per shadow ray (a pair walk's visit serves two rays)
                            "nodes_per_shadow_ray": round(tot["occ_node_visits"] / max(1, tot["shadow_rays"]), 2),
                            "spill_pushes_per_closest_ray": round(tot["spill_pushes"] / max(1, tot["primary_rays"] +
                                                                                            tot["bounce_rays"]), 4),
                            "spill_pushes_per_shadow_ray": round(tot["occ_spill_pushes"] /
                                                                 max(1, tot["shadow_rays"]), 4),
                            "node_width": 4, "stack_overflow": overflow}}
    fd.close()

    if not args.no_serve:
        t_s = time.perf_counter()
        result["serving"] = serving_benchmark(ctx, n_requests=args.serve_requests)
        log(f"serving benchmark {time.perf_counter() - t_s:.1f}s")

    if not args.no_batch:
        t_b = time.perf_counter()
        result["batched"] = batched_benchmark(ctx, args.batch_steps)
        log(f"batched benchmark {time.perf_counter() - t_b:.1f}s")

    if not args.no_incoherent:
        t_i = time.perf_counter()
        result["incoherent"] = incoherent_benchmark(ctx)
        log(f"incoherent benchmark {time.perf_counter() - t_i:.1f}s")

    if not args.no_cpu:
        # GPU frame 0 with primary hits, for the full-size parity check (the
        # scene's TLAS as built holds frame 0's transforms)
        ds.refit(ds.instances)
        g_out = ds.make_outputs(hits=True)
        ds.trace_image(frame_index=0, outputs=g_out)
        torch.cuda.synchronize()
        threads = len(os.sched_getaffinity(0))
        hs, (o_rgba, o_rgb, o_hits), cpu_mrays, cpu_s, cpu_n = cpu_baseline(scene, threads)
        g_hits = hits_to_numpy(g_out[2])
        hits_equal = bool(np.array_equal(g_hits.view(np.uint32), o_hits.view(np.uint32)))
        rgb_err = float(np.abs(g_out[1].cpu().numpy().reshape(-1, 3) - o_rgb).max())
        rgba_equal = bool(np.array_equal(g_out[0].cpu().numpy().view(np.uint32), o_rgba))
        result["cpu_baseline"] = {"value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"frame 0 of the same scene, all {args.width}x{args.height} pixels, "
                                            f"traced {cpu_n} times in {cpu_s:.1f} s"}
        result["parity"] = {"frame": 0, "primary_hits_bit_exact": hits_equal, "rgba8_exact": rgba_equal,
                            "rgb_max_abs_err": rgb_err}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
