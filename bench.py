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
    entered. The work of the bounce stream (the side_* counters) is left out:
    it is timed apart."""
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


def measured_l2_hit_rate(kernel):
    """L2 (TCC) hit rate of `kernel`'s launches from the committed rocprofv3
    summary of this benchmark (TCC_HIT / (TCC_HIT + TCC_MISS)), or None."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_trace_kernels.json")
    try:
        with open(path) as f:
            return json.load(f).get("l2_hit_rate", {}).get(kernel)
    except (OSError, ValueError):
        return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration) from the committed rocprofv3 summary of this
    benchmark, or None."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_trace_kernels.json")
    try:
        v = json.load(open(path))["per_launch_bytes"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return int(v["fetch"] + v["write"])


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


# (requests per batch, batches in flight, samples per batch): plain = one
# request per launch, one at a time; served = continuous batching with up to
# 5 batches in flight of at most 1 M samples each (the setting with the best
# throughput and p99 in a sweep on MI355X with 16 hardware queues; larger
# batches delay small requests)
SERVE_MODES = {"plain": (1, 1, 0), "served": (32, 5, 1 << 20)}


def serving_benchmark(ctx, n_scenes=32, n_requests=160, seed=0x5EED0004, modes=None):
    """Served stream: n_scenes small-tori scenes resident in HBM; requests of
    256x256, 1280x720 or 1920x1080 with 1-8 samples per pixel, each with a
    random instance-transform update (a refit of a private TLAS copy).
    Plain = one request per launch, one at a time; served = continuous
    batching (up to 32 requests and 1 M samples per batch, 4 batches in
    flight). Returns
    saturation throughput (all requests at t=0) and p50/p99 latency at a
    Poisson arrival rate of 0.6x the plain path's saturation throughput."""
    import numpy as np
    import torch
    from rtg import scenes
    from rtg.device import DeviceScene
    from rtg.serve import Server, make_request

    t0 = time.perf_counter()
    resident = [DeviceScene(ctx, scenes.small_scene(seed=seed + 1 + k, width=256, height=256))
                for k in range(n_scenes)]
    rng = scenes.Rng(seed)
    sizes = [(256, 256), (1280, 720), (1920, 1080)]
    reqs, keep = [], []
    paths = 0
    for i in range(n_requests):
        sid = int(rng.integers(1, 0, n_scenes)[0])
        ds = resident[sid]
        w, h = sizes[int(rng.integers(1, 0, 3)[0])]
        spp = int(rng.integers(1, 1, 9)[0])
        xf = [ins.transform.copy() for ins in ds.scene.instances]
        rot = rng.uniform(3 * len(xf)).reshape(-1, 3)
        for k in range(1, len(xf)):
            xf[k] = scenes.affine(xf[k][:, 3], scenes.rotation_from_uniform(*rot[k]), 1.0)
        buf = ds.instance_buffer(xf)
        out = torch.empty(w * h, dtype=torch.int32, device="cuda:0")
        r = make_request(sid, ds, w, h, spp, 1, 3, frame=i, instances=buf, outputs=(out, None, None), tag=i)
        reqs.append(r)
        keep.append((buf, out))
        paths += w * h * spp
    torch.cuda.synchronize()
    log(f"serving setup {time.perf_counter() - t0:.1f}s: {n_scenes} scenes, {n_requests} requests, "
        f"{paths / 1e6:.1f} M samples")

    def replay(srv, arrivals):
        comp = srv.replay(reqs[:len(arrivals)], arrivals)
        assert all(c.status == 0 for c in comp)
        lat = np.array([c.done_s - c.submit_s for c in comp]) * 1e3
        span = max(c.done_s for c in comp) - min(c.submit_s for c in comp)
        return lat, span, max(c.batch_size for c in comp), len({c.start_s for c in comp})

    # each mode: one server, warmed (arenas sized, code loaded), then all
    # requests at t=0 (saturation), then the same requests as Poisson
    # arrivals at 0.6x the plain path's saturation throughput
    modes = dict(modes or SERVE_MODES)
    assert next(iter(modes)) == "plain"
    sat, poisson, rate, arrivals = {}, {}, None, None
    for name, (mb, mi, ms) in modes.items():
        srv = Server(ctx, budget_bytes=48 << 30, max_batch=mb, max_inflight=mi, max_samples=ms)
        for ds in resident:
            srv.add_scene(ds)
        replay(srv, np.zeros(min(16, n_requests)))
        lat, span, maxb, nb = replay(srv, np.zeros(n_requests))
        sat[name] = {"throughput_rps": round(n_requests / span, 2), "makespan_s": round(span, 3),
                     "max_batch": maxb, "batches": nb}
        if rate is None:
            rate = 0.6 * sat["plain"]["throughput_rps"]
            arrivals = np.cumsum(-np.log(1.0 - scenes.Rng(seed + 99).uniform(n_requests)) / rate)
        lat, span, maxb, nb = replay(srv, arrivals)
        poisson[name] = {"p50_ms": round(float(np.percentile(lat, 50)), 2),
                         "p99_ms": round(float(np.percentile(lat, 99)), 2),
                         "throughput_rps": round(n_requests / span, 2), "max_batch": maxb}
        srv.close()
    # exact ray count of the request set (counters of a plain replay)
    ctx.set_counters(True)
    rays = 0
    srv = Server(ctx, budget_bytes=48 << 30, max_batch=1, max_inflight=1)
    for ds in resident:
        srv.add_scene(ds)
    for r in reqs:
        st, _ = srv.submit(r)
        srv.poll(wait_all=True)
        c = ctx.counters()
        rays += c.primary_rays + c.bounce_rays + c.shadow_rays
    srv.close()
    ctx.set_counters(False)
    for name in modes:
        sat[name]["mrays_s"] = round(rays / sat[name]["makespan_s"] / 1e6, 1)
    for d in resident:
        d.destroy()
    return {"scenes": n_scenes, "requests": n_requests, "samples": paths, "rays": int(rays),
            "resolutions": ["256x256", "1280x720", "1920x1080"], "spp": "1-8",
            "saturation": sat, "poisson_rate_rps": round(rate, 2), "poisson": poisson,
            "served_vs_plain_throughput": (round(sat["served"]["throughput_rps"] / sat["plain"]["throughput_rps"], 3)
                                           if "served" in sat and "plain" in sat else None)}


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
    best, tuned_ms = autotune(ctx, jobs, repeats=1)
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
        traffic = measured_traffic(f"k_trace_rays_{name}")
        out[name] = {"ms": round(ms, 3), "mrays_s": round(n_rays / (ms / 1e3) / 1e6, 1),
                     "nodes_per_ray": round(nodes / n_rays, 2), "tris_per_ray": round(trisv / n_rays, 2),
                     "algorithmic_gb_s": round(nb / (ms / 1e3) / 1e9, 1),
                     "measured_traffic_gb_s": round(traffic / (ms / 1e3) / 1e9, 1) if traffic else None,
                     "l2_hit_rate": measured_l2_hit_rate(f"k_trace_rays_{name}")}
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
    bufs = [ds.instance_buffer(scene.transforms_for_frame(f)) for f in range(W + K)]
    # frames in flight: frame f uses slot f % FIF (a TLAS copy, a stream and
    # output images of its own); the library keeps one frame state per stream
    FIF = max(1, args.frames_in_flight)
    slot_tlas = [ds.clone_tlas() for _ in range(FIF)] if FIF > 1 else [None]
    slot_stream = [torch.cuda.Stream() for _ in range(FIF)] if FIF > 1 else [None]
    slot_out = [ds.make_outputs(hits=False) for _ in range(FIF)]
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t_setup:.1f}s: {len(scene.instances)} instances, "
        f"{scene.instanced_triangles} instanced triangles, {args.width}x{args.height}, {FIF} frames in flight")

    def frame(f):
        k = f % FIF
        ds.refit(bufs[f], tlas=slot_tlas[k], stream=slot_stream[k])
        ds.trace_image(frame_index=f, outputs=slot_out[k], stream=slot_stream[k], tlas=slot_tlas[k])

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

    # the dominant traversal kernel: the one whose frame-stream launches move
    # the most algorithmic bytes (the level-0 shadow rays, three quarters of
    # the rays); by time the two kernels are close, and which one is longer
    # depends on how the two frames in flight interleave
    kinds = {"closest": ktime["trace_closest"], "occlusion": ktime["trace_occlusion"]}
    dom = max(kinds, key=lambda k: algorithmic_bytes(tot, k))
    ms, launches = kinds[dom]
    bytes_total = algorithmic_bytes(tot, dom)
    achieved = bytes_total / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    per_launch_bytes = bytes_total / max(1, launches)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(f"k_trace_{dom}"),
                "l2_hit_rate": measured_l2_hit_rate(f"k_trace_{dom}"),
                "kernel": f"k_trace_{'closest' if dom == 'closest' else 'occlusion'}",
                "avg_launch_ms": round(ms / max(1, launches), 4),
                "algorithmic_bytes_per_launch": int(per_launch_bytes)}
    kernel_ms = {k: round(v[0] / max(1, K), 4) for k, v in ktime.items()}

    result = {"metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": K, "warmup": W,
              "ms_per_step": round(elapsed / K * 1e3, 3), "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32", "data": "synthetic",
              "config": {"workload": "1080p animated scene: 1024 objects (tori 48x24, level-4 icospheres, "
                                     "lognormal triangle soups) + 64 alpha-clipped fences + 32 mirrors + ground; "
                                     "4 blue-noise soft-shadow rays per hit, 2 deg cone; up to 3 traces per path; "
                                     "TLAS refit every frame",
                         "width": args.width, "height": args.height, "instances": len(scene.instances),
                         "instanced_triangles": scene.instanced_triangles, "spp": 1,
                         "rays_per_frame": int(rays / K), "frames_in_flight": FIF, "seed": scene.seed},
              "roofline": roofline, "kernel_ms_per_frame": kernel_ms,
              "traversal": {"nodes_per_closest_ray": round(tot["node_visits"] / max(1, tot["primary_rays"] +
                                                                                    tot["bounce_rays"]), 2),
                            "nodes_per_shadow_ray": round(tot["occ_node_visits"] / max(1, tot["shadow_rays"]), 2),
                            "node_width": 4, "stack_overflow": overflow}}

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
        # GPU frame 0 with primary hits, for the full-size parity check
        ds.refit(bufs[0])
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
