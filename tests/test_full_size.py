"""The batched and served workloads at their full size, on the GPU:

* batched: the 64 scenes of 1k-500k triangles at 128^2 to 1080p with 1-4
  samples traced by one graph launch give, bit for bit, what each job gives
  traced alone; a job of more than 100k triangles equals the oracle on a
  seeded pixel subset;
* served: the 160-request stream over 32 resident scenes (256^2, 720p and
  1080p, 1-8 samples, a transform update each) served with continuous
  batching gives, bit for bit, what each request gives on the plain path
  (TLAS copy, refit, one rtg_trace_image); a sample of served requests
  equals the oracle on seeded pixel subsets.
"""
import copy
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_host as oh
from rtg import abi, scenes
from rtg.batch import Batch
from rtg.device import DeviceScene, hits_to_numpy
from rtg.lib import check, load
from rtg.serve import Server, make_request

pytestmark = pytest.mark.gpu

RGB_TOL = 1e-4


def oracle_pixels(host, job_uniforms, width, height, opts, pixels):
    """The oracle's RGBA8, linear RGB and primary hits of `pixels` of a job."""
    n = width * height
    rgba = np.zeros(n, np.uint32)
    rgb = np.zeros(n * 3, np.float32)
    hits = np.zeros(n * opts.spp, abi.HIT_DTYPE)
    j = abi.ImageJob()
    j.tlas = host.tlas.ptr
    j.scene = host.bindings
    j.uniforms = job_uniforms
    j.width, j.height = width, height
    j.opts = opts
    j.out_rgba8, j.out_rgb, j.out_primary_hits = rgba.ctypes.data, rgb.ctypes.data, hits.ctypes.data
    px = np.ascontiguousarray(pixels, np.uint32)
    st = oh.lib().orc_trace_image(C.byref(j), px.ctypes.data, len(px), 0, None)
    assert st == 0
    return rgba, rgb.reshape(-1, 3), hits


def check_subset(g_rgba, g_rgb, g_hits, o, pixels, spp):
    o_rgba, o_rgb, o_hits = o
    assert np.array_equal(g_rgba[pixels], o_rgba[pixels])
    assert float(np.abs(g_rgb[pixels] - o_rgb[pixels]).max()) <= RGB_TOL
    if g_hits is not None:
        samples = (pixels[:, None] * spp + np.arange(spp)[None, :]).reshape(-1)
        assert np.array_equal(g_hits[samples].view(np.uint32), o_hits[samples].view(np.uint32))


def test_batched_workload_full_size(ctx):
    ws = scenes.batch_workload()
    assert len(ws) == 64
    print(f"batched workload: {sum(s.instanced_triangles for s in ws)} triangles, "
          f"{sum(s.width * s.height * s.spp for s in ws)} samples", flush=True)
    dss = [DeviceScene(ctx, s) for s in ws]
    alone = [ds.make_outputs() for ds in dss]
    for ds, o in zip(dss, alone):
        j = ds.job(o, frame_index=0)
        check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
    outs = [ds.make_outputs() for ds in dss]
    b = Batch(ctx, [ds.job(o, frame_index=0) for ds, o in zip(dss, outs)])
    for replay in range(2):
        for o in outs:
            for t in o:
                t.fill_(-1)
        b.launch()
        torch.cuda.synchronize()
        for k, (a, g) in enumerate(zip(alone, outs)):
            for x, y in zip(a, g):
                assert torch.equal(x, y), (replay, k)
    b.close()
    # a job of more than 100k triangles against the oracle
    big = [k for k, s in enumerate(ws) if s.instanced_triangles > 100_000]
    assert big
    k = max(big, key=lambda i: ws[i].width * ws[i].height)
    s = ws[k]
    rgba, rgb, ph = (t.cpu().numpy() for t in outs[k])
    pixels = np.sort(np.random.default_rng(41).choice(s.width * s.height, 3000, replace=False))
    host = oh.HostScene(s)
    o = oracle_pixels(host, s.uniforms(0), s.width, s.height, s.trace_opts(), pixels)
    check_subset(rgba.view(np.uint32), rgb.reshape(-1, 3), hits_to_numpy(outs[k][2]), o, pixels, s.spp)
    for ds in dss:
        ds.destroy()


def test_served_stream_full_size(ctx):
    resident_scenes, stream = scenes.served_stream()
    print(f"served stream: {len(resident_scenes)} scenes, {len(stream)} requests, "
          f"{sum(r['width'] * r['height'] * r['spp'] for r in stream)} samples", flush=True)
    resident = [DeviceScene(ctx, s) for s in resident_scenes]
    sample = {1, 3, 20, 50}  # requests also checked against the oracle (with primary hits)
    reqs, outs, bufs = [], [], []
    for i, q in enumerate(stream):
        ds = resident[q["scene"]]
        n = q["width"] * q["height"]
        o = (torch.full((n,), -1, dtype=torch.int32, device="cuda:0"),
             torch.full((n * 3,), -1.0, dtype=torch.float32, device="cuda:0"),
             torch.full((n * q["spp"] * 6,), -1, dtype=torch.int32, device="cuda:0") if i in sample else None)
        buf = ds.instance_buffer(q["transforms"])
        reqs.append(make_request(q["scene"], ds, q["width"], q["height"], q["spp"], 1, 3, frame=q["frame"],
                                 instances=buf, outputs=o, tag=i))
        outs.append(o)
        bufs.append(buf)
    srv = Server(ctx, budget_bytes=48 << 30, max_batch=32, max_inflight=5, max_samples=1 << 20)
    for ds in resident:
        srv.add_scene(ds)
    comp = srv.replay(reqs, np.zeros(len(reqs)))
    stats = srv.stats()
    srv.close()
    assert all(c.status == 0 for c in comp)
    assert stats["max_batch"] > 1 and stats["max_inflight_seen"] > 1
    torch.cuda.synchronize()
    # each served result equals the plain path's: a TLAS copy refitted with
    # the request's transforms and one rtg_trace_image
    L = load()
    for i, (r, q) in enumerate(zip(reqs, stream)):
        ds = resident[q["scene"]]
        clone = ds.clone_tlas()
        ds.refit(bufs[i], tlas=clone)
        n = q["width"] * q["height"]
        p = (torch.zeros(n, dtype=torch.int32, device="cuda:0"), torch.zeros(n * 3, dtype=torch.float32,
                                                                              device="cuda:0"),
             torch.zeros(n * q["spp"] * 6, dtype=torch.int32, device="cuda:0") if i in sample else None)
        j = ds.job(p, tlas=clone, width=q["width"], height=q["height"], uniforms=r.uniforms, opts=r.opts)
        check(L.rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
        torch.cuda.synchronize()
        for x, y in zip(outs[i], p):
            if x is not None:
                assert torch.equal(x, y), i
        L.rtg_tlas_destroy(ctx.ptr, clone)
    # served results against the oracle on pixel subsets
    for i in sorted(sample):
        q = stream[i]
        s = copy.copy(resident_scenes[q["scene"]])
        host = oh.HostScene(s)
        host.refit(q["transforms"])
        rgba, rgb, ph = outs[i]
        pixels = np.sort(np.random.default_rng(43 + i).choice(q["width"] * q["height"], 2000, replace=False))
        o = oracle_pixels(host, reqs[i].uniforms, q["width"], q["height"], reqs[i].opts, pixels)
        check_subset(rgba.cpu().numpy().view(np.uint32), rgb.cpu().numpy().reshape(-1, 3), hits_to_numpy(ph), o,
                     pixels, q["spp"])
    for ds in resident:
        ds.destroy()
