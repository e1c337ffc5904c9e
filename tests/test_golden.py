"""Golden fixtures (tests/golden/*.npz, written by make_golden.py): the oracle
reproduces them bit for bit on the CPU; the HIP path reproduces them on the GPU
(hit records and RGBA8 exactly, linear RGB within 1e-4)."""
import os

import numpy as np
import pytest

import oracle_host as oh
from rtg import abi, scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FRAMES = ["small_frame.npz", "large_mini.npz", "small_spp.npz", "default_portal.npz"]


def load(name):
    a = dict(np.load(os.path.join(HERE, name)))
    return scenes.Scene.from_arrays(a), a


@pytest.mark.parametrize("name", FRAMES)
def test_oracle_reproduces_frame(name):
    s, a = load(name)
    hs = oh.HostScene(s)
    rgba, rgb, hits, cnt = hs.trace_image(frame_index=int(a["frame"][0]))
    assert np.array_equal(hits.view(np.uint32).reshape(-1, 6), a["out_hits"])
    assert np.array_equal(rgba, a["out_rgba8"])
    assert np.array_equal(rgb, a["out_rgb"])
    # the fixtures hold the counters the oracle fills (the leading fields);
    # the bounce-stream subset stays zero on the host
    got = [getattr(cnt, f) for f, _ in abi.TraceCounters._fields_]
    n = len(a["out_counters"])
    assert got[:n] == list(a["out_counters"]) and not any(got[n:])


def test_oracle_reproduces_rays():
    s, a = load("rays_small.npz")
    hs = oh.HostScene(s)
    rays = a["rays"].view(abi.RAY_DTYPE).reshape(-1)
    hits, _ = hs.trace_rays(rays)
    assert np.array_equal(hits.view(np.uint32).reshape(-1, 6), a["out_hits"])
    occ, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION)
    assert np.array_equal(occ, a["out_occluded"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", FRAMES)
def test_hip_reproduces_frame(ctx, name):
    import torch
    from rtg.device import DeviceScene, hits_to_numpy
    s, a = load(name)
    ds = DeviceScene(ctx, s)
    ctx.set_counters(True)
    try:
        rgba, rgb, ph = ds.trace_image(frame_index=int(a["frame"][0]))
        cnt = ctx.counters()
    finally:
        ctx.set_counters(False)
    torch.cuda.synchronize()
    # ray and traversal counts (stack spills included) equal the oracle's
    for f, v in zip((f for f, _ in abi.TraceCounters._fields_), a["out_counters"]):
        if not f.startswith("side_"):
            assert getattr(cnt, f) == v, f
    assert np.array_equal(hits_to_numpy(ph).view(np.uint32).reshape(-1, 6), a["out_hits"])
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), a["out_rgba8"])
    err = float(np.abs(rgb.cpu().numpy() - a["out_rgb"].reshape(-1)).max())
    assert err <= 1e-4, err
    ds.destroy()


@pytest.mark.gpu
def test_hip_reproduces_rays(ctx):
    import torch
    from rtg.device import DeviceScene, hits_to_numpy, rays_to_device
    s, a = load("rays_small.npz")
    ds = DeviceScene(ctx, s)
    rays = a["rays"].view(abi.RAY_DTYPE).reshape(-1)
    d = rays_to_device(rays)
    out = ds.trace_rays(d, len(rays))
    occ = ds.trace_rays(d, len(rays), mode=abi.RTG_QUERY_OCCLUSION)
    torch.cuda.synchronize()
    assert np.array_equal(hits_to_numpy(out).view(np.uint32).reshape(-1, 6), a["out_hits"])
    assert np.array_equal(occ.cpu().numpy().astype(np.uint32), a["out_occluded"])
    ds.destroy()
