"""The frame driver: frames in flight (each with its own TLAS copy, mapped
instance array and frame state) give, frame for frame, what the plain path
gives (refit of a TLAS copy with that frame's transforms, then one trace),
which the parity tests pin to the oracle; and one frame equals the oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_host as oh
from rtg import abi, scenes
from rtg.device import DeviceScene, hits_to_numpy
from rtg.frames import FrameDriver
from rtg.lib import check, load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def animated(ctx):
    s = scenes.large_scene(width=96, height=64, n_objects=40, n_fences=6, n_mirrors=4, n_soup_meshes=2,
                           soup_tris=200)
    return s, DeviceScene(ctx, s)


def _plain(ctx, ds, frame):
    s = ds.scene
    clone = ds.clone_tlas()
    ds.refit(ds.instance_buffer(s.transforms_for_frame(frame)), tlas=clone)
    o = ds.make_outputs()
    j = ds.job(o, frame_index=frame, tlas=clone)
    check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
    torch.cuda.synchronize()
    out = tuple(t.cpu().numpy().copy() for t in o)
    load().rtg_tlas_destroy(ctx.ptr, clone)
    return out


@pytest.mark.parametrize("in_flight", [1, 2, 3])
def test_frames_in_flight_equal_plain_frames(ctx, animated, in_flight):
    s, ds = animated
    drv = FrameDriver(ds, in_flight)
    streams = [torch.cuda.Stream() for _ in range(in_flight)]
    outs = []
    for f in range(6):
        inst, slot_tlas = drv.begin(f)
        assert slot_tlas != ds.tlas
        inst["transform"] = np.stack(s.transforms_for_frame(f)).astype(np.float32)
        o = ds.make_outputs()
        torch.cuda.synchronize()  # outputs allocated on the current stream
        outs.append(o)
        drv.submit(f, ds.job(o, frame_index=f), stream=streams[f % in_flight])
    torch.cuda.synchronize()
    got = [tuple(t.cpu().numpy().copy() for t in o) for o in outs]
    drv.close()
    for f in range(6):
        want = _plain(ctx, ds, f)
        for a, b in zip(got[f], want):
            assert np.array_equal(a, b), f


def test_frame_matches_oracle(ctx, animated):
    s, ds = animated
    drv = FrameDriver(ds, 2)
    frame = 4
    inst, _ = drv.begin(frame)
    inst["transform"] = np.stack(s.transforms_for_frame(frame)).astype(np.float32)
    o = ds.make_outputs()
    drv.submit(frame, ds.job(o, frame_index=frame))
    torch.cuda.synchronize()
    drv.close()
    s.set_frame(frame)
    rgba, rgb, hits, _ = oh.HostScene(s).trace_image(frame_index=frame)
    s.set_frame(0)
    assert np.array_equal(hits_to_numpy(o[2]).view(np.uint32), hits.view(np.uint32))
    assert np.array_equal(o[0].cpu().numpy().view(np.uint32), rgba)
    assert np.abs(o[1].cpu().numpy().reshape(-1, 3) - rgb).max() <= 1e-4


def test_plain_calls_on_two_streams_run_as_frames_in_flight(ctx, animated):
    """rtg_trace_image on two streams at once (a TLAS copy each): the context
    keeps one frame state per stream, so each frame equals the same frame
    traced alone."""
    s, ds = animated
    streams = [torch.cuda.Stream() for _ in range(2)]
    clones = [ds.clone_tlas() for _ in range(2)]
    bufs = [ds.instance_buffer(s.transforms_for_frame(f)) for f in range(6)]
    outs = [ds.make_outputs() for _ in range(6)]
    torch.cuda.synchronize()
    for f in range(6):
        k = f % 2
        ds.refit(bufs[f], tlas=clones[k], stream=streams[k])
        ds.trace_image(frame_index=f, outputs=outs[f], stream=streams[k], tlas=clones[k])
    torch.cuda.synchronize()
    got = [tuple(t.cpu().numpy().copy() for t in o) for o in outs]
    for c in clones:
        load().rtg_tlas_destroy(ctx.ptr, c)
    for f in range(6):
        want = _plain(ctx, ds, f)
        for a, b in zip(got[f], want):
            assert np.array_equal(a, b), f
