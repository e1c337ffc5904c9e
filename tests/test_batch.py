"""Graph-captured batches and launch tuning: a batch of uneven scenes traced
by one graph launch gives, bit for bit, what each job gives alone (which the
parity tests pin to the oracle); every replay gives the same; updates of the
job table take effect; tuning settings never change results."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_host as oh
from rtg import abi, scenes
from rtg.batch import Batch, autotune, trace_batch
from rtg.device import DeviceScene, hits_to_numpy
from rtg.lib import check, load

pytestmark = pytest.mark.gpu

RESOLUTIONS = ((16, 16), (40, 24), (64, 48), (33, 17))


@pytest.fixture(scope="module")
def mini(ctx):
    ws = scenes.batch_workload(n_scenes=8, seed=0x5EED0103, tri_range=(1_000, 30_000), resolutions=RESOLUTIONS)
    return ws, [DeviceScene(ctx, s) for s in ws]


def _jobs(dss, frame=None):
    outs, jobs = [], []
    for ds in dss:
        o = ds.make_outputs()
        outs.append(o)
        jobs.append(ds.job(o, frame_index=frame))
    return outs, jobs


def _host(outs):
    torch.cuda.synchronize()
    return [tuple(t.cpu().numpy().copy() for t in o) for o in outs]


def _alone(ctx, dss, frame=None):
    outs, jobs = _jobs(dss, frame)
    for j in jobs:
        check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
    return _host(outs)


def _equal(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for p, q in zip(x, y):
            assert np.array_equal(p, q)


def test_graph_batch_equals_jobs_alone(ctx, mini):
    ws, dss = mini
    alone = _alone(ctx, dss)
    outs, jobs = _jobs(dss)
    b = Batch(ctx, jobs)
    for replay in range(3):
        for o in outs:
            for t in o:
                t.zero_()
        torch.cuda.synchronize()
        b.launch()
        _equal(alone, _host(outs))
    b.close()


def test_graph_batch_job_matches_oracle(ctx, mini):
    ws, dss = mini
    k = int(np.argmin([s.instanced_triangles for s in ws]))
    outs, jobs = _jobs(dss)
    b = Batch(ctx, jobs)
    b.launch()
    torch.cuda.synchronize()
    rgba, rgb, ph = outs[k]
    o_rgba, o_rgb, o_hits, _ = oh.HostScene(ws[k]).trace_image()
    assert np.array_equal(hits_to_numpy(ph).view(np.uint32), o_hits.view(np.uint32))
    assert np.abs(rgb.cpu().numpy().reshape(-1, 3) - o_rgb).max() <= 1e-4
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), o_rgba)
    b.close()


def test_graph_batch_update(ctx, mini):
    ws, dss = mini
    outs, jobs = _jobs(dss, frame=0)
    b = Batch(ctx, jobs)
    b.launch()
    _equal(_alone(ctx, dss, frame=0), _host(outs))
    _, jobs7 = _jobs(dss, frame=7)
    # same outputs, new uniforms
    for j, o in zip(jobs7, outs):
        j.out_rgba8, j.out_rgb, j.out_primary_hits = (o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr())
    b.update(jobs7)
    b.launch()
    _equal(_alone(ctx, dss, frame=7), _host(outs))
    # a changed size is refused
    bad = list(jobs7)
    bad[0] = dss[0].job(outs[0], width=dss[0].scene.width // 2)
    with pytest.raises(Exception):
        b.update(bad)
    b.close()


def test_graph_batch_counters_equal_eager(ctx, mini):
    ws, dss = mini
    outs, jobs = _jobs(dss)
    ctx.set_counters(True)
    try:
        trace_batch(ctx, jobs)
        eager = ctx.counters().as_dict()
        b = Batch(ctx, jobs)
        b.launch()
        graph = ctx.counters().as_dict()
        b.close()
    finally:
        ctx.set_counters(False)
    assert eager == graph
    assert eager["primary_rays"] == sum(s.width * s.height * s.spp for s in ws)


SETTINGS = [
    dict(),
    dict(blocks_per_cu=8, feed_chunk=256),
    dict(refill_closest=0, refill_occlusion=0),
    dict(refill_closest=32, refill_occlusion=48, path_order=0),
    dict(blocks_per_cu=1, feed_chunk=1024, refill_closest=1, refill_occlusion=1),
    dict(bounce_blocks=0),
    dict(bounce_blocks=1),
    dict(blocks_per_cu=2, bounce_blocks=4096),
    dict(stack_lds_entries=8),
    dict(stack_lds_entries=32, blocks_per_cu=5),
    dict(stack_lds_entries=12, refill_closest=0, refill_occlusion=0),
]


def test_tuning_never_changes_results(ctx, mini):
    ws, dss = mini
    saved = ctx.tuning()
    try:
        ref = None
        for kw in SETTINGS:
            ctx.set_tuning(abi.Tuning.make(**kw))
            outs, jobs = _jobs(dss)
            trace_batch(ctx, jobs)
            got = _host(outs)
            if ref is None:
                ref = got
            else:
                _equal(ref, got)
    finally:
        ctx.set_tuning(saved)


def test_bad_tuning_is_refused(ctx):
    saved = ctx.tuning()
    for kw in [dict(blocks_per_cu=0), dict(blocks_per_cu=9), dict(feed_chunk=96), dict(refill_closest=65),
               dict(path_order=2), dict(bounce_blocks=4097), dict(stack_lds_entries=4),
               dict(stack_lds_entries=36), dict(stack_lds_entries=10)]:
        with pytest.raises(Exception):
            ctx.set_tuning(abi.Tuning.make(**kw))
    assert ctx.tuning().as_dict() == saved.as_dict()
    # zero asks for the default depth
    try:
        ctx.set_tuning(abi.Tuning.make(**{**saved.as_dict(), "stack_lds_entries": 0}))
        assert ctx.tuning().stack_lds_entries == 16
    finally:
        ctx.set_tuning(saved)


def test_autotune(ctx, mini):
    ws, dss = mini
    saved = ctx.tuning()
    before = _alone(ctx, dss)
    try:
        outs, jobs = _jobs(dss)
        best, ms = autotune(ctx, jobs, repeats=1)
        assert ms > 0
        assert ctx.tuning().as_dict() == best.as_dict()
        assert best.blocks_per_cu in (4, 5, 6, 8) and best.path_order in (0, 1)
        assert best.stack_lds_entries in (8, 12, 16, 24, 32)
        _equal(before, _host(outs))
        _equal(before, _alone(ctx, dss))
    finally:
        ctx.set_tuning(saved)


def test_jobs_with_different_shadow_ray_counts(ctx, mini):
    """One batch whose jobs cast 1, 3, 2 and 5 shadow rays per textured hit
    (odd counts leave a padding slot after a path's rays, and the occlusion
    kernel looks each pair's job up): every job equals itself traced alone
    and the oracle, and the batch's shadow-ray count is the jobs' sum."""
    ws, dss = mini
    counts = [1, 3, 2, 5]
    sel = list(range(len(counts)))
    opts = []
    for k, n in zip(sel, counts):
        o = ws[k].trace_opts()
        o.shadow_rays = n
        opts.append(o)

    def jobs_of():
        outs, jobs = [], []
        for k, o in zip(sel, opts):
            out = dss[k].make_outputs(spp=o.spp)
            outs.append(out)
            jobs.append(dss[k].job(out, frame_index=3, opts=o))
        return outs, jobs

    outs, jobs = jobs_of()
    for j in jobs:
        check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
    alone = _host(outs)
    outs, jobs = jobs_of()
    ctx.set_counters(True)
    try:
        trace_batch(ctx, jobs)
        c = ctx.counters()
    finally:
        ctx.set_counters(False)
    batched = _host(outs)
    _equal(alone, batched)
    shadow = 0
    for k, o, got in zip(sel, opts, batched):
        hs = oh.HostScene(ws[k])
        rgba, rgb, hits, oc = hs.trace_image(frame_index=3, opts=o)
        assert np.array_equal(got[0].view(np.uint32), rgba)
        assert np.array_equal(got[2].view(np.uint32).reshape(-1), hits.view(np.uint32).reshape(-1))
        shadow += oc.shadow_rays
    assert c.shadow_rays == shadow > 0
