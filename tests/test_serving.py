"""Serving (continuous batching over resident scenes): every result is the
same alone and in the mix, batches really form, the HBM budget rejects what
can never fit."""
import numpy as np
import pytest
import torch

from rtg import abi, scenes
from rtg.device import DeviceScene
from rtg.serve import Server, make_request

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident(ctx):
    out = []
    for k in range(3):
        s = scenes.small_scene(seed=0x5EED0000 + 100 + k, width=64, height=64, n_tori=6)
        out.append(DeviceScene(ctx, s))
    return out


def request_set(resident, n=10, seed=3):
    rng = scenes.Rng(seed)
    sizes = [(48, 32), (64, 64), (96, 40), (33, 17)]
    reqs, outs, bufs = [], [], []
    for i in range(n):
        sid = int(rng.integers(1, 0, len(resident))[0])
        ds = resident[sid]
        w, h = sizes[int(rng.integers(1, 0, len(sizes))[0])]
        spp = int(rng.integers(1, 1, 4)[0])
        # a random transform update of every torus
        xf = [ins.transform.copy() for ins in ds.scene.instances]
        for k in range(1, len(xf)):
            xf[k] = scenes.affine(xf[k][:, 3], scenes.rotation_from_uniform(*rng.uniform(3)), 1.0)
        buf = ds.instance_buffer(xf)
        o = (torch.zeros(w * h, dtype=torch.int32, device="cuda:0"),
             torch.zeros(w * h * 3, dtype=torch.float32, device="cuda:0"),
             torch.zeros(w * h * spp * 6, dtype=torch.int32, device="cuda:0"))
        r = make_request(sid, ds, w, h, spp, 2, 3, frame=i, instances=buf, outputs=o, tag=i)
        reqs.append(r)
        outs.append(o)
        bufs.append((xf, buf))
    return reqs, outs, bufs


def snapshot(outs):
    torch.cuda.synchronize()
    return [tuple(t.cpu().numpy().copy() for t in o) for o in outs]


def test_results_alone_equal_results_in_the_mix(ctx, resident):
    reqs, outs, bufs = request_set(resident)
    # alone: one request per launch, one launch at a time, arriving far apart
    srv = Server(ctx, budget_bytes=1 << 30, max_batch=1, max_inflight=1)
    ids = [srv.add_scene(ds) for ds in resident]
    assert ids == [0, 1, 2]
    comp = srv.replay(reqs, np.arange(len(reqs)) * 0.02)
    assert all(c.status == 0 and c.batch_size == 1 for c in comp)
    srv.close()
    alone = snapshot(outs)
    for o in outs:
        for t in o:
            t.zero_()
    # in the mix: all arrive at once, batched and overlapping
    srv = Server(ctx, budget_bytes=1 << 30, max_batch=8, max_inflight=2)
    for ds in resident:
        srv.add_scene(ds)
    comp = srv.replay(reqs, np.zeros(len(reqs)))
    stats = srv.stats()
    srv.close()
    assert all(c.status == 0 for c in comp)
    assert stats["max_batch"] > 1 and stats["batches"] < len(reqs)
    mixed = snapshot(outs)
    for a, b in zip(alone, mixed):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    # and equal to the plain path: refit a clone, trace one image
    for r, (xf, buf), a in zip(reqs, bufs, alone):
        ds = resident[r.scene]
        clone = ds.clone_tlas()
        ds.refit(buf, tlas=clone)
        o = (torch.zeros(r.width * r.height, dtype=torch.int32, device="cuda:0"),
             torch.zeros(r.width * r.height * 3, dtype=torch.float32, device="cuda:0"),
             torch.zeros(r.width * r.height * r.opts.spp * 6, dtype=torch.int32, device="cuda:0"))
        j = ds.job(o, tlas=clone, width=r.width, height=r.height, uniforms=r.uniforms, opts=r.opts)
        from rtg.lib import check, load
        import ctypes as C
        check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
        b = snapshot([o])[0]
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        load().rtg_tlas_destroy(ctx.ptr, clone)


def spill_area_bytes(ctx):
    """Device bytes of one slot's traversal spill area: one 704-byte stack
    spill per lane of the persistent grid (blocks_per_cu workgroups of 256
    lanes per CU)."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return cus * ctx.tuning().blocks_per_cu * 256 * 704


def test_budget_must_hold_every_slots_spill_area(ctx):
    from rtg.lib import RtgError
    spill = spill_area_bytes(ctx)
    with pytest.raises(RtgError, match=r"\(-6\)"):
        Server(ctx, budget_bytes=max(64 << 20, 2 * spill - (1 << 20)), max_batch=4, max_inflight=2)
    Server(ctx, budget_bytes=2 * spill + (64 << 20), max_batch=4, max_inflight=2).close()


def test_device_memory_stays_within_the_budget(ctx, resident):
    spill = spill_area_bytes(ctx)
    budget = 3 * spill + (48 << 20)
    reqs, outs, bufs = request_set(resident, n=24, seed=9)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    srv = Server(ctx, budget_bytes=budget, max_batch=8, max_inflight=3)
    for ds in resident:
        srv.add_scene(ds)
    comp = srv.replay(reqs, np.zeros(len(reqs)))
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    stats = srv.stats()
    srv.close()
    assert all(c.status == 0 for c in comp)
    assert stats["max_inflight_seen"] > 1
    # what the slots held, by the server's count and by the device's
    assert 3 * spill <= stats["peak_bytes"] <= budget
    assert free0 - free1 <= budget + (8 << 20)


def test_budget_rejects_requests_that_never_fit(ctx, resident):
    srv = Server(ctx, budget_bytes=2 * spill_area_bytes(ctx) + (64 << 20), max_batch=4, max_inflight=2)
    srv.add_scene(resident[0])
    big = make_request(0, resident[0], 1920, 1080, 8, 4, 3, frame=0)
    st, _ = srv.submit(big)
    assert st == abi.RTG_ERROR_BUDGET
    assert srv.stats()["rejected"] == 1
    small = make_request(0, resident[0], 32, 32, 1, 1, 3, frame=0)
    st, _ = srv.submit(small)
    assert st == 0
    done = srv.poll(wait_all=True)
    assert len(done) == 1 and done[0].status == 0
    srv.close()


def test_sample_cap_bounds_batches(ctx, resident):
    reqs, outs, bufs = request_set(resident, n=12, seed=5)
    cap = 6000
    srv = Server(ctx, budget_bytes=1 << 30, max_batch=8, max_inflight=2, max_samples=cap)
    for ds in resident:
        srv.add_scene(ds)
    comp = srv.replay(reqs, np.zeros(len(reqs)))
    srv.close()
    assert all(c.status == 0 for c in comp)
    batches = {}
    for c, r in zip(comp, reqs):
        batches.setdefault(c.start_s, []).append(r.width * r.height * r.opts.spp)
    assert len(batches) < len(reqs)  # some batching still happens
    for samples in batches.values():
        assert len(samples) == 1 or sum(samples) <= cap
    capped = snapshot(outs)
    for o in outs:
        for t in o:
            t.zero_()
    srv = Server(ctx, budget_bytes=1 << 30, max_batch=1, max_inflight=1)
    for ds in resident:
        srv.add_scene(ds)
    srv.replay(reqs, np.arange(len(reqs)) * 0.01)
    srv.close()
    for a, b in zip(snapshot(outs), capped):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_cancel_and_deadline(ctx, resident):
    """A pending request can be cancelled, and one whose deadline passes
    before its batch launches times out; neither runs (outputs untouched)
    and both complete at once; the others complete as usual."""
    srv = Server(ctx, budget_bytes=8 << 30, max_batch=1, max_inflight=1)
    srv.add_scene(resident[0])

    def outputs(w, h, spp):
        return (torch.full((w * h,), -1, dtype=torch.int32, device="cuda:0"),
                torch.full((w * h * 3,), -1.0, dtype=torch.float32, device="cuda:0"),
                torch.full((w * h * spp * 6,), -1, dtype=torch.int32, device="cuda:0"))

    # two large requests keep the single slot busy while the others queue
    big = [make_request(0, resident[0], 1024, 1024, 4, 4, 3, frame=k, outputs=outputs(1024, 1024, 4), tag=k)
           for k in range(2)]
    ob, oc, od = outputs(40, 30, 1), outputs(40, 30, 1), outputs(40, 30, 1)
    b = make_request(0, resident[0], 40, 30, 1, 2, 3, frame=5, outputs=ob, tag=10)
    # a deadline already past (the server clock runs from its creation)
    c = make_request(0, resident[0], 40, 30, 1, 2, 3, frame=6, outputs=oc, tag=11, deadline_s=1e-6)
    d = make_request(0, resident[0], 40, 30, 1, 2, 3, frame=7, outputs=od, tag=12,
                     deadline_s=srv.clock() + 600.0)
    tickets = []
    for r in big + [b, c, d]:
        st, t = srv.submit(r)
        assert st == 0
        tickets.append(t)
    assert srv.cancel(tickets[2]) == 0
    assert srv.cancel(tickets[2]) == abi.RTG_ERROR_NOT_FOUND
    assert srv.cancel(10 ** 9) == abi.RTG_ERROR_NOT_FOUND
    done = {x.tag: x for x in srv.poll(wait_all=True)}
    srv.close()
    assert sorted(done) == [0, 1, 10, 11, 12]
    assert done[0].status == 0 and done[1].status == 0 and done[12].status == 0
    assert done[10].status == abi.RTG_ERROR_CANCELLED and done[10].batch_size == 0
    assert done[11].status == abi.RTG_ERROR_TIMEOUT and done[11].batch_size == 0
    assert done[10].start_s == done[10].done_s and done[11].start_s == done[11].done_s
    torch.cuda.synchronize()
    for o in (ob, oc):
        assert bool((o[0] == -1).all()) and bool((o[1] == -1.0).all()) and bool((o[2] == -1).all())
    assert not bool((od[0] == -1).all())
