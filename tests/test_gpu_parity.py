"""HIP path vs the CPU oracle: acceleration-structure bytes, hit records,
occlusion and shaded colour. Integer and index results are compared bit for
bit; shaded linear RGB within 1e-4 absolute."""
import numpy as np
import pytest
import torch

import oracle_host as oh
from rtg import abi, scenes
from rtg.device import DeviceScene, hits_to_numpy, rays_to_device

pytestmark = pytest.mark.gpu

RGB_TOL = 1e-4


def tlas_bytes_masked(blob, n):
    """TLAS blob bytes with the BLAS address fields zeroed (they are device
    addresses on the GPU and host pointers in the oracle)."""
    b = np.frombuffer(blob, dtype=np.uint8).copy()
    hdr = b[:128].view(np.uint64)
    prims_off, extra_off = int(hdr[9]), int(hdr[10])
    recs = b[prims_off:prims_off + 64 * n].reshape(n, 64)
    recs[:, 48:56] = 0
    src = b[extra_off:extra_off + 64 * n].reshape(n, 64)
    src[:, 56:64] = 0
    return b


@pytest.fixture(scope="module")
def small(ctx):
    s = scenes.small_scene(width=128, height=96)
    return s, DeviceScene(ctx, s), oh.HostScene(s)


def test_blas_bytes_match_oracle(ctx, small):
    s, ds, hs = small
    for k, m in enumerate(s.meshes):
        g = ds.as_bytes(ds.blas[k])
        o = np.frombuffer(hs.blas[k].bytes(), dtype=np.uint8)
        assert g.shape == o.shape, (m.name, g.shape, o.shape)
        assert np.array_equal(g, o), m.name


@pytest.mark.parametrize("mesh", ["torus", "icosphere", "soup", "one", "empty", "multi"])
def test_blas_shapes(ctx, mesh):
    if mesh == "torus":
        m = scenes.torus(48, 24)
    elif mesh == "icosphere":
        m = scenes.icosphere(3)
    elif mesh == "soup":
        m = scenes.soup(5000, 0.05, rng=scenes.Rng(5))
    elif mesh == "one":
        m = scenes.Mesh("one", [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 1]] * 3, [[0, 0]] * 3,
                        [(np.array([[0, 1, 2]]), True)])
    elif mesh == "empty":
        m = scenes.Mesh("empty", [[0, 0, 0]], [[0, 0, 1]], [[0, 0]], [(np.zeros((0, 3)), True)])
    else:
        t = scenes.torus(12, 6)
        ix = t.geometries[0][0]
        m = scenes.Mesh("multi", t.positions, t.normals, t.uvs, [(ix[:50], True), (ix[50:], False)])
    s = scenes.Scene("blas")
    s.meshes.append(m)
    s.models.append(scenes.Model(0, [(0, 0, -1)]))
    s.instances.append(scenes.InstanceData(scenes.affine((0, 0, 0)), 0, 0))
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    g = ds.as_bytes(ds.blas[0])
    o = np.frombuffer(hs.blas[0].bytes(), dtype=np.uint8)
    assert np.array_equal(g, o)
    ds.destroy()


def test_tlas_bytes_and_refit_match_oracle(ctx, small):
    s, ds, hs = small
    n = len(s.instances)
    g = tlas_bytes_masked(ds.as_bytes(ds.tlas), n)
    o = tlas_bytes_masked(hs.tlas.bytes(), n)
    assert np.array_equal(g, o)
    # refit a clone with new rotations (src == dst UPDATE semantics)
    rng = scenes.Rng(99)
    xf = [i.transform.copy() for i in s.instances]
    for k in range(1, n):
        xf[k] = scenes.affine(xf[k][:, 3], scenes.rotation_from_uniform(*rng.uniform(3)), 1.3)
    clone = ds.clone_tlas()
    buf = ds.instance_buffer(xf)
    ds.refit(buf, tlas=clone)
    torch.cuda.synchronize()
    hs2 = oh.HostScene(s)
    hs2.refit(xf)
    g2 = tlas_bytes_masked(ds.as_bytes(clone), n)
    o2 = tlas_bytes_masked(hs2.tlas.bytes(), n)
    assert np.array_equal(g2, o2)


def _compare_rays(ds, hs, rays, mode, **kw):
    d = rays_to_device(rays)
    out = ds.trace_rays(d, len(rays), mode=mode, **kw)
    torch.cuda.synchronize()
    ref, _ = hs.trace_rays(rays, mode=mode, **kw)
    if mode == abi.RTG_QUERY_CLOSEST:
        got = hits_to_numpy(out)
        assert np.array_equal(got.view(np.uint32).reshape(-1, 6), ref.view(np.uint32).reshape(-1, 6))
    else:
        got = out.cpu().numpy().astype(np.uint32)
        assert np.array_equal(got, ref)
    return ref


@pytest.mark.parametrize("mode", [abi.RTG_QUERY_CLOSEST, abi.RTG_QUERY_OCCLUSION])
def test_trace_rays_small(ctx, small, mode):
    s, ds, hs = small
    rays = scenes.incoherent_rays(50000, (-10, -1, -10), (10, 4, 10), seed=11)
    ref = _compare_rays(ds, hs, rays, mode)
    if mode == abi.RTG_QUERY_CLOSEST:
        assert (ref["instance"] != abi.RTG_NO_HIT).mean() > 0.2


def test_trace_rays_counters_match_oracle(ctx, small):
    """Traversal counts of both query kinds, stack pushes past the LDS part
    (16 closest-hit, 32 occlusion entries) included, equal the oracle's."""
    s, ds, hs = small
    rays = scenes.incoherent_rays(20000, (-10, -1, -10), (10, 4, 10), seed=12)
    d = rays_to_device(rays)
    for mode, names in ((abi.RTG_QUERY_CLOSEST, ("node_visits", "triangle_tests", "instance_visits", "spill_pushes")),
                        (abi.RTG_QUERY_OCCLUSION, ("occ_node_visits", "occ_triangle_tests", "occ_instance_visits",
                                                   "occ_spill_pushes"))):
        ctx.set_counters(True)
        try:
            ds.trace_rays(d, len(rays), mode)
            c = ctx.counters()
        finally:
            ctx.set_counters(False)
        _, oc = hs.trace_rays(rays, mode=mode)
        for f in names:
            assert getattr(c, f) == getattr(oc, f), f


@pytest.mark.parametrize("entries", [8, 12, 24, 32])
def test_stack_lds_entries_change_only_the_spill_counts(ctx, small, entries):
    """With every LDS stack depth the queries give the oracle's hits, and
    the pushes past the LDS part are the oracle's for that depth."""
    s, ds, hs = small
    rays = scenes.incoherent_rays(20000, (-10, -1, -10), (10, 4, 10), seed=13)
    d = rays_to_device(rays)
    saved = ctx.tuning()
    t = abi.Tuning.make(**{**saved.as_dict(), "stack_lds_entries": entries})
    ctx.set_tuning(t)
    oh.lib().orc_set_stack_lds(entries)
    try:
        for mode, names in ((abi.RTG_QUERY_CLOSEST, ("node_visits", "spill_pushes")),
                            (abi.RTG_QUERY_OCCLUSION, ("occ_node_visits", "occ_spill_pushes"))):
            ctx.set_counters(True)
            try:
                out = ds.trace_rays(d, len(rays), mode)
                c = ctx.counters()
            finally:
                ctx.set_counters(False)
            ref, oc = hs.trace_rays(rays, mode=mode)
            if mode == abi.RTG_QUERY_CLOSEST:
                assert np.array_equal(hits_to_numpy(out).view(np.uint32), ref.view(np.uint32))
            else:
                assert np.array_equal(out.cpu().numpy().astype(np.uint32), ref)
            for f in names:
                assert getattr(c, f) == getattr(oc, f), f
            if entries <= 12 and mode == abi.RTG_QUERY_CLOSEST:
                assert c.spill_pushes > 0
    finally:
        oh.lib().orc_set_stack_lds(16)
        ctx.set_tuning(saved)


def _compare_image(ds, hs, frame=None, opts=None, rgb_tol=RGB_TOL):
    outs = ds.trace_image(frame_index=frame, opts=opts)
    torch.cuda.synchronize()
    rgba, rgb, ph = outs
    o_rgba, o_rgb, o_hits, _ = hs.trace_image(frame_index=frame, opts=opts)
    got_hits = hits_to_numpy(ph)
    assert np.array_equal(got_hits["instance"], o_hits["instance"])
    assert np.array_equal(got_hits["primitive"], o_hits["primitive"])
    assert np.array_equal(got_hits["geometry"], o_hits["geometry"])
    assert np.array_equal(got_hits.view(np.uint32), o_hits.view(np.uint32))
    g_rgb = rgb.cpu().numpy().reshape(-1, 3)
    err = np.abs(g_rgb - o_rgb).max()
    assert err <= rgb_tol, err
    g_rgba = rgba.cpu().numpy().view(np.uint32)
    assert np.array_equal(g_rgba, o_rgba)
    return o_rgb


@pytest.mark.parametrize("nsh,refill_occlusion", [(3, 64), (5, 0), (1, 64), (2, 0)])
def test_shadow_ray_pairs_of_any_count(ctx, nsh, refill_occlusion):
    """The occlusion kernels trace a path's shadow rays in pairs (an odd
    count's last alone, after a padding slot in the queue), persistent
    (refill 64) or one pair per lane and round (refill 0): the image and
    every traversal counter equal the oracle's pair walk."""
    s = scenes.large_scene(width=96, height=64, n_objects=120, n_fences=12, n_mirrors=6)
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    o = s.trace_opts()
    o.shadow_rays = nsh
    o.spp = 2
    saved = ctx.tuning()
    ctx.set_tuning(abi.Tuning.make(**{**saved.as_dict(), "refill_occlusion": refill_occlusion}))
    try:
        ctx.set_counters(True)
        rgba, rgb, ph = ds.trace_image(frame_index=4, opts=o)
        gc = ctx.counters().as_dict()
        ctx.set_counters(False)
        torch.cuda.synchronize()
        o_rgba, o_rgb, o_hits, oc = hs.trace_image(frame_index=4, opts=o)
    finally:
        ctx.set_counters(False)
        ctx.set_tuning(saved)
    assert np.array_equal(hits_to_numpy(ph).view(np.uint32), o_hits.view(np.uint32))
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), o_rgba)
    assert float(np.abs(rgb.cpu().numpy().reshape(-1, 3) - o_rgb).max()) <= RGB_TOL
    assert oc.shadow_rays > 0 and oc.occ_node_visits > 0
    for f in ("shadow_rays", "occ_node_visits", "occ_triangle_tests", "occ_instance_visits", "occ_anyhit_calls",
              "occ_spill_pushes", "node_visits", "anyhit_calls"):
        assert gc[f] == getattr(oc, f), f
    ds.destroy()


@pytest.mark.parametrize("entries", [8, 24])
def test_frame_counters_with_other_stack_depths(ctx, entries):
    """A frame of the reduced large scene (bounce levels, soft shadows,
    alpha fences) traced with 8 or 24 LDS stack entries: the image and every
    traversal counter, spill pushes included, equal the oracle's at that
    depth."""
    s = scenes.large_scene(width=160, height=90)
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    saved = ctx.tuning()
    ctx.set_tuning(abi.Tuning.make(**{**saved.as_dict(), "stack_lds_entries": entries}))
    oh.lib().orc_set_stack_lds(entries)
    try:
        ctx.set_counters(True)
        rgba, rgb, ph = ds.trace_image(frame_index=2)
        gc = ctx.counters().as_dict()
        ctx.set_counters(False)
        torch.cuda.synchronize()
        o_rgba, o_rgb, o_hits, oc = hs.trace_image(frame_index=2)
    finally:
        ctx.set_counters(False)
        oh.lib().orc_set_stack_lds(16)
        ctx.set_tuning(saved)
    assert np.array_equal(hits_to_numpy(ph).view(np.uint32), o_hits.view(np.uint32))
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), o_rgba)
    assert float(np.abs(rgb.cpu().numpy().reshape(-1, 3) - o_rgb).max()) <= RGB_TOL
    # the bounce stream keeps its own subset of the counters; the totals are
    # what the oracle counts
    for f in ("primary_rays", "bounce_rays", "shadow_rays", "node_visits", "triangle_tests", "instance_visits",
              "anyhit_calls", "occ_node_visits", "occ_triangle_tests", "occ_instance_visits", "occ_anyhit_calls",
              "spill_pushes", "occ_spill_pushes"):
        assert gc[f] == getattr(oc, f), f
    ds.destroy()


def test_trace_image_small(ctx, small):
    s, ds, hs = small
    _compare_image(ds, hs)


def test_trace_image_large_reduced(ctx):
    s = scenes.large_scene(width=192, height=108)
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    _compare_image(ds, hs)
    ds.destroy()


def test_trace_image_multisample(ctx, small):
    s, ds, hs = small
    o = s.trace_opts()
    o.spp = 3
    o.shadow_rays = 2
    _compare_image(ds, hs, frame=5, opts=o)


def test_incoherent_full_scene(ctx):
    """The incoherent workload's full 1,000,000-triangle BLAS: bytes equal to
    the oracle's, and 262,144 of its rays (origins in the cube, directions on
    the sphere, tmax = inf) give identical hits, occlusion and traversal
    counts in both query kinds."""
    s = scenes.incoherent_scene()
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    assert s.instanced_triangles == 1_000_000
    g = ds.as_bytes(ds.blas[0])
    o = np.frombuffer(hs.blas[0].bytes(), dtype=np.uint8)
    assert np.array_equal(g, o)
    rays = scenes.incoherent_rays(1 << 18, seed=0x5EED0102)
    assert np.isinf(rays["tmax"]).all()
    d = rays_to_device(rays)
    for mode in (abi.RTG_QUERY_CLOSEST, abi.RTG_QUERY_OCCLUSION):
        ctx.set_counters(True)
        try:
            out = ds.trace_rays(d, len(rays), mode)
            c = ctx.counters()
        finally:
            ctx.set_counters(False)
        ref, oc = hs.trace_rays(rays, mode=mode)
        if mode == abi.RTG_QUERY_CLOSEST:
            got = hits_to_numpy(out)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
            assert (ref["instance"] != abi.RTG_NO_HIT).mean() > 0.5
            names = ("node_visits", "triangle_tests", "instance_visits", "spill_pushes")
        else:
            assert np.array_equal(out.cpu().numpy().astype(np.uint32), ref)
            names = ("occ_node_visits", "occ_triangle_tests", "occ_instance_visits", "occ_spill_pushes")
        for f in names:
            assert getattr(c, f) == getattr(oc, f), f
    ds.destroy()


def test_trace_rays_refill_threshold_never_changes_results(ctx, monkeypatch):
    """Ray queries refill a wavefront once RTG_REFILL_RAYS of its lanes are
    idle (read when a context is created): whole-wave, any-idle and the
    default threshold give the same hits and occlusion as the oracle."""
    from rtg.lib import Context
    s = scenes.incoherent_scene(n_tris=50_000)
    hs = oh.HostScene(s)
    rays = scenes.incoherent_rays(1 << 16, seed=0x5EED0103)
    ref_hits, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_CLOSEST)
    ref_occ, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION)
    for refill in ("64", "0", None):
        if refill is None:
            monkeypatch.delenv("RTG_REFILL_RAYS", raising=False)
        else:
            monkeypatch.setenv("RTG_REFILL_RAYS", refill)
        c = Context(0)
        try:
            ds = DeviceScene(c, s)
            d = rays_to_device(rays)
            got = hits_to_numpy(ds.trace_rays(d, len(rays), abi.RTG_QUERY_CLOSEST))
            assert np.array_equal(got.view(np.uint32), ref_hits.view(np.uint32)), refill
            occ = ds.trace_rays(d, len(rays), abi.RTG_QUERY_OCCLUSION).cpu().numpy().astype(np.uint32)
            assert np.array_equal(occ, ref_occ), refill
            ds.destroy()
        finally:
            c.close()


def test_bgra_output_and_png(ctx, small, tmp_path):
    """The output step: BGRA8 byte order as the reference's swapchain format,
    equal to the oracle's, and a PNG of the frame that decodes to it."""
    from rtg import image
    s, ds, hs = small
    o = s.trace_opts()
    o.flags = abi.RTG_TRACE_OUTPUT_BGRA8
    outs = ds.trace_image(opts=o)
    torch.cuda.synchronize()
    o_rgba, _, _, _ = hs.trace_image(opts=o)
    assert np.array_equal(outs[0].cpu().numpy().view(np.uint32), o_rgba)
    img = image.readback(outs[0], s.width, s.height, o.flags)
    p = tmp_path / "frame.png"
    image.write_png(str(p), img)
    assert np.array_equal(image.read_png(str(p)), image.unpack(o_rgba, s.width, s.height, bgra=True))


def test_unknown_trace_flags_are_refused(ctx, small):
    s, ds, hs = small
    o = s.trace_opts()
    o.flags = 0x100
    with pytest.raises(Exception):
        ds.trace_image(opts=o)


def test_trace_image_denser_scene_frames(ctx):
    """A denser reduced animated scene (200 objects, 16 alpha fences, 8
    mirrors, 4 shadow rays): frame 2 and a two-sample frame 5 equal the
    oracle's."""
    s = scenes.large_scene(width=160, height=90, n_objects=200, n_fences=16, n_mirrors=8)
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    s.set_frame(2)
    ds.refit(ds.instance_buffer(s.transforms_for_frame(2)))
    _compare_image(ds, hs_refit(hs, s), frame=2)
    o = s.trace_opts()
    o.spp = 2
    _compare_image(ds, hs, frame=5, opts=o)
    ds.destroy()


def test_bounce_levels_beside_shadow_rays(ctx):
    """The mirror bounce levels traced on a second stream beside the first
    level's shadow rays (bounce_blocks > 0, down to one workgroup and up to
    half the grid) give the same image, hits and traversal counts as every
    level in sequence (bounce_blocks = 0), and equal the oracle."""
    s = scenes.large_scene(width=160, height=90, n_objects=200, n_fences=16, n_mirrors=8)
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    saved = ctx.tuning()
    results = []
    try:
        for bb in (0, 1, 128, 4096):
            ctx.set_tuning(abi.Tuning.make(bounce_blocks=bb))
            ctx.set_counters(True)
            rgba, rgb, ph = ds.trace_image()
            torch.cuda.synchronize()
            c = ctx.counters().as_dict()
            ctx.set_counters(False)
            results.append((rgba.cpu().numpy().view(np.uint32), rgb.cpu().numpy(),
                            hits_to_numpy(ph).view(np.uint32), c))
            if bb == 128:
                _compare_image(ds, hs)
    finally:
        ctx.set_counters(False)
        ctx.set_tuning(saved)
    assert results[0][3]["bounce_rays"] > 0
    totals = lambda c: {k: v for k, v in c.items() if not k.startswith("side_")}
    assert not any(v for k, v in results[0][3].items() if k.startswith("side_"))
    assert results[2][3]["side_occ_node_visits"] > 0 and results[2][3]["side_node_visits"] > 0
    assert results[2][3]["side_shadow_rays"] > 0
    for r in results[1:]:
        assert np.array_equal(r[0], results[0][0])
        assert np.array_equal(r[1], results[0][1])
        assert np.array_equal(r[2], results[0][2])
        assert totals(r[3]) == totals(results[0][3])
    ds.destroy()


def hs_refit(hs, s):
    hs.refit([i.transform for i in s.instances])
    return hs
