"""The traversal and intersection rules pinned on the oracle in
test_oracle_trace.py, replayed through the HIP path with rtg_trace_rays:
watertightness through vertices and edges, inclusive t range, the closest-hit
tie-break, any-hit (alpha clip) under every combination of ray flags,
instance flags and hit-group offsets, cull masks, axis-aligned and signed-zero
directions, empty ranges, an empty TLAS and rays with non-finite or degenerate
components. Hit records and occlusion must be
bit-identical to the oracle's, and the expected outcome of each rule is checked
as well, so that a rule both sides got wrong still fails."""
import math

import numpy as np
import pytest
import torch

import oracle_host as oh
from rtg import abi, scenes
from rtg.device import DeviceScene, hits_to_numpy, rays_to_device
from test_oracle_trace import fence_rays, fence_scene, one_mesh_scene, rays_from

pytestmark = pytest.mark.gpu

MODES = (abi.RTG_QUERY_CLOSEST, abi.RTG_QUERY_OCCLUSION)


class Pair:
    """One scene on the GPU and on the oracle."""

    def __init__(self, ctx, s):
        self.ds = DeviceScene(ctx, s)
        self.hs = oh.HostScene(s)

    def trace(self, rays, mode=abi.RTG_QUERY_CLOSEST, **kw):
        d = rays_to_device(rays)
        g = self.ds.trace_rays(d, len(rays), mode=mode, **kw)
        torch.cuda.synchronize()
        o, _ = self.hs.trace_rays(rays, mode=mode, **kw)
        if mode == abi.RTG_QUERY_CLOSEST:
            got = hits_to_numpy(g)
            assert np.array_equal(got.view(np.uint32).reshape(-1, 6), o.view(np.uint32).reshape(-1, 6)), kw
        else:
            got = g.cpu().numpy().astype(np.uint32)
            assert np.array_equal(got, o), kw
        return o

    def close(self):
        self.ds.destroy()


def test_watertight_through_vertices_and_edges(ctx):
    mesh = scenes.torus(24, 12, 1.0, 0.35)
    p = Pair(ctx, one_mesh_scene(mesh))
    u = 0.37
    o = np.array([math.cos(u), 0.0, math.sin(u)], np.float32)
    ix = mesh.geometries[0][0]
    pos = mesh.positions
    tg = np.concatenate([pos, (pos[ix[:, 0]] + pos[ix[:, 1]]) * 0.5, (pos[ix[:, 1]] + pos[ix[:, 2]]) * 0.5,
                         (pos[ix[:, 2]] + pos[ix[:, 0]]) * 0.5]).astype(np.float32)
    d = tg - o
    rays = rays_from(o, d[np.linalg.norm(d, axis=1) < 0.36])
    hits = p.trace(rays)
    assert np.all(hits["instance"] != abi.RTG_NO_HIT)
    assert np.all(p.trace(rays, abi.RTG_QUERY_OCCLUSION) == 1)
    p.close()


def test_t_range_is_inclusive(ctx):
    quad = scenes.quad(2.0, 2.0)
    p = Pair(ctx, one_mesh_scene(quad, [scenes.affine((0, 5, 0))]))
    f = np.float32(5.0)
    below = np.nextafter(f, np.float32(0))
    above = np.nextafter(f, np.float32(10))
    for tmin, tmax, expect in [(0.0, f, 1), (0.0, below, 0), (f, 10.0, 1), (above, 10.0, 0), (f, f, 1),
                               (6.0, 4.0, 0)]:
        r = rays_from((0.1, 0.0, 0.2), (0.0, 1.0, 0.0), tmin, tmax)
        h = p.trace(r)
        assert (h["instance"][0] != abi.RTG_NO_HIT) == bool(expect), (tmin, tmax)
        if expect:
            assert h["t"][0] == f
        assert p.trace(r, abi.RTG_QUERY_OCCLUSION)[0] == expect, (tmin, tmax)
    p.close()


def test_axis_aligned_and_signed_zero_directions(ctx):
    quad = scenes.quad(2.0, 2.0)
    p = Pair(ctx, one_mesh_scene(quad, [scenes.affine((0, 1, 0))]))
    z = np.float32(0.0)
    nz = np.float32(-0.0)
    dirs = [(z, 1, z), (nz, 1, nz), (z, 1, nz), (nz, 1, z)]
    o = [(0.25, 0.0, -0.5), (-0.99, 0.0, 0.99), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0)]  # inside, near and on the edge
    r = rays_from(np.repeat(np.asarray(o, np.float32), len(dirs), 0), np.tile(np.asarray(dirs, np.float32), (len(o), 1)))
    h = p.trace(r)
    assert np.all(h["instance"][:8] == 0)
    p.trace(r, abi.RTG_QUERY_OCCLUSION)
    # parallel to the quad's plane: no hit
    r = rays_from((0.0, 1.0, -3.0), (0.0, nz, 1.0))
    p.trace(r)
    p.close()


def test_closest_tie_break(ctx):
    quad = scenes.quad(2.0, 2.0)
    p = Pair(ctx, one_mesh_scene(quad, [scenes.affine((0, 1, 0))] * 3))
    rays = scenes.incoherent_rays(500, (-0.9, -1, -0.9), (0.9, 0.5, 0.9), seed=5)
    rays["direction"] = [0, 1, 0]
    h = p.trace(rays)
    assert np.all(h["instance"] == 0)
    p.close()


@pytest.mark.parametrize("inst_flags", [0, abi.RTG_INSTANCE_FORCE_OPAQUE, abi.RTG_INSTANCE_FORCE_NO_OPAQUE])
def test_any_hit_under_ray_and_instance_flags(ctx, inst_flags):
    p = Pair(ctx, fence_scene(inst_flags))
    r, opaque = fence_rays()
    ray_flags = [abi.RTG_RAY_FLAG_NONE, abi.RTG_RAY_FLAG_OPAQUE, abi.RTG_RAY_FLAG_NO_OPAQUE,
                 abi.RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT, abi.RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER]
    for rf in ray_flags:
        for sbt in (0, 1, 3):
            for mode in MODES:
                res = p.trace(r, mode, ray_flags=rf, sbt_offset=sbt)
                hit = (res["instance"] != abi.RTG_NO_HIT) if mode == abi.RTG_QUERY_CLOSEST else res.astype(bool)
                # the any-hit stage runs for non-opaque triangles of a hit
                # group inside the table; it drops the clipped texels
                non_opaque = (inst_flags != abi.RTG_INSTANCE_FORCE_OPAQUE or rf == abi.RTG_RAY_FLAG_NO_OPAQUE)
                non_opaque = non_opaque and rf != abi.RTG_RAY_FLAG_OPAQUE
                if non_opaque and sbt < abi.RTG_HIT_GROUP_COUNT:
                    assert np.array_equal(hit, opaque), (rf, sbt, mode)
                else:
                    assert np.all(hit), (rf, sbt, mode)
    p.close()


def test_cull_mask(ctx):
    quad = scenes.quad(2.0, 2.0)
    p = Pair(ctx, one_mesh_scene(quad, [scenes.affine((0, 1, 0)), scenes.affine((0, 2, 0)),
                                        scenes.affine((0, 3, 0))], masks=[0x01, 0x02, 0x00]))
    r = rays_from((0.1, 0.0, 0.1), (0, 1, 0))
    for mask, expect in [(0xFF, 0), (0x01, 0), (0x02, 1), (0x04, abi.RTG_NO_HIT), (0x00, abi.RTG_NO_HIT)]:
        h = p.trace(r, cull_mask=mask)
        assert h["instance"][0] == expect, mask
        assert p.trace(r, abi.RTG_QUERY_OCCLUSION, cull_mask=mask)[0] == (expect != abi.RTG_NO_HIT), mask
    p.close()


def test_empty_tlas_and_empty_query(ctx):
    s = one_mesh_scene(scenes.quad(2.0, 2.0), [])
    s.instances.clear()
    p = Pair(ctx, s)
    rays = scenes.incoherent_rays(1000, (-2, -2, -2), (2, 2, 2), seed=3)
    h = p.trace(rays)
    assert np.all(h["instance"] == abi.RTG_NO_HIT)
    assert not p.trace(rays, abi.RTG_QUERY_OCCLUSION).any()
    # zero rays is a no-op
    d = rays_to_device(rays[:1])
    out = p.ds.trace_rays(d, 0)
    torch.cuda.synchronize()
    assert out.numel() == 0
    p.close()


def test_non_finite_and_degenerate_rays(ctx):
    """Rays with NaN or infinite components, a zero direction, tmin > tmax,
    tmax = 0, a negative tmin or denormal direction components: each query
    ends (no hang, no stack overflow) with the oracle's bits, and the oracle
    answers them as the mathematics does (no hit from NaN rays or empty
    ranges)."""
    s = scenes.small_scene(width=8, height=8, n_tori=6)
    p = Pair(ctx, s)
    nan, inf, den = np.float32(np.nan), np.float32(np.inf), np.float32(1e-40)
    cases = [
        ((nan, 1.0, 0.0), (0.0, -1.0, 0.0), 0.0, 100.0),
        ((0.0, 3.0, 0.0), (nan, -1.0, 0.0), 0.0, 100.0),
        ((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), nan, 100.0),
        ((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), 0.0, nan),
        ((0.0, 3.0, 0.0), (0.0, 0.0, 0.0), 0.0, 100.0),
        ((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), 5.0, 1.0),
        ((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), 0.0, 0.0),
        ((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), -4.0, 100.0),
        ((0.0, 3.0, 0.0), (den, -1.0, den), 0.0, inf),
        ((inf, 3.0, 0.0), (-1.0, 0.0, 0.0), 0.0, inf),
        ((0.0, 3.0, 0.0), (inf, -1.0, 0.0), 0.0, inf),
        ((0.0, 1e30, 0.0), (0.0, -1.0, 0.0), 0.0, inf),
        ((0.0, 3.0, 0.0), (0.0, -1e-30, 0.0), 0.0, inf),
    ]
    o = np.array([c[0] for c in cases], np.float32)
    d = np.array([c[1] for c in cases], np.float32)
    rays = np.zeros(len(cases), abi.RAY_DTYPE)
    rays["origin"], rays["direction"] = o, d
    rays["tmin"] = [c[2] for c in cases]
    rays["tmax"] = [c[3] for c in cases]
    ctx.status(clear=True)
    h = p.trace(rays)
    occ = p.trace(rays, abi.RTG_QUERY_OCCLUSION)
    assert ctx.status(clear=True) == 0
    # NaN components and empty ranges hit nothing
    for k in (0, 1, 2, 3, 5, 6):
        assert h["instance"][k] == abi.RTG_NO_HIT and occ[k] == 0, k
    # the plain downward ray from y = 3 hits the ground from a negative tmin
    # too, and with a tiny or denormal sideways drift
    assert h["instance"][7] != abi.RTG_NO_HIT and h["instance"][8] != abi.RTG_NO_HIT
    p.close()
