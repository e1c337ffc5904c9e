"""Pins the oracle's traversal and intersection semantics: BVH traversal vs
brute force over every triangle, watertightness, t-range boundaries, the
closest-hit tie-break, any-hit (alpha clip) rules, ray and instance flags,
masks (Vulkan ray-tracing rules as the reference relies on them)."""
import math

import numpy as np
import pytest

import oracle_host as oh
from rtg import abi, scenes


def rays_from(o, d, tmin=0.0, tmax=np.inf):
    o = np.atleast_2d(np.asarray(o, np.float32))
    d = np.atleast_2d(np.asarray(d, np.float32))
    n = max(len(o), len(d))
    r = np.zeros(n, dtype=abi.RAY_DTYPE)
    r["origin"] = np.broadcast_to(o, (n, 3))
    r["direction"] = np.broadcast_to(d, (n, 3))
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def one_mesh_scene(mesh, transforms=None, hit_group=abi.RTG_HIT_TEXTURED, flags=0, masks=None, material=None):
    s = scenes.Scene("test")
    scenes._base_textures(s, 1)
    s.meshes.append(mesh)
    mat = material or s.material((0.5, 0.5, 0.5), 0.5)
    s.models.append(scenes.Model(0, [mat] * len(mesh.geometries)))
    transforms = transforms or [scenes.affine((0, 0, 0))]
    for k, xf in enumerate(transforms):
        s.instances.append(scenes.InstanceData(xf, 0, 0, hit_group, flags, (masks or [0xFF] * len(transforms))[k]))
    return s


@pytest.mark.parametrize("which", ["small", "large"])
def test_bvh_traversal_equals_brute_force(which):
    if which == "small":
        s = scenes.small_scene(width=16, height=16)
        lo, hi = (-10, -1, -10), (10, 4, 10)
    else:
        s = scenes.large_scene(width=16, height=16, n_objects=60, n_fences=20, n_mirrors=6, n_soup_meshes=3,
                               soup_tris=300)
        lo, hi = (-40, -1, -40), (40, 4, 40)
    hs = oh.HostScene(s)
    rays = scenes.incoherent_rays(4000, lo, hi, seed=21)
    rays["tmax"] = 1e4
    for mode in (abi.RTG_QUERY_CLOSEST, abi.RTG_QUERY_OCCLUSION):
        a, _ = hs.trace_rays(rays, mode=mode)
        b, _ = hs.trace_rays(rays, mode=mode, brute=True)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), mode
    # shadow-style rays (sbt offset 1, terminate on first hit) through the fences
    a, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION, sbt_offset=1,
                         ray_flags=abi.RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT)
    b, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION, sbt_offset=1, brute=True,
                         ray_flags=abi.RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT)
    assert np.array_equal(a, b)


def test_watertight_through_vertices_and_edges():
    mesh = scenes.torus(24, 12, 1.0, 0.35)
    hs = oh.HostScene(one_mesh_scene(mesh))
    # origin on the tube's centre circle: every ray must leave the closed tube
    u = 0.37
    o = np.array([math.cos(u), 0.0, math.sin(u)], np.float32)
    targets = [mesh.positions]
    ix = mesh.geometries[0][0]
    p = mesh.positions
    targets.append((p[ix[:, 0]] + p[ix[:, 1]]) * 0.5)
    targets.append((p[ix[:, 1]] + p[ix[:, 2]]) * 0.5)
    targets.append((p[ix[:, 2]] + p[ix[:, 0]]) * 0.5)
    tg = np.concatenate(targets).astype(np.float32)
    d = tg - o
    near = np.linalg.norm(d, axis=1) < 0.36  # the near side of the tube only
    rays = rays_from(o, d[near])
    hits, _ = hs.trace_rays(rays)
    assert np.all(hits["instance"] != abi.RTG_NO_HIT)
    occ, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION)
    assert np.all(occ == 1)


def test_t_range_is_inclusive():
    quad = scenes.quad(2.0, 2.0)  # y = 0 plane
    hs = oh.HostScene(one_mesh_scene(quad, [scenes.affine((0, 5, 0))]))
    f = np.float32(5.0)
    below = np.nextafter(f, np.float32(0))
    above = np.nextafter(f, np.float32(10))
    cases = [(0.0, f, 1), (0.0, below, 0), (f, 10.0, 1), (above, 10.0, 0)]
    for tmin, tmax, expect in cases:
        r = rays_from((0.1, 0.0, 0.2), (0.0, 1.0, 0.0), tmin, tmax)
        h, _ = hs.trace_rays(r)
        assert (h["instance"][0] != abi.RTG_NO_HIT) == bool(expect), (tmin, tmax)
        if expect:
            assert h["t"][0] == f
    # barycentrics: u weights vertex 1, v vertex 2 (hitAttributeEXT)
    r = rays_from((0.5, 0.0, -0.5), (0.0, 1.0, 0.0))
    h, _ = hs.trace_rays(r)
    tri = quad.geometries[0][0][h["primitive"][0]]
    pv = quad.positions[tri]
    point = (1 - h["u"][0] - h["v"][0]) * pv[0] + h["u"][0] * pv[1] + h["v"][0] * pv[2]
    np.testing.assert_allclose(point[[0, 2]], [0.5, -0.5], atol=1e-6)


def test_closest_tie_break_is_order_independent():
    quad = scenes.quad(2.0, 2.0)
    # three coincident copies: the hit must name the lowest instance index
    s = one_mesh_scene(quad, [scenes.affine((0, 1, 0))] * 3)
    hs = oh.HostScene(s)
    rays = scenes.incoherent_rays(500, (-0.9, -1, -0.9), (0.9, 0.5, 0.9), seed=5)
    rays["direction"] = [0, 1, 0]
    h, _ = hs.trace_rays(rays)
    assert np.all(h["instance"] == 0)
    b, _ = hs.trace_rays(rays, brute=True)
    assert np.array_equal(h.view(np.uint32), b.view(np.uint32))


def fence_scene(flags=0):
    s = scenes.Scene("fence")
    scenes._base_textures(s, 1)
    checker = s.add_texture(scenes.checker_texture(256, 32))
    s.meshes.append(scenes.quad(2.0, 2.0, 1.0, "fence", opaque=False, vertical=True))
    s.models.append(scenes.Model(0, [(checker, 0, -1)]))
    s.instances.append(scenes.InstanceData(scenes.affine((0, 0, 2)), 0, 0, abi.RTG_HIT_TEXTURED, flags))
    return s


def fence_rays():
    # the quad spans x in [-1, 1], y in [0, 2] at z = 2; uv = ((x+1)/2, y/2)
    xs = (np.arange(16) + 0.5) / 16 * 2 - 1
    ys = (np.arange(16) + 0.5) / 16 * 2
    X, Y = np.meshgrid(xs, ys)
    o = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    r = rays_from(o, (0, 0, 1))
    u = (X.ravel() + 1) / 2
    v = Y.ravel() / 2
    opaque = (((np.floor(u * 256) // 32) + (np.floor(v * 256) // 32)) % 2) == 0
    return r, opaque


def test_alpha_clip_any_hit():
    s = fence_scene()
    hs = oh.HostScene(s)
    r, opaque = fence_rays()
    h, _ = hs.trace_rays(r)
    assert np.array_equal(h["instance"] != abi.RTG_NO_HIT, opaque)
    occ, _ = hs.trace_rays(r, mode=abi.RTG_QUERY_OCCLUSION, sbt_offset=1)
    assert np.array_equal(occ.astype(bool), opaque)
    # OPAQUE ray flag skips any-hit: every ray hits
    h, _ = hs.trace_rays(r, ray_flags=abi.RTG_RAY_FLAG_OPAQUE)
    assert np.all(h["instance"] != abi.RTG_NO_HIT)
    # a hit group beyond the table has no any-hit stage (portal instance + shadow offset)
    h, _ = hs.trace_rays(r, sbt_offset=3)
    assert np.all(h["instance"] != abi.RTG_NO_HIT)


def test_instance_force_opaque_flag():
    hs = oh.HostScene(fence_scene(abi.RTG_INSTANCE_FORCE_OPAQUE))
    r, _ = fence_rays()
    h, _ = hs.trace_rays(r)
    assert np.all(h["instance"] != abi.RTG_NO_HIT)
    # the ray's NO_OPAQUE flag wins over the instance flag
    h, _ = hs.trace_rays(r, ray_flags=abi.RTG_RAY_FLAG_NO_OPAQUE)
    _, opaque = fence_rays()
    assert np.array_equal(h["instance"] != abi.RTG_NO_HIT, opaque)


def test_cull_mask():
    quad = scenes.quad(2.0, 2.0)
    s = one_mesh_scene(quad, [scenes.affine((0, 1, 0)), scenes.affine((0, 2, 0))], masks=[0x01, 0x02])
    hs = oh.HostScene(s)
    r = rays_from((0.1, 0.0, 0.1), (0, 1, 0))
    for mask, expect in [(0xFF, 0), (0x01, 0), (0x02, 1), (0x04, abi.RTG_NO_HIT)]:
        h, _ = hs.trace_rays(r, cull_mask=mask)
        assert h["instance"][0] == expect, mask


def test_shadow_rays_of_the_small_scene_match_brute_force():
    s = scenes.small_scene(width=16, height=16)
    hs = oh.HostScene(s)
    rng = np.random.default_rng(8)
    o = np.stack([rng.uniform(-9, 9, 3000), np.full(3000, 0.001), rng.uniform(-9, 9, 3000)], 1)
    rays = rays_from(o, s.sun_dir, 0.001, 10000.0)
    a, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION, sbt_offset=1)
    b, _ = hs.trace_rays(rays, mode=abi.RTG_QUERY_OCCLUSION, sbt_offset=1, brute=True)
    assert np.array_equal(a, b) and 0.02 < a.mean() < 0.9
