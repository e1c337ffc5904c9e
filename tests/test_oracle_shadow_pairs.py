"""Pins the oracle's pair walk of an image's shadow rays (two rays of one
path per walk, the union of their nodes visited once) against the
single-ray occlusion query, which the brute-force and golden checks pin.

Per ray the pair walk tests the nodes and triangles the ray overlaps in the
single-ray query's order, up to its first accepted hit, so:
  * every image is the same bit for bit, for any shadow-ray count;
  * any-hit calls are equal (each ray runs the same any-hits);
  * with one shadow ray (every walk a lone ray) nodes, triangles and
    instances are equal;
  * with pairs, each count lies between half the single-ray count (a union
    holds at least either ray's nodes) and the single-ray count.
A dropped mask bit, a mask not narrowed at a pop or a hit that ends both
rays fails one of these."""
import ctypes as C

import numpy as np
import pytest

import oracle_host as oh
from rtg import scenes

L = oh.lib()

WALK = ("occ_node_visits", "occ_triangle_tests", "occ_instance_visits")


@pytest.fixture(scope="module")
def scene():
    s = scenes.large_scene(width=64, height=40, n_objects=40, n_fences=10, n_mirrors=4, n_soup_meshes=2,
                           soup_tris=200)
    s.set_frame(2)
    return s, oh.HostScene(s)


def trace(hs, s, nsh, pairs):
    o = s.trace_opts()
    o.shadow_rays = nsh
    L.orc_set_shadow_pairs(pairs)
    try:
        rgba, rgb, hits, cnt = hs.trace_image(frame_index=2, opts=o)
    finally:
        L.orc_set_shadow_pairs(1)
    return rgba.view(np.uint32).copy(), rgb.copy(), hits.view(np.uint32).copy(), cnt


@pytest.mark.parametrize("nsh", [1, 2, 3, 5])
def test_pairs_give_the_single_ray_image(scene, nsh):
    s, hs = scene
    a = trace(hs, s, nsh, 1)
    b = trace(hs, s, nsh, 0)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    ca, cb = a[3], b[3]
    assert ca.shadow_rays == cb.shadow_rays > 0
    assert ca.occ_anyhit_calls == cb.occ_anyhit_calls
    if nsh > 2:
        assert ca.occ_anyhit_calls > 0
    for f in WALK:
        pa, sa = getattr(ca, f), getattr(cb, f)
        if nsh == 1:
            assert pa == sa, f
        else:
            assert (sa + 1) // 2 <= pa < sa, f


def visits(hs, o, d0, d1, pair):
    cap = 1 << 16
    log = np.zeros(3 * cap, dtype=np.uint64)
    n = np.zeros(1, dtype=np.uint32)
    f3 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    o, d0, d1 = f3(o), f3(d0), f3(d1)
    r = L.orc_debug_visits(hs.tlas.ptr, C.byref(hs.bindings), o.ctypes.data, d0.ctypes.data, d1.ctypes.data,
                           int(pair), log.ctypes.data, cap, n.ctypes.data)
    assert r >= 0
    return r, [tuple(int(x) for x in log[3 * k:3 * k + 3]) for k in range(int(n[0]))]


def test_pair_walk_visits_the_union_of_the_single_walks(scene):
    """Per ray the pair walk follows the single-ray walk's order up to that
    ray's first accepted hit, so it visits exactly the union of the two
    single-ray walks' nodes, triangles and instance entries, each once, and
    reports each ray occluded as its single walk does. Pairs of shadow rays
    from random points toward the sun's cone, and pairs of unrelated
    directions."""
    s, hs = scene
    rng = np.random.default_rng(11)
    sun = np.array(s.uniforms(2).sun_dir[:3], dtype=np.float64)
    sun /= np.linalg.norm(sun)
    lo, hi = np.array([-20.0, 0.02, -20.0]), np.array([20.0, 3.0, 20.0])
    n_pairs, occl_mixed = 300, 0
    for k in range(n_pairs):
        o = lo + (hi - lo) * rng.random(3)
        if k % 3 == 2:
            d = rng.normal(size=(2, 3))
        else:
            d = sun + 0.05 * rng.normal(size=(2, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        r0, v0 = visits(hs, o, d[0], d[1], 0)
        r1, v1 = visits(hs, o, d[1], d[0], 0)
        rp, vp = visits(hs, o, d[0], d[1], 1)
        assert len(set(vp)) == len(vp)
        assert set(vp) == set(v0) | set(v1)
        assert rp == (r0 & 1) | ((r1 & 1) << 1)
        occl_mixed += rp in (1, 2)
    # some pairs end with one ray occluded and the other not
    assert occl_mixed > 10
