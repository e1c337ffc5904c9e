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
