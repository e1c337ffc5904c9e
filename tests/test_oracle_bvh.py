"""Pins the oracle's LBVH build and TLAS refit with structural invariants and a
brute-force construction of the binary radix tree (Karras 2012) from its
definition: over keys sorted by (Morton code, primitive index), the node
covering leaves [i, j] splits where the common prefix of the 62-bit augmented
keys of i and j ends."""
import ctypes as C

import numpy as np
import pytest

import oracle_host as oh
from blob import EMPTY, FIRST, LEAF, Blob
from rtg import scenes

L = oh.lib()


def tri_boxes(mesh):
    boxes = []
    for ix, _ in mesh.geometries:
        v = mesh.positions[ix]  # (T, 3, 3)
        boxes.append(np.concatenate([v.min(1), v.max(1)], 1))
    return np.concatenate(boxes, 0) if boxes else np.zeros((0, 6), np.float32)


def morton_codes(boxes):
    """Morton codes of box centres, quantised in float32 (the build's rule)."""
    c = ((boxes[:, :3] + boxes[:, 3:]) * np.float32(0.5) + np.float32(0.0)).astype(np.float32)
    fin = np.isfinite(c)
    lo = np.array([c[fin[:, a], a].min() if fin[:, a].any() else np.inf for a in range(3)], np.float32)
    hi = np.array([c[fin[:, a], a].max() if fin[:, a].any() else -np.inf for a in range(3)], np.float32)
    codes = np.zeros(len(c), np.uint64)
    # per-point code by orc_morton3, itself pinned by test_morton_interleave
    for i in range(len(c)):
        codes[i] = L.orc_morton3(float(c[i, 0]), float(c[i, 1]), float(c[i, 2]),
                                 lo.ctypes.data_as(C.POINTER(C.c_float)), hi.ctypes.data_as(C.POINTER(C.c_float)))
    return codes


def spread_py(v):
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def test_morton_interleave():
    lo = np.zeros(3, np.float32)
    hi = np.ones(3, np.float32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    # corners and a few interior points: code = x bits << 2 | y << 1 | z
    for x, y, z in [(0, 0, 0), (1, 1, 1), (0.5, 0.25, 0.75), (0.999, 0.0, 0.5), (0.1, 0.9, 0.3)]:
        qx, qy, qz = [min(1023, int(np.float32(t) * np.float32(1024.0))) for t in (x, y, z)]
        expect = (spread_py(qx) << 2) | (spread_py(qy) << 1) | spread_py(qz)
        assert L.orc_morton3(x, y, z, p(lo), p(hi)) == expect


def brute_radix_tree(aug):
    """Splits of the binary radix tree over sorted distinct 62-bit keys."""
    splits = {}

    def prefix(a, b):
        return 64 - int(a ^ b).bit_length()

    def build(i, j):
        if i == j:
            return
        p = prefix(aug[i], aug[j])
        s = i
        while s + 1 <= j and prefix(aug[i], aug[s + 1]) > p:
            s += 1
        splits[(i, j)] = s
        build(i, s)
        build(s + 1, j)

    build(0, len(aug) - 1)
    return splits


@pytest.mark.parametrize("kind", ["torus", "soup", "dups"])
def test_blas_is_the_radix_tree_of_sorted_morton_keys(kind):
    if kind == "torus":
        mesh = scenes.torus(24, 12)
    elif kind == "soup":
        mesh = scenes.soup(700, 0.05, rng=scenes.Rng(3))
    else:  # many duplicate centres: the index tie-break decides the tree
        base = scenes.torus(6, 4)
        ix = np.concatenate([base.geometries[0][0]] * 5)
        mesh = scenes.Mesh("dups", base.positions, base.normals, base.uvs, [(ix, True)])
    blob = Blob(oh.build_blas(mesh).bytes())
    boxes = tri_boxes(mesh)
    n = len(boxes)
    order = [int(t["prim"]) for t in blob.prims]  # single geometry: prim = original index
    assert sorted(order) == list(range(n))        # each primitive exactly once
    codes = morton_codes(boxes)
    keyed = [(int(codes[p]) << 32) | p for p in order]
    assert keyed == sorted(keyed)                  # stable sort: leaves in (Morton, index) order
    # the tree's tie-break for equal codes is the position in the sorted array
    aug = [(int(codes[p]) << 32) | k for k, p in enumerate(order)]
    splits = brute_radix_tree(aug)
    ranges = blob.leaf_ranges()
    assert len(ranges) == n - 1
    for node, (first, last, left_end) in ranges.items():
        assert splits[(first, last)] == left_end, node


def test_blas_boxes_are_exact_unions():
    mesh = scenes.icosphere(2)
    blob = Blob(oh.build_blas(mesh).bytes())
    boxes = tri_boxes(mesh)
    for node, s, ref, _ in blob.walk():
        lo, hi = blob.child_box(node, s)
        if ref & LEAF:
            t = blob.prims[ref & FIRST]
            v = np.stack([t["v0"], t["v1"], t["v2"]])
            assert np.array_equal(lo, v.min(0)) and np.array_equal(hi, v.max(0))
        else:
            a = blob.child_box(ref, 0)
            b = blob.child_box(ref, 1)
            assert np.array_equal(lo, np.minimum(a[0], b[0])) and np.array_equal(hi, np.maximum(a[1], b[1]))
    assert np.array_equal(blob.h["lo"], boxes[:, :3].min(0))
    assert np.array_equal(blob.h["hi"], boxes[:, 3:].max(0))
    assert int(blob.h["max_depth"]) == max(d for _, _, _, d in blob.walk())


def test_blas_geometry_ids_and_opacity():
    t = scenes.torus(8, 4)
    ix = t.geometries[0][0]
    mesh = scenes.Mesh("two", t.positions, t.normals, t.uvs, [(ix[:20], True), (ix[20:], False)])
    blob = Blob(oh.build_blas(mesh).bytes())
    seen = set()
    for tr in blob.prims:
        g = int(tr["geom"]) & 0x7FFFFFFF
        nonopaque = bool(int(tr["geom"]) & 0x80000000)
        assert nonopaque == (g == 1)
        src = mesh.geometries[g][0][int(tr["prim"])]
        assert np.array_equal(tr["v0"], mesh.positions[src[0]])
        assert np.array_equal(tr["v2"], mesh.positions[src[2]])
        seen.add((g, int(tr["prim"])))
    assert len(seen) == len(ix)


@pytest.mark.parametrize("n", [0, 1])
def test_tiny_blas(n):
    mesh = scenes.Mesh("tiny", [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 1]] * 3, [[0, 0]] * 3,
                       [(np.array([[0, 1, 2]] * n).reshape(-1, 3), True)])
    blob = Blob(oh.build_blas(mesh).bytes())
    nd = blob.nodes[0]
    assert int(nd["child"][1]) == EMPTY
    assert int(nd["child"][0]) == (LEAF if n == 1 else EMPTY)
    assert int(blob.h["node_count"]) == 1


def test_bad_index_is_rejected():
    mesh = scenes.Mesh("bad", [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 1]] * 3, [[0, 0]] * 3,
                       [(np.array([[0, 1, 3]]), True)])
    with pytest.raises(RuntimeError):
        oh.build_blas(mesh)


def instance_world_box_f64(xf, lo, hi):
    corners = np.array([[(lo, hi)[c & 1][0], (lo, hi)[(c >> 1) & 1][1], (lo, hi)[(c >> 2) & 1][2]]
                        for c in range(8)], np.float64)
    w = corners @ xf[:, :3].astype(np.float64).T + xf[:, 3].astype(np.float64)
    return w.min(0), w.max(0)


def test_tlas_refit_equals_fresh_build_bounds():
    s = scenes.small_scene(width=32, height=32)
    hs = oh.HostScene(s)
    rng = scenes.Rng(4)
    xf = [i.transform.copy() for i in s.instances]
    for k in range(1, len(xf)):
        xf[k] = scenes.affine(xf[k][:, 3] + rng.uniform(3, -2, 2), scenes.rotation_from_uniform(*rng.uniform(3)),
                              float(rng.uniform(1, 0.5, 2)[0]))
    hs.refit(xf)
    refit = Blob(hs.tlas.bytes())
    # a fresh build with the new transforms
    s2 = scenes.small_scene(width=32, height=32)
    for k, x in enumerate(xf):
        s2.instances[k].transform = x
    fresh = Blob(oh.HostScene(s2).tlas.bytes())
    by_inst_refit = {int(r["index"]): refit.leafbox[k] for k, r in enumerate(refit.prims)}
    by_inst_fresh = {int(r["index"]): fresh.leafbox[k] for k, r in enumerate(fresh.prims)}
    for i in by_inst_refit:
        assert np.array_equal(by_inst_refit[i], by_inst_fresh[i])
    assert np.array_equal(refit.h["lo"], fresh.h["lo"]) and np.array_equal(refit.h["hi"], fresh.h["hi"])
    # every node box is the union of its children, leaves are their instance boxes
    for node, slot, ref, _ in refit.walk():
        lo, hi = refit.child_box(node, slot)
        if ref & LEAF:
            assert np.array_equal(np.concatenate([lo, hi]), refit.leafbox[ref & FIRST])
        else:
            a, b = refit.child_box(ref, 0), refit.child_box(ref, 1)
            assert np.array_equal(lo, np.minimum(a[0], b[0])) and np.array_equal(hi, np.maximum(a[1], b[1]))
    # containment: every instance's transformed BLAS box lies in its leaf box
    blas_info = [b.info() for b in hs.blas]
    for k, r in enumerate(refit.prims):
        ins = s.instances[int(r["index"])]
        bi = blas_info[ins.mesh]
        wlo, whi = instance_world_box_f64(xf[int(r["index"])], np.array(bi.bounds[:3]), np.array(bi.bounds[3:]))
        box = refit.leafbox[k].astype(np.float64)
        assert np.all(box[:3] <= wlo + 1e-5 * (1 + abs(wlo))) and np.all(box[3:] >= whi - 1e-5 * (1 + abs(whi)))
    # the world-to-object records invert the transforms
    for r in refit.prims:
        m = np.eye(4)
        m[:3] = xf[int(r["index"])]
        w = np.eye(4)
        w[:3] = r["w2o"].reshape(3, 4)
        np.testing.assert_allclose(w @ m, np.eye(4), atol=2e-5)
