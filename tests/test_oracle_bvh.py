"""Pins the oracle's BVH build and TLAS refit with structural invariants and a
brute-force construction from the definitions: the binary radix tree of
Karras 2012 (over keys sorted by (Morton code, primitive index), the node
covering leaves [i, j] splits where the common prefix of the 62-bit augmented
keys of i and j ends), collapsed top-down into 4-wide nodes by the
largest-surface-area rule, breadth first, with subtrees of at most
BLAS_LEAF_MAX primitives as leaves."""
import ctypes as C

import numpy as np
import pytest

import oracle_host as oh
from blob import EMPTY, FIRST, LEAF, Blob, leaf_span
from rtg import scenes

L = oh.lib()
BLAS_LEAF_MAX = 1  # primitives per BLAS leaf at most


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


def f32_area(lo, hi):
    e = [np.float32(hi[a]) - np.float32(lo[a]) for a in range(3)]
    return np.float32(np.float32(e[0] * e[1]) + np.float32(e[1] * e[2])) + np.float32(e[2] * e[0])


def collapse_reference(splits, leaf_boxes, leaf_max):
    """Child refs of the 4-wide nodes, breadth first, from the binary tree's
    splits: a node rooted at binary range r takes r's two halves and, while
    it has fewer than four slots, opens the slot of largest box area (first
    slot on ties) among subtrees of more than leaf_max leaves."""
    n = len(leaf_boxes)
    if n <= max(1, leaf_max):
        return [[LEAF | ((n - 1) << 27) if n else EMPTY, EMPTY, EMPTY, EMPTY]]

    def halves(r):
        i, j = r
        s = splits[(i, j)]
        return [(i, s), (s + 1, j)]

    def area(r):
        i, j = r
        return f32_area(leaf_boxes[i:j + 1, :3].min(0), leaf_boxes[i:j + 1, 3:].max(0))

    def openable(r):
        return r[1] - r[0] + 1 > leaf_max

    out, queue, head = [], [(0, n - 1)], 0
    while head < len(queue):
        r = queue[head]
        head += 1
        slots = halves(r)
        while len(slots) < 4:
            best = None
            for k, c in enumerate(slots):
                if openable(c) and (best is None or area(c) > area(slots[best])):
                    best = k
            if best is None:
                break
            slots[best:best + 1] = halves(slots[best])
        refs = []
        for c in slots:
            if openable(c):
                refs.append(len(queue))
                queue.append(c)
            else:
                refs.append(LEAF | ((c[1] - c[0]) << 27) | c[0])
        out.append(refs + [EMPTY] * (4 - len(refs)))
    return out


@pytest.mark.parametrize("kind", ["torus", "soup", "dups"])
def test_blas_is_the_collapsed_radix_tree_of_sorted_morton_keys(kind):
    if kind == "torus":
        mesh = scenes.torus(24, 12)
    elif kind == "soup":
        mesh = scenes.soup(700, 0.05, rng=scenes.Rng(3))
    else:  # many duplicate centres: the index tie-break decides the tree
        base = scenes.torus(6, 4)
        ix = np.concatenate([base.geometries[0][0]] * 5)
        mesh = scenes.Mesh("dups", base.positions, base.normals, base.uvs, [(ix, True)])
    blob = Blob(oh.build_blas(mesh, flags=0).bytes())  # the plain LBVH
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
    expect = collapse_reference(splits, boxes[order].astype(np.float32), BLAS_LEAF_MAX)
    assert int(blob.h["node_count"]) == len(expect)
    got = [[int(c) for c in nd["child"]] for nd in blob.nodes]
    assert got == expect
    assert int(blob.h["node_width"]) == 4
    assert int(blob.h["max_depth"]) == max(d for _, _, _, d in blob.walk())  # walk() numbers the root level 1


def test_blas_boxes_are_exact_unions():
    mesh = scenes.icosphere(2)
    blob = Blob(oh.build_blas(mesh).bytes())
    boxes = tri_boxes(mesh)
    covered = []
    for node, s, ref, _ in blob.walk():
        lo, hi = blob.child_box(node, s)
        if ref == EMPTY:
            assert np.all(lo == np.inf) and np.all(hi == -np.inf)
        elif ref & LEAF:
            f, c = leaf_span(ref)
            assert 1 <= c <= BLAS_LEAF_MAX
            covered.extend(range(f, f + c))
            t = blob.prims[f:f + c]
            v = np.concatenate([t["v0"], t["v1"], t["v2"]])
            assert np.array_equal(lo, v.min(0)) and np.array_equal(hi, v.max(0))
        else:
            kids = [blob.child_box(ref, k) for k in range(4)]
            assert np.array_equal(lo, np.min([k[0] for k in kids], 0))
            assert np.array_equal(hi, np.max([k[1] for k in kids], 0))
    assert sorted(covered) == list(range(len(boxes)))  # every leaf position exactly once
    assert np.array_equal(blob.h["lo"], boxes[:, :3].min(0))
    assert np.array_equal(blob.h["hi"], boxes[:, 3:].max(0))


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


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 17])
def test_tiny_blas(n):
    """n copies of one triangle: equal Morton codes, so the radix tree splits
    the sorted positions at their highest differing bit."""
    mesh = scenes.Mesh("tiny", [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 1]] * 3, [[0, 0]] * 3,
                       [(np.array([[0, 1, 2]] * n).reshape(-1, 3), True)])
    blob = Blob(oh.build_blas(mesh, flags=0).bytes())  # the plain LBVH
    assert [int(t["prim"]) for t in blob.prims] == list(range(n))  # equal keys keep the input order
    if n <= BLAS_LEAF_MAX or n <= 1:
        # one node, one leaf of all primitives (or none)
        assert [int(c) for c in blob.nodes[0]["child"]] == [LEAF | ((n - 1) << 27) if n else EMPTY, EMPTY, EMPTY,
                                                             EMPTY]
        assert int(blob.h["node_count"]) == 1
        return
    splits = brute_radix_tree(list(range(n)))
    expect = collapse_reference(splits, tri_boxes(mesh).astype(np.float32), BLAS_LEAF_MAX)
    assert [[int(c) for c in nd["child"]] for nd in blob.nodes] == expect
    if n == 5:
        # 0-3 | 4, then the first of the equal-area halves is opened first
        assert expect[0] == [LEAF | 0, LEAF | 1, 1, LEAF | 4]


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
        if ref == EMPTY:
            continue
        if ref & LEAF:
            assert leaf_span(ref)[1] == 1
            assert np.array_equal(np.concatenate([lo, hi]), refit.leafbox[ref & FIRST])
        else:
            kids = [refit.child_box(ref, k) for k in range(4)]
            assert np.array_equal(lo, np.min([k[0] for k in kids], 0))
            assert np.array_equal(hi, np.max([k[1] for k in kids], 0))
    # the refit keeps the topology and the level data of the build
    assert [list(nd["child"]) for nd in refit.nodes] == [list(nd["child"]) for nd in Blob(
        oh.HostScene(s).tlas.bytes()).nodes]
    assert refit.level_off[0] == 0 and refit.level_off[-1] == int(refit.h["node_count"])
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


# --- treelet restructuring (RTG_BUILD_PREFER_FAST_TRACE) --------------------

FAST = oh.BLAS_FLAGS


def _tree_leaves(lc, rc, node, out):
    for c in (int(lc[node]), int(rc[node])):
        if c & LEAF:
            out.append(c & ~LEAF)
        else:
            _tree_leaves(lc, rc, c, out)
    return out


def _internal_area_sum(lc, rc, boxes_sorted):
    """Summed surface area (float64, of the float32 boxes) of all internal
    nodes of a binary tree, and each node's leaf list."""
    total = 0.0
    for i in range(len(lc)):
        leaves = _tree_leaves(lc, rc, i, [])
        b = boxes_sorted[leaves]
        e = b[:, 3:].max(0).astype(np.float64) - b[:, :3].min(0).astype(np.float64)
        total += e[0] * e[1] + e[1] * e[2] + e[2] * e[0]
    return total


def _all_trees(items):
    """Every full binary tree over the leaf set `items` (nested pairs)."""
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    for mask in range(1 << len(rest)):
        if mask == (1 << len(rest)) - 1:
            continue  # the other side would be empty
        left = [first] + [x for k, x in enumerate(rest) if mask >> k & 1]
        right = [x for k, x in enumerate(rest) if not mask >> k & 1]
        for tl in _all_trees(left):
            for tr in _all_trees(right):
                yield (tl, tr)


def _tree_cost(t, boxes):
    """Summed float64 area of the internal nodes of a nested-pair tree."""
    def walk(t):
        if not isinstance(t, tuple):
            return [t], 0.0
        la, ca = walk(t[0])
        lb, cb = walk(t[1])
        leaves = la + lb
        b = boxes[leaves]
        e = b[:, 3:].max(0).astype(np.float64) - b[:, :3].min(0).astype(np.float64)
        return leaves, ca + cb + e[0] * e[1] + e[1] * e[2] + e[2] * e[0]
    return walk(t)[1]


def _random_boxes(n, seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-4, 4, (n, 3)).astype(np.float32)
    size = rng.lognormal(-1.0, 1.0, (n, 3)).astype(np.float32)
    return np.concatenate([lo, lo + size], 1).astype(np.float32)


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (5, 3), (6, 4), (7, 5), (7, 6)])
def test_treelet_tree_is_sah_optimal_up_to_seven_leaves(n, seed):
    """With at most seven primitives the root's treelet holds them all, so the
    restructured tree must have the least summed internal-node area of all
    (2n-3)!! binary trees over them (found here by enumeration)."""
    boxes = _random_boxes(n, seed)
    lc, rc, order = oh.binary_tree(boxes, FAST)
    sb = boxes[order]
    got = _internal_area_sum(lc, rc, sb)
    best = min(_tree_cost(t, sb) for t in _all_trees(list(range(n))))
    assert got <= best * (1 + 1e-6)
    plain = _internal_area_sum(*oh.binary_tree(boxes, 0)[:2], sb)
    assert got <= plain * (1 + 1e-6)


@pytest.mark.parametrize("kind", ["soup", "torus", "grid"])
def test_treelet_tree_is_valid_and_never_costlier(kind):
    if kind == "soup":
        boxes = tri_boxes(scenes.soup(700, 0.05, rng=scenes.Rng(3))).astype(np.float32)
    elif kind == "torus":
        boxes = tri_boxes(scenes.torus(24, 12)).astype(np.float32)
    else:  # equal boxes on a grid: ties everywhere
        g = np.stack(np.meshgrid(*[np.arange(6, dtype=np.float32)] * 3), -1).reshape(-1, 3)
        boxes = np.concatenate([g, g + 0.5], 1).astype(np.float32)
    n = len(boxes)
    lc, rc, order = oh.binary_tree(boxes, FAST)
    assert sorted(order.tolist()) == list(range(n))
    # every internal node but the root has one parent; every leaf appears once
    kids = [int(c) for c in np.concatenate([lc, rc])]
    internal = sorted(c for c in kids if not c & LEAF)
    assert internal == list(range(1, n - 1))
    assert sorted(c & ~LEAF for c in kids if c & LEAF) == list(range(n))
    sb = boxes[order]
    plain_lc, plain_rc, plain_order = oh.binary_tree(boxes, 0)
    assert np.array_equal(order, plain_order)  # the same sorted primitives
    assert _internal_area_sum(lc, rc, sb) <= _internal_area_sum(plain_lc, plain_rc, sb) * (1 + 1e-6)


def test_treelet_structures_give_the_same_frame():
    """Restructured BLASes and TLAS change which nodes are visited (fewer),
    never what is hit: the frame, hit records included, equals the plain
    LBVH's bit for bit."""
    s = scenes.large_scene(width=64, height=40, n_objects=40, n_fences=6, n_mirrors=4, n_soup_meshes=2,
                           soup_tris=300)
    fast = oh.HostScene(s).trace_image()
    plain = oh.HostScene(s, blas_flags=0, tlas_flags=oh.abi.RTG_BUILD_ALLOW_UPDATE).trace_image()
    assert np.array_equal(fast[0], plain[0])
    assert np.array_equal(fast[1], plain[1])
    assert np.array_equal(fast[2].view(np.uint32), plain[2].view(np.uint32))
    assert fast[3].node_visits < plain[3].node_visits
    assert fast[3].occ_node_visits < plain[3].occ_node_visits


TREELET_LEAVES = 7
TREELET_PASSES = 5


def treelet_reference(lc, rc, leaf_boxes, passes=TREELET_PASSES):
    """Treelet restructuring (Karras & Aila 2013, section 4) restated from the
    method's description, applied to the plain LBVH's binary tree (lc, rc over
    internal nodes 0..n-2, LEAF | i for sorted primitive i):

    * each pass visits every internal node after all nodes below it;
    * the treelet of node r starts as r's two children; while it has fewer
      than seven leaves, its internal leaf of largest float32 box area (the
      first in treelet order on ties) is opened: its left child takes its
      place and its right child is appended;
    * the treelet gets the topology of least summed internal-node area: over
      leaf subsets S (bit k = treelet leaf k), cost(S) = area(S) + the least
      cost(P) + cost(S \\ P) over subsets P holding S's lowest leaf, tried in
      decreasing P and keeping the first least; single leaves cost 0; all in
      float32;
    * node numbers are handed out from a stack: r for the whole set, then the
      opened nodes in the order they were opened, the left part of a split
      before the right, the left part's subtree first;
    * boxes of the treelet's nodes are then recomputed children first.
    Returns the restructured (lc, rc)."""
    lc = [int(x) for x in lc]
    rc = [int(x) for x in rc]
    n_int = len(lc)
    f = np.float32

    def box_of(c, boxes):
        return leaf_boxes[c & ~LEAF] if c & LEAF else boxes[c]

    def union(a, b):
        return np.concatenate([np.minimum(a[:3], b[:3]), np.maximum(a[3:], b[3:])]).astype(np.float32)

    def area(b):
        return f32_area(b[:3], b[3:])

    # boxes of the internal nodes, children first
    boxes = [None] * n_int

    def fill(i):
        for c in (lc[i], rc[i]):
            if not c & LEAF:
                fill(c)
        boxes[i] = union(box_of(lc[i], boxes), box_of(rc[i], boxes))

    fill(0)

    def postorder(i, out):
        for c in (lc[i], rc[i]):
            if not c & LEAF:
                postorder(c, out)
        out.append(i)
        return out

    for _ in range(passes):
        for r in postorder(0, []):
            leaves, inner = [lc[r], rc[r]], []
            while len(leaves) < TREELET_LEAVES:
                best = None
                for k, c in enumerate(leaves):
                    if not c & LEAF and (best is None or area(boxes[c]) > area(boxes[leaves[best]])):
                        best = k
                if best is None:
                    break
                c = leaves[best]
                inner.append(c)
                leaves[best] = lc[c]
                leaves.append(rc[c])
            m = len(leaves)
            if m < 3:
                continue
            lb = [box_of(c, boxes) for c in leaves]
            cost, split = {}, {}
            for S in range(1, 1 << m):
                if S & (S - 1) == 0:
                    cost[S] = f(0.0)
                    continue
                b = None
                for k in range(m):
                    if S >> k & 1:
                        b = lb[k] if b is None else union(b, lb[k])
                low = S & -S
                best_c, best_p = None, None
                P = (S - 1) & S
                while P:
                    if P & low:
                        c = f(cost[P] + cost[S ^ P])
                        if best_c is None or c < best_c:
                            best_c, best_p = c, P
                    P = (P - 1) & S
                cost[S] = f(area(b) + best_c)
                split[S] = best_p
            stack, order, nxt = [((1 << m) - 1, r)], [], 0
            while stack:
                S, node = stack.pop()
                order.append(node)
                parts = [split[S], S ^ split[S]]
                refs = []
                for p in parts:
                    if p & (p - 1):
                        refs.append(inner[nxt])
                        nxt += 1
                    else:
                        refs.append(leaves[p.bit_length() - 1])
                lc[node], rc[node] = refs
                for h in (1, 0):
                    if parts[h] & (parts[h] - 1):
                        stack.append((parts[h], refs[h]))
            for node in reversed(order):
                boxes[node] = union(box_of(lc[node], boxes), box_of(rc[node], boxes))
    return np.array(lc, np.uint32), np.array(rc, np.uint32)


@pytest.mark.parametrize("kind,n,seed", [("random", 24, 1), ("random", 61, 2), ("random", 90, 3),
                                         ("grid", 64, 0), ("torus", 0, 0)])
def test_treelet_passes_match_a_restatement_node_for_node(kind, n, seed):
    """The oracle's five treelet passes give, node for node, the tree that an
    independent restatement of the method (treelet_reference) builds from the
    oracle's plain LBVH over the same primitives: which nodes form each
    treelet, its optimal topology, the tie-breaks and the node numbering."""
    if kind == "random":
        boxes = _random_boxes(n, seed)
    elif kind == "grid":  # equal boxes: ties in every area comparison
        g = np.stack(np.meshgrid(*[np.arange(4, dtype=np.float32)] * 3), -1).reshape(-1, 3)
        boxes = np.concatenate([g, g + 0.5], 1).astype(np.float32)
    else:
        boxes = tri_boxes(scenes.torus(8, 6)).astype(np.float32)
    plain_lc, plain_rc, order = oh.binary_tree(boxes, 0)
    lc, rc, order_fast = oh.binary_tree(boxes, FAST)
    assert np.array_equal(order, order_fast)
    ref_lc, ref_rc = treelet_reference(plain_lc, plain_rc, boxes[order])
    assert np.array_equal(lc, ref_lc)
    assert np.array_equal(rc, ref_rc)
