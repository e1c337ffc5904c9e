"""Parsing of acceleration-structure blobs (layout: include/rtg.h and
csrc/rtg_device.hpp) for structural tests."""
import numpy as np

LEAF = 0x80000000
EMPTY = 0xFFFFFFFF
FIRST = 0x07FFFFFF

HEADER = np.dtype([("magic", "<u4"), ("kind", "<u4"), ("node_count", "<u4"), ("prim_count", "<u4"),
                   ("geom_count", "<u4"), ("max_depth", "<u4"), ("node_width", "<u4"), ("flags", "<u4"),
                   ("lo", "<f4", 3), ("pad0", "<f4"), ("hi", "<f4", 3), ("pad1", "<f4"),
                   ("nodes_off", "<u8"), ("prims_off", "<u8"), ("extra_off", "<u8"), ("level_off", "<u8"),
                   ("total_size", "<u8"), ("leafbox_off", "<u8"), ("reserved", "<u8", 2)])
NODE = np.dtype([("x", "<f4", 4), ("y", "<f4", 4), ("z", "<f4", 4), ("child", "<u4", 2), ("pad", "<u4", 2)])
TRI = np.dtype([("v0", "<f4", 3), ("prim", "<u4"), ("v1", "<f4", 3), ("geom", "<u4"), ("v2", "<f4", 3),
                ("pad", "<u4")])
INST = np.dtype([("w2o", "<f4", 12), ("blas", "<u8"), ("index", "<u4"), ("mask_flags", "<u4")])
assert HEADER.itemsize == 128 and NODE.itemsize == 64 and TRI.itemsize == 48 and INST.itemsize == 64


class Blob:
    def __init__(self, data):
        b = np.frombuffer(bytes(data), dtype=np.uint8)
        self.raw = b
        self.h = b[:128].view(HEADER)[0]
        n = int(self.h["node_count"])
        self.nodes = b[int(self.h["nodes_off"]):int(self.h["nodes_off"]) + 64 * n].view(NODE)
        p = int(self.h["prim_count"])
        if int(self.h["kind"]) == 0:
            self.prims = b[int(self.h["prims_off"]):int(self.h["prims_off"]) + 48 * p].view(TRI)
        else:
            self.prims = b[int(self.h["prims_off"]):int(self.h["prims_off"]) + 64 * p].view(INST)
            self.leafbox = b[int(self.h["leafbox_off"]):int(self.h["leafbox_off"]) + 24 * p].view("<f4").reshape(-1, 6)

    def child_box(self, node, s):
        nd = self.nodes[node]
        return (np.array([nd["x"][2 * s], nd["y"][2 * s], nd["z"][2 * s]]),
                np.array([nd["x"][2 * s + 1], nd["y"][2 * s + 1], nd["z"][2 * s + 1]]))

    def walk(self):
        """Yields (node, slot, ref, depth) for every child slot, depth-first."""
        stack = [(0, 1)]
        while stack:
            node, depth = stack.pop()
            for s in range(2):
                ref = int(self.nodes[node]["child"][s])
                yield node, s, ref, depth
                if ref != EMPTY and not ref & LEAF:
                    stack.append((ref, depth + 1))

    def leaf_ranges(self):
        """Maps each internal node to the (first, last) leaf index below it."""
        out = {}

        def rng(ref):
            if ref & LEAF:
                f = ref & FIRST
                c = ((ref >> 27) & 0xF) + 1
                return f, f + c - 1
            a = rng(int(self.nodes[ref]["child"][0]))
            b = rng(int(self.nodes[ref]["child"][1]))
            out[ref] = (a[0], b[1], a[1])  # first, last, end of left child
            return a[0], b[1]

        if int(self.h["prim_count"]) >= 2:
            rng(0)
        return out
