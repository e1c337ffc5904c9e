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
NODE = np.dtype([("lo_x", "<f4", 4), ("hi_x", "<f4", 4), ("lo_y", "<f4", 4), ("hi_y", "<f4", 4),
                 ("lo_z", "<f4", 4), ("hi_z", "<f4", 4), ("child", "<u4", 4), ("pad", "<u4", 4)])
WIDTH = 4
TRI = np.dtype([("v0", "<f4", 3), ("prim", "<u4"), ("v1", "<f4", 3), ("geom", "<u4"), ("v2", "<f4", 3),
                ("pad", "<u4")])
INST = np.dtype([("w2o", "<f4", 12), ("blas", "<u8"), ("index", "<u4"), ("mask_flags", "<u4")])
assert HEADER.itemsize == 128 and NODE.itemsize == 128 and TRI.itemsize == 48 and INST.itemsize == 64


class Blob:
    def __init__(self, data):
        b = np.frombuffer(bytes(data), dtype=np.uint8)
        self.raw = b
        self.h = b[:128].view(HEADER)[0]
        n = int(self.h["node_count"])
        self.nodes = b[int(self.h["nodes_off"]):int(self.h["nodes_off"]) + 128 * n].view(NODE)
        p = int(self.h["prim_count"])
        if int(self.h["kind"]) == 0:
            self.prims = b[int(self.h["prims_off"]):int(self.h["prims_off"]) + 48 * p].view(TRI)
        else:
            self.prims = b[int(self.h["prims_off"]):int(self.h["prims_off"]) + 64 * p].view(INST)
            self.leafbox = b[int(self.h["leafbox_off"]):int(self.h["leafbox_off"]) + 24 * p].view("<f4").reshape(-1, 6)
            lo = int(self.h["level_off"])
            nl = int(b[lo:lo + 4].view("<u4")[0])
            self.level_off = [int(x) for x in b[lo + 4:lo + 8 + 4 * nl].view("<u4")]

    def child_box(self, node, s):
        nd = self.nodes[node]
        return (np.array([nd["lo_x"][s], nd["lo_y"][s], nd["lo_z"][s]]),
                np.array([nd["hi_x"][s], nd["hi_y"][s], nd["hi_z"][s]]))

    def walk(self):
        """Yields (node, slot, ref, depth) for every child slot, depth-first."""
        stack = [(0, 1)]
        while stack:
            node, depth = stack.pop()
            for s in range(WIDTH):
                ref = int(self.nodes[node]["child"][s])
                yield node, s, ref, depth
                if ref != EMPTY and not ref & LEAF:
                    stack.append((ref, depth + 1))


def leaf_span(ref):
    """(first, count) of a leaf ref."""
    return ref & FIRST, ((ref >> 27) & 0xF) + 1
