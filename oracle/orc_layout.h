/*
 * orc_layout.h -- byte layout of acceleration-structure blobs as the oracle
 * writes them (TEST INFRASTRUCTURE). The GPU build writes the same bytes;
 * the parity tests compare the two.
 *
 * Blob = header (128 B) | nodes (128 B each, breadth-first) | prims | extra.
 *   BLAS prims: 48-B triangle records in leaf order:
 *       v0.xyz, primitive id | v1.xyz, geometry id (bit 31 = non-opaque) |
 *       v2.xyz, 0
 *   BLAS extra: per-geometry flags, u32 each.
 *   TLAS prims: 64-B instance leaf records in leaf order:
 *       world-to-object 3x4 (row-major, 48 B), BLAS address (8 B),
 *       instance index (4 B), mask | instance flags << 8 (4 B)
 *   TLAS extra: copy of the source instance records (64 B each, original
 *       order), then level data for the refit: level_count (u32) and
 *       level_offsets[level_count + 1] (u32; level L = nodes
 *       [off[L], off[L+1]) of the breadth-first order, root level 0), then
 *       the world box of each leaf (6 floats, leaf order).
 * Box planes are combined with selects (b < a ? b : a), so that ties between
 *       +0 and -0 resolve to the first operand on every device; a slot box is
 *       the fold, from the empty box, of its leaf's primitive boxes in leaf
 *       order or of its child node's four slot boxes in slot order.
 * Node (128 B), four child slots: lo_x[4], hi_x[4], lo_y[4], hi_y[4],
 *       lo_z[4], hi_z[4], child refs (4 x u32), 4 x u32 zero. Unused slots
 *       hold EMPTY and the empty box (lo = +inf, hi = -inf).
 * Child ref: internal node index (< 2^31), or RTG leaf = 0x80000000 |
 *       (count - 1) << 27 | first (count <= 16 primitives first..first+count-1
 *       of the leaf order), or EMPTY = 0xFFFFFFFF.
 * Wide tree (the build): the binary radix tree of Karras 2012 collapsed to
 *       4-wide nodes top-down in breadth-first order; a node rooted at binary
 *       node r takes r's two children and, while it has fewer than four
 *       slots, replaces the slot whose binary subtree has the largest box
 *       surface area (first slot on ties; subtrees of at most leaf_max
 *       primitives are not opened) by that subtree's two children, in place.
 *       A subtree of at most leaf_max primitives is a leaf.
 */
#ifndef ORC_LAYOUT_H
#define ORC_LAYOUT_H
#include <stdint.h>

#define ORC_MAGIC 0x53415452u /* 'RTAS' */
#define ORC_KIND_BLAS 0u
#define ORC_KIND_TLAS 1u
#define ORC_LEAF_BIT 0x80000000u
#define ORC_EMPTY 0xFFFFFFFFu
#define ORC_NONOPAQUE_BIT 0x80000000u

typedef struct orc_header {
    uint32_t magic;
    uint32_t kind;
    uint32_t node_count;
    uint32_t prim_count;
    uint32_t geom_count;
    uint32_t max_depth;
    uint32_t node_width;
    uint32_t flags;
    float lo[3];
    float pad0;
    float hi[3];
    float pad1;
    uint64_t nodes_off;
    uint64_t prims_off;
    uint64_t extra_off;
    uint64_t level_off; /* TLAS: offset of the level data; BLAS: 0 */
    uint64_t total_size;
    uint64_t leafbox_off; /* TLAS: world box of each leaf (refit scratch); BLAS: 0 */
    uint64_t reserved[2];
} orc_header;

/* primitives per BLAS leaf at most (TLAS leaves hold one instance) */
#define ORC_BLAS_LEAF_MAX 1u
#define ORC_TLAS_LEAF_MAX 1u

typedef struct orc_node {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t pad[4];
} orc_node;

typedef struct orc_tri {
    float v0[3];
    uint32_t prim;
    float v1[3];
    uint32_t geom;
    float v2[3];
    uint32_t pad;
} orc_tri;

typedef struct orc_inst {
    float w2o[12];
    uint64_t blas;
    uint32_t index;
    uint32_t mask_flags;
} orc_inst;

#ifdef __cplusplus
static_assert(sizeof(orc_header) == 128, "header");
static_assert(sizeof(orc_node) == 128, "node");
static_assert(sizeof(orc_tri) == 48, "tri");
static_assert(sizeof(orc_inst) == 64, "inst");
#else
_Static_assert(sizeof(orc_header) == 128, "header");
_Static_assert(sizeof(orc_node) == 128, "node");
_Static_assert(sizeof(orc_tri) == 48, "tri");
_Static_assert(sizeof(orc_inst) == 64, "inst");
#endif

static inline uint64_t orc_align(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

#endif
