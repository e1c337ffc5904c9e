/*
 * orc_layout.h -- byte layout of acceleration-structure blobs as the oracle
 * writes them (TEST INFRASTRUCTURE). The GPU build writes the same bytes;
 * the parity tests compare the two.
 *
 * Blob = header (128 B) | nodes (64 B each) | prims | extra.
 *   BLAS prims: 48-B triangle records in leaf order:
 *       v0.xyz, primitive id | v1.xyz, geometry id (bit 31 = non-opaque) |
 *       v2.xyz, 0
 *   BLAS extra: per-geometry flags, u32 each.
 *   TLAS prims: 64-B instance leaf records in leaf order:
 *       world-to-object 3x4 (row-major, 48 B), BLAS address (8 B),
 *       instance index (4 B), mask | instance flags << 8 (4 B)
 *   TLAS extra: copy of the source instance records (64 B each, original
 *       order), then level data for the refit: level_count (u32),
 *       level_offsets[level_count + 1] (u32), level_nodes[internal nodes]
 *       (u32; deepest level first, ascending node index within a level),
 *       sized for the deepest possible tree, then the world box of each
 *       leaf (6 floats, leaf order).
 * Box planes are combined with selects (b < a ? b : a), so that ties between
 *       +0 and -0 resolve to the first operand on every device.
 * Node (64 B): x planes (lo0, hi0, lo1, hi1), y planes, z planes,
 *       child refs (2 x u32), 2 x u32 zero.
 * Child ref: internal node index (< 2^31), or RTG leaf = 0x80000000 |
 *       (count - 1) << 27 | first, or EMPTY = 0xFFFFFFFF.
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

typedef struct orc_node {
    float x[4];
    float y[4];
    float z[4];
    uint32_t child[2];
    uint32_t pad[2];
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
static_assert(sizeof(orc_node) == 64, "node");
static_assert(sizeof(orc_tri) == 48, "tri");
static_assert(sizeof(orc_inst) == 64, "inst");
#else
_Static_assert(sizeof(orc_header) == 128, "header");
_Static_assert(sizeof(orc_node) == 64, "node");
_Static_assert(sizeof(orc_tri) == 48, "tri");
_Static_assert(sizeof(orc_inst) == 64, "inst");
#endif

static inline uint64_t orc_align(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

#endif
