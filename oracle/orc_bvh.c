/*
 * orc_bvh.c -- CPU oracle: LBVH build and TLAS refit (TEST INFRASTRUCTURE).
 *
 * Restates, for the host, the acceleration-structure operations that the
 * reference delegates to the Vulkan driver:
 *   BLAS build  ≙ AccelerationStructure::build_blas  src/util_structs.rs:140-224
 *   TLAS build  ≙ build_tlas                         src/util_functions.rs:453-510
 *   TLAS UPDATE ≙ AccelerationStructure::update_tlas src/util_structs.rs:285-357
 * The driver's own BVH is closed; the method restated here is the linear BVH
 * of Karras 2012 (HPG, "Maximizing Parallelism in the Construction of BVHs,
 * Octrees, and k-d Trees"): 30-bit Morton codes of primitive box centres,
 * a stable sort, the radix-tree split search, bottom-up box merging.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"
#include "orc_layout.h"
#include "orc_math.h"

typedef struct {
    float lo[3], hi[3];
} box3;

static box3 box_empty(void) {
    box3 b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = INFINITY;
        b.hi[a] = -INFINITY;
    }
    return b;
}
/* min/max of box planes as selects: the first operand wins ties. */
static inline float mins(float a, float b) { return b < a ? b : a; }
static inline float maxs(float a, float b) { return b > a ? b : a; }

static box3 box_union(box3 a, box3 b) {
    box3 r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = mins(a.lo[k], b.lo[k]);
        r.hi[k] = maxs(a.hi[k], b.hi[k]);
    }
    return r;
}

/* Spreads the low 10 bits of v to every third bit. */
static uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

/* 10 bits per axis: q = clamp(((c - lo) / (hi - lo)) * 1024, 0, 1023) truncated;
 * an axis of zero extent (or a non-finite centre) gives 0. */
static uint32_t quantize10(float c, float lo, float hi) {
    float ext = hi - lo;
    float f = ext > 0.0f ? (c - lo) / ext : 0.0f;
    float q = fminf(fmaxf(f * 1024.0f, 0.0f), 1023.0f);
    return (uint32_t)q;
}

uint32_t orc_morton3(float cx, float cy, float cz, const float lo[3], const float hi[3]) {
    return (spread3(quantize10(cx, lo[0], hi[0])) << 2) | (spread3(quantize10(cy, lo[1], hi[1])) << 1) |
           spread3(quantize10(cz, lo[2], hi[2]));
}

/* Stable LSD radix sort of (key, value) by key, 8-bit digits. */
static void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n) {
    uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* v2 = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (int pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        uint32_t count[257];
        memset(count, 0, sizeof(count));
        for (uint32_t i = 0; i < n; ++i) count[((keys[i] >> shift) & 0xffu) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t d = (keys[i] >> shift) & 0xffu;
            uint32_t dst = count[d]++;
            k2[dst] = keys[i];
            v2[dst] = vals[i];
        }
        memcpy(keys, k2, sizeof(uint32_t) * n);
        memcpy(vals, v2, sizeof(uint32_t) * n);
    }
    free(k2);
    free(v2);
}

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* delta(i, j): common-prefix length of keys i and j, with the index as a
 * tie-break for equal keys (Karras 2012, section 4), -1 out of range. */
static int delta(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

/* Radix-tree node i over n sorted keys: children as refs (leaf bit set for
 * leaves). Karras 2012, Figure 4. */
static void karras_node(const uint32_t* keys, int n, int i, uint32_t* left, uint32_t* right) {
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j;
    int hi = i < j ? j : i;
    *left = (lo == gamma) ? (ORC_LEAF_BIT | (uint32_t)gamma) : (uint32_t)gamma;
    *right = (hi == gamma + 1) ? (ORC_LEAF_BIT | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
}

/* Output of the generic LBVH over n primitive boxes. */
typedef struct {
    uint32_t n;
    uint32_t node_count;
    uint32_t* order;    /* order[k] = primitive of leaf k */
    orc_node* nodes;    /* node_count */
    box3 root;
    uint32_t max_depth;
    uint32_t* depth;    /* internal node depth (root = 0) */
} lbvh;

static void set_child(orc_node* nd, int slot, uint32_t ref, box3 b) {
    nd->x[2 * slot] = b.lo[0];
    nd->x[2 * slot + 1] = b.hi[0];
    nd->y[2 * slot] = b.lo[1];
    nd->y[2 * slot + 1] = b.hi[1];
    nd->z[2 * slot] = b.lo[2];
    nd->z[2 * slot + 1] = b.hi[2];
    nd->child[slot] = ref;
}

static box3 node_box(const orc_node* nd) {
    box3 b;
    b.lo[0] = mins(nd->x[0], nd->x[2]);
    b.hi[0] = maxs(nd->x[1], nd->x[3]);
    b.lo[1] = mins(nd->y[0], nd->y[2]);
    b.hi[1] = maxs(nd->y[1], nd->y[3]);
    b.lo[2] = mins(nd->z[0], nd->z[2]);
    b.hi[2] = maxs(nd->z[1], nd->z[3]);
    return b;
}

static int lbvh_build(const box3* prim_boxes, uint32_t n, lbvh* out) {
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->node_count = n >= 2 ? n - 1 : 1;
    out->order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    out->nodes = (orc_node*)calloc(out->node_count, sizeof(orc_node));
    out->depth = (uint32_t*)calloc(out->node_count, sizeof(uint32_t));
    if (!out->order || !out->nodes || !out->depth) return RTG_ERROR_OUT_OF_MEMORY;

    if (n <= 1) {
        box3 e = box_empty();
        set_child(&out->nodes[0], 0, n == 1 ? ORC_LEAF_BIT : ORC_EMPTY, n == 1 ? prim_boxes[0] : e);
        set_child(&out->nodes[0], 1, ORC_EMPTY, e);
        if (n == 1) out->order[0] = 0;
        out->root = n == 1 ? prim_boxes[0] : e;
        out->max_depth = 1;
        return RTG_OK;
    }

    /* centre bounds over finite centres */
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float* cen = (float*)malloc(sizeof(float) * 3 * n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            /* + 0.0f turns a -0 centre into +0 */
            float c = (prim_boxes[i].lo[a] + prim_boxes[i].hi[a]) * 0.5f + 0.0f;
            cen[3 * i + a] = c;
            if (c == c && fabsf(c) != INFINITY) {
                clo[a] = fminf(clo[a], c);
                chi[a] = fmaxf(chi[a], c);
            }
        }
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) {
        keys[i] = orc_morton3(cen[3 * i], cen[3 * i + 1], cen[3 * i + 2], clo, chi);
        out->order[i] = i;
    }
    free(cen);
    radix_sort_pairs(keys, out->order, n);

    uint32_t nint = n - 1;
    uint32_t* lc = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    uint32_t* rc = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    uint32_t* parent_int = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    uint32_t* parent_leaf = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < nint; ++i) {
        karras_node(keys, (int)n, (int)i, &lc[i], &rc[i]);
        uint32_t c[2] = {lc[i], rc[i]};
        for (int s = 0; s < 2; ++s) {
            if (c[s] & ORC_LEAF_BIT) parent_leaf[c[s] & ~ORC_LEAF_BIT] = i;
            else parent_int[c[s]] = i;
        }
    }
    parent_int[0] = 0xffffffffu;
    free(keys);

    /* bottom-up boxes: a node is finished when both children are */
    box3* ibox = (box3*)malloc(sizeof(box3) * nint);
    uint32_t* height = (uint32_t*)calloc(nint, sizeof(uint32_t));
    uint8_t* arrivals = (uint8_t*)calloc(nint, 1);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t node = parent_leaf[k];
        while (node != 0xffffffffu) {
            if (arrivals[node]++ == 0) break; /* first child to arrive stops */
            box3 cb[2];
            uint32_t ch[2];
            uint32_t hmax = 0;
            uint32_t c[2] = {lc[node], rc[node]};
            for (int s = 0; s < 2; ++s) {
                if (c[s] & ORC_LEAF_BIT) {
                    cb[s] = prim_boxes[out->order[c[s] & ~ORC_LEAF_BIT]];
                    ch[s] = 0;
                } else {
                    cb[s] = ibox[c[s]];
                    ch[s] = height[c[s]];
                }
                if (ch[s] > hmax) hmax = ch[s];
            }
            ibox[node] = box_union(cb[0], cb[1]);
            height[node] = hmax + 1;
            set_child(&out->nodes[node], 0, c[0], cb[0]);
            set_child(&out->nodes[node], 1, c[1], cb[1]);
            node = parent_int[node];
        }
    }
    out->root = ibox[0];
    out->max_depth = height[0];
    /* depth of each internal node (root = 0); parents have smaller indices? no:
       walk up the parent chain */
    for (uint32_t i = 0; i < nint; ++i) {
        uint32_t dpt = 0, p = parent_int[i];
        while (p != 0xffffffffu) {
            ++dpt;
            p = parent_int[p];
        }
        out->depth[i] = dpt;
    }
    free(lc);
    free(rc);
    free(parent_int);
    free(parent_leaf);
    free(ibox);
    free(height);
    free(arrivals);
    return RTG_OK;
}

static void lbvh_free(lbvh* b) {
    free(b->order);
    free(b->nodes);
    free(b->depth);
}

static void write_header(orc_header* h, uint32_t kind, const lbvh* b, uint32_t geom_count) {
    memset(h, 0, sizeof(*h));
    h->magic = ORC_MAGIC;
    h->kind = kind;
    h->node_count = b->node_count;
    h->prim_count = b->n;
    h->geom_count = geom_count;
    h->max_depth = b->max_depth;
    h->node_width = 2;
    for (int a = 0; a < 3; ++a) {
        h->lo[a] = b->root.lo[a];
        h->hi[a] = b->root.hi[a];
    }
}

static v3 load_vertex(const rtg_geometry* g, uint32_t idx) {
    const float* p = (const float*)((const uint8_t*)(uintptr_t)g->vertex_data + (uint64_t)idx * g->vertex_stride);
    return V3(p[0], p[1], p[2]);
}

int orc_blas_build(const rtg_geometry* geoms, uint32_t geom_count, void** out_blob, uint64_t* out_size) {
    uint64_t total = 0;
    for (uint32_t g = 0; g < geom_count; ++g) {
        if (geoms[g].index_count % 3 != 0 || geoms[g].vertex_stride < 12 || geoms[g].vertex_stride % 4)
            return RTG_ERROR_INVALID_ARGUMENT;
        total += geoms[g].index_count / 3;
    }
    if (total >= (1u << 27)) return RTG_ERROR_LIMIT;
    uint32_t n = (uint32_t)total;
    box3* boxes = (box3*)malloc(sizeof(box3) * (n ? n : 1));
    orc_tri* tris = (orc_tri*)malloc(sizeof(orc_tri) * (n ? n : 1));
    uint32_t p = 0;
    for (uint32_t g = 0; g < geom_count; ++g) {
        const uint32_t* idx = (const uint32_t*)(uintptr_t)geoms[g].index_data;
        uint32_t nt = geoms[g].index_count / 3;
        for (uint32_t t = 0; t < nt; ++t, ++p) {
            uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
            if (i0 > geoms[g].max_vertex || i1 > geoms[g].max_vertex || i2 > geoms[g].max_vertex) {
                free(boxes);
                free(tris);
                return RTG_ERROR_INVALID_ARGUMENT;
            }
            v3 a = load_vertex(&geoms[g], i0), b = load_vertex(&geoms[g], i1), c = load_vertex(&geoms[g], i2);
            orc_tri* tr = &tris[p];
            tr->v0[0] = a.x; tr->v0[1] = a.y; tr->v0[2] = a.z;
            tr->v1[0] = b.x; tr->v1[1] = b.y; tr->v1[2] = b.z;
            tr->v2[0] = c.x; tr->v2[1] = c.y; tr->v2[2] = c.z;
            tr->prim = t;
            tr->geom = g | ((geoms[g].flags & RTG_GEOMETRY_OPAQUE) ? 0u : ORC_NONOPAQUE_BIT);
            tr->pad = 0;
            box3 bx;
            bx.lo[0] = mins(mins(a.x, b.x), c.x); bx.hi[0] = maxs(maxs(a.x, b.x), c.x);
            bx.lo[1] = mins(mins(a.y, b.y), c.y); bx.hi[1] = maxs(maxs(a.y, b.y), c.y);
            bx.lo[2] = mins(mins(a.z, b.z), c.z); bx.hi[2] = maxs(maxs(a.z, b.z), c.z);
            boxes[p] = bx;
        }
    }
    lbvh b;
    int st = lbvh_build(boxes, n, &b);
    if (st) {
        free(boxes);
        free(tris);
        lbvh_free(&b);
        return st;
    }
    uint64_t nodes_off = 128;
    uint64_t prims_off = orc_align(nodes_off + (uint64_t)b.node_count * 64, 64);
    uint64_t extra_off = orc_align(prims_off + (uint64_t)n * 48, 64);
    uint64_t size = orc_align(extra_off + (uint64_t)geom_count * 4, 64);
    uint8_t* blob = (uint8_t*)calloc(1, size);
    orc_header* h = (orc_header*)blob;
    write_header(h, ORC_KIND_BLAS, &b, geom_count);
    h->nodes_off = nodes_off;
    h->prims_off = prims_off;
    h->extra_off = extra_off;
    h->total_size = size;
    memcpy(blob + nodes_off, b.nodes, (size_t)b.node_count * 64);
    orc_tri* dst = (orc_tri*)(blob + prims_off);
    for (uint32_t k = 0; k < n; ++k) dst[k] = tris[b.order[k]];
    uint32_t* gflags = (uint32_t*)(blob + extra_off);
    for (uint32_t g = 0; g < geom_count; ++g) gflags[g] = geoms[g].flags;
    lbvh_free(&b);
    free(boxes);
    free(tris);
    *out_blob = blob;
    *out_size = size;
    return RTG_OK;
}

/* World box of an instance: the 8 corners of its BLAS root box through the
 * object-to-world transform; empty if the BLAS is empty. */
static box3 instance_box(const rtg_instance* in) {
    const orc_header* bh = (const orc_header*)(uintptr_t)in->acceleration_structure_reference;
    box3 r = box_empty();
    if (!bh || bh->prim_count == 0) return r;
    const float* m = (const float*)in->transform;
    for (int c = 0; c < 8; ++c) {
        v3 p = V3((c & 1) ? bh->hi[0] : bh->lo[0], (c & 2) ? bh->hi[1] : bh->lo[1], (c & 4) ? bh->hi[2] : bh->lo[2]);
        v3 w = m34_point(m, p);
        r.lo[0] = mins(r.lo[0], w.x); r.hi[0] = maxs(r.hi[0], w.x);
        r.lo[1] = mins(r.lo[1], w.y); r.hi[1] = maxs(r.hi[1], w.y);
        r.lo[2] = mins(r.lo[2], w.z); r.hi[2] = maxs(r.hi[2], w.z);
    }
    return r;
}

static void fill_inst_record(orc_inst* rec, const rtg_instance* in, uint32_t index) {
    m34_inverse((const float*)in->transform, rec->w2o);
    rec->blas = in->acceleration_structure_reference;
    rec->index = index;
    rec->mask_flags = (in->custom_index_and_mask >> 24) | ((in->sbt_offset_and_flags >> 24) << 8);
}

int orc_tlas_build(const rtg_instance* inst, uint32_t count, void** out_blob, uint64_t* out_size) {
    if (count >= (1u << 27)) return RTG_ERROR_LIMIT;
    box3* boxes = (box3*)calloc(count ? count : 1, sizeof(box3));
    for (uint32_t i = 0; i < count; ++i) boxes[i] = instance_box(&inst[i]);
    lbvh b;
    int st = lbvh_build(boxes, count, &b);
    free(boxes);
    if (st) {
        lbvh_free(&b);
        return st;
    }
    uint32_t nint = count >= 2 ? count - 1 : 0;
    uint32_t levels = 0;
    for (uint32_t i = 0; i < nint; ++i)
        if (b.depth[i] + 1 > levels) levels = b.depth[i] + 1;
    uint64_t nodes_off = 128;
    uint64_t prims_off = orc_align(nodes_off + (uint64_t)b.node_count * 64, 64);
    uint64_t extra_off = orc_align(prims_off + (uint64_t)count * 64, 64);
    uint64_t level_off = orc_align(extra_off + (uint64_t)count * 64, 64);
    uint64_t leafbox_off = orc_align(level_off + 4ull * (1 + (nint + 1) + nint), 64);
    uint64_t size = orc_align(leafbox_off + (uint64_t)count * 24, 64);
    uint8_t* blob = (uint8_t*)calloc(1, size);
    orc_header* h = (orc_header*)blob;
    write_header(h, ORC_KIND_TLAS, &b, 0);
    h->nodes_off = nodes_off;
    h->prims_off = prims_off;
    h->extra_off = extra_off;
    h->level_off = level_off;
    h->leafbox_off = leafbox_off;
    h->total_size = size;
    memcpy(blob + nodes_off, b.nodes, (size_t)b.node_count * 64);
    orc_inst* recs = (orc_inst*)(blob + prims_off);
    box3* leafbox = (box3*)(blob + leafbox_off);
    for (uint32_t k = 0; k < count; ++k) {
        fill_inst_record(&recs[k], &inst[b.order[k]], b.order[k]);
        leafbox[k] = instance_box(&inst[b.order[k]]);
    }
    memcpy(blob + extra_off, inst, (size_t)count * 64);
    /* level lists: deepest level first, ascending node index within a level */
    uint32_t* lv = (uint32_t*)(blob + level_off);
    lv[0] = levels;
    uint32_t* offs = lv + 1;
    uint32_t* list = lv + 2 + levels;
    uint32_t pos = 0;
    for (uint32_t L = 0; L < levels; ++L) {
        uint32_t depth = levels - 1 - L;
        offs[L] = pos;
        for (uint32_t i = 0; i < nint; ++i)
            if (b.depth[i] == depth) list[pos++] = i;
    }
    offs[levels] = pos;
    lbvh_free(&b);
    *out_blob = blob;
    *out_size = size;
    return RTG_OK;
}

int orc_tlas_refit(void* blob_v, const rtg_instance* inst, uint32_t count) {
    uint8_t* blob = (uint8_t*)blob_v;
    orc_header* h = (orc_header*)blob;
    if (h->magic != ORC_MAGIC || h->kind != ORC_KIND_TLAS || h->prim_count != count) return RTG_ERROR_INVALID_ARGUMENT;
    orc_node* nodes = (orc_node*)(blob + h->nodes_off);
    orc_inst* recs = (orc_inst*)(blob + h->prims_off);
    memcpy(blob + h->extra_off, inst, (size_t)count * 64);
    box3* leaf = (box3*)(blob + h->leafbox_off);
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t src = recs[k].index;
        fill_inst_record(&recs[k], &inst[src], src);
        leaf[k] = instance_box(&inst[src]);
    }
    box3 root;
    if (count <= 1) {
        box3 e = box_empty();
        if (count == 1) set_child(&nodes[0], 0, ORC_LEAF_BIT, leaf[0]);
        root = count == 1 ? leaf[0] : e;
    } else {
        const uint32_t* lv = (const uint32_t*)(blob + h->level_off);
        uint32_t levels = lv[0];
        const uint32_t* offs = lv + 1;
        const uint32_t* list = lv + 2 + levels;
        for (uint32_t L = 0; L < levels; ++L)
            for (uint32_t q = offs[L]; q < offs[L + 1]; ++q) {
                orc_node* nd = &nodes[list[q]];
                for (int s = 0; s < 2; ++s) {
                    uint32_t c = nd->child[s];
                    box3 cb = (c & ORC_LEAF_BIT) ? leaf[c & 0x07ffffffu] : node_box(&nodes[c]);
                    set_child(nd, s, c, cb);
                }
            }
        root = node_box(&nodes[0]);
    }
    for (int a = 0; a < 3; ++a) {
        h->lo[a] = root.lo[a];
        h->hi[a] = root.hi[a];
    }
    return RTG_OK;
}

void orc_free(void* blob) { free(blob); }

int orc_as_get_info(const void* blob, rtg_as_info* out) {
    const orc_header* h = (const orc_header*)blob;
    if (!h || h->magic != ORC_MAGIC) return RTG_ERROR_INVALID_ARGUMENT;
    out->kind = h->kind;
    out->node_count = h->node_count;
    out->prim_count = h->prim_count;
    out->geometry_count = h->geom_count;
    out->max_depth = h->max_depth;
    out->node_width = h->node_width;
    for (int a = 0; a < 3; ++a) {
        out->bounds[a] = h->lo[a];
        out->bounds[3 + a] = h->hi[a];
    }
    out->size_bytes = h->total_size;
    return RTG_OK;
}
