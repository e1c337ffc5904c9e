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
 * a stable sort, the radix-tree split search, bottom-up box merging; with
 * RTG_BUILD_PREFER_FAST_TRACE, treelet restructuring (Karras & Aila 2013)
 * of the binary tree before its collapse into 4-wide nodes.
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
 * leaves) and the leaf range the node covers. Karras 2012, Figure 4. */
static void karras_node(const uint32_t* keys, int n, int i, uint32_t* left, uint32_t* right, uint32_t* first,
                        uint32_t* count) {
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
    *first = (uint32_t)lo;
    *count = (uint32_t)(hi - lo + 1);
}

/* Surface area of a box (half of it: the heuristic only compares). */
static float box_area(box3 b) {
    float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
    return (ex * ey + ey * ez) + ez * ex;
}

/* The binary radix tree over n >= 2 sorted primitives. */
typedef struct {
    uint32_t* lc;     /* n - 1 internal nodes: child refs */
    uint32_t* rc;
    uint32_t* first;  /* leaf range */
    uint32_t* count;
    float* area;      /* box surface area */
} bintree;

/* The 4-wide tree of a build. */
typedef struct {
    uint32_t n;
    uint32_t* order;      /* order[k] = primitive of leaf k */
    orc_node* nodes;      /* breadth-first */
    uint32_t node_count;
    uint32_t* level_off;  /* level_count + 1 */
    uint32_t level_count;
    box3 root;
} wtree;

static void wtree_free(wtree* w) {
    free(w->order);
    free(w->nodes);
    free(w->level_off);
}

static void set_slot(orc_node* nd, int s, uint32_t ref, box3 b) {
    nd->lo_x[s] = b.lo[0];
    nd->hi_x[s] = b.hi[0];
    nd->lo_y[s] = b.lo[1];
    nd->hi_y[s] = b.hi[1];
    nd->lo_z[s] = b.lo[2];
    nd->hi_z[s] = b.hi[2];
    nd->child[s] = ref;
}

static box3 slot_box(const orc_node* nd, int s) {
    box3 b;
    b.lo[0] = nd->lo_x[s];
    b.hi[0] = nd->hi_x[s];
    b.lo[1] = nd->lo_y[s];
    b.hi[1] = nd->hi_y[s];
    b.lo[2] = nd->lo_z[s];
    b.hi[2] = nd->hi_z[s];
    return b;
}

/* Box of a whole node: its four slot boxes folded in slot order. */
static box3 node_box(const orc_node* nd) {
    box3 b = box_empty();
    for (int s = 0; s < 4; ++s) b = box_union(b, slot_box(nd, s));
    return b;
}

/* Slot box of child ref c: its leaf's primitive boxes folded in leaf order,
 * or its node's box. */
static box3 ref_box(const orc_node* nodes, uint32_t c, const box3* leaf_boxes) {
    box3 b = box_empty();
    if (c == ORC_EMPTY) return b;
    if (c & ORC_LEAF_BIT) {
        uint32_t f = c & 0x07ffffffu, cnt = ((c >> 27) & 0xfu) + 1u;
        for (uint32_t k = f; k < f + cnt; ++k) b = box_union(b, leaf_boxes[k]);
        return b;
    }
    return node_box(&nodes[c]);
}

/* Slots of the wide node rooted at binary node r: r's children, then the
 * largest-area openable slot replaced by its two children while fewer than
 * four. Returns the slot count. */
static int expand_slots(const bintree* t, uint32_t r, uint32_t leaf_max, uint32_t s[4]) {
    int m = 2;
    s[0] = t->lc[r];
    s[1] = t->rc[r];
    while (m < 4) {
        int best = -1;
        float best_area = 0.0f;
        for (int k = 0; k < m; ++k) {
            uint32_t c = s[k];
            if ((c & ORC_LEAF_BIT) || t->count[c] <= leaf_max) continue;
            if (best < 0 || t->area[c] > best_area) {
                best = k;
                best_area = t->area[c];
            }
        }
        if (best < 0) break;
        uint32_t c = s[best];
        for (int k = m; k > best + 1; --k) s[k] = s[k - 1];
        s[best] = t->lc[c];
        s[best + 1] = t->rc[c];
        ++m;
    }
    return m;
}

/* Treelet restructuring of the binary tree (Karras & Aila 2013, HPG, "Fast
 * Parallel Construction of High-Quality Bounding Volume Hierarchies",
 * section 4), with single-primitive leaves. The treelet of internal node r:
 * r's two children, then, while fewer than ORC_TREELET_LEAVES, the treelet
 * leaf with the largest area that is an internal node replaced by its left
 * child in place and its right child appended. Its topology is replaced by
 * the one that minimises the summed surface area of its internal nodes
 * (the SAH cost with fixed leaves), found by dynamic programming over the
 * subsets of treelet leaves: cost(S) = area(S) + min over partitions (P, S\P),
 * P holding S's lowest leaf, of cost(P) + cost(S\P), partitions tried in
 * decreasing order of P with the first minimum kept. The new internal nodes
 * reuse the treelet's node numbers: r for the whole set, then the expanded
 * nodes in expansion order, handed out as subsets are split from a stack
 * (P's subtree first). */
#define ORC_TREELET_LEAVES 7

static box3 bin_box(const box3* ibox, const box3* leaf_boxes, uint32_t c) {
    return (c & ORC_LEAF_BIT) ? leaf_boxes[c & ~ORC_LEAF_BIT] : ibox[c];
}

static void treelet_optimize(bintree* t, box3* ibox, const box3* leaf_boxes, uint32_t r) {
    uint32_t leaf[ORC_TREELET_LEAVES], inner[ORC_TREELET_LEAVES];
    int nl = 2, ni = 0;
    leaf[0] = t->lc[r];
    leaf[1] = t->rc[r];
    while (nl < ORC_TREELET_LEAVES) {
        int best = -1;
        float best_area = 0.0f;
        for (int k = 0; k < nl; ++k) {
            uint32_t c = leaf[k];
            if (c & ORC_LEAF_BIT) continue;
            if (best < 0 || t->area[c] > best_area) {
                best = k;
                best_area = t->area[c];
            }
        }
        if (best < 0) break;
        uint32_t c = leaf[best];
        inner[ni++] = c;
        leaf[best] = t->lc[c];
        leaf[nl++] = t->rc[c];
    }
    if (nl < 3) return;
    const uint32_t full = (1u << nl) - 1u;
    float cost[1 << ORC_TREELET_LEAVES];
    uint32_t split[1 << ORC_TREELET_LEAVES];
    for (uint32_t S = 1; S <= full; ++S) {
        if ((S & (S - 1)) == 0) {
            cost[S] = 0.0f;
            split[S] = 0;
            continue;
        }
        box3 b = box_empty();
        for (int k = 0; k < nl; ++k)
            if (S & (1u << k)) b = box_union(b, bin_box(ibox, leaf_boxes, leaf[k]));
        const uint32_t low = S & (0u - S);
        float best = 0.0f;
        uint32_t bp = 0;
        for (uint32_t P = (S - 1u) & S; P != 0; P = (P - 1u) & S) {
            if (!(P & low)) continue;
            float c = cost[P] + cost[S ^ P];
            if (bp == 0 || c < best) {
                best = c;
                bp = P;
            }
        }
        cost[S] = box_area(b) + best;
        split[S] = bp;
    }
    /* topology: subsets split from a stack, node numbers handed out in order */
    uint32_t stk_set[ORC_TREELET_LEAVES], stk_id[ORC_TREELET_LEAVES], order[ORC_TREELET_LEAVES];
    int sp = 0, no = 0, next = 0;
    stk_set[sp] = full;
    stk_id[sp++] = r;
    while (sp > 0) {
        --sp;
        uint32_t S = stk_set[sp], id = stk_id[sp];
        order[no++] = id;
        uint32_t part[2] = {split[S], S ^ split[S]};
        uint32_t ref[2];
        for (int h = 0; h < 2; ++h)
            ref[h] = (part[h] & (part[h] - 1u)) ? inner[next++] : leaf[__builtin_ctz(part[h])];
        t->lc[id] = ref[0];
        t->rc[id] = ref[1];
        for (int h = 1; h >= 0; --h)
            if (part[h] & (part[h] - 1u)) {
                stk_set[sp] = part[h];
                stk_id[sp++] = ref[h];
            }
    }
    /* boxes, areas and leaf counts, children before parents */
    for (int k = no - 1; k >= 0; --k) {
        uint32_t id = order[k], c[2] = {t->lc[id], t->rc[id]};
        ibox[id] = box_union(bin_box(ibox, leaf_boxes, c[0]), bin_box(ibox, leaf_boxes, c[1]));
        t->area[id] = box_area(ibox[id]);
        uint32_t cnt = 0;
        for (int h = 0; h < 2; ++h) cnt += (c[h] & ORC_LEAF_BIT) ? 1u : t->count[c[h]];
        t->count[id] = cnt;
    }
}

/* Passes of treelet restructuring of a build: five with
 * RTG_BUILD_PREFER_FAST_TRACE, none otherwise (the plain LBVH). */
static int treelet_rounds(uint32_t flags) { return (flags & RTG_BUILD_PREFER_FAST_TRACE) ? 5 : 0; }

/* Builds the wide tree over n primitive boxes, after `rounds` passes of
 * treelet restructuring. With `keep_lc`/`keep_rc` (n - 1 entries each) the
 * final binary tree is also returned (a test hook). */
static int wide_build(const box3* prim_boxes, uint32_t n, uint32_t leaf_max, int rounds, wtree* out,
                      uint32_t* keep_lc, uint32_t* keep_rc) {
    memset(out, 0, sizeof(*out));
    out->n = n;
    uint32_t cap = n >= 2 ? n - 1 : 1;
    out->order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    out->nodes = (orc_node*)calloc(cap, sizeof(orc_node));
    out->level_off = (uint32_t*)calloc(cap + 2, sizeof(uint32_t));
    box3* leaf_boxes = (box3*)malloc(sizeof(box3) * (n ? n : 1));
    if (!out->order || !out->nodes || !out->level_off || !leaf_boxes) {
        free(leaf_boxes);
        return RTG_ERROR_OUT_OF_MEMORY;
    }
    for (uint32_t i = 0; i < n; ++i) out->order[i] = i;

    if (n <= leaf_max || n <= 1) {
        /* one node with one leaf (or none) */
        box3 e = box_empty();
        for (int s = 0; s < 4; ++s) set_slot(&out->nodes[0], s, ORC_EMPTY, e);
        if (n > 0) {
            uint32_t ref = ORC_LEAF_BIT | ((n - 1) << 27);
            for (uint32_t k = 0; k < n; ++k) leaf_boxes[k] = prim_boxes[k];
            set_slot(&out->nodes[0], 0, ref, ref_box(out->nodes, ref, leaf_boxes));
        }
        out->node_count = 1;
        out->level_count = 1;
        out->level_off[0] = 0;
        out->level_off[1] = 1;
        out->root = node_box(&out->nodes[0]);
        free(leaf_boxes);
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
    for (uint32_t i = 0; i < n; ++i) keys[i] = orc_morton3(cen[3 * i], cen[3 * i + 1], cen[3 * i + 2], clo, chi);
    free(cen);
    radix_sort_pairs(keys, out->order, n);
    for (uint32_t k = 0; k < n; ++k) leaf_boxes[k] = prim_boxes[out->order[k]];

    /* binary radix tree and the boxes of its internal nodes */
    uint32_t nint = n - 1;
    bintree t;
    t.lc = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    t.rc = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    t.first = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    t.count = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    t.area = (float*)malloc(sizeof(float) * nint);
    uint32_t* parent_int = (uint32_t*)malloc(sizeof(uint32_t) * nint);
    uint32_t* parent_leaf = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < nint; ++i) {
        karras_node(keys, (int)n, (int)i, &t.lc[i], &t.rc[i], &t.first[i], &t.count[i]);
        uint32_t c[2] = {t.lc[i], t.rc[i]};
        for (int s = 0; s < 2; ++s) {
            if (c[s] & ORC_LEAF_BIT) parent_leaf[c[s] & ~ORC_LEAF_BIT] = i;
            else parent_int[c[s]] = i;
        }
    }
    parent_int[0] = 0xffffffffu;
    free(keys);
    /* bottom-up: a node is finished when both children are */
    box3* ibox = (box3*)malloc(sizeof(box3) * nint);
    uint8_t* arrivals = (uint8_t*)calloc(nint, 1);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t node = parent_leaf[k];
        while (node != 0xffffffffu) {
            if (arrivals[node]++ == 0) break; /* the first child to arrive stops */
            uint32_t c[2] = {t.lc[node], t.rc[node]};
            box3 cb[2];
            for (int s = 0; s < 2; ++s)
                cb[s] = (c[s] & ORC_LEAF_BIT) ? leaf_boxes[c[s] & ~ORC_LEAF_BIT] : ibox[c[s]];
            ibox[node] = box_union(cb[0], cb[1]);
            t.area[node] = box_area(ibox[node]);
            node = parent_int[node];
        }
    }
    {
        for (int round = 0; round < rounds; ++round) {
            memset(arrivals, 0, nint);
            for (uint32_t k = 0; k < n; ++k) {
                uint32_t node = parent_leaf[k];
                while (node != 0xffffffffu) {
                    if (arrivals[node]++ == 0) break;
                    treelet_optimize(&t, ibox, leaf_boxes, node);
                    node = parent_int[node];
                }
            }
            /* parents of the restructured tree */
            for (uint32_t i = 0; i < nint; ++i) {
                uint32_t c[2] = {t.lc[i], t.rc[i]};
                for (int s2 = 0; s2 < 2; ++s2) {
                    if (c[s2] & ORC_LEAF_BIT) parent_leaf[c[s2] & ~ORC_LEAF_BIT] = i;
                    else parent_int[c[s2]] = i;
                }
            }
        }
    }
    free(ibox);
    free(arrivals);
    free(parent_int);
    free(parent_leaf);
    if (keep_lc && keep_rc) {
        memcpy(keep_lc, t.lc, sizeof(uint32_t) * nint);
        memcpy(keep_rc, t.rc, sizeof(uint32_t) * nint);
    }

    /* top-down collapse, breadth first */
    uint32_t* root_bin = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    root_bin[0] = 0;
    uint32_t lb = 0, le = 1, levels = 0;
    while (lb < le) {
        out->level_off[levels++] = lb;
        uint32_t next = le;
        for (uint32_t w = lb; w < le; ++w) {
            uint32_t s[4];
            int m = expand_slots(&t, root_bin[w], leaf_max, s);
            for (int k = 0; k < 4; ++k) {
                uint32_t ref = ORC_EMPTY;
                if (k < m) {
                    uint32_t c = s[k];
                    if (c & ORC_LEAF_BIT) {
                        ref = c;
                    } else if (t.count[c] <= leaf_max) {
                        ref = ORC_LEAF_BIT | ((t.count[c] - 1) << 27) | t.first[c];
                    } else {
                        ref = next;
                        root_bin[next++] = c;
                    }
                }
                out->nodes[w].child[k] = ref;
            }
        }
        lb = le;
        le = next;
    }
    out->level_off[levels] = le;
    out->level_count = levels;
    out->node_count = le;
    free(root_bin);
    free(t.lc);
    free(t.rc);
    free(t.first);
    free(t.count);
    free(t.area);

    /* slot boxes, deepest level first */
    for (uint32_t L = levels; L-- > 0;)
        for (uint32_t w = out->level_off[L]; w < out->level_off[L + 1]; ++w)
            for (int k = 0; k < 4; ++k) {
                uint32_t c = out->nodes[w].child[k];
                set_slot(&out->nodes[w], k, c, ref_box(out->nodes, c, leaf_boxes));
            }
    out->root = node_box(&out->nodes[0]);
    free(leaf_boxes);
    return RTG_OK;
}

static void write_header(orc_header* h, uint32_t kind, const wtree* w, uint32_t geom_count) {
    memset(h, 0, sizeof(*h));
    h->magic = ORC_MAGIC;
    h->kind = kind;
    h->node_count = w->node_count;
    h->prim_count = w->n;
    h->geom_count = geom_count;
    h->max_depth = w->level_count;
    h->node_width = 4;
    for (int a = 0; a < 3; ++a) {
        h->lo[a] = w->root.lo[a];
        h->hi[a] = w->root.hi[a];
    }
}

static v3 load_vertex(const rtg_geometry* g, uint32_t idx) {
    const float* p = (const float*)((const uint8_t*)(uintptr_t)g->vertex_data + (uint64_t)idx * g->vertex_stride);
    return V3(p[0], p[1], p[2]);
}

int orc_blas_build(const rtg_geometry* geoms, uint32_t geom_count, uint32_t flags, void** out_blob,
                   uint64_t* out_size) {
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
    wtree w;
    int st = wide_build(boxes, n, ORC_BLAS_LEAF_MAX, treelet_rounds(flags), &w, NULL, NULL);
    if (st) {
        free(boxes);
        free(tris);
        wtree_free(&w);
        return st;
    }
    uint64_t nodes_off = 128;
    uint64_t prims_off = orc_align(nodes_off + (uint64_t)w.node_count * sizeof(orc_node), 64);
    uint64_t extra_off = orc_align(prims_off + (uint64_t)n * 48, 64);
    uint64_t size = orc_align(extra_off + (uint64_t)geom_count * 4, 64);
    uint8_t* blob = (uint8_t*)calloc(1, size);
    orc_header* h = (orc_header*)blob;
    write_header(h, ORC_KIND_BLAS, &w, geom_count);
    h->nodes_off = nodes_off;
    h->prims_off = prims_off;
    h->extra_off = extra_off;
    h->total_size = size;
    memcpy(blob + nodes_off, w.nodes, (size_t)w.node_count * sizeof(orc_node));
    orc_tri* dst = (orc_tri*)(blob + prims_off);
    for (uint32_t k = 0; k < n; ++k) dst[k] = tris[w.order[k]];
    uint32_t* gflags = (uint32_t*)(blob + extra_off);
    for (uint32_t g = 0; g < geom_count; ++g) gflags[g] = geoms[g].flags;
    wtree_free(&w);
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

int orc_tlas_build(const rtg_instance* inst, uint32_t count, uint32_t flags, void** out_blob, uint64_t* out_size) {
    if (count >= (1u << 27)) return RTG_ERROR_LIMIT;
    box3* boxes = (box3*)calloc(count ? count : 1, sizeof(box3));
    for (uint32_t i = 0; i < count; ++i) boxes[i] = instance_box(&inst[i]);
    wtree w;
    int st = wide_build(boxes, count, ORC_TLAS_LEAF_MAX, treelet_rounds(flags), &w, NULL, NULL);
    free(boxes);
    if (st) {
        wtree_free(&w);
        return st;
    }
    uint64_t nodes_off = 128;
    uint64_t prims_off = orc_align(nodes_off + (uint64_t)w.node_count * sizeof(orc_node), 64);
    uint64_t extra_off = orc_align(prims_off + (uint64_t)count * 64, 64);
    uint64_t level_off = orc_align(extra_off + (uint64_t)count * 64, 64);
    uint64_t leafbox_off = orc_align(level_off + 4ull * (2 + w.level_count), 64);
    uint64_t size = orc_align(leafbox_off + (uint64_t)count * 24, 64);
    uint8_t* blob = (uint8_t*)calloc(1, size);
    orc_header* h = (orc_header*)blob;
    write_header(h, ORC_KIND_TLAS, &w, 0);
    h->nodes_off = nodes_off;
    h->prims_off = prims_off;
    h->extra_off = extra_off;
    h->level_off = level_off;
    h->leafbox_off = leafbox_off;
    h->total_size = size;
    memcpy(blob + nodes_off, w.nodes, (size_t)w.node_count * sizeof(orc_node));
    orc_inst* recs = (orc_inst*)(blob + prims_off);
    box3* leafbox = (box3*)(blob + leafbox_off);
    for (uint32_t k = 0; k < count; ++k) {
        fill_inst_record(&recs[k], &inst[w.order[k]], w.order[k]);
        leafbox[k] = instance_box(&inst[w.order[k]]);
    }
    memcpy(blob + extra_off, inst, (size_t)count * 64);
    uint32_t* lv = (uint32_t*)(blob + level_off);
    lv[0] = w.level_count;
    for (uint32_t L = 0; L <= w.level_count; ++L) lv[1 + L] = w.level_off[L];
    wtree_free(&w);
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
    /* same topology: slot boxes again, deepest level first */
    const uint32_t* lv = (const uint32_t*)(blob + h->level_off);
    uint32_t levels = lv[0];
    const uint32_t* offs = lv + 1;
    for (uint32_t L = levels; L-- > 0;)
        for (uint32_t q = offs[L]; q < offs[L + 1]; ++q)
            for (int s = 0; s < 4; ++s) {
                uint32_t c = nodes[q].child[s];
                set_slot(&nodes[q], s, c, ref_box(nodes, c, leaf));
            }
    box3 root = node_box(&nodes[0]);
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

int orc_debug_binary_tree(const float* boxes, uint32_t n, uint32_t flags, uint32_t* lc, uint32_t* rc,
                          uint32_t* order) {
    if (n < 2) return RTG_ERROR_INVALID_ARGUMENT;
    box3* b = (box3*)malloc(sizeof(box3) * n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            b[i].lo[a] = boxes[6 * i + a];
            b[i].hi[a] = boxes[6 * i + 3 + a];
        }
    wtree w;
    int st = wide_build(b, n, 1u, treelet_rounds(flags), &w, lc, rc);
    if (st == RTG_OK) memcpy(order, w.order, sizeof(uint32_t) * n);
    wtree_free(&w);
    free(b);
    return st;
}
