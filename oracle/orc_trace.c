/*
 * orc_trace.c -- CPU oracle: traversal, the reference's shader stages and
 * the per-pixel trace step (TEST INFRASTRUCTURE; never part of the product).
 *
 * Stage semantics restated here, each from the file:line named at the
 * function. Operation order follows the reference's source expressions,
 * left to right, one rounding per written operation (no FMA).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "orc.h"
#include "orc_layout.h"
#include "orc_math.h"

#define STACK_MAX 512u
#define SENTINEL 0xFFFFFFFEu
/* 1 + 2*gamma(3) rounded up: far-plane inflation of the robust slab test
 * (Ize 2013, "Robust BVH Ray Traversal", JCGT 2(2)). */
#define FAR_SCALE_BITS 0x3f800003u

typedef struct {
    float t, u, v;
    uint32_t inst, prim, geom;
} hitrec;

typedef struct {
    const orc_header* th;
    const orc_node* tnodes;
    const orc_inst* tinst;
    const rtg_instance* tsrc;
    const rtg_model_info* models;
    uint32_t model_count;
    const rtg_texture* textures;
    uint32_t texture_count;
    float lut[256];
    uint32_t ray_flags, cull_mask, sbt_offset;
    rtg_trace_counters cnt;
    /* test hook (orc_debug_visits): when set, every node visit, triangle
       test and instance entry appends (kind, instance record or 0, record) */
    uint64_t* vlog;
    uint32_t vlog_n, vlog_cap;
} qctx;

static void visit_log(qctx* q, uint32_t kind, const void* inst, const void* rec) {
    if (!q->vlog || q->vlog_n >= q->vlog_cap) return;
    uint64_t* e = q->vlog + 3u * q->vlog_n++;
    e[0] = kind;
    e[1] = (uint64_t)(uintptr_t)inst;
    e[2] = (uint64_t)(uintptr_t)rec;
}

/* ---------------------------------------------------------------- */
/* Textures: the sampler2D lookups of the hit shaders (LOD 0, REPEAT). */
/* ---------------------------------------------------------------- */

void orc_srgb_lut(float lut[256]) {
    /* IEC 61966-2-1 decode, evaluated in double, rounded once to float. */
    for (int i = 0; i < 256; ++i) {
        double c = i / 255.0;
        double l = c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4);
        lut[i] = (float)l;
    }
}

static int wrap_coord(float f, uint32_t size) {
    if (!(fabsf(f) < 16777216.0f)) return 0; /* non-finite or huge: texel 0 */
    long long i = (long long)f;
    long long m = i % (long long)size;
    if (m < 0) m += size;
    return (int)m;
}

static void fetch_texel(const rtg_texture* t, const float* lut, int x, int y, float out[4]) {
    uint64_t idx = (uint64_t)y * t->width + (uint64_t)x;
    if (t->format == RTG_FORMAT_RGBA32_FLOAT) {
        const float* p = (const float*)(uintptr_t)t->data + 4 * idx;
        out[0] = p[0]; out[1] = p[1]; out[2] = p[2]; out[3] = p[3];
        return;
    }
    const uint8_t* p = (const uint8_t*)(uintptr_t)t->data + 4 * idx;
    if (t->format == RTG_FORMAT_RGBA8_SRGB) {
        out[0] = lut[p[0]]; out[1] = lut[p[1]]; out[2] = lut[p[2]];
    } else {
        out[0] = (float)p[0] / 255.0f; out[1] = (float)p[1] / 255.0f; out[2] = (float)p[2] / 255.0f;
    }
    out[3] = (float)p[3] / 255.0f;
}

/* texture(textures[i], uv) with the image's own sampler: NEAREST takes texel
 * floor(uv * size); LINEAR weights the four texels around uv * size - 0.5. */
static void sample_tex(const rtg_texture* t, const float* lut, float u, float v, float out[4]) {
    if (t->width == 0 || t->height == 0) {
        out[0] = out[1] = out[2] = 0.0f; out[3] = 1.0f;
        return;
    }
    if (t->filter == RTG_FILTER_NEAREST) {
        int x = wrap_coord(floorf(u * (float)t->width), t->width);
        int y = wrap_coord(floorf(v * (float)t->height), t->height);
        fetch_texel(t, lut, x, y, out);
        return;
    }
    float x = u * (float)t->width - 0.5f;
    float y = v * (float)t->height - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = wrap_coord(fx, t->width), y0 = wrap_coord(fy, t->height);
    int x1 = x0 + 1 == (int)t->width ? 0 : x0 + 1;
    int y1 = y0 + 1 == (int)t->height ? 0 : y0 + 1;
    float t00[4], t10[4], t01[4], t11[4];
    fetch_texel(t, lut, x0, y0, t00);
    fetch_texel(t, lut, x1, y0, t10);
    fetch_texel(t, lut, x0, y1, t01);
    fetch_texel(t, lut, x1, y1, t11);
    float ia = 1.0f - a, ib = 1.0f - b;
    for (int c = 0; c < 4; ++c) out[c] = (t00[c] * ia + t10[c] * a) * ib + (t01[c] * ia + t11[c] * a) * b;
}

static void sample_index(const qctx* q, uint32_t index, float u, float v, float out[4]) {
    if (index >= q->texture_count) {
        out[0] = out[1] = out[2] = 0.0f; out[3] = 1.0f;
        return;
    }
    sample_tex(&q->textures[index], q->lut, u, v, out);
}

void orc_sample_texture(const rtg_texture* tex, const float lut[256], float u, float v, float out[4]) {
    sample_tex(tex, lut, u, v, out);
}

/* ---------------------------------------------------------------- */
/* Ray/triangle: watertight test (Woop, Benthin, Wald 2013, JCGT 2(1)) */
/* ---------------------------------------------------------------- */

typedef struct {
    v3 o, d, inv;
    int kx, ky, kz;
    float sx, sy, sz;
} oray;

static void oray_setup(oray* r, v3 o, v3 d) {
    r->o = o;
    r->d = d;
    r->inv = V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dv[3] = {d.x, d.y, d.z};
    if (dv[kz] < 0.0f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r->kx = kx;
    r->ky = ky;
    r->kz = kz;
    /* sz = 1/d[kz] (the inverse already computed), sx = d[kx]*sz, sy = d[ky]*sz */
    float iv[3] = {r->inv.x, r->inv.y, r->inv.z};
    r->sz = iv[kz];
    r->sx = dv[kx] * r->sz;
    r->sy = dv[ky] * r->sz;
}

/* Returns 1 on a hit with tmin <= t <= tmax; (u, v) weight vertices 1, 2. */
static int tri_test(const oray* r, const float* p0, const float* p1, const float* p2, float tmin, float tmax,
                    float* t_out, float* u_out, float* v_out) {
    float A[3] = {p0[0] - r->o.x, p0[1] - r->o.y, p0[2] - r->o.z};
    float B[3] = {p1[0] - r->o.x, p1[1] - r->o.y, p1[2] - r->o.z};
    float C[3] = {p2[0] - r->o.x, p2[1] - r->o.y, p2[2] - r->o.z};
    float Ax = A[r->kx] - r->sx * A[r->kz], Ay = A[r->ky] - r->sy * A[r->kz];
    float Bx = B[r->kx] - r->sx * B[r->kz], By = B[r->ky] - r->sy * B[r->kz];
    float Cx = C[r->kx] - r->sx * C[r->kz], Cy = C[r->ky] - r->sy * C[r->kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = (U + V) + W;
    if (det == 0.0f) return 0;
    float Az = r->sz * A[r->kz], Bz = r->sz * B[r->kz], Cz = r->sz * C[r->kz];
    float T = (U * Az + V * Bz) + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return 0;
    *t_out = t;
    *u_out = V / det;
    *v_out = W / det;
    return 1;
}

void orc_watertight(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                    float tmin, float tmax, int* hit, float* t, float* u, float* v) {
    oray r;
    oray_setup(&r, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]));
    *hit = tri_test(&r, v0, v1, v2, tmin, tmax, t, u, v);
}

/* Robust slab test of child s of a node; returns 1 on overlap of
 * [tmin, tcur]; *tn receives the entry distance. */
static int box_test(const orc_node* nd, int s, const oray* r, float tmin, float tcur, float* tn) {
    /* near/far planes by the sign bit of the inverse direction (1/-0 = -inf
       selects like a negative direction; for a NaN component both planes
       give NaN, so the choice does not matter) */
    float nx = signbit(r->inv.x) ? nd->hi_x[s] : nd->lo_x[s];
    float fx = signbit(r->inv.x) ? nd->lo_x[s] : nd->hi_x[s];
    float ny = signbit(r->inv.y) ? nd->hi_y[s] : nd->lo_y[s];
    float fy = signbit(r->inv.y) ? nd->lo_y[s] : nd->hi_y[s];
    float nz = signbit(r->inv.z) ? nd->hi_z[s] : nd->lo_z[s];
    float fz = signbit(r->inv.z) ? nd->lo_z[s] : nd->hi_z[s];
    float tnx = (nx - r->o.x) * r->inv.x, tfx = (fx - r->o.x) * r->inv.x;
    float tny = (ny - r->o.y) * r->inv.y, tfy = (fy - r->o.y) * r->inv.y;
    float tnz = (nz - r->o.z) * r->inv.z, tfz = (fz - r->o.z) * r->inv.z;
    float tnear = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
    float tfar = fminf(fminf(fminf(tfx, tfy), tfz) * u2f(FAR_SCALE_BITS), tcur);
    *tn = tnear;
    return tnear <= tfar;
}

static int is_opaque(uint32_t geom_word, uint32_t inst_flags, uint32_t ray_flags) {
    int opaque = (geom_word & ORC_NONOPAQUE_BIT) == 0;
    if (inst_flags & RTG_INSTANCE_FORCE_OPAQUE) opaque = 1;
    else if (inst_flags & RTG_INSTANCE_FORCE_NO_OPAQUE) opaque = 0;
    if (ray_flags & RTG_RAY_FLAG_OPAQUE) opaque = 1;
    else if (ray_flags & RTG_RAY_FLAG_NO_OPAQUE) opaque = 0;
    return opaque;
}

/* any_hit_alpha_clip (shaders/any_hit_alpha_clip.glsl:11-29): returns 1 if
 * the candidate is kept (alpha >= 0.5). */
static int alpha_clip_keep(qctx* q, uint64_t* anyhit_counter, uint32_t inst_index, uint32_t geom, uint32_t prim,
                           float u, float v) {
    ++*anyhit_counter;
    if (!q->models) return 1;
    uint32_t custom = q->tsrc[inst_index].custom_index_and_mask & 0xffffffu;
    if (custom >= q->model_count) return 1;
    const rtg_model_info* mi = &q->models[custom];
    const rtg_geometry_info* gi = (const rtg_geometry_info*)(uintptr_t)mi->geometry_info_address + geom;
    const uint32_t* idx = (const uint32_t*)(uintptr_t)gi->index_buffer_address + 3ull * prim;
    const float* uvs = (const float*)(uintptr_t)mi->uv_buffer_address;
    v3 bc = V3((1.0f - u) - v, u, v);
    v2 a = {uvs[2 * idx[0]], uvs[2 * idx[0] + 1]};
    v2 b = {uvs[2 * idx[1]], uvs[2 * idx[1] + 1]};
    v2 c = {uvs[2 * idx[2]], uvs[2 * idx[2] + 1]};
    v2 uv = v2_interp(a, b, c, bc);
    float texel[4];
    sample_index(q, gi->diffuse_image_index, uv.x, uv.y, texel);
    return !(texel[3] < 0.5f);
}

static int key_less(uint32_t ia, uint32_t ga, uint32_t pa, uint32_t ib, uint32_t gb, uint32_t pb) {
    if (ia != ib) return ia < ib;
    if (ga != gb) return ga < gb;
    return pa < pb;
}

/* closest-hit stack entries the device keeps in LDS per lane */
static uint32_t g_stack_lds = 16;
void orc_set_stack_lds(uint32_t entries) { g_stack_lds = entries; }
/* shadow rays of an image in pairs (1) or one by one (0) */
static int g_shadow_pairs = 1;
void orc_set_shadow_pairs(int on) { g_shadow_pairs = on; }

/* Two-level traversal. Returns 1 if an accepted hit was found. */
static int traverse(qctx* q, v3 o, v3 d, float tmin, float tmax, int occlusion, hitrec* best) {
    uint32_t stack_ref[STACK_MAX];
    float stack_tn[STACK_MAX];
    uint32_t sp = 0;
    oray world, obj;
    oray_setup(&world, o, d);
    const oray* cr = &world;
    const orc_node* nodes = q->tnodes;
    const orc_tri* tris = NULL;
    const orc_inst* cur_inst = NULL;
    int in_blas = 0;
    float tcur = tmax;
    int found = 0;
    best->t = tmax;
    best->u = best->v = 0.0f;
    best->inst = best->prim = best->geom = RTG_NO_HIT;
    /* counters of this query kind: nodes, triangles, instances, any-hits */
    uint64_t* kc = occlusion ? &q->cnt.occ_node_visits : &q->cnt.node_visits;
    /* pushes at stack positions the device keeps in its per-lane spill area
       (beyond g_stack_lds closest-hit or twice as many occlusion entries in
       LDS); the push order, hence every position, is the device's */
    uint64_t* ksp = occlusion ? &q->cnt.occ_spill_pushes : &q->cnt.spill_pushes;
    const uint32_t lds_depth = occlusion ? 2u * g_stack_lds : g_stack_lds;
    uint32_t cur = 0;
    int have = 1;
    while (1) {
        if (have) {
            if (!(cur & ORC_LEAF_BIT)) {
                const orc_node* nd = &nodes[cur];
                kc[0]++;
                visit_log(q, 0, in_blas ? cur_inst : NULL, nd);
                /* the overlapped slots in ascending (entry distance, slot)
                   order for closest-hit queries, in slot order for
                   occlusion queries (any hit ends them, so no order is
                   needed for the result): the first is visited next, the
                   others are pushed so that the second is on top */
                uint32_t ref[4];
                float tn[4];
                int nh = 0;
                for (int s = 0; s < 4; ++s) {
                    float t = 0.0f;
                    if (nd->child[s] == ORC_EMPTY || !box_test(nd, s, cr, tmin, tcur, &t)) continue;
                    int k = nh++;
                    while (!occlusion && k > 0 && t < tn[k - 1]) {
                        ref[k] = ref[k - 1];
                        tn[k] = tn[k - 1];
                        --k;
                    }
                    ref[k] = nd->child[s];
                    tn[k] = t;
                }
                if (nh == 0) {
                    have = 0;
                    continue;
                }
                for (int k = nh - 1; k >= 1; --k) {
                    if (sp >= STACK_MAX) return -1;
                    if (sp >= lds_depth) ++*ksp;
                    stack_ref[sp] = ref[k];
                    stack_tn[sp] = tn[k];
                    ++sp;
                }
                cur = ref[0];
                continue;
            }
            if (!in_blas) {
                const orc_inst* rec = &q->tinst[cur & 0x07ffffffu];
                have = 0;
                if (((rec->mask_flags & 0xffu) & q->cull_mask) == 0) continue;
                const orc_header* bh = (const orc_header*)(uintptr_t)rec->blas;
                if (!bh) continue;
                kc[2]++;
                visit_log(q, 2, NULL, rec);
                if (sp >= STACK_MAX) return -1;
                if (sp >= lds_depth) ++*ksp;
                stack_ref[sp] = SENTINEL;
                stack_tn[sp] = -INFINITY;
                ++sp;
                oray_setup(&obj, m34_point(rec->w2o, o), m34_dir(rec->w2o, d));
                cr = &obj;
                nodes = (const orc_node*)((const uint8_t*)bh + bh->nodes_off);
                tris = (const orc_tri*)((const uint8_t*)bh + bh->prims_off);
                cur_inst = rec;
                in_blas = 1;
                cur = 0;
                have = 1;
                continue;
            }
            uint32_t first = cur & 0x07ffffffu;
            uint32_t count = ((cur >> 27) & 0xfu) + 1u;
            for (uint32_t k = first; k < first + count; ++k) {
                const orc_tri* tr = &tris[k];
                kc[1]++;
                visit_log(q, 1, cur_inst, tr);
                float t, u, v;
                if (!tri_test(cr, tr->v0, tr->v1, tr->v2, tmin, tcur, &t, &u, &v)) continue;
                uint32_t geom = tr->geom & ~ORC_NONOPAQUE_BIT;
                uint32_t inst = cur_inst->index;
                if (!occlusion &&
                    !(t < best->t || (t == best->t && key_less(inst, geom, tr->prim, best->inst, best->geom, best->prim))))
                    continue;
                uint32_t iflags = (cur_inst->mask_flags >> 8) & 0xffu;
                if (!is_opaque(tr->geom, iflags, q->ray_flags)) {
                    uint32_t group = (q->tsrc[inst].sbt_offset_and_flags & 0xffffffu) + q->sbt_offset;
                    if (group < RTG_HIT_GROUP_COUNT && !alpha_clip_keep(q, &kc[3], inst, geom, tr->prim, u, v)) continue;
                }
                found = 1;
                best->t = t;
                best->u = u;
                best->v = v;
                best->inst = inst;
                best->geom = geom;
                best->prim = tr->prim;
                if (occlusion || (q->ray_flags & RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT)) return 1;
                tcur = t;
            }
            have = 0;
            continue;
        }
        /* pop */
        while (sp > 0) {
            --sp;
            uint32_t r = stack_ref[sp];
            if (r == SENTINEL) {
                in_blas = 0;
                cr = &world;
                nodes = q->tnodes;
                continue;
            }
            if (stack_tn[sp] > tcur) continue;
            cur = r;
            have = 1;
            break;
        }
        if (!have) break;
    }
    return found;
}

/* Two shadow rays of one path (same origin and interval) walked together:
 * the union of the nodes either ray overlaps, each visited once, stack
 * entries carrying the mask of the rays that overlap them (8-byte entries on
 * the device, so g_stack_lds of them fit in LDS). Children are taken in slot
 * order (the lowest overlapped one next, the others pushed highest first), a
 * ray leaves the mask once it has an accepted hit, and the walk ends when
 * both have one or the stack is empty. Per ray the result is the single-ray
 * occlusion query's (every node and triangle it overlaps is still tested,
 * and a hit is accepted per ray); the counters count visits of the pair.
 * m = 3 for a pair, 1 for ray 0 alone. Returns the mask of occluded rays,
 * or -1 if the stack overflowed. The shadow rays' flags and hit-group
 * offset are q's. */
static int traverse_pair(qctx* q, v3 o, v3 d0, v3 d1, uint32_t m, float tmin, float tmax) {
    uint32_t stack_ref[STACK_MAX];
    uint32_t stack_mask[STACK_MAX];
    uint32_t sp = 0;
    oray world[2], obj[2];
    oray_setup(&world[0], o, d0);
    oray_setup(&world[1], o, d1);
    const oray* cr = world;
    const orc_node* nodes = q->tnodes;
    const orc_tri* tris = NULL;
    const orc_inst* cur_inst = NULL;
    int in_blas = 0;
    uint64_t* kc = &q->cnt.occ_node_visits;
    uint32_t alive = m, mask = m, cur = 0;
    int have = 1;
    while (1) {
        if (have) {
            if (!(cur & ORC_LEAF_BIT)) {
                const orc_node* nd = &nodes[cur];
                kc[0]++;
                visit_log(q, 0, in_blas ? cur_inst : NULL, nd);
                uint32_t ref[4], msk[4];
                int nh = 0;
                for (int s = 0; s < 4; ++s) {
                    if (nd->child[s] == ORC_EMPTY) continue;
                    uint32_t c = 0;
                    for (int r = 0; r < 2; ++r) {
                        float t;
                        if ((mask >> r) & 1u && box_test(nd, s, &cr[r], tmin, tmax, &t)) c |= 1u << r;
                    }
                    if (!c) continue;
                    ref[nh] = nd->child[s];
                    msk[nh] = c;
                    ++nh;
                }
                if (nh == 0) {
                    have = 0;
                    continue;
                }
                for (int k = nh - 1; k >= 1; --k) {
                    if (sp >= STACK_MAX) return -1;
                    if (sp >= g_stack_lds) q->cnt.occ_spill_pushes++;
                    stack_ref[sp] = ref[k];
                    stack_mask[sp] = msk[k];
                    ++sp;
                }
                cur = ref[0];
                mask = msk[0];
                continue;
            }
            if (!in_blas) {
                const orc_inst* rec = &q->tinst[cur & 0x07ffffffu];
                have = 0;
                if (((rec->mask_flags & 0xffu) & q->cull_mask) == 0) continue;
                const orc_header* bh = (const orc_header*)(uintptr_t)rec->blas;
                if (!bh) continue;
                kc[2]++;
                visit_log(q, 2, NULL, rec);
                if (sp >= STACK_MAX) return -1;
                if (sp >= g_stack_lds) q->cnt.occ_spill_pushes++;
                stack_ref[sp] = SENTINEL;
                stack_mask[sp] = 0;
                ++sp;
                oray_setup(&obj[0], m34_point(rec->w2o, o), m34_dir(rec->w2o, d0));
                oray_setup(&obj[1], m34_point(rec->w2o, o), m34_dir(rec->w2o, d1));
                cr = obj;
                nodes = (const orc_node*)((const uint8_t*)bh + bh->nodes_off);
                tris = (const orc_tri*)((const uint8_t*)bh + bh->prims_off);
                cur_inst = rec;
                in_blas = 1;
                cur = 0;
                have = 1;
                continue;
            }
            uint32_t first = cur & 0x07ffffffu;
            uint32_t count = ((cur >> 27) & 0xfu) + 1u;
            for (uint32_t k = first; k < first + count; ++k) {
                const orc_tri* tr = &tris[k];
                kc[1]++;
                visit_log(q, 1, cur_inst, tr);
                for (int r = 0; r < 2; ++r) {
                    if (!((mask & alive) >> r & 1u)) continue;
                    float t, u, v;
                    if (!tri_test(&cr[r], tr->v0, tr->v1, tr->v2, tmin, tmax, &t, &u, &v)) continue;
                    uint32_t geom = tr->geom & ~ORC_NONOPAQUE_BIT;
                    uint32_t inst = cur_inst->index;
                    uint32_t iflags = (cur_inst->mask_flags >> 8) & 0xffu;
                    if (!is_opaque(tr->geom, iflags, q->ray_flags)) {
                        uint32_t group = (q->tsrc[inst].sbt_offset_and_flags & 0xffffffu) + q->sbt_offset;
                        if (group < RTG_HIT_GROUP_COUNT && !alpha_clip_keep(q, &kc[3], inst, geom, tr->prim, u, v))
                            continue;
                    }
                    alive &= ~(1u << r);
                }
                if (alive == 0) return (int)m;
            }
            have = 0;
            continue;
        }
        /* pop: a sentinel leaves the BLAS; entries no live ray overlaps are
           dropped */
        while (sp > 0) {
            --sp;
            uint32_t r = stack_ref[sp];
            if (r == SENTINEL) {
                in_blas = 0;
                cr = world;
                nodes = q->tnodes;
                continue;
            }
            if ((stack_mask[sp] & alive) == 0) continue;
            cur = r;
            mask = stack_mask[sp] & alive;
            have = 1;
            break;
        }
        if (!have) break;
    }
    return (int)(m & ~alive);
}

static int qctx_init(qctx* q, const void* tlas, const rtg_scene_bindings* scene) {
    memset(q, 0, sizeof(*q));
    const orc_header* h = (const orc_header*)tlas;
    if (!h || h->magic != ORC_MAGIC || h->kind != ORC_KIND_TLAS) return RTG_ERROR_INVALID_ARGUMENT;
    q->th = h;
    q->tnodes = (const orc_node*)((const uint8_t*)tlas + h->nodes_off);
    q->tinst = (const orc_inst*)((const uint8_t*)tlas + h->prims_off);
    q->tsrc = (const rtg_instance*)((const uint8_t*)tlas + h->extra_off);
    if (scene) {
        q->models = (const rtg_model_info*)(uintptr_t)scene->model_info;
        q->model_count = scene->model_count;
        q->textures = (const rtg_texture*)(uintptr_t)scene->textures;
        q->texture_count = scene->texture_count;
    }
    orc_srgb_lut(q->lut);
    q->cull_mask = 0xffu;
    return RTG_OK;
}

static void add_counters(rtg_trace_counters* dst, const rtg_trace_counters* s) {
    dst->primary_rays += s->primary_rays;
    dst->bounce_rays += s->bounce_rays;
    dst->shadow_rays += s->shadow_rays;
    dst->node_visits += s->node_visits;
    dst->triangle_tests += s->triangle_tests;
    dst->instance_visits += s->instance_visits;
    dst->anyhit_calls += s->anyhit_calls;
    dst->occ_node_visits += s->occ_node_visits;
    dst->occ_triangle_tests += s->occ_triangle_tests;
    dst->occ_instance_visits += s->occ_instance_visits;
    dst->occ_anyhit_calls += s->occ_anyhit_calls;
    dst->spill_pushes += s->spill_pushes;
    dst->occ_spill_pushes += s->occ_spill_pushes;
    dst->textured_hits += s->textured_hits;
}

/* ---------------------------------------------------------------- */
/* Thread pool over work items                                       */
/* ---------------------------------------------------------------- */
typedef struct {
    void (*fn)(void* arg, qctx* q, uint32_t item);
    void* arg;
    uint32_t items;
    volatile uint32_t next;
    pthread_mutex_t mu;
    const void* tlas;
    const rtg_scene_bindings* scene;
    uint32_t ray_flags, cull_mask, sbt_offset;
    rtg_trace_counters total;
    int status;
} pool;

static void* pool_worker(void* p_) {
    pool* p = (pool*)p_;
    qctx* q = (qctx*)malloc(sizeof(qctx));
    if (qctx_init(q, p->tlas, p->scene) != RTG_OK) {
        free(q);
        return NULL;
    }
    q->ray_flags = p->ray_flags;
    q->cull_mask = p->cull_mask;
    q->sbt_offset = p->sbt_offset;
    while (1) {
        uint32_t it = __sync_fetch_and_add(&p->next, 1u);
        if (it >= p->items) break;
        p->fn(p->arg, q, it);
    }
    pthread_mutex_lock(&p->mu);
    add_counters(&p->total, &q->cnt);
    pthread_mutex_unlock(&p->mu);
    free(q);
    return NULL;
}

static int run_pool(pool* p, int threads) {
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > p->items) threads = p->items ? (int)p->items : 1;
    pthread_mutex_init(&p->mu, NULL);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 1; i < threads; ++i) pthread_create(&th[i], NULL, pool_worker, p);
    pool_worker(p);
    for (int i = 1; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&p->mu);
    return RTG_OK;
}

/* ---------------------------------------------------------------- */
/* Ray queries                                                       */
/* ---------------------------------------------------------------- */
typedef struct {
    const rtg_ray* rays;
    uint32_t count, mode;
    rtg_hit* hits;
    uint32_t* occluded;
} rays_job;

#define RAY_CHUNK 256u

static void rays_item(void* a, qctx* q, uint32_t item) {
    rays_job* j = (rays_job*)a;
    uint32_t b = item * RAY_CHUNK, e = b + RAY_CHUNK < j->count ? b + RAY_CHUNK : j->count;
    for (uint32_t i = b; i < e; ++i) {
        const rtg_ray* r = &j->rays[i];
        hitrec h;
        int occl = j->mode == RTG_QUERY_OCCLUSION;
        int f = traverse(q, V3(r->origin[0], r->origin[1], r->origin[2]),
                         V3(r->direction[0], r->direction[1], r->direction[2]), r->tmin, r->tmax, occl, &h);
        if (occl) {
            j->occluded[i] = f > 0 ? 1u : 0u;
        } else {
            rtg_hit* o = &j->hits[i];
            o->t = h.t;
            o->instance = h.inst;
            o->primitive = h.prim;
            o->geometry = h.geom;
            o->u = h.u;
            o->v = h.v;
        }
    }
}

int orc_trace_rays(const void* tlas, const rtg_scene_bindings* scene, const rtg_ray* rays, uint32_t count,
                   uint32_t mode, uint32_t ray_flags, uint32_t cull_mask, uint32_t sbt_offset, rtg_hit* hits,
                   uint32_t* occluded, rtg_trace_counters* counters, int threads) {
    if (mode > RTG_QUERY_OCCLUSION) return RTG_ERROR_INVALID_ARGUMENT;
    if (mode == RTG_QUERY_CLOSEST && !hits) return RTG_ERROR_INVALID_ARGUMENT;
    if (mode == RTG_QUERY_OCCLUSION && !occluded) return RTG_ERROR_INVALID_ARGUMENT;
    qctx probe;
    if (qctx_init(&probe, tlas, scene) != RTG_OK) return RTG_ERROR_INVALID_ARGUMENT;
    rays_job j = {rays, count, mode, hits, occluded};
    pool p;
    memset(&p, 0, sizeof(p));
    p.fn = rays_item;
    p.arg = &j;
    p.items = (count + RAY_CHUNK - 1) / RAY_CHUNK;
    p.tlas = tlas;
    p.scene = scene;
    p.ray_flags = ray_flags;
    p.cull_mask = cull_mask & 0xffu;
    p.sbt_offset = sbt_offset;
    run_pool(&p, threads);
    if (counters) {
        memset(counters, 0, sizeof(*counters));
        add_counters(counters, &p.total);
        counters->primary_rays = count;
    }
    return RTG_OK;
}

int orc_trace_rays_brute(const void* tlas, const rtg_scene_bindings* scene, const rtg_ray* rays, uint32_t count,
                         uint32_t mode, uint32_t ray_flags, uint32_t cull_mask, uint32_t sbt_offset, rtg_hit* hits,
                         uint32_t* occluded) {
    qctx q;
    if (qctx_init(&q, tlas, scene) != RTG_OK) return RTG_ERROR_INVALID_ARGUMENT;
    q.ray_flags = ray_flags;
    q.cull_mask = cull_mask & 0xffu;
    q.sbt_offset = sbt_offset;
    uint32_t ninst = q.th->prim_count;
    for (uint32_t i = 0; i < count; ++i) {
        const rtg_ray* r = &rays[i];
        v3 o = V3(r->origin[0], r->origin[1], r->origin[2]);
        v3 d = V3(r->direction[0], r->direction[1], r->direction[2]);
        hitrec best;
        best.t = r->tmax;
        best.u = best.v = 0.0f;
        best.inst = best.prim = best.geom = RTG_NO_HIT;
        int found = 0;
        for (uint32_t k = 0; k < ninst && !(found && mode == RTG_QUERY_OCCLUSION); ++k) {
            const orc_inst* rec = &q.tinst[k];
            if (((rec->mask_flags & 0xffu) & q.cull_mask) == 0 || !rec->blas) continue;
            const orc_header* bh = (const orc_header*)(uintptr_t)rec->blas;
            const orc_tri* tris = (const orc_tri*)((const uint8_t*)bh + bh->prims_off);
            oray obj;
            oray_setup(&obj, m34_point(rec->w2o, o), m34_dir(rec->w2o, d));
            for (uint32_t t_i = 0; t_i < bh->prim_count; ++t_i) {
                const orc_tri* tr = &tris[t_i];
                float t, u, v;
                if (!tri_test(&obj, tr->v0, tr->v1, tr->v2, r->tmin, best.t, &t, &u, &v)) continue;
                uint32_t geom = tr->geom & ~ORC_NONOPAQUE_BIT;
                if (mode == RTG_QUERY_CLOSEST &&
                    !(t < best.t || (t == best.t && key_less(rec->index, geom, tr->prim, best.inst, best.geom, best.prim))))
                    continue;
                uint32_t iflags = (rec->mask_flags >> 8) & 0xffu;
                if (!is_opaque(tr->geom, iflags, ray_flags)) {
                    uint32_t group = (q.tsrc[rec->index].sbt_offset_and_flags & 0xffffffu) + sbt_offset;
                    if (group < RTG_HIT_GROUP_COUNT && !alpha_clip_keep(&q, &q.cnt.anyhit_calls, rec->index, geom, tr->prim, u, v))
                        continue;
                }
                found = 1;
                if (mode == RTG_QUERY_OCCLUSION) break;
                best.t = t;
                best.u = u;
                best.v = v;
                best.inst = rec->index;
                best.geom = geom;
                best.prim = tr->prim;
            }
        }
        if (mode == RTG_QUERY_OCCLUSION) {
            occluded[i] = (uint32_t)found;
        } else {
            rtg_hit* h = &hits[i];
            h->t = best.t;
            h->instance = best.inst;
            h->primitive = best.prim;
            h->geometry = best.geom;
            h->u = best.u;
            h->v = best.v;
        }
    }
    return RTG_OK;
}

/* ---------------------------------------------------------------- */
/* Shading stages                                                    */
/* ---------------------------------------------------------------- */

/* pbr.glsl:56-65 and shaders/ray-tracing/src/pbr.rs:41-50 */
float orc_v_smith_ggx_correlated(float NoV, float NoL, float roughness) {
    float a2 = roughness * roughness;
    float ggxv = NoL * sqrtf((NoV * NoV) * (1.0f - a2) + a2);
    float ggxl = NoV * sqrtf((NoL * NoL) * (1.0f - a2) + a2);
    return 0.5f / (ggxv + ggxl);
}

/* pbr.glsl:44-51 */
float orc_d_ggx(float NoH, float roughness) {
    float a = NoH * roughness;
    float k = roughness / ((1.0f - NoH * NoH) + a * a);
    return (k * k) * (1.0f / ORC_PI_F);
}

/* pbr.glsl:174-211 (brdf) with calculate_dot_params :30-39, F_Schlick
 * :69-76, compute_f90/Fd_Burley :88-103. pow(x, 5.0) = ((x*x)*(x*x))*x. */
static v3 brdf(v3 n, v3 view, v3 l, float pr, v3 base, float metallic, float reflectance, float intensity) {
    v3 h = v3_normalize(v3_add(view, l));
    float rough = pr * pr;
    float NoV = clampf(v3_dot(n, view), 1e-9f, 1.0f);
    float NoH = clampf(v3_dot(n, h), 0.0f, 1.0f);
    float NoL = clampf(v3_dot(n, l), 0.0f, 1.0f);
    float LoH = clampf(v3_dot(l, h), 0.0f, 1.0f);
    float D = orc_d_ggx(NoH, rough);
    float dielectric_f0 = (0.16f * reflectance) * reflectance;
    v3 f0 = v3_add(V3(dielectric_f0 * (1.0f - metallic), dielectric_f0 * (1.0f - metallic),
                      dielectric_f0 * (1.0f - metallic)),
                   v3_scale(base, metallic));
    float f90 = 0.5f + ((2.0f * rough) * LoH) * LoH;
    float p5 = orc_pow5f(1.0f - LoH);
    v3 F = V3(f0.x + (f90 - f0.x) * p5, f0.y + (f90 - f0.y) * p5, f0.z + (f90 - f0.z) * p5);
    float Vis = orc_v_smith_ggx_correlated(NoV, NoL, rough);
    float DV = D * Vis;
    v3 spec = v3_scale(F, DV);
    float ls = 1.0f + (f90 - 1.0f) * orc_pow5f(1.0f - NoL);
    float vs = 1.0f + (f90 - 1.0f) * orc_pow5f(1.0f - NoV);
    float fd = (ls * vs) * (1.0f / ORC_PI_F);
    v3 diffuse = v3_scale(base, fd);
    v3 comb = v3_add(diffuse, spec);
    float in_nol = intensity * NoL;
    return v3_scale(comb, in_nol);
}

void orc_brdf(const float normal[3], const float view[3], const float light[3], float perceptual_roughness,
              const float base_colour[3], float metallic, float reflectance, float light_intensity, float out[3]) {
    v3 r = brdf(V3(normal[0], normal[1], normal[2]), V3(view[0], view[1], view[2]), V3(light[0], light[1], light[2]),
                perceptual_roughness, V3(base_colour[0], base_colour[1], base_colour[2]), metallic, reflectance,
                light_intensity);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

/* shaders/ray-tracing/src/lib.rs:84-92 */
static float srgb1(float x) {
    if (x <= 0.0031308f) return x * 12.92f;
    return 1.055f * orc_powf(x, 1.0f / 2.4f) - 0.055f;
}
void orc_linear_to_srgb(const float in[3], float out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = srgb1(in[i]);
}

static float smoothstep_(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}

/* shaders/ray-tracing/src/heatmap.rs:5-38 */
void orc_heatmap_temperature(float heat, float out[3]) {
    static const float col[10][3] = {
        {0.0f, 2.0f, 91.0f}, {0.0f, 108.0f, 251.0f}, {0.0f, 221.0f, 221.0f}, {51.0f, 221.0f, 0.0f},
        {255.0f, 252.0f, 0.0f}, {255.0f, 180.0f, 0.0f}, {255.0f, 104.0f, 0.0f}, {226.0f, 22.0f, 0.0f},
        {191.0f, 0.0f, 83.0f}, {145.0f, 0.0f, 65.0f}};
    float h = clampf(heat, 0.0f, 1.0f) * 10.0f;
    int hi = (int)h;
    int cur = hi, prv = hi - 1 > 0 ? hi - 1 : 0, nxt = hi + 1 < 9 ? hi + 1 : 9;
    if (cur > 9) cur = 9; /* heat == 1.0 indexes one past the table in the reference */
    float hf = floorf(h), hc = ceilf(h);
    float blur = 0.8f;
    float wc = smoothstep_(hf - blur, hf + blur, h) * (1.0f - smoothstep_(hc - blur, hc + blur, h));
    float wp = 1.0f - smoothstep_(hf - blur, hf + blur, h);
    float wn = smoothstep_(hc - blur, hc + blur, h);
    for (int c = 0; c < 3; ++c) {
        float r = (wc * (col[cur][c] / 255.0f) + wp * (col[prv][c] / 255.0f)) + wn * (col[nxt][c] / 255.0f);
        out[c] = clampf(r, 0.0f, 1.0f);
    }
}

/* closest_hit_textured.glsl:77-94 */
static v3 sample_dir_light(float rx, float ry, v3 c, float radius) {
    float rad = sqrtf(rx);
    float angle = (ry * 2.0f) * ORC_PI_F;
    float s, co;
    orc_sincosf(angle, &s, &co);
    float px = (rad * co) * radius, py = (rad * s) * radius;
    v3 t = v3_normalize(v3_cross(c, V3(0.0f, 1.0f, 0.0f)));
    v3 b = v3_normalize(v3_cross(t, c));
    return v3_normalize(v3_add(v3_add(c, v3_scale(t, px)), v3_scale(b, py)));
}
void orc_sample_directional_light(const float rng[2], const float centre[3], float radius, float out[3]) {
    v3 r = sample_dir_light(rng[0], rng[1], V3(centre[0], centre[1], centre[2]), radius);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

float orc_cos(float x) { return orc_cosf(x); }
float orc_sin(float x) { return orc_sinf(x); }
float orc_pow(float x, float y) { return orc_powf(x, y); }
float orc_log2(float x) { return orc_log2f(x); }
float orc_exp2(float y) { return orc_exp2f(y); }

typedef struct {
    const rtg_image_job* job;
    const uint32_t* pixels;
    uint32_t pixel_count;
    float cos_sun;
} image_ctx;

static v3 load_v3(uint64_t base, uint32_t i) {
    const float* p = (const float*)(uintptr_t)base + 3ull * i;
    return V3(p[0], p[1], p[2]);
}
static v2 load_v2(uint64_t base, uint32_t i) {
    const float* p = (const float*)(uintptr_t)base + 2ull * i;
    v2 r = {p[0], p[1]};
    return r;
}

/* ---------------------------------------------------------------- */
/* The stage functions of the trace step. Each is used by the per-pixel */
/* path below and exported (orc_*) for the tests that pin it.           */
/* ---------------------------------------------------------------- */

/* ray_generation's camera ray (shaders/ray-tracing/src/lib.rs:126-142):
 * pixel point (px + jx, py + jy) / launch size * 2 - 1 = render coordinates
 * (x, y); origin = view_inverse * (0, 0, 0, 1); direction = view_inverse *
 * (normalize((proj_inverse * (x, y, 1, 1)).xyz), 0). No perspective divide:
 * the w of proj_inverse * (x, y, 1, 1) is dropped. */
static void camera_ray(const rtg_uniforms* U, uint32_t width, uint32_t height, uint32_t px, uint32_t py, float jx,
                       float jy, v3* origin, v3* dir) {
    float pcx = (float)px + jx, pcy = (float)py + jy;
    float tcx = pcx / (float)width, tcy = pcy / (float)height;
    float rcx = tcx * 2.0f - 1.0f, rcy = tcy * 2.0f - 1.0f;
    float zero_one[4] = {0.0f, 0.0f, 0.0f, 1.0f}, tmp[4];
    mat4_mul_vec4(U->view_inverse, zero_one, tmp);
    *origin = V3(tmp[0], tmp[1], tmp[2]);
    float tv[4] = {rcx, rcy, 1.0f, 1.0f};
    mat4_mul_vec4(U->proj_inverse, tv, tmp);
    v3 ld = v3_normalize(V3(tmp[0], tmp[1], tmp[2]));
    float dv[4] = {ld.x, ld.y, ld.z, 0.0f};
    mat4_mul_vec4(U->view_inverse, dv, tmp);
    *dir = V3(tmp[0], tmp[1], tmp[2]);
}
void orc_camera_ray(const rtg_uniforms* U, uint32_t width, uint32_t height, uint32_t px, uint32_t py, float jx,
                    float jy, float origin[3], float dir[3]) {
    v3 o, d;
    camera_ray(U, width, height, px, py, jx, jy, &o, &d);
    origin[0] = o.x; origin[1] = o.y; origin[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

/* primary_ray_miss (lib.rs:38-51): the sun disk is white where the ray lies
 * within sun_radius of sun_dir (dot > cos(sun_radius)), the sky is
 * (0, 0, 0.05). cos_sun = cos(sun_radius). */
static v3 primary_miss(v3 dir, v3 sun, float cos_sun) {
    return v3_dot(dir, sun) > cos_sun ? V3(1.0f, 1.0f, 1.0f) : V3(0.0f, 0.0f, 0.05f);
}
void orc_primary_miss(const float dir[3], const float sun_dir[3], float sun_radius, float out[3]) {
    v3 c = primary_miss(V3(dir[0], dir[1], dir[2]), V3(sun_dir[0], sun_dir[1], sun_dir[2]), orc_cosf(sun_radius));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* get_shadow_terminator_fix_shadow_origin (closest_hit_textured.glsl:13-39;
 * Ray Tracing Gems II ch. 4.3), model space: the hit point P = interpolate(
 * positions, bc) plus the interpolation of each vertex's offset
 * (P - P_i) - min(0, (P - P_i) . n_i) n_i. */
static v3 terminator_fix(const v3 P[3], const v3 N[3], v3 bc) {
    v3 pos = v3_interp(P[0], P[1], P[2], bc);
    v3 off[3];
    for (int k = 0; k < 3; ++k) {
        v3 vtp = v3_sub(pos, P[k]);
        float dp = fminf(0.0f, v3_dot(vtp, N[k]));
        off[k] = v3_sub(vtp, v3_scale(N[k], dp));
    }
    return v3_add(pos, v3_interp(off[0], off[1], off[2], bc));
}
void orc_terminator_fix(const float pos[9], const float nrm[9], const float bc[3], float out[3]) {
    v3 P[3], N[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = V3(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
        N[k] = V3(nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2]);
    }
    v3 r = terminator_fix(P, N, V3(bc[0], bc[1], bc[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* sample_blue_noise (closest_hit_textured.glsl:96-111) and
 * animated_blue_noise (:113-120): the two numbers of shadow ray i of pixel
 * (px, py): the .r of the 64x64 mask at (px, py) + 2i (13, 41) and at
 * (px, py) + (2i + 1)(13, 41), over 64, each plus (frame % 32) times the
 * golden ratio's fraction, modulo 1. `bn` NULL samples as (0, 0, 0, 1). */
static void blue_noise_pair(const rtg_texture* bn, const float* lut, uint32_t px, uint32_t py, uint32_t i,
                            uint32_t frame, float* rx, float* ry) {
    static const rtg_texture none = {0, 0, 0, 0, 0};
    if (!bn) bn = &none;
    uint32_t o1x = i * 2u * 13u, o1y = i * 2u * 41u;
    uint32_t o2x = (i * 2u + 1u) * 13u, o2y = (i * 2u + 1u) * 41u;
    float t1[4], t2[4];
    sample_tex(bn, lut, (float)(px + o1x) / 64.0f, (float)(py + o1y) / 64.0f, t1);
    sample_tex(bn, lut, (float)(px + o2x) / 64.0f, (float)(py + o2y) / 64.0f, t2);
    float anim = (float)(frame % 32u) * 0.618033988749f;
    *rx = fractf(t1[0] + anim);
    *ry = fractf(t2[0] + anim);
}
void orc_blue_noise(const rtg_texture* bn, const float lut[256], uint32_t px, uint32_t py, uint32_t i,
                    uint32_t frame, float out[2]) {
    blue_noise_pair(bn, lut, px, py, i, frame, &out[0], &out[1]);
}

/* calculate_normal (closest_hit_textured.glsl:141-157): without a normal map
 * the interpolated normal rotated into world space by
 * mat3(gl_WorldToObject3x4EXT) (:69-72) and normalised; with one, the map's
 * rgb * 2 - 1, normalised, through the cotangent frame of the triangle
 * (compute_cotangent_frame :123-139) first. map_rgb NULL: no normal map. */
static v3 shading_normal(const float o2w[12], v3 nrm, const v3 P[3], const v2 UV[3], const float* map_rgb) {
    float w2o[12];
    m34_inverse(o2w, w2o);
    if (!map_rgb) return v3_normalize(m34_transpose_dir(w2o, nrm));
    v3 mapn = V3(map_rgb[0] * 2.0f - 1.0f, map_rgb[1] * 2.0f - 1.0f, map_rgb[2] * 2.0f - 1.0f);
    v3 dp1 = v3_sub(P[1], P[0]), dp2 = v3_sub(P[2], P[0]);
    float du1x = UV[1].x - UV[0].x, du1y = UV[1].y - UV[0].y, du2x = UV[2].x - UV[0].x, du2y = UV[2].y - UV[0].y;
    v3 dp2perp = v3_cross(dp2, nrm), dp1perp = v3_cross(nrm, dp1);
    v3 T = v3_add(v3_scale(dp2perp, du1x), v3_scale(dp1perp, du2x));
    v3 B = v3_add(v3_scale(dp2perp, du1y), v3_scale(dp1perp, du2y));
    float invmax = 1.0f / sqrtf(fmaxf(v3_dot(T, T), v3_dot(B, B)));
    v3 tt = v3_scale(T, invmax), bb = v3_scale(B, invmax);
    v3 m = v3_normalize(mapn);
    v3 local = v3_add(v3_add(v3_scale(tt, m.x), v3_scale(bb, m.y)), v3_scale(nrm, m.z));
    return v3_normalize(m34_transpose_dir(w2o, local));
}
void orc_shading_normal(const float o2w[12], const float nrm[3], const float pos[9], const float uv[6],
                        const float* map_rgb, float out[3]) {
    v3 P[3];
    v2 UV[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = V3(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
        UV[k].x = uv[2 * k];
        UV[k].y = uv[2 * k + 1];
    }
    v3 r = shading_normal(o2w, V3(nrm[0], nrm[1], nrm[2]), P, UV, map_rgb);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* closest_hit_textured's colour (closest_hit_textured.glsl:203-225): the
 * BRDF towards the centre sun direction with light intensity sun_factor (the
 * mean visibility of the shadow rays) and reflectance 0.5, plus the ambient
 * term 0.1 * base colour. */
static v3 textured_colour(v3 normal, v3 ray_dir, v3 sun, float roughness, v3 base, float metallic,
                          float sun_factor) {
    v3 col = brdf(normal, v3_neg(ray_dir), sun, roughness, base, metallic, 0.5f, sun_factor);
    v3 amb = V3(0.1f * base.x, 0.1f * base.y, 0.1f * base.z);
    return v3_add(col, amb);
}
void orc_textured_colour(const float normal[3], const float ray_dir[3], const float sun_dir[3], float roughness,
                         const float base[3], float metallic, float sun_factor, float out[3]) {
    v3 r = textured_colour(V3(normal[0], normal[1], normal[2]), V3(ray_dir[0], ray_dir[1], ray_dir[2]),
                           V3(sun_dir[0], sun_dir[1], sun_dir[2]), roughness, V3(base[0], base[1], base[2]),
                           metallic, sun_factor);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* closest_hit_mirror (closest_hit_mirror.glsl:11-29): the continuation ray
 * leaves the hit point o + d t along reflect(d, n), n = the interpolated
 * normal rotated by mat3(gl_WorldToObject3x4EXT) and normalised;
 * reflect(d, n) = d - 2 (n . d) n. */
static void mirror_stage(const float o2w[12], v3 n_interp, v3 o, v3 d, float t, v3* new_o, v3* new_d) {
    float w2o[12];
    m34_inverse(o2w, w2o);
    v3 n = v3_normalize(m34_transpose_dir(w2o, n_interp));
    float k = 2.0f * v3_dot(n, d);
    *new_d = v3_sub(d, v3_scale(n, k));
    *new_o = v3_add(o, v3_scale(d, t));
}
void orc_mirror(const float o2w[12], const float normal[3], const float origin[3], const float dir[3], float t,
                float new_origin[3], float new_dir[3]) {
    v3 no, nd;
    mirror_stage(o2w, V3(normal[0], normal[1], normal[2]), V3(origin[0], origin[1], origin[2]),
                 V3(dir[0], dir[1], dir[2]), t, &no, &nd);
    new_origin[0] = no.x; new_origin[1] = no.y; new_origin[2] = no.z;
    new_dir[0] = nd.x; new_dir[1] = nd.y; new_dir[2] = nd.z;
}

/* closest_hit_portal (shaders/ray-tracing/src/lib.rs:300-312): the ray
 * continues in its direction from the hit point o + d t moved by (0, 5, 0). */
static void portal_stage(v3 o, v3 d, float t, v3* new_o, v3* new_d) {
    *new_d = d;
    *new_o = v3_add(v3_add(o, v3_scale(d, t)), V3(0.0f, 5.0f, 0.0f));
}
void orc_portal(const float origin[3], const float dir[3], float t, float new_origin[3], float new_dir[3]) {
    v3 no, nd;
    portal_stage(V3(origin[0], origin[1], origin[2]), V3(dir[0], dir[1], dir[2]), t, &no, &nd);
    new_origin[0] = no.x; new_origin[1] = no.y; new_origin[2] = no.z;
    new_dir[0] = nd.x; new_dir[1] = nd.y; new_dir[2] = nd.z;
}

/* The heatmap branch of ray_generation (lib.rs:174-186): colour =
 * heatmap_temperature(cycles / 1e6) + colour * 1e-6, per pixel, for the
 * traversal cycle counts a device measured (n pixels, rgb in and out). */
void orc_heatmap_pixels(const uint32_t* cycles, const float* rgb, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        float hm[3];
        orc_heatmap_temperature((float)cycles[i] / 1000000.0f, hm);
        for (int c = 0; c < 3; ++c) out[3 * i + c] = hm[c] + rgb[3 * i + c] * 0.000001f;
    }
}

/* closest_hit_textured main (closest_hit_textured.glsl:174-226). Returns the
 * final colour given the hit; traces its own shadow rays. */
static v3 shade_textured(qctx* q, const rtg_image_job* job, uint32_t px, uint32_t py, uint32_t frame,
                         const hitrec* h, v3 ray_dir) {
    const rtg_instance* src = &q->tsrc[h->inst];
    uint32_t custom = src->custom_index_and_mask & 0xffffffu;
    if (custom >= q->model_count) return V3(0.0f, 0.0f, 0.0f);
    q->cnt.textured_hits++;
    const rtg_model_info* mi = &q->models[custom];
    const rtg_geometry_info* gi = (const rtg_geometry_info*)(uintptr_t)mi->geometry_info_address + h->geom;
    const uint32_t* idx = (const uint32_t*)(uintptr_t)gi->index_buffer_address + 3ull * h->prim;
    uint32_t i0 = idx[0], i1 = idx[1], i2 = idx[2];
    v3 P[3] = {load_v3(mi->position_buffer_address, i0), load_v3(mi->position_buffer_address, i1),
               load_v3(mi->position_buffer_address, i2)};
    v3 N[3] = {load_v3(mi->normal_buffer_address, i0), load_v3(mi->normal_buffer_address, i1),
               load_v3(mi->normal_buffer_address, i2)};
    v2 UV[3] = {load_v2(mi->uv_buffer_address, i0), load_v2(mi->uv_buffer_address, i1),
                load_v2(mi->uv_buffer_address, i2)};
    v3 bc = V3((1.0f - h->u) - h->v, h->u, h->v);
    v3 nrm = v3_interp(N[0], N[1], N[2], bc);
    v2 uv = v2_interp(UV[0], UV[1], UV[2], bc);
    const float* o2w = (const float*)src->transform;
    v3 shadow_origin = m34_point(o2w, terminator_fix(P, N, bc));

    const rtg_uniforms* U = &job->uniforms;
    v3 sun = V3(U->sun_dir[0], U->sun_dir[1], U->sun_dir[2]);
    const rtg_texture* bn =
        U->blue_noise_texture_index < q->texture_count ? &q->textures[U->blue_noise_texture_index] : NULL;
    uint32_t nsh = job->opts.shadow_rays;
    float vis = 0.0f;
    uint32_t saved_flags = q->ray_flags, saved_sbt = q->sbt_offset;
    q->ray_flags = RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER;
    q->sbt_offset = 1;
    /* the rays in pairs (2j, 2j + 1), an odd count's last alone, as the
       device's occlusion lanes take them; a stack overflow leaves a pair's
       rays unresolved, counted as not visible */
    for (uint32_t i = 0; i < nsh; i += g_shadow_pairs ? 2 : 1) {
        float rx, ry;
        blue_noise_pair(bn, q->lut, px, py, i, frame, &rx, &ry);
        v3 sd0 = sample_dir_light(rx, ry, sun, U->sun_radius), sd1 = sd0;
        if (!g_shadow_pairs) {
            hitrec sh;
            q->cnt.shadow_rays++;
            vis = vis + (traverse(q, shadow_origin, sd0, 0.001f, 10000.0f, 1, &sh) > 0 ? 0.0f : 1.0f);
            continue;
        }
        uint32_t m = 1;
        if (i + 1 < nsh) {
            blue_noise_pair(bn, q->lut, px, py, i + 1, frame, &rx, &ry);
            sd1 = sample_dir_light(rx, ry, sun, U->sun_radius);
            m = 3;
        }
        q->cnt.shadow_rays += m == 3 ? 2 : 1;
        int occl = traverse_pair(q, shadow_origin, sd0, sd1, m, 0.001f, 10000.0f);
        if (occl < 0) occl = (int)m;
        vis = vis + ((occl & 1) ? 0.0f : 1.0f);
        if (m == 3) vis = vis + ((occl & 2) ? 0.0f : 1.0f);
    }
    q->ray_flags = saved_flags;
    q->sbt_offset = saved_sbt;
    float sun_factor = vis / (float)nsh;

    /* read_material_from_textures (:47-63) */
    float dt[4], mr[4];
    sample_index(q, gi->diffuse_image_index, uv.x, uv.y, dt);
    sample_index(q, gi->metallic_roughness_image_index, uv.x, uv.y, mr);
    v3 base = V3(dt[0], dt[1], dt[2]);
    float metallic = mr[2], roughness = mr[1];

    float mt[4];
    const float* map_rgb = NULL;
    if (gi->normal_map_image_index >= 0) {
        sample_index(q, (uint32_t)gi->normal_map_image_index, uv.x, uv.y, mt);
        map_rgb = mt;
    }
    v3 normal = shading_normal(o2w, nrm, P, UV, map_rgb);
    return textured_colour(normal, ray_dir, sun, roughness, base, metallic, sun_factor);
}

/* The normal of the hit triangle interpolated in model space (the mirror
 * stage's interpolate(read_vec3_triangle(normals), bc),
 * closest_hit_mirror.glsl:21). */
static v3 hit_normal(const qctx* q, const hitrec* h) {
    const rtg_instance* src = &q->tsrc[h->inst];
    const rtg_model_info* mi = &q->models[src->custom_index_and_mask & 0xffffffu];
    const rtg_geometry_info* gi = (const rtg_geometry_info*)(uintptr_t)mi->geometry_info_address + h->geom;
    const uint32_t* idx = (const uint32_t*)(uintptr_t)gi->index_buffer_address + 3ull * h->prim;
    v3 na = load_v3(mi->normal_buffer_address, idx[0]), nb = load_v3(mi->normal_buffer_address, idx[1]),
       nc = load_v3(mi->normal_buffer_address, idx[2]);
    v3 bc = V3((1.0f - h->u) - h->v, h->u, h->v);
    return v3_interp(na, nb, nc, bc);
}

/* One sample of one pixel: ray_generation (lib.rs:94-191). */
static v3 trace_sample(qctx* q, const image_ctx* ic, uint32_t px, uint32_t py, uint32_t s, rtg_hit* primary) {
    const rtg_image_job* job = ic->job;
    const rtg_uniforms* U = &job->uniforms;
    uint32_t spp = job->opts.spp;
    float jx = 0.5f, jy = 0.5f;
    if (spp > 1) {
        jx = fractf(0.5f + (float)s * 0.754877669f);
        jy = fractf(0.5f + (float)s * 0.569840296f);
    }
    uint32_t frame = U->frame_index * spp + s;
    v3 origin, dir;
    camera_ray(U, job->width, job->height, px, py, jx, jy, &origin, &dir);
    v3 colour = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t tr = 0; tr < job->opts.max_traces; ++tr) {
        colour = V3(0.0f, 0.0f, 0.0f);
        v3 no = V3(0.0f, 0.0f, 0.0f), nd = V3(0.0f, 0.0f, 0.0f);
        hitrec h;
        q->ray_flags = RTG_RAY_FLAG_NONE;
        q->sbt_offset = 0;
        if (tr == 0) q->cnt.primary_rays++;
        else q->cnt.bounce_rays++;
        int f = traverse(q, origin, dir, 0.01f, 10000.0f, 0, &h);
        if (tr == 0 && primary) {
            primary->t = h.t;
            primary->instance = h.inst;
            primary->primitive = h.prim;
            primary->geometry = h.geom;
            primary->u = h.u;
            primary->v = h.v;
        }
        if (f <= 0) {
            v3 sun = V3(U->sun_dir[0], U->sun_dir[1], U->sun_dir[2]);
            colour = primary_miss(dir, sun, ic->cos_sun);
        } else {
            const rtg_instance* src = &q->tsrc[h.inst];
            uint32_t group = src->sbt_offset_and_flags & 0xffffffu;
            int has_model = (src->custom_index_and_mask & 0xffffffu) < q->model_count;
            if (group == RTG_HIT_TEXTURED) {
                colour = shade_textured(q, job, px, py, frame, &h, dir);
            } else if (group == RTG_HIT_MIRROR && has_model) {
                mirror_stage((const float*)src->transform, hit_normal(q, &h), origin, dir, h.t, &no, &nd);
            } else if (group == RTG_HIT_PORTAL) {
                portal_stage(origin, dir, h.t, &no, &nd);
            }
        }
        if (nd.x == 0.0f && nd.y == 0.0f && nd.z == 0.0f) break;
        origin = no;
        dir = nd;
    }
    return colour;
}

#define TILE 16u

static void image_item(void* a, qctx* q, uint32_t item) {
    image_ctx* ic = (image_ctx*)a;
    const rtg_image_job* job = ic->job;
    uint32_t first, last;
    if (ic->pixels) {
        first = item * 64u;
        last = first + 64u < ic->pixel_count ? first + 64u : ic->pixel_count;
    } else {
        first = 0;
        last = 0;
    }
    uint32_t tiles_x = (job->width + TILE - 1) / TILE;
    uint32_t n_work = ic->pixels ? last - first : TILE * TILE;
    for (uint32_t w = 0; w < n_work; ++w) {
        uint32_t px, py;
        if (ic->pixels) {
            uint32_t p = ic->pixels[first + w];
            px = p % job->width;
            py = p / job->width;
        } else {
            px = (item % tiles_x) * TILE + w % TILE;
            py = (item / tiles_x) * TILE + w / TILE;
            if (px >= job->width || py >= job->height) continue;
        }
        uint32_t pix = py * job->width + px;
        v3 acc = V3(0.0f, 0.0f, 0.0f);
        for (uint32_t s = 0; s < job->opts.spp; ++s) {
            rtg_hit* ph = job->out_primary_hits ? &job->out_primary_hits[(uint64_t)pix * job->opts.spp + s] : NULL;
            v3 c = trace_sample(q, ic, px, py, s, ph);
            acc = v3_add(acc, c);
        }
        v3 col = job->opts.spp == 1 ? acc : v3_scale(acc, 1.0f / (float)job->opts.spp);
        if (job->out_rgb) {
            job->out_rgb[3ull * pix] = col.x;
            job->out_rgb[3ull * pix + 1] = col.y;
            job->out_rgb[3ull * pix + 2] = col.z;
        }
        if (job->out_rgba8) {
            float lin[3] = {col.x, col.y, col.z}, sr[3];
            orc_linear_to_srgb(lin, sr);
            uint32_t r = (uint32_t)(clampf(sr[0], 0.0f, 1.0f) * 255.0f + 0.5f);
            uint32_t g = (uint32_t)(clampf(sr[1], 0.0f, 1.0f) * 255.0f + 0.5f);
            uint32_t b = (uint32_t)(clampf(sr[2], 0.0f, 1.0f) * 255.0f + 0.5f);
            if (job->opts.flags & RTG_TRACE_OUTPUT_BGRA8) {
                uint32_t t = r;
                r = b;
                b = t;
            }
            job->out_rgba8[pix] = r | (g << 8) | (b << 16) | (255u << 24);
        }
    }
}

int orc_trace_image(const rtg_image_job* job, const uint32_t* pixel_list, uint32_t pixel_count, int threads,
                    rtg_trace_counters* counters) {
    if (!job || job->width == 0 || job->height == 0 || job->opts.spp == 0 || job->opts.max_traces == 0 ||
        job->opts.shadow_rays == 0)
        return RTG_ERROR_INVALID_ARGUMENT;
    image_ctx ic;
    ic.job = job;
    ic.pixels = pixel_list;
    ic.pixel_count = pixel_count;
    ic.cos_sun = orc_cosf(job->uniforms.sun_radius);
    pool p;
    memset(&p, 0, sizeof(p));
    p.fn = image_item;
    p.arg = &ic;
    if (pixel_list) p.items = (pixel_count + 63u) / 64u;
    else p.items = ((job->width + TILE - 1) / TILE) * ((job->height + TILE - 1) / TILE);
    p.tlas = (const void*)(uintptr_t)job->tlas;
    p.scene = &job->scene;
    p.cull_mask = 0xffu;
    run_pool(&p, threads);
    if (counters) {
        memset(counters, 0, sizeof(*counters));
        add_counters(counters, &p.total);
    }
    return RTG_OK;
}

/* Test hook: the visits of one shadow ray's walk (pair = 0: ray d0 alone,
 * the single-ray occlusion walk) or of the pair walk of d0 and d1 (pair = 1),
 * with the image's shadow-ray settings (tmin 0.001, tmax 10000,
 * TerminateOnFirstHit | SkipClosestHit, hit-group offset 1). log receives
 * cap triples (kind 0 node / 1 triangle / 2 instance entry, the instance
 * record of the BLAS walked or 0, the record); *n their count. Returns the
 * walk's result (occluded: 1 / mask of occluded rays) or a negative error. */
int orc_debug_visits(const void* tlas, const rtg_scene_bindings* scene, const float o[3], const float d0[3],
                     const float d1[3], int pair, uint64_t* log, uint32_t cap, uint32_t* n) {
    qctx* q = (qctx*)malloc(sizeof(qctx));
    if (!q) return RTG_ERROR_OUT_OF_MEMORY;
    if (qctx_init(q, tlas, scene) != RTG_OK) {
        free(q);
        return RTG_ERROR_INVALID_ARGUMENT;
    }
    q->ray_flags = RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER;
    q->sbt_offset = 1;
    q->vlog = log;
    q->vlog_cap = cap;
    v3 vo = V3(o[0], o[1], o[2]), va = V3(d0[0], d0[1], d0[2]), vb = V3(d1[0], d1[1], d1[2]);
    int r;
    if (pair) {
        r = traverse_pair(q, vo, va, vb, 3u, 0.001f, 10000.0f);
    } else {
        hitrec h;
        r = traverse(q, vo, va, 0.001f, 10000.0f, 1, &h);
    }
    *n = q->vlog_n;
    int full = q->vlog_n >= cap;
    free(q);
    return full ? RTG_ERROR_LIMIT : r;
}
