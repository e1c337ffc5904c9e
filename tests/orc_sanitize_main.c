/*
 * Drives every entry point of the oracle (TEST INFRASTRUCTURE) over a small
 * procedural scene so that a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile, target `sanitize`) checks its
 * memory accesses and arithmetic: BLAS builds (plain and PREFER_FAST_TRACE,
 * opaque and alpha-tested geometry), a TLAS build and refit, closest-hit and
 * occlusion ray queries (BVH and brute force, one and several threads),
 * image jobs with every hit group, textures of every format and filter, the
 * heatmap mode and a pixel subset. Exits non-zero on a failed check; the
 * sanitizers abort on the first error they find.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../oracle/orc.h"

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                     \
        }                                                                \
    } while (0)

static uint64_t rng_state = 0x5EED0300u;
static float frand(void) { /* SplitMix64 → [0, 1) */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)(z >> 40) / 16777216.0f;
}

typedef struct {
    float* pos;
    float* nrm;
    float* uv;
    uint32_t* idx;
    uint32_t nv, nt;
} mesh;

/* An n x n grid of quads in the xz plane with a random height field. */
static mesh grid_mesh(uint32_t n, float size, float bump) {
    mesh m;
    m.nv = (n + 1) * (n + 1);
    m.nt = 2 * n * n;
    m.pos = malloc(sizeof(float) * 3 * m.nv);
    m.nrm = malloc(sizeof(float) * 3 * m.nv);
    m.uv = malloc(sizeof(float) * 2 * m.nv);
    m.idx = malloc(sizeof(uint32_t) * 3 * m.nt);
    for (uint32_t j = 0; j <= n; ++j)
        for (uint32_t i = 0; i <= n; ++i) {
            uint32_t v = j * (n + 1) + i;
            m.pos[3 * v] = size * ((float)i / n - 0.5f);
            m.pos[3 * v + 1] = bump * frand();
            m.pos[3 * v + 2] = size * ((float)j / n - 0.5f);
            m.nrm[3 * v] = 0.0f;
            m.nrm[3 * v + 1] = 1.0f;
            m.nrm[3 * v + 2] = 0.0f;
            m.uv[2 * v] = 3.0f * (float)i / n - 1.0f; /* negative and > 1: wrapped */
            m.uv[2 * v + 1] = 2.5f * (float)j / n;
        }
    uint32_t t = 0;
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t a = j * (n + 1) + i, b = a + 1, c = a + n + 1, d = c + 1;
            m.idx[3 * t] = a, m.idx[3 * t + 1] = c, m.idx[3 * t + 2] = b, ++t;
            m.idx[3 * t] = b, m.idx[3 * t + 1] = c, m.idx[3 * t + 2] = d, ++t;
        }
    return m;
}

static void free_mesh(mesh* m) {
    free(m->pos);
    free(m->nrm);
    free(m->uv);
    free(m->idx);
}

static void set_transform(rtg_instance* in, float angle, float tx, float ty, float tz, float s) {
    float c = cosf(angle) * s, sn = sinf(angle) * s;
    float m[3][4] = {{c, 0.0f, sn, tx}, {0.0f, s, 0.0f, ty}, {-sn, 0.0f, c, tz}};
    memcpy(in->transform, m, sizeof(m));
}

int main(void) {
    /* meshes: a bumpy 24x24 grid (opaque) and a 3x3 grid (alpha tested) */
    mesh a = grid_mesh(24, 4.0f, 0.3f), b = grid_mesh(3, 2.0f, 0.0f);
    rtg_geometry ga = {(uint64_t)(uintptr_t)a.pos, 12, a.nv - 1, (uint64_t)(uintptr_t)a.idx, 3 * a.nt,
                       RTG_GEOMETRY_OPAQUE};
    rtg_geometry gb = {(uint64_t)(uintptr_t)b.pos, 12, b.nv - 1, (uint64_t)(uintptr_t)b.idx, 3 * b.nt, 0};
    void *blas_a = NULL, *blas_a2 = NULL, *blas_b = NULL;
    uint64_t sz = 0;
    CHECK(orc_blas_build(&ga, 1, RTG_BUILD_PREFER_FAST_TRACE, &blas_a, &sz) == 0 && sz > 0);
    CHECK(orc_blas_build(&ga, 1, 0, &blas_a2, &sz) == 0);
    CHECK(orc_blas_build(&gb, 1, RTG_BUILD_PREFER_FAST_TRACE, &blas_b, &sz) == 0);
    rtg_as_info info;
    CHECK(orc_as_get_info(blas_a, &info) == 0);

    /* textures: 1x1 RGBA32F, a 5x3 sRGB checker with alpha (linear), a 7x6
     * UNORM noise (nearest), a 16x16 UNORM blue noise */
    float white[4] = {0.8f, 0.7f, 0.6f, 1.0f}, mr[4] = {1.0f, 0.4f, 0.2f, 1.0f};
    uint32_t checker[15], noise[42], bn[256];
    for (int k = 0; k < 15; ++k) checker[k] = (k & 1) ? 0xFF3060C0u : 0x10A0F020u;
    for (int k = 0; k < 42; ++k) noise[k] = (uint32_t)(frand() * 4294967295.0f);
    for (int k = 0; k < 256; ++k) bn[k] = (uint32_t)k * 0x01010101u;
    rtg_texture tex[6] = {
        {(uint64_t)(uintptr_t)white, 1, 1, RTG_FORMAT_RGBA32_FLOAT, RTG_FILTER_NEAREST},
        {(uint64_t)(uintptr_t)mr, 1, 1, RTG_FORMAT_RGBA32_FLOAT, RTG_FILTER_NEAREST},
        {(uint64_t)(uintptr_t)bn, 16, 16, RTG_FORMAT_RGBA8_UNORM, RTG_FILTER_NEAREST},
        {(uint64_t)(uintptr_t)checker, 5, 3, RTG_FORMAT_RGBA8_SRGB, RTG_FILTER_LINEAR},
        {(uint64_t)(uintptr_t)noise, 7, 6, RTG_FORMAT_RGBA8_UNORM, RTG_FILTER_NEAREST},
        {0, 0, 0, RTG_FORMAT_RGBA8_UNORM, RTG_FILTER_NEAREST}, /* empty entry */
    };
    rtg_geometry_info gi_a = {(uint64_t)(uintptr_t)a.idx, 0, 1, -1, 0};
    rtg_geometry_info gi_b = {(uint64_t)(uintptr_t)b.idx, 3, 4, 4, 0}; /* normal map */
    rtg_model_info models[2] = {
        {(uint64_t)(uintptr_t)a.pos, (uint64_t)(uintptr_t)a.nrm, (uint64_t)(uintptr_t)a.uv,
         (uint64_t)(uintptr_t)&gi_a},
        {(uint64_t)(uintptr_t)b.pos, (uint64_t)(uintptr_t)b.nrm, (uint64_t)(uintptr_t)b.uv,
         (uint64_t)(uintptr_t)&gi_b},
    };
    rtg_scene_bindings scene = {(uint64_t)(uintptr_t)models, (uint64_t)(uintptr_t)tex, 2, 6};

    /* instances: ground, bumps, fences, a mirror, a portal and one that
     * the cull mask hides; one names a model index past the table */
    enum { NI = 12 };
    rtg_instance inst[NI];
    memset(inst, 0, sizeof(inst));
    for (int k = 0; k < NI; ++k) {
        int fence = k % 3 == 2;
        uint32_t group = k == 5 ? RTG_HIT_MIRROR : (k == 7 ? RTG_HIT_PORTAL : RTG_HIT_TEXTURED);
        uint32_t flags = fence ? RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE : (k == 4 ? RTG_INSTANCE_FORCE_NO_OPAQUE : 0);
        uint32_t model = k == 9 ? 7u : (fence ? 1u : 0u);
        uint32_t mask = k == 10 ? 0x02u : 0xFFu;
        set_transform(&inst[k], 0.7f * k, 2.5f * (k % 4) - 4.0f, fence ? 0.0f : 0.2f * k, 2.0f * (k / 4) - 2.0f,
                      k == 0 ? 3.0f : 1.0f);
        inst[k].custom_index_and_mask = model | (mask << 24);
        inst[k].sbt_offset_and_flags = group | (flags << 24);
        inst[k].acceleration_structure_reference = (uint64_t)(uintptr_t)(fence ? blas_b : (k == 6 ? blas_a2 : blas_a));
    }
    void* tlas = NULL;
    CHECK(orc_tlas_build(inst, NI, RTG_BUILD_ALLOW_UPDATE | RTG_BUILD_PREFER_FAST_TRACE, &tlas, &sz) == 0);

    /* ray queries: BVH against brute force, with every ray flag */
    enum { NR = 1500 };
    rtg_ray* rays = malloc(sizeof(rtg_ray) * NR);
    for (int k = 0; k < NR; ++k) {
        rtg_ray* r = &rays[k];
        r->origin[0] = 12.0f * frand() - 6.0f;
        r->origin[1] = 3.0f * frand() - 0.5f;
        r->origin[2] = 12.0f * frand() - 6.0f;
        float dx = frand() - 0.5f, dy = frand() - 0.5f, dz = frand() - 0.5f;
        if (k % 97 == 0) dx = dz = 0.0f; /* axis-parallel rays */
        r->direction[0] = dx, r->direction[1] = dy, r->direction[2] = dz;
        r->tmin = k % 5 == 0 ? 0.0f : 0.001f;
        r->tmax = k % 7 == 0 ? INFINITY : 20.0f;
    }
    rtg_hit *h1 = malloc(sizeof(rtg_hit) * NR), *h2 = malloc(sizeof(rtg_hit) * NR);
    uint32_t *o1 = malloc(sizeof(uint32_t) * NR), *o2 = malloc(sizeof(uint32_t) * NR);
    const uint32_t flag_sets[4] = {RTG_RAY_FLAG_NONE, RTG_RAY_FLAG_OPAQUE, RTG_RAY_FLAG_NO_OPAQUE,
                                   RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER};
    for (int f = 0; f < 4; ++f)
        for (uint32_t mode = 0; mode < 2; ++mode) {
            rtg_trace_counters c;
            memset(&c, 0, sizeof(c));
            CHECK(orc_trace_rays(tlas, &scene, rays, NR, mode, flag_sets[f], 0xFF, f & 1, h1, o1, &c, f == 0 ? 1 : 3) == 0);
            CHECK(orc_trace_rays_brute(tlas, &scene, rays, NR, mode, flag_sets[f], 0xFF, f & 1, h2, o2) == 0);
            /* the first accepted hit of a terminating closest-hit query
               depends on the visit order; the others do not */
            const int first_hit = (flag_sets[f] & RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT) != 0;
            for (int k = 0; k < NR; ++k) {
                if (mode == RTG_QUERY_OCCLUSION) CHECK(o1[k] == o2[k]);
                else if (!first_hit) CHECK(h1[k].instance == h2[k].instance && h1[k].primitive == h2[k].primitive);
                else CHECK((h1[k].instance == RTG_NO_HIT) == (h2[k].instance == RTG_NO_HIT));
            }
        }

    /* image jobs: all pixels on one and on several threads, a subset, the
     * heatmap mode, several samples; then a refit and the same frame */
    rtg_image_job job;
    memset(&job, 0, sizeof(job));
    job.tlas = (rtg_tlas)(uintptr_t)tlas;
    job.scene = scene;
    float view[16] = {1, 0, 0, 0, 0, 0.8f, -0.6f, 0, 0, 0.6f, 0.8f, 0, 0, 6.0f, 8.0f, 1}; /* at (0, 6, 8), looking down -z */
    float proj[16] = {0.8f, 0, 0, 0, 0, 0.6f, 0, 0, 0, 0, 0, -10.0f, 0, 0, -1.0f, 10.0f};
    memcpy(job.uniforms.view_inverse, view, sizeof(view));
    memcpy(job.uniforms.proj_inverse, proj, sizeof(proj));
    job.uniforms.sun_dir[0] = 0.3f, job.uniforms.sun_dir[1] = 0.9f, job.uniforms.sun_dir[2] = 0.3f;
    job.uniforms.sun_radius = 0.02f;
    job.uniforms.blue_noise_texture_index = 2;
    job.uniforms.frame_index = 3;
    job.width = 48, job.height = 32;
    job.opts.shadow_rays = 2, job.opts.max_traces = 3, job.opts.spp = 1, job.opts.flags = 0;
    const uint32_t npx = job.width * job.height;
    uint32_t* rgba = malloc(sizeof(uint32_t) * npx * 4);
    float* rgb = malloc(sizeof(float) * 3 * npx * 4);
    rtg_hit* ph = malloc(sizeof(rtg_hit) * npx * 4);
    job.out_rgba8 = rgba, job.out_rgb = rgb, job.out_primary_hits = ph;
    rtg_trace_counters c;
    memset(&c, 0, sizeof(c));
    CHECK(orc_trace_image(&job, NULL, 0, 1, &c) == 0);
    CHECK(c.primary_rays == npx && c.shadow_rays > 0 && c.bounce_rays > 0 && c.anyhit_calls > 0);
    uint32_t* ref = malloc(sizeof(uint32_t) * npx);
    memcpy(ref, rgba, sizeof(uint32_t) * npx);
    CHECK(orc_trace_image(&job, NULL, 0, 4, NULL) == 0);
    CHECK(memcmp(ref, rgba, sizeof(uint32_t) * npx) == 0);
    uint32_t subset[5] = {0, 7, npx / 2, npx - 2, npx - 1};
    CHECK(orc_trace_image(&job, subset, 5, 2, NULL) == 0);
    job.opts.flags = RTG_TRACE_HEATMAP | RTG_TRACE_OUTPUT_BGRA8;
    CHECK(orc_trace_image(&job, NULL, 0, 2, NULL) == 0);
    job.opts.flags = 0;
    job.opts.spp = 4;
    job.out_primary_hits = NULL;
    CHECK(orc_trace_image(&job, NULL, 0, 3, NULL) == 0);
    for (int k = 0; k < NI; ++k) set_transform(&inst[k], 0.3f * k + 1.0f, 2.4f * (k % 4) - 4.0f, 0.1f * k,
                                              2.1f * (k / 4) - 2.0f, k == 0 ? 3.0f : 1.1f);
    CHECK(orc_tlas_refit(tlas, inst, NI) == 0);
    job.opts.spp = 1;
    CHECK(orc_trace_image(&job, NULL, 0, 2, NULL) == 0);

    free(ref);
    free(rgba);
    free(rgb);
    free(ph);
    free(h1);
    free(h2);
    free(o1);
    free(o2);
    free(rays);
    orc_free(tlas);
    orc_free(blas_a);
    orc_free(blas_a2);
    orc_free(blas_b);
    free_mesh(&a);
    free_mesh(&b);
    printf("oracle sanitize run: ok\n");
    return 0;
}
