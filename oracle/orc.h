/*
 * orc.h -- the CPU oracle of the trace step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liborc.so. The product never links, imports or calls it.
 *
 * The oracle restates, on the host, in plain C:
 *   - the LBVH build of bottom- and top-level acceleration structures
 *     (Karras 2012, "Maximizing Parallelism in the Construction of BVHs,
 *     Octrees, and k-d Trees"), producing byte-identical blobs to the GPU
 *     build (the address field of TLAS instance records aside);
 *   - the TLAS refit (src/util_structs.rs:285-357 semantics: UPDATE in place);
 *   - two-level traversal with the watertight ray/triangle test (Woop,
 *     Benthin and Wald 2013, JCGT 2(1)) and the robust slab test (Ize 2013),
 *     closest hit with a (t, instance, geometry, primitive) tie-break, and
 *     any-hit occlusion, with any_hit_alpha_clip on non-opaque geometry
 *     (shaders/any_hit_alpha_clip.glsl:11-29);
 *   - the seven shader stages of the reference's pipeline
 *     (src/main.rs:243-305) run per pixel: ray_generation
 *     (shaders/ray-tracing/src/lib.rs:94-191), primary_ray_miss (:40-51),
 *     shadow_ray_miss (:33-36), closest_hit_portal (:300-312),
 *     closest_hit_textured (shaders/closest_hit_textured.glsl:174-226 with
 *     shaders/pbr.glsl:174-211), closest_hit_mirror
 *     (shaders/closest_hit_mirror.glsl:11-29).
 *
 * Device addresses in the ABI structs (rtg_model_info, rtg_geometry_info,
 * rtg_texture, rtg_geometry, instance BLAS references) are host pointers here.
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>
#include "../include/rtg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Builds a BLAS blob in host memory (free with orc_free). flags: the
 * RTG_BUILD_* flags of rtg_blas_build (PREFER_FAST_TRACE adds treelet
 * restructuring to the LBVH). */
int orc_blas_build(const rtg_geometry* geoms, uint32_t geom_count, uint32_t flags, void** out_blob,
                   uint64_t* out_size);
/* Builds a TLAS blob; instance BLAS references are orc BLAS blob pointers. */
int orc_tlas_build(const rtg_instance* instances, uint32_t count, uint32_t flags, void** out_blob,
                   uint64_t* out_size);
/* Test hook: the binary tree a build over n >= 2 primitive boxes (6 floats
 * each: lo xyz, hi xyz) collapses, after its treelet passes: internal node i
 * has children lc[i], rc[i] (bit 31 set: leaf k, the primitive order[k]). */
int orc_debug_binary_tree(const float* boxes, uint32_t n, uint32_t flags, uint32_t* lc, uint32_t* rc,
                          uint32_t* order);
/* In-place refit of a TLAS blob with new instance records (same count). */
int orc_tlas_refit(void* tlas_blob, const rtg_instance* instances, uint32_t count);
void orc_free(void* blob);
int orc_as_get_info(const void* blob, rtg_as_info* out);

/* Stack entries per lane the device keeps in LDS (its
   rtg_tuning.stack_lds_entries; 16 by default; the single-ray occlusion
   queries keep twice as many): sets which pushes the spill counters count. */
void orc_set_stack_lds(uint32_t entries);
/* 1 (the default): an image's shadow rays are walked in pairs of one path,
   as the device traces them; 0: each alone with the single-ray occlusion
   query. The visibility, hence the image, is the same either way; only the
   occlusion counters differ. */
void orc_set_shadow_pairs(int on);
/* The visits of one shadow ray's walk, or of a pair walk (test hook; see
   orc_trace.c). */
int orc_debug_visits(const void* tlas, const rtg_scene_bindings* scene, const float o[3], const float d0[3],
                     const float d1[3], int pair, uint64_t* log, uint32_t cap, uint32_t* n);
/* Ray queries (rtg_trace_rays semantics). threads <= 0 means all cores. */
int orc_trace_rays(const void* tlas, const rtg_scene_bindings* scene, const rtg_ray* rays,
                   uint32_t count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                   uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded,
                   rtg_trace_counters* counters, int threads);
/* The same queries by testing every triangle of every instance (no BVH). */
int orc_trace_rays_brute(const void* tlas, const rtg_scene_bindings* scene, const rtg_ray* rays,
                         uint32_t count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                         uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded);

/* One frame of one image job (rtg_trace_image semantics). If pixel_list is
 * non-NULL only those pixels (y * width + x) are rendered. threads <= 0 means
 * all cores. */
int orc_trace_image(const rtg_image_job* job, const uint32_t* pixel_list, uint32_t pixel_count,
                    int threads, rtg_trace_counters* counters);

/* Exposed shading helpers for known-answer tests. */
float orc_v_smith_ggx_correlated(float NoV, float NoL, float roughness);
float orc_d_ggx(float NoH, float roughness);
void orc_brdf(const float normal[3], const float view[3], const float light[3],
              float perceptual_roughness, const float base_colour[3], float metallic,
              float reflectance, float light_intensity, float out[3]);
void orc_linear_to_srgb(const float in[3], float out[3]);
void orc_heatmap_temperature(float heat, float out[3]);
void orc_sample_directional_light(const float rng[2], const float centre[3], float radius,
                                  float out[3]);
void orc_sample_texture(const rtg_texture* tex, const float srgb_lut[256], float u, float v,
                        float out[4]);
void orc_srgb_lut(float lut[256]);
float orc_cos(float x);
float orc_sin(float x);
float orc_pow(float x, float y);
float orc_log2(float x);
float orc_exp2(float y);
void orc_watertight(const float o[3], const float d[3], const float v0[3], const float v1[3],
                    const float v2[3], float tmin, float tmax, int* hit, float* t, float* u,
                    float* v);
uint32_t orc_morton3(float cx, float cy, float cz, const float lo[3], const float hi[3]);

/* The stage functions the per-pixel trace uses (exported for the tests that
 * pin each one). Matrices are the 3x4 row-major object-to-world transform of
 * the instance; points and vectors are 3 floats. */
/* ray_generation's camera ray for pixel (px, py) + (jx, jy) (0.5 = centre). */
void orc_camera_ray(const rtg_uniforms* uniforms, uint32_t width, uint32_t height, uint32_t px, uint32_t py,
                    float jx, float jy, float origin[3], float dir[3]);
/* primary_ray_miss: sun disk or sky colour. */
void orc_primary_miss(const float dir[3], const float sun_dir[3], float sun_radius, float out[3]);
/* The shadow-terminator-fixed hit point in model space (pos, nrm: the
 * triangle's three vertex positions and normals; bc: barycentrics). */
void orc_terminator_fix(const float pos[9], const float nrm[9], const float bc[3], float out[3]);
/* The two animated blue-noise numbers of shadow ray i of pixel (px, py). */
void orc_blue_noise(const rtg_texture* blue_noise, const float srgb_lut[256], uint32_t px, uint32_t py,
                    uint32_t i, uint32_t frame, float out[2]);
/* The world-space shading normal (map_rgb: the normal map's sample, or NULL). */
void orc_shading_normal(const float o2w[12], const float nrm[3], const float pos[9], const float uv[6],
                        const float* map_rgb, float out[3]);
/* closest_hit_textured's colour: BRDF * sun_factor + 0.1 * base. */
void orc_textured_colour(const float normal[3], const float ray_dir[3], const float sun_dir[3], float roughness,
                         const float base[3], float metallic, float sun_factor, float out[3]);
/* closest_hit_mirror and closest_hit_portal continuation rays. */
void orc_mirror(const float o2w[12], const float normal[3], const float origin[3], const float dir[3], float t,
                float new_origin[3], float new_dir[3]);
void orc_portal(const float origin[3], const float dir[3], float t, float new_origin[3], float new_dir[3]);
/* Heatmap mode: heatmap_temperature(cycles / 1e6) + rgb * 1e-6 per pixel. */
void orc_heatmap_pixels(const uint32_t* cycles, const float* rgb, uint32_t n, float* out);

#ifdef __cplusplus
}
#endif
#endif
