/*
 * rtg.h -- C ABI of the MI355X trace step of expenses/ray-tracing-gallery.
 *
 * The reference builds its acceleration structures and traces rays through
 * the Vulkan KHR ray-tracing extensions, called from Rust through `ash`.
 * This header is the drop-in boundary that replaces those calls:
 *
 *   vkGetAccelerationStructureBuildSizesKHR  src/util_structs.rs:197-206,321-330;
 *                                            src/util_functions.rs:486-495
 *   vkCmdBuildAccelerationStructuresKHR      src/util_structs.rs:269-274 (BUILD),
 *                                            src/util_structs.rs:345-354 (UPDATE)
 *   vkCmdCopyAccelerationStructureKHR        src/util_structs.rs:389-395 (CLONE)
 *   vkGetBufferDeviceAddress (AS address)    src/gpu_structs.rs:50;
 *                                            src/command_buffer_recording.rs:112
 *   vkCmdTraceRaysKHR(raygen,miss,hit,W,H,1) src/command_buffer_recording.rs:116-126
 *
 * together with the seven shader stages that the trace call runs
 * (src/main.rs:243-305): ray_generation, primary_ray_miss, shadow_ray_miss,
 * closest_hit_portal (shaders/ray-tracing/src/lib.rs), closest_hit_textured,
 * closest_hit_mirror and any_hit_alpha_clip (the GLSL files under shaders/).
 *
 * Conventions
 *  - Every entry point returns RTG_OK (0) or a negative rtg_status; the
 *    message of the last failure on the calling thread is rtg_last_error().
 *  - Device pointers are plain 64-bit device addresses (hipMalloc memory).
 *    Host arrays are borrowed for the duration of the call only.
 *  - Work is stream ordered on the hipStream_t passed as `stream`
 *    (NULL = the null stream). Build entry points may synchronise their
 *    stream; refit, clone and trace never synchronise.
 *  - A context is used by one host thread at a time, as the reference records
 *    every command on its event-loop thread (src/main.rs:728).
 *  - The library owns acceleration-structure storage and its scratch arena;
 *    callers own vertex/index/texture/instance/output buffers.
 */
#ifndef RTG_H
#define RTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTG_ABI_VERSION 1u

typedef enum rtg_status {
    RTG_OK = 0,
    RTG_ERROR_INVALID_ARGUMENT = -1,
    RTG_ERROR_OUT_OF_MEMORY = -2,
    RTG_ERROR_DEVICE = -3,
    RTG_ERROR_LIMIT = -4,
    RTG_ERROR_NOT_FOUND = -5,
    RTG_ERROR_BUDGET = -6,
    RTG_ERROR_STATE = -7,
    /* a served request's trace ran but set RTG_STATUS_STACK_OVERFLOW or
       RTG_STATUS_QUEUE_RANGE: its outputs are not valid (rtg_completion.status) */
    RTG_ERROR_RESULT = -8,
    /* a served request was cancelled before its batch launched (rtg_completion.status) */
    RTG_ERROR_CANCELLED = -9,
    /* a served request's deadline passed before its batch launched (rtg_completion.status) */
    RTG_ERROR_TIMEOUT = -10
} rtg_status;

/* ------------------------------------------------------------------------ */
/* ABI structs shared with the reference (layouts are fixed; see asserts).   */
/* ------------------------------------------------------------------------ */

/* vk::AccelerationStructureInstanceKHR twin, 64 bytes
 * (src/gpu_structs.rs:17-25, ::new :27-53, transpose_matrix_for_instance
 * :55-58). `transform` is the row-major 3x4 object-to-world matrix.
 * custom_index_and_mask = custom_index (low 24 bits) | mask << 24
 * (vk::Packed24_8, gpu_structs.rs:45).
 * sbt_offset_and_flags = hit-group offset (low 24) | instance flags << 24
 * (gpu_structs.rs:35-49; HitShader Textured=0, Mirror=1, Portal=2,
 * src/main.rs:46-50).
 * acceleration_structure_reference = the rtg_blas address. */
typedef struct rtg_instance {
    float transform[3][4];
    uint32_t custom_index_and_mask;
    uint32_t sbt_offset_and_flags;
    uint64_t acceleration_structure_reference;
} rtg_instance;

#define RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE 0x1u
#define RTG_INSTANCE_TRIANGLE_FLIP_FACING 0x2u
#define RTG_INSTANCE_FORCE_OPAQUE 0x4u
#define RTG_INSTANCE_FORCE_NO_OPAQUE 0x8u

/* Hit groups of the reference pipeline (src/main.rs:46-50, :289-305). */
#define RTG_HIT_TEXTURED 0u
#define RTG_HIT_MIRROR 1u
#define RTG_HIT_PORTAL 2u
#define RTG_HIT_GROUP_COUNT 3u

/* shared-structs/src/lib.rs:7-19 (Uniforms), GLSL mirror
 * shaders/hit_shader_common.glsl:3-13. Column-major 4x4 matrices. 176 B. */
typedef struct rtg_uniforms {
    float view_inverse[16];
    float proj_inverse[16];
    float sun_dir[3];
    uint32_t _padding0;
    float sun_radius;
    uint32_t blue_noise_texture_index;
    uint32_t ggx_lut_texture_index;
    uint32_t frame_index;
    uint8_t show_heatmap;
    uint8_t _padding1[15];
} rtg_uniforms;

/* shared-structs/src/lib.rs:21-29 (ModelInfo). Device addresses of the
 * model's shared vertex streams (vec3 positions/normals, vec2 uvs, tightly
 * packed) and of its rtg_geometry_info table. 32 B. */
typedef struct rtg_model_info {
    uint64_t position_buffer_address;
    uint64_t normal_buffer_address;
    uint64_t uv_buffer_address;
    uint64_t geometry_info_address;
} rtg_model_info;

/* shared-structs/src/lib.rs:31-47 (GeometryInfo + GeometryImages). The
 * index buffer holds one uvec3 of u32 vertex indices per triangle. A
 * negative normal_map_image_index means "no normal map". 24 B. */
typedef struct rtg_geometry_info {
    uint64_t index_buffer_address;
    uint32_t diffuse_image_index;
    uint32_t metallic_roughness_image_index;
    int32_t normal_map_image_index;
    uint32_t _padding;
} rtg_geometry_info;

/* ------------------------------------------------------------------------ */
/* Textures: one entry of the bindless sampler2D array                      */
/* (hit_shader_common.glsl:67; ImageManager src/util_structs.rs:1296-1377). */
/* ------------------------------------------------------------------------ */
#define RTG_FORMAT_RGBA8_UNORM 0u  /* R8G8B8A8_UNORM: blue noise, LUT, normal maps */
#define RTG_FORMAT_RGBA8_SRGB 1u   /* R8G8B8A8_SRGB: diffuse and metallic-roughness */
#define RTG_FORMAT_RGBA32_FLOAT 2u /* R32G32B32A32_SFLOAT: 1x1 fallback colours */

#define RTG_FILTER_NEAREST 0u /* the default sampler (util_structs.rs:1311-1315) */
#define RTG_FILTER_LINEAR 1u  /* glTF magFilter != NEAREST (util_structs.rs:954-955) */

typedef struct rtg_texture {
    uint64_t data;   /* device address of width*height texels, row-major, no padding */
    uint32_t width;
    uint32_t height;
    uint32_t format; /* RTG_FORMAT_* */
    uint32_t filter; /* RTG_FILTER_*; wrap is always REPEAT, LOD 0 */
} rtg_texture;

/* ------------------------------------------------------------------------ */
/* Context                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct rtg_ctx rtg_ctx;

/* Creates a context on HIP device `device` (makes it current). */
int rtg_create(rtg_ctx** out, int device);
void rtg_destroy(rtg_ctx* ctx);
/* Message of the last failing call on this thread ("" if none). */
const char* rtg_last_error(void);
/* RTG_ABI_VERSION of the built library. */
uint32_t rtg_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Bottom-level acceleration structures                                     */
/* ≙ AccelerationStructure::build_blas (src/util_structs.rs:140-224)       */
/* ------------------------------------------------------------------------ */

/* One triangle geometry (vk::AccelerationStructureGeometryTrianglesDataKHR,
 * util_structs.rs:160-176): R32G32B32_SFLOAT vertices with `vertex_stride`
 * bytes between vertices, UINT32 indices (3 per triangle). */
typedef struct rtg_geometry {
    uint64_t vertex_data;   /* device address */
    uint32_t vertex_stride; /* bytes, >= 12 and a multiple of 4 */
    uint32_t max_vertex;    /* highest valid vertex index (num_vertices - 1, or 0) */
    uint64_t index_data;    /* device address */
    uint32_t index_count;   /* multiple of 3 */
    uint32_t flags;         /* RTG_GEOMETRY_OPAQUE */
} rtg_geometry;

#define RTG_GEOMETRY_OPAQUE 0x1u /* vk::GeometryFlagsKHR::OPAQUE (util_structs.rs:154-158) */

/* Build flags (vk::BuildAccelerationStructureFlagsKHR). PREFER_FAST_TRACE is
 * what the reference passes (util_structs.rs:194, util_functions.rs:480-483):
 * it adds five passes of treelet restructuring (SAH-optimal topology of
 * 7-leaf treelets) to the LBVH, about a quarter fewer node visits per ray on
 * instanced scenes; without it the build is the plain LBVH. ALLOW_UPDATE is
 * accepted for TLASes, which can always be refitted. */
#define RTG_BUILD_ALLOW_UPDATE 0x1u
#define RTG_BUILD_PREFER_FAST_TRACE 0x4u

typedef uint64_t rtg_blas; /* device address of the BLAS (instance.acceleration_structure_reference) */
typedef uint64_t rtg_tlas; /* device address of the TLAS (PushConstantBufferAddresses.acceleration_structure) */

/* ≙ vkGetAccelerationStructureBuildSizesKHR. */
typedef struct rtg_build_sizes {
    uint64_t acceleration_structure_size;
    uint64_t build_scratch_size;
    uint64_t update_scratch_size;
} rtg_build_sizes;

int rtg_blas_query_sizes(rtg_ctx* ctx, const rtg_geometry* geometries, uint32_t geometry_count,
                         rtg_build_sizes* out);
/* Builds a BLAS over `geometry_count` geometries (host array of descriptors
 * whose buffers are on the device). Geometry index = position in the array,
 * primitive index = triangle index within its geometry. */
int rtg_blas_build(rtg_ctx* ctx, const rtg_geometry* geometries, uint32_t geometry_count,
                   uint32_t build_flags, rtg_blas* out, void* stream);
int rtg_blas_destroy(rtg_ctx* ctx, rtg_blas blas);

/* ------------------------------------------------------------------------ */
/* Top-level acceleration structures                                        */
/* ≙ build_tlas (src/util_functions.rs:453-510),                            */
/*   AccelerationStructure::update_tlas (src/util_structs.rs:285-357),      */
/*   AccelerationStructure::clone (src/util_structs.rs:359-405)             */
/* ------------------------------------------------------------------------ */

/* `instances` is a device array of `count` rtg_instance records. */
int rtg_tlas_query_sizes(rtg_ctx* ctx, uint32_t count, rtg_build_sizes* out);
int rtg_tlas_build(rtg_ctx* ctx, const rtg_instance* instances, uint32_t count,
                   uint32_t build_flags, rtg_tlas* out, void* stream);
/* In-place UPDATE (src == dst, util_structs.rs:318-319): new transforms and
 * BLAS references with the same instance count; topology is kept. One kernel. */
int rtg_tlas_refit(rtg_ctx* ctx, rtg_tlas tlas, const rtg_instance* instances, uint32_t count,
                   void* stream);
/* CLONE mode copy into a new TLAS (src/main.rs:622). */
int rtg_tlas_clone(rtg_ctx* ctx, rtg_tlas src, rtg_tlas* out, void* stream);
int rtg_tlas_destroy(rtg_ctx* ctx, rtg_tlas tlas);

/* Statistics of a built structure (host copy of its header; synchronises). */
typedef struct rtg_as_info {
    uint32_t kind;        /* 0 = BLAS, 1 = TLAS */
    uint32_t node_count;  /* 128-byte 4-wide nodes */
    uint32_t prim_count;  /* triangles (BLAS) or instances (TLAS) */
    uint32_t geometry_count;
    uint32_t max_depth;   /* longest root-to-leaf path, in nodes */
    uint32_t node_width;  /* children per node */
    float bounds[6];      /* root box lo.xyz, hi.xyz */
    uint64_t size_bytes;
} rtg_as_info;
int rtg_as_get_info(rtg_ctx* ctx, uint64_t as_address, rtg_as_info* out);

/* ------------------------------------------------------------------------ */
/* Ray queries (the traversal core without shading)                         */
/* ------------------------------------------------------------------------ */

typedef struct rtg_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} rtg_ray; /* 32 B */

/* Closest-hit record. instance = gl_InstanceID (index into the TLAS's
 * instance array), geometry = gl_GeometryIndexEXT, primitive = gl_PrimitiveID,
 * (u, v) = hitAttributeEXT barycentrics of vertices 1 and 2, t = gl_HitTEXT.
 * A miss has instance == RTG_NO_HIT and t == tmax of the ray. */
typedef struct rtg_hit {
    float t;
    uint32_t instance;
    uint32_t primitive;
    uint32_t geometry;
    float u;
    float v;
} rtg_hit; /* 24 B */

#define RTG_NO_HIT 0xFFFFFFFFu

/* gl_RayFlags*EXT that the traversal honours. */
#define RTG_RAY_FLAG_NONE 0x00u
#define RTG_RAY_FLAG_OPAQUE 0x01u
#define RTG_RAY_FLAG_NO_OPAQUE 0x02u
#define RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT 0x04u
#define RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER 0x08u

/* Scene bindings of a trace: ≙ the push constants
 * (PushConstantBufferAddresses, shared-structs/src/lib.rs:49-56) and the
 * bindless texture descriptor set (src/main.rs:416-460). Needed by the
 * any-hit stage (alpha clip) and by the hit shaders. */
typedef struct rtg_scene_bindings {
    uint64_t model_info;   /* device address of rtg_model_info[model_count] */
    uint64_t textures;     /* device address of rtg_texture[texture_count] */
    uint32_t model_count;
    uint32_t texture_count;
} rtg_scene_bindings;

/* Traces `count` rays (device array) against `tlas`.
 * mode RTG_QUERY_CLOSEST writes `hits[count]`; RTG_QUERY_OCCLUSION writes
 * `occluded[count]` (1 = some accepted hit in [tmin, tmax], else 0).
 * ray_flags as gl_RayFlags; cull_mask as traceRayEXT's cullMask;
 * sbt_offset selects the hit group (instance offset + sbt_offset) whose
 * any-hit stage (alpha clip on non-opaque geometry) runs. `scene` may be
 * NULL when every geometry is opaque. */
#define RTG_QUERY_CLOSEST 0u
#define RTG_QUERY_OCCLUSION 1u
#define RTG_MAX_QUERY_RAYS 0x80000000u /* rays per call (RTG_ERROR_LIMIT beyond) */
int rtg_trace_rays(rtg_ctx* ctx, rtg_tlas tlas, const rtg_scene_bindings* scene,
                   const rtg_ray* rays, uint32_t count, uint32_t mode, uint32_t ray_flags,
                   uint32_t cull_mask, uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded,
                   void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rtg_trace_rays with the same arguments, results and
 * errors, whose scratch (the rays' feed counter and the spill stacks of the
 * persistent traversal workgroups) is a file-scope device array of
 * RTG_STATIC_SCRATCH_BYTES that each GPU holds until the process ends,
 * instead of memory the context allocates. It calls no hipMalloc, hipFree or
 * allocating routine and is never timed. Its grid holds at most
 * RTG_STATIC_QUERY_BLOCKS workgroups of 256 lanes (the spill stacks the array
 * has room for); the launch shape never changes results. Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two threads, two contexts) race and must be serialised by the caller. */
#define RTG_STATIC_QUERY_BLOCKS 512u
#define RTG_STATIC_SCRATCH_BYTES 92274944u /* 256 + 512 * 256 lanes * 704 B of spill stack */
int rtg_trace_rays_static(rtg_ctx* ctx, rtg_tlas tlas, const rtg_scene_bindings* scene,
                          const rtg_ray* rays, uint32_t count, uint32_t mode, uint32_t ray_flags,
                          uint32_t cull_mask, uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded,
                          void* stream);

/* ------------------------------------------------------------------------ */
/* The trace step: ≙ vkCmdTraceRaysKHR(W, H, 1) with the reference's        */
/* seven stages (PerFrameResources::record,                                 */
/* src/command_buffer_recording.rs:79-126).                                  */
/* ------------------------------------------------------------------------ */

typedef struct rtg_trace_opts {
    uint32_t shadow_rays; /* cone shadow rays per textured hit; reference: 2 (closest_hit_textured.glsl:195) */
    uint32_t max_traces;  /* traces per path; reference: 3 (shaders/ray-tracing/src/lib.rs:144) */
    uint32_t spp;         /* samples per pixel; 1 = pixel centre as in lib.rs:126 */
    uint32_t flags;       /* RTG_TRACE_* */
} rtg_trace_opts;

/* Heatmap mode: colour = heatmap_temperature(cycles / 1e6) (lib.rs:174-186);
 * also enabled by uniforms.show_heatmap. */
#define RTG_TRACE_HEATMAP 0x1u
/* out_rgba8 in B8G8R8A8 byte order instead of R8G8B8A8: the storage image
 * takes the surface format, B8G8R8A8_UNORM when the surface offers it
 * (src/util_functions.rs:67-73, src/main.rs:606-611), and is copied to the
 * swapchain image as is (src/command_buffer_recording.rs:166-179). */
#define RTG_TRACE_OUTPUT_BGRA8 0x2u

typedef struct rtg_image_job {
    rtg_tlas tlas;
    rtg_scene_bindings scene;
    rtg_uniforms uniforms;
    uint32_t width;
    uint32_t height;
    rtg_trace_opts opts;
    uint32_t* out_rgba8;       /* device, width*height packed RGBA8 (sRGB-encoded, A = 255); may be NULL */
    float* out_rgb;            /* device, width*height*3 linear colour before encoding; may be NULL */
    rtg_hit* out_primary_hits; /* device, width*height*spp first-trace hits; may be NULL */
} rtg_image_job;

/* One frame of one image (the plain single-request path). */
int rtg_trace_image(rtg_ctx* ctx, const rtg_image_job* job, void* stream);
/* Many images of uneven size (different scenes, resolutions, spp) in one set
 * of persistent launches. Results equal tracing each job alone. */
int rtg_trace_batch(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t job_count, void* stream);

/* Ray and traversal counters of the last rtg_trace_image / rtg_trace_batch /
 * rtg_trace_rays call made with counters enabled (synchronises). */
typedef struct rtg_trace_counters {
    uint64_t primary_rays;     /* first-trace rays */
    uint64_t bounce_rays;      /* mirror/portal continuation rays */
    uint64_t shadow_rays;      /* occlusion rays */
    /* closest-hit traversals */
    uint64_t node_visits;      /* 128-byte node records fetched */
    uint64_t triangle_tests;   /* triangle records tested */
    uint64_t instance_visits;  /* TLAS leaf -> BLAS transitions */
    uint64_t anyhit_calls;     /* alpha-clip evaluations */
    /* occlusion traversals: rtg_trace_rays counts per ray; an image's
       shadow rays are walked in pairs of one path's rays (rays 2j and 2j+1,
       an odd count's last alone), and a pair's node visit, triangle test or
       instance entry counts once for both rays (one record fetch);
       any-hit calls count per ray */
    uint64_t occ_node_visits;
    uint64_t occ_triangle_tests;
    uint64_t occ_instance_visits;
    uint64_t occ_anyhit_calls;
    /* of the shadow-ray and occlusion counts above, those traced on the
       bounce stream beside the first level's shadow rays
       (rtg_tuning.bounce_blocks; zero when every level runs in sequence) */
    uint64_t side_shadow_rays;
    uint64_t side_occ_node_visits;
    uint64_t side_occ_triangle_tests;
    uint64_t side_occ_instance_visits;
    uint64_t side_occ_anyhit_calls;
    /* of the closest-hit counts above, those of the bounce levels traced on
       the bounce stream (their rays are bounce_rays) */
    uint64_t side_node_visits;
    uint64_t side_triangle_tests;
    uint64_t side_instance_visits;
    uint64_t side_anyhit_calls;
    /* traversal-stack pushes that went past the part of the stack in LDS
       (rtg_tuning.stack_lds_entries 8-byte entries per lane, 16 by default;
       twice as many 4-byte entries for rtg_trace_rays occlusion queries) to
       the lane's spill area in device memory; the side_* pair is the bounce
       stream's share */
    uint64_t spill_pushes;
    uint64_t occ_spill_pushes;
    uint64_t side_spill_pushes;
    uint64_t side_occ_spill_pushes;
    /* textured hits shaded (closest_hit_textured runs; each casts the job's
       shadow rays), of them on the bounce stream, and the continuation
       (bounce) rays that the bounce stream's shading adds */
    uint64_t textured_hits;
    uint64_t side_textured_hits;
    uint64_t side_bounce_rays;
} rtg_trace_counters;
/* Enables traversal counters (slower kernels) for subsequent traces. */
int rtg_set_counters(rtg_ctx* ctx, int enable);
int rtg_get_counters(rtg_ctx* ctx, rtg_trace_counters* out);

/* Sticky device status bits of the context, read after all work queued so
 * far has finished (synchronises); clear != 0 resets them. A trace that sets
 * RTG_STATUS_STACK_OVERFLOW has rays whose traversal stack (96 entries) ran
 * out: their results are not valid. */
#define RTG_STATUS_BAD_INDEX 0x1u      /* BLAS build: an index above max_vertex */
#define RTG_STATUS_UNKNOWN_BLAS 0x2u   /* TLAS build/refit: a reference that is not a live BLAS */
#define RTG_STATUS_STACK_OVERFLOW 0x4u /* trace: traversal stack exhausted */
#define RTG_STATUS_QUEUE_RANGE 0x8u    /* trace: a ray-queue size or entry out of range (one frame
                                          state driven by two traces at once); dropped, results invalid */
int rtg_get_status(rtg_ctx* ctx, uint32_t* bits, int clear);

/* Per-kernel-class GPU time (HIP events on the launch stream), accumulated
 * while timing is enabled. kind: RTG_KERNEL_*; returns total ms and launches. */
#define RTG_KERNEL_RAYGEN 0u
#define RTG_KERNEL_TRACE_CLOSEST 1u
#define RTG_KERNEL_SHADE 2u
#define RTG_KERNEL_TRACE_OCCLUSION 3u
#define RTG_KERNEL_RESOLVE 4u
#define RTG_KERNEL_REFIT 5u
#define RTG_KERNEL_BATCH_GRAPH 6u /* one replay of an rtg_batch graph (all its kernels) */
#define RTG_KERNEL_BOUNCE_SIDE 7u /* the bounce levels and their shadow rays on the bounce stream,
                                     beside the first level's shadow rays (one span per frame) */
#define RTG_KERNEL_COUNT 8u
int rtg_set_timing(rtg_ctx* ctx, int enable);
int rtg_get_kernel_time(rtg_ctx* ctx, uint32_t kind, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------------ */
/* Frame driver: frames in flight over one scene.                           */
/* ≙ the reference's two Frame resources, each with a TLAS clone, a mapped  */
/* instance buffer and an output image (src/main.rs:600-668), and its frame */
/* loop: wait for the frame's fence, write the instance transforms, refit   */
/* that frame's TLAS, trace (src/main.rs:916-996, src/scene.rs:167-204).    */
/* ------------------------------------------------------------------------ */
typedef struct rtg_frames rtg_frames;
/* frames_in_flight (1..8) slots, each a copy of `tlas` with its own pinned
 * host instance array (seeded with the TLAS's instances) and frame state. */
int rtg_frames_create(rtg_ctx* ctx, rtg_tlas tlas, uint32_t frames_in_flight, rtg_frames** out);
/* Waits until the slot of `frame` (frame % frames_in_flight) has finished its
 * last frame and returns its host instance array (write this frame's
 * transforms there) and its TLAS. */
int rtg_frames_begin(rtg_frames* frames, uint64_t frame, rtg_instance** host_instances, rtg_tlas* slot_tlas);
/* On `stream`: uploads the slot's instances, refits its TLAS and traces `job`
 * with the slot's TLAS in place of job->tlas. */
int rtg_frames_submit(rtg_frames* frames, uint64_t frame, const rtg_image_job* job, void* stream);
int rtg_frames_destroy(rtg_frames* frames);

/* ------------------------------------------------------------------------ */
/* Launch tuning and the autotuner.                                         */
/* No counterpart in the reference (the Vulkan driver picks its own launch  */
/* shape for vkCmdTraceRaysKHR, command_buffer_recording.rs:116-126). These */
/* parameters change speed only: every setting gives identical results.     */
/* ------------------------------------------------------------------------ */
typedef struct rtg_tuning {
    uint32_t blocks_per_cu;    /* persistent traversal workgroups (256 lanes) per CU, 1..8 */
    uint32_t feed_chunk;       /* queue entries a wavefront takes per atomic, 64..1024, multiple of 64 */
    uint32_t refill_closest;   /* idle lanes (1..64) before a wavefront takes new closest-hit rays;
                                  0 = non-persistent kernel (grid-stride rounds) */
    uint32_t refill_occlusion; /* the same for occlusion rays */
    uint32_t path_order;       /* 0 = row-major paths, 1 = 8x8 pixel blocks per wavefront */
    uint32_t bounce_blocks;    /* persistent workgroups (0..4096) of a second stream that traces the
                                  bounce levels (on at most 128 of them) and the shadow rays they add
                                  (on all) beside the first level's shadow rays (taken from the
                                  blocks_per_cu grid, at most half of it); 192 by default;
                                  0 = every level in sequence on the caller's stream */
    uint32_t stack_lds_entries; /* traversal stack entries per lane kept in LDS (8..32, multiple of 4;
                                  0 = 16): a workgroup takes entries * 2 KB of LDS, deeper entries
                                  spill to global memory; the occlusion queries of rtg_trace_rays
                                  keep twice as many (half-size) entries */
    uint32_t reserved;         /* zero */
} rtg_tuning;
int rtg_get_tuning(rtg_ctx* ctx, rtg_tuning* out);
int rtg_set_tuning(rtg_ctx* ctx, const rtg_tuning* tuning);
/* Times the jobs (as one rtg_trace_batch, `repeats` timed runs per setting
 * after one warm-up) over a coordinate search of the tuning space starting
 * from the current setting, makes the fastest setting current and returns it
 * with its median time in ms. Writes the jobs' outputs (synchronises). */
int rtg_autotune(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t job_count, uint32_t repeats, rtg_tuning* best,
                 double* best_ms);

/* ------------------------------------------------------------------------ */
/* Graph-captured batches.                                                  */
/* No counterpart in the reference (one vkCmdTraceRaysKHR per frame,        */
/* command_buffer_recording.rs:116-126). A batch is a fixed set of jobs     */
/* (scenes and images of uneven size) whose whole trace step -- ray         */
/* generation, persistent traversal, shading, occlusion, resolve -- is      */
/* captured once into a HIP graph with the context's tuning and counter     */
/* setting of that moment and replayed by one graph launch. The batch owns  */
/* its frame state; it does not share it with other calls.                  */
/* ------------------------------------------------------------------------ */
typedef struct rtg_batch rtg_batch;
int rtg_batch_create(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t job_count, rtg_batch** out);
/* Replaces the job table (TLAS, bindings, uniforms, outputs) for the next
 * launches; width, height and opts of every job must be unchanged. The copy
 * is ordered on `stream`. */
int rtg_batch_update(rtg_batch* batch, const rtg_image_job* jobs, uint32_t job_count, void* stream);
int rtg_batch_launch(rtg_batch* batch, void* stream);
int rtg_batch_destroy(rtg_batch* batch);

/* ------------------------------------------------------------------------ */
/* Serving: render requests over resident scenes with continuous batching.  */
/* No counterpart in the reference, which renders one frame at a time       */
/* (src/main.rs:916-996); this is the request loop a renderer service puts  */
/* around the trace step. A worker thread forms batches from pending        */
/* requests whenever a slot (a stream with its own frame state) is free and */
/* the batch fits the slot's share of the HBM budget; batches on different  */
/* slots overlap on the device. Results do not depend on the batch a        */
/* request lands in. While a server runs, the context must not be used from */
/* other threads except for rtg_server_* calls.                             */
/* ------------------------------------------------------------------------ */
typedef struct rtg_server rtg_server;

typedef struct rtg_server_config {
    uint64_t hbm_budget_bytes;  /* device memory of all slots: each of the max_inflight slots holds the spill area
                                   of the persistent traversal grid (cu_count * blocks_per_cu * 256 lanes * 704 B)
                                   and the frame state and TLAS copies of its batches within its share
                                   (budget / max_inflight); rtg_server_create refuses (RTG_ERROR_BUDGET) a budget
                                   whose share cannot hold a spill area */
    uint32_t max_batch_jobs;    /* requests per batch; 1 = one request per launch */
    uint32_t max_inflight;      /* batches in flight at once (1..8) */
    uint64_t max_batch_samples; /* samples (pixels x spp) per batch, the first request excepted; 0 = no cap.
                                   Every request of a batch completes when the batch does, so a cap bounds
                                   how long small requests wait behind large ones. */
} rtg_server_config;

typedef struct rtg_request {
    uint32_t scene;                 /* id from rtg_server_add_scene */
    uint32_t width;
    uint32_t height;
    rtg_trace_opts opts;
    rtg_uniforms uniforms;
    const rtg_instance* instances;  /* device array with this request's transforms (a private TLAS copy is
                                       refitted with it, ≙ write_resources, src/scene.rs:167-204); NULL = as built */
    uint32_t* out_rgba8;            /* device outputs as in rtg_image_job */
    float* out_rgb;
    rtg_hit* out_primary_hits;
    uint64_t tag;                   /* returned in the completion */
    double deadline_s;              /* server clock (rtg_server_clock) by which the request's batch must
                                       launch, else it completes with RTG_ERROR_TIMEOUT without running;
                                       0 = none */
} rtg_request;

typedef struct rtg_completion {
    uint64_t ticket;
    uint64_t tag;
    int32_t status;    /* RTG_OK, RTG_ERROR_RESULT (the trace overflowed a traversal stack or dropped rays),
                          RTG_ERROR_CANCELLED, RTG_ERROR_TIMEOUT (neither ran: outputs untouched,
                          batch_size 0, start_s = done_s) or the launch error */
    uint32_t batch_size;
    double submit_s;   /* server clock (seconds since rtg_server_create) */
    double start_s;    /* its batch was launched */
    double done_s;     /* the worker saw the batch finish */
} rtg_completion;

typedef struct rtg_server_stats {
    uint64_t requests;
    uint64_t batches;
    uint64_t rejected;
    uint64_t max_batch;
    uint64_t max_inflight_seen;
    uint64_t peak_bytes; /* most device bytes the slots held at once (frame state, spill areas, TLAS copies) */
} rtg_server_stats;

/* The server traces with the context's launch tuning of this moment. */
int rtg_server_create(rtg_ctx* ctx, const rtg_server_config* cfg, rtg_server** out);
/* Registers a resident scene: its TLAS (the template copied per request) and
 * its bindings. Scenes stay resident until the server is destroyed. */
int rtg_server_add_scene(rtg_server* srv, rtg_tlas tlas, const rtg_scene_bindings* scene, uint32_t* out_id);
/* Queues a request (never blocks on the device). RTG_ERROR_BUDGET if it can
 * never fit one slot of the budget. */
int rtg_server_submit(rtg_server* srv, const rtg_request* req, uint64_t* out_ticket);
/* Cancels a request that is still pending (its batch not launched): it
 * completes at once with RTG_ERROR_CANCELLED and its outputs are untouched.
 * RTG_ERROR_NOT_FOUND if the ticket is not pending (launched, completed or
 * unknown). */
int rtg_server_cancel(rtg_server* srv, uint64_t ticket);
/* Copies up to `capacity` completions; with wait_all, first waits until every
 * submitted request has completed. */
int rtg_server_poll(rtg_server* srv, rtg_completion* out, uint32_t capacity, uint32_t* count, int wait_all);
/* Submits requests[i] at arrival_s[i] seconds after the call (a native
 * arrival clock), waits for all of them and writes their completions in
 * request order. */
int rtg_server_replay(rtg_server* srv, const rtg_request* requests, const double* arrival_s, uint32_t count,
                      rtg_completion* out);
double rtg_server_clock(const rtg_server* srv);
int rtg_server_get_stats(const rtg_server* srv, rtg_server_stats* out);
int rtg_server_destroy(rtg_server* srv);

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(rtg_instance) == 64, "instance record is 64 bytes");
static_assert(offsetof(rtg_instance, custom_index_and_mask) == 48, "instance layout");
static_assert(offsetof(rtg_instance, acceleration_structure_reference) == 56, "instance layout");
static_assert(sizeof(rtg_uniforms) == 176, "Uniforms is 176 bytes");
static_assert(offsetof(rtg_uniforms, sun_dir) == 128, "Uniforms layout");
static_assert(offsetof(rtg_uniforms, sun_radius) == 144, "Uniforms layout");
static_assert(offsetof(rtg_uniforms, frame_index) == 156, "Uniforms layout");
static_assert(offsetof(rtg_uniforms, show_heatmap) == 160, "Uniforms layout");
static_assert(sizeof(rtg_model_info) == 32, "ModelInfo is 32 bytes");
static_assert(sizeof(rtg_geometry_info) == 24, "GeometryInfo is 24 bytes");
static_assert(sizeof(rtg_ray) == 32, "ray is 32 bytes");
static_assert(sizeof(rtg_hit) == 24, "hit is 24 bytes");
static_assert(sizeof(rtg_texture) == 24, "texture descriptor is 24 bytes");
static_assert(sizeof(rtg_tuning) == 32, "tuning is 32 bytes");
static_assert(sizeof(rtg_server_config) == 24, "server config is 24 bytes");
#else
_Static_assert(sizeof(rtg_instance) == 64, "instance record is 64 bytes");
_Static_assert(sizeof(rtg_uniforms) == 176, "Uniforms is 176 bytes");
_Static_assert(sizeof(rtg_model_info) == 32, "ModelInfo is 32 bytes");
_Static_assert(sizeof(rtg_geometry_info) == 24, "GeometryInfo is 24 bytes");
_Static_assert(sizeof(rtg_ray) == 32, "ray is 32 bytes");
_Static_assert(sizeof(rtg_hit) == 24, "hit is 24 bytes");
#endif

#endif /* RTG_H */
