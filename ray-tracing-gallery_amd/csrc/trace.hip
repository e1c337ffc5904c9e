// trace.hip -- the trace step on gfx950: ray queries and the wavefront
// pipeline behind rtg_trace_image / rtg_trace_batch.
//
// vkCmdTraceRaysKHR(W, H, 1) (src/command_buffer_recording.rs:116-126) runs
// ray_generation once per pixel; the pixel's bounce loop (up to 3 traces,
// shaders/ray-tracing/src/lib.rs:144-172) calls the hit and miss stages, and
// closest_hit_textured traces its shadow rays (closest_hit_textured.glsl:195).
// Here each stage is a kernel over a queue of rays:
//   raygen -> [trace closest -> shade] x max_traces -> trace occlusion -> resolve
// Shade appends continuation rays (mirror, portal) and shadow rays to queues
// with one atomic per wavefront (ballot/prefix-sum compaction). A path's
// result depends only on its own rays, so queue order never changes results.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "rtg_device.hpp"
#include "rtg_internal.hpp"

namespace rtg {

constexpr int kTraceThreads = 256;
constexpr int kCountSlots = 16;
constexpr int kFeedLevel = 8;    // counts[kFeedLevel + level]: feed counter of a trace level
constexpr int kFeedShadow = 12;  // counts[kFeedShadow]: feed counter of the shadow queue
constexpr int kShadowFirst = 5;      // counts[5]: shadow rays of trace level 0
constexpr int kFeedShadowRest = 13;  // counts[13]: feed counter of the shadow rays after them
constexpr int kStatusSlot = 14;      // counts[14]: status bits of the trace (merged into the context's)
constexpr int kPartAll = 0, kPartFirst = 1, kPartRest = 2;  // parts of the shadow queue
// path of a shadow-queue padding slot (after an odd count of a path's rays)
constexpr uint32_t kNoPath = 0xFFFFFFFFu;

// Order of an image's pixels along its paths: bands of 8 rows, each band
// walked column by column, so that the 64 lanes of a wavefront cover an 8x8
// block of pixels (coherent rays). A bijection of [0, W*H) for every W, H.
// order 0 is plain row-major.
__device__ inline void order_to_pixel(uint32_t t, uint32_t W, uint32_t H, uint32_t& x, uint32_t& y,
                                      uint32_t order) {
    if (order == 0) {
        x = t % W;
        y = t / W;
        return;
    }
    uint32_t full = (H / 8) * 8;
    uint32_t fb = full * W;
    if (t < fb) {
        uint32_t band = t / (8 * W);
        uint32_t i = t - band * 8 * W;
        x = i / 8;
        y = band * 8 + (i & 7);
    } else {
        uint32_t h = H - full;
        uint32_t i = t - fb;
        x = i / h;
        y = full + i % h;
    }
}

// A job's path index local = t * spp + s: pixel-order index t and sample s.
// Path indices are 32-bit (queue entries hold them), so 32-bit arithmetic.
__device__ inline void split_sample(uint32_t local, uint32_t spp, uint32_t& t, uint32_t& s) {
    if (spp == 1) {
        t = local;
        s = 0;
        return;
    }
    t = local / spp;
    s = local - t * spp;
}

struct JobDev {
    const uint8_t* tlas;
    SceneView scene;
    float view_inverse[16];
    float proj_inverse[16];
    float sun[3];
    float sun_radius;
    uint32_t bn_index;
    uint32_t frame_index;
    uint32_t width, height, spp, shadow_rays, max_traces, flags;
    uint64_t sample_base;  // first path of the job
    uint64_t pixel_base;   // first pixel of the job in the batch's pixel numbering
    uint32_t* out_rgba8;
    float* out_rgb;
    rtg_hit* out_hits;
};

struct Queue {
    float4* o;  // origin.xyz, tmin
    float4* d;  // direction.xyz, tmax
    uint32_t* path;
};

struct Frame {
    const JobDev* jobs;
    uint32_t njobs;
    uint64_t paths;
    float4* path_rgb;    // final colour of paths that did not end on a textured hit
    float4* shade_a;     // NoL, combined.xyz of a textured hit
    float4* shade_b;     // ambient.xyz; w: bits of 1 + the path's first shadow ray if textured, else 0
    uint8_t* sflag;      // per shadow-queue entry: 1 if the shadow ray saw the sun
    uint32_t* cycles;    // heatmap cycle counts
    Queue q[2];
    uint4* hit_a;        // t bits, instance, primitive, geometry
    float2* hit_b;       // u, v
    Queue sq;            // shadow rays
    uint32_t* counts;    // [0..3] queue sizes per trace level, [4] shadow rays, [5] shadow
                         // rays of level 0, [8..11] feed counters per trace level,
                         // [12], [13] shadow feed counters
    uint64_t* counters;  // rtg_trace_counters
    uint2* spill;
    uint32_t order;           // pixel order along paths (0 row-major, 1 8x8 bands)
    int refill_closest;       // idle lanes that trigger a refill (0: rounds kernel)
    int refill_occlusion;
    uint32_t feed_chunk;      // queue entries a wave takes per atomic
    int stack_lds;            // closest-hit stack entries per lane in LDS (the launches' LDS holds them)
    uint32_t side;            // 1 on the bounce stream (counts its shadow rays apart)
    // capacities of the path and shadow queues: sizes read from the device
    // are clamped to them and entries checked, so that a frame state driven
    // by two traces at once drops rays (kStatusQueueRange) instead of
    // reading or writing outside its buffers
    uint32_t path_cap, shadow_cap;
    uint32_t* status;
};

__device__ inline void queue_range_error(const Frame& F) { atomicOr(F.status, (uint32_t)kStatusQueueRange); }

__device__ inline uint32_t find_job(const JobDev* jobs, uint32_t njobs, uint64_t p, bool by_pixel) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        uint64_t b = by_pixel ? jobs[mid].pixel_base : jobs[mid].sample_base;
        if (b <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Wavefront-aggregated allocation: every lane of the wave calls it (lanes
// that need nothing pass 0); one atomic per wave.
__device__ inline uint32_t wave_alloc(uint32_t want, uint32_t* counter) {
    const int lane = threadIdx.x & 63;
    uint32_t incl = want;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += v;
    }
    uint32_t total = (uint32_t)__shfl((int)incl, 63);
    uint32_t base = 0;
    if (lane == 63 && total) base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, 63);
    return base + incl - want;
}

__device__ inline void wave_add_counter(uint64_t* c, uint32_t v) {
    uint64_t x = v;
    for (int off = 32; off > 0; off >>= 1) x += (uint64_t)__shfl_xor((long long)x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd((unsigned long long*)c, (unsigned long long)x);
}

// Traversal counters of rtg_trace_counters: slots 3..6 for closest-hit
// queries, 7..10 for occlusion queries; the bounce stream counts its shadow
// rays in slot 11 (and slot 2), their occlusion queries in 12..15 and its
// closest-hit queries in 16..19 (added to 7..10 and 3..6 when read).
constexpr int kCountersClosest = 3;
constexpr int kCountersOcclusion = 7;
constexpr int kCounterSideShadow = 11;
constexpr int kCountersOcclusionSide = 12;
constexpr int kCountersClosestSide = 16;
// stack pushes past the LDS part: slots 20 (closest), 21 (occlusion) and,
// for the bounce stream, 22 (closest) and 23 (occlusion); shaded textured
// hits: slot 24 (all), 25 (bounce stream); slot 26: continuation rays the
// bounce stream's shading adds
constexpr int kCounterTextured = 24;
constexpr int kCounterSlots = 27;

__device__ inline int spill_slot(int base) {
    return base == kCountersClosest ? 20 : base == kCountersOcclusion ? 21 : base == kCountersClosestSide ? 22 : 23;
}

__device__ inline void flush_counters(uint64_t* counters, const TravCounters& c, int base) {
    wave_add_counter(&counters[base + 0], c.nodes);
    wave_add_counter(&counters[base + 1], c.tris);
    wave_add_counter(&counters[base + 2], c.insts);
    wave_add_counter(&counters[base + 3], c.anyhits);
    wave_add_counter(&counters[spill_slot(base)], c.spills);
}

// Lanes of a persistent traversal kernel keep tracing rays until the queue is
// empty: a lane that finishes asks the wave's feeder for the next entry once
// at least `refill_min` lanes are idle (or all are).
__device__ inline bool want_refill(uint64_t idle, const WaveFeeder& f, int refill_min) {
    return idle != 0 && !f.exhausted && ((int)__popcll(idle) >= refill_min || idle == ~0ull);
}

// ---------------------------------------------------------------------------
// Ray queries: rtg_trace_rays
// ---------------------------------------------------------------------------
template <bool OCC, bool COUNT>
__global__ __launch_bounds__(kTraceThreads) void k_trace_rays(const uint8_t* tlas, SceneView S, const rtg_ray* rays,
                                                              uint32_t n, uint32_t ray_flags, uint32_t cull_mask,
                                                              uint32_t sbt_offset, rtg_hit* hits, uint32_t* occ,
                                                              uint2* spill, uint32_t* feed_counter,
                                                              uint64_t* counters, uint32_t* status, int refill_min,
                                                              uint32_t feed_chunk, int stack_lds) {
    extern __shared__ uint2 s_stack[];
    const TlasView T = tlas_view(tlas);
    const uint32_t gtid = blockIdx.x * kTraceThreads + threadIdx.x;
    Stack<OCC> st;
    st.init(s_stack, stack_lds, reinterpret_cast<uint8_t*>(spill), gtid, threadIdx.x, kTraceThreads);
    TravCounters cnt = {0, 0, 0, 0, 0};
    Traverser<OCC, COUNT> tr;
    WaveFeeder feeder;
    feeder.init(feed_chunk);
    bool active = false;
    uint32_t ray = 0;
    uint32_t it = 0;  // iteration parity: leaf visits on even, instance entries on odd iterations
    while (true) {
        ++it;
        uint64_t idle = __ballot(!active);
        if (want_refill(idle, feeder, refill_min)) {
            uint32_t id = feeder.feed(idle, !active, feed_counter, n);
            if (id != ~0u) {
                const float4* rp = reinterpret_cast<const float4*>(rays + id);
                float4 a = rp[0], b = rp[1];
                tr.start(T, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, ray_flags, cull_mask, sbt_offset, st);
                ray = id;
                active = true;
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            int res = tr.step(T, S, st, cnt, (it & 1u) == 0u, (it & 1u) != 0u);
            if (res != kRayActive) {
                if (res == kRayOverflow) atomicOr(status, (uint32_t)kStatusStackOverflow);
                if (OCC) {
                    occ[ray] = res == kRayHit ? 1u : 0u;
                } else {
                    const HitRec& h = tr.best;
                    rtg_hit o;
                    o.t = h.t;
                    o.instance = h.inst;
                    o.primitive = h.prim;
                    o.geometry = h.geom;
                    o.u = h.u;
                    o.v = h.v;
                    hits[ray] = o;
                }
                active = false;
            }
        }
    }
    if (COUNT) flush_counters(counters, cnt, OCC ? kCountersOcclusion : kCountersClosest);
}

// ---------------------------------------------------------------------------
// Wavefront pipeline
// ---------------------------------------------------------------------------
__global__ void k_frame_begin(uint32_t* counts, uint64_t* counters, bool reset_counters) {
    if (threadIdx.x < kCountSlots) counts[threadIdx.x] = 0;
    if (reset_counters && threadIdx.x < kCounterSlots) counters[threadIdx.x] = 0;
}

// ORs a trace's status bits into the context's sticky ones.
__global__ void k_status_merge(const uint32_t* trace_status, uint32_t* ctx_status) {
    if (threadIdx.x == 0 && *trace_status) atomicOr(ctx_status, *trace_status);
}

__global__ void k_shadow_snapshot(uint32_t* counts) {
    if (threadIdx.x == 0) counts[kShadowFirst] = counts[4];
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void k_reset_counters(uint64_t* counters, uint64_t primary) {
    if (threadIdx.x < kCounterSlots) counters[threadIdx.x] = threadIdx.x == 0 ? primary : 0;
}

// ray_generation (shaders/ray-tracing/src/lib.rs:94-142): one path per
// (pixel, sample). Sample s > 0 of a multi-sample pixel moves the pixel
// centre by the R2 sequence; sample 0 of one sample is the reference's +0.5.
__global__ __launch_bounds__(256) void k_raygen(Frame F, bool count) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < F.paths; p += stride) {
        uint32_t j = find_job(F.jobs, F.njobs, p, false);
        const JobDev& J = F.jobs[j];
        uint32_t t, s, px, py;
        split_sample((uint32_t)(p - J.sample_base), J.spp, t, s);
        order_to_pixel(t, J.width, J.height, px, py, F.order);
        float jx = 0.5f, jy = 0.5f;
        if (J.spp > 1) {
            jx = fractf(0.5f + (float)s * 0.754877669f);
            jy = fractf(0.5f + (float)s * 0.569840296f);
        }
        float pcx = (float)px + jx, pcy = (float)py + jy;
        float tcx = pcx / (float)J.width, tcy = pcy / (float)J.height;
        float rcx = tcx * 2.0f - 1.0f, rcy = tcy * 2.0f - 1.0f;
        float zo[4] = {0.0f, 0.0f, 0.0f, 1.0f}, tmp[4];
        mat4_mul(J.view_inverse, zo, tmp);
        V3 origin = v3(tmp[0], tmp[1], tmp[2]);
        float tv[4] = {rcx, rcy, 1.0f, 1.0f};
        mat4_mul(J.proj_inverse, tv, tmp);
        V3 ld = normalize(v3(tmp[0], tmp[1], tmp[2]));
        float dv[4] = {ld.x, ld.y, ld.z, 0.0f};
        mat4_mul(J.view_inverse, dv, tmp);
        F.q[0].o[p] = make_float4(origin.x, origin.y, origin.z, 0.01f);
        F.q[0].d[p] = make_float4(tmp[0], tmp[1], tmp[2], 10000.0f);
        F.q[0].path[p] = (uint32_t)p;
        F.path_rgb[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        F.shade_b[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        F.cycles[p] = 0;
    }
    if (count && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&F.counters[0], F.paths);
}

// Issue priority of a kernel's waves (s_setprio: the order in which a SIMD
// picks among its ready waves, not a queue priority). With two frames in
// flight each frame slot runs a chain: closest hits, shading, the bounce
// levels with the shadow rays they add, resolve, then the slot's next frame.
// The first level's shadow rays run beside the bounce levels with time to
// spare and keep priority 0. The bounce stream's kernels take 2: its levels
// hold few rays, are bound by the latency of their longest ones and form the
// longest part of the chain (on the 1080p scene 1.81-1.90 instead of
// 2.06-2.26 ms per frame). The first level's closest hits and shading take
// 1: they share the GPU with the other frame's shadow rays, and ahead of
// them the two frames settled less often into the slower way of overlapping
// (1 of 16 runs on three boxes, against 3 of 12; the others 0.7 % faster).
__device__ __forceinline__ void bounce_stream_priority() { __builtin_amdgcn_s_setprio(2); }
__device__ __forceinline__ void frame_chain_priority() { __builtin_amdgcn_s_setprio(1); }

// Closest-hit traversal of trace level b (count from the device unless fixed).
// SIDE instances run the bounce levels on the bounce stream. SINGLE instances
// serve launches of one job: its TLAS view and scene stay in scalar registers
// instead of per lane, which lets k_trace_closest_one fit 96 VGPRs without
// spills, five waves per SIMD like the occlusion kernel, so any mix of the
// two frames' traversal waves fills a SIMD.
template <bool COUNT, bool SIDE, bool SINGLE>
__device__ __forceinline__ void trace_closest(Frame F, int level, const uint32_t* count_ptr, uint32_t fixed_count,
                                              uint32_t* status, uint2* s_stack) {
    if (SIDE) bounce_stream_priority();
    else frame_chain_priority();
    const uint32_t n = min(count_ptr ? *count_ptr : fixed_count, F.path_cap);
    const Queue Q = F.q[level & 1];
    const uint32_t gtid = blockIdx.x * kTraceThreads + threadIdx.x;
    Stack<false> st;
    st.init(s_stack, F.stack_lds, reinterpret_cast<uint8_t*>(F.spill), gtid, threadIdx.x, kTraceThreads);
    TravCounters cnt = {0, 0, 0, 0, 0};
    Traverser<false, COUNT> tr;
    WaveFeeder feeder;
    feeder.init(F.feed_chunk);
    bool active = false;
    uint32_t slot = 0;
    const JobDev* J = SINGLE ? &F.jobs[0] : nullptr;
    TlasView T;
    if (SINGLE) T = tlas_view(J->tlas);
    uint32_t c0 = 0;
    uint32_t it = 0;  // iteration parity: leaf visits on even, instance entries on odd iterations
    while (true) {
        ++it;
        uint64_t idle = __ballot(!active);
        if (want_refill(idle, feeder, F.refill_closest)) {
            uint32_t id = feeder.feed(idle, !active, &F.counts[kFeedLevel + level], n);
            if (id != ~0u) {
                slot = id;
                uint32_t path = Q.path[id];
                if (path >= F.paths) {
                    queue_range_error(F);
                    path = 0;  // traced as path 0's ray; the shade stage drops it
                }
                if (!SINGLE) {
                    J = &F.jobs[find_job(F.jobs, F.njobs, path, false)];
                    T = tlas_view(J->tlas);
                }
                float4 a = Q.o[id], b = Q.d[id];
                tr.start(T, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, RTG_RAY_FLAG_NONE, 0xffu, 0u, st);
                if (J->flags & RTG_TRACE_HEATMAP) c0 = (uint32_t)__builtin_readcyclecounter();
                active = true;
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            int res = tr.step(T, J->scene, st, cnt, (it & 1u) == 0u, (it & 1u) != 0u);
            if (res != kRayActive) {
                if (res == kRayOverflow) atomicOr(status, (uint32_t)kStatusStackOverflow);
                if (J->flags & RTG_TRACE_HEATMAP) {
                    // the path is read again rather than kept live through the traversal
                    const uint32_t path = Q.path[slot];
                    atomicAdd(&F.cycles[path < F.paths ? path : 0u], (uint32_t)__builtin_readcyclecounter() - c0);
                }
                const HitRec& h = tr.best;
                F.hit_a[slot] = make_uint4(f2u(h.t), h.inst, h.prim, h.geom);
                F.hit_b[slot] = make_float2(h.u, h.v);
                active = false;
            }
        }
    }
    if (COUNT) flush_counters(F.counters, cnt, SIDE ? kCountersClosestSide : kCountersClosest);
}

// Launches of one job (held to 96 VGPRs, five waves per SIMD).
template <bool COUNT, bool SIDE>
__global__ __launch_bounds__(kTraceThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_trace_closest_one(
    Frame F, int level, const uint32_t* count_ptr, uint32_t fixed_count, uint32_t* status) {
    extern __shared__ uint2 s_stack[];
    trace_closest<COUNT, SIDE, true>(F, level, count_ptr, fixed_count, status, s_stack);
}

// Launches of several jobs (batches): the job of each ray is looked up; at
// 96 VGPRs this variant would spill, so it keeps four waves per SIMD.
template <bool COUNT, bool SIDE>
__global__ __launch_bounds__(kTraceThreads) void k_trace_closest(Frame F, int level, const uint32_t* count_ptr,
                                                                 uint32_t fixed_count, uint32_t* status) {
    extern __shared__ uint2 s_stack[];
    trace_closest<COUNT, SIDE, false>(F, level, count_ptr, fixed_count, status, s_stack);
}

// Non-persistent variant: lanes take rays in grid-stride rounds and run each
// to completion (a wave waits for its slowest lane).
template <bool COUNT, bool SIDE>
__global__ __launch_bounds__(kTraceThreads) void k_trace_closest_rounds(Frame F, int level, const uint32_t* count_ptr,
                                                                        uint32_t fixed_count, uint32_t* status) {
    extern __shared__ uint2 s_stack[];
    if (SIDE) bounce_stream_priority();
    else frame_chain_priority();
    const uint32_t n = min(count_ptr ? *count_ptr : fixed_count, F.path_cap);
    const Queue Q = F.q[level & 1];
    const uint32_t gtid = blockIdx.x * kTraceThreads + threadIdx.x;
    Stack<false> st;
    st.init(s_stack, F.stack_lds, reinterpret_cast<uint8_t*>(F.spill), gtid, threadIdx.x, kTraceThreads);
    TravCounters cnt = {0, 0, 0, 0, 0};
    const uint32_t stride = gridDim.x * kTraceThreads;
    const uint32_t rounds = (n + stride - 1) / stride;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint32_t i = r * stride + gtid;
        if (i < n) {
            uint32_t p = Q.path[i];
            if (p >= F.paths) {
                queue_range_error(F);
                p = 0;  // traced as path 0's ray; the shade stage drops it
            }
            const JobDev& J = F.jobs[find_job(F.jobs, F.njobs, p, false)];
            const TlasView T = tlas_view(J.tlas);
            float4 a = Q.o[i], b = Q.d[i];
            HitRec h;
            int res = traverse<false, COUNT>(T, J.scene, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w,
                                             RTG_RAY_FLAG_NONE, 0xffu, 0u, h, st, cnt);
            if (res < 0) atomicOr(status, (uint32_t)kStatusStackOverflow);
            F.hit_a[i] = make_uint4(f2u(h.t), h.inst, h.prim, h.geom);
            F.hit_b[i] = make_float2(h.u, h.v);
        }
    }
    if (COUNT) flush_counters(F.counters, cnt, SIDE ? kCountersClosestSide : kCountersClosest);
}

__device__ inline V3 ld3(uint64_t base, uint32_t i) {
    const float* p = reinterpret_cast<const float*>(base) + 3ull * i;
    return v3(gld_f32(p), gld_f32(p + 1), gld_f32(p + 2));
}
__device__ inline V2 ld2(uint64_t base, uint32_t i) {
    const float* p = reinterpret_cast<const float*>(base) + 2ull * i;
    return V2{gld_f32(p), gld_f32(p + 1)};
}
// ModelInfo / GeometryInfo / index-buffer reads of the hit shaders
// (hit_shader_common.glsl:83-154)
struct HitGeom {
    uint64_t pos, nrm, uv;
    uint32_t i0, i1, i2;
    uint32_t diffuse, mr;
    int32_t normal_map;
};
__device__ inline HitGeom load_hit_geom(const SceneView& S, uint32_t custom, uint32_t geom, uint32_t prim) {
    const rtg_model_info* mi = S.models + custom;
    HitGeom g;
    g.pos = gld_u64(&mi->position_buffer_address);
    g.nrm = gld_u64(&mi->normal_buffer_address);
    g.uv = gld_u64(&mi->uv_buffer_address);
    const rtg_geometry_info* gi = reinterpret_cast<const rtg_geometry_info*>(gld_u64(&mi->geometry_info_address)) + geom;
    const uint32_t* idx = reinterpret_cast<const uint32_t*>(gld_u64(&gi->index_buffer_address)) + 3ull * prim;
    g.i0 = gld_u32(idx);
    g.i1 = gld_u32(idx + 1);
    g.i2 = gld_u32(idx + 2);
    g.diffuse = gld_u32(&gi->diffuse_image_index);
    g.mr = gld_u32(&gi->metallic_roughness_image_index);
    g.normal_map = (int32_t)gld_u32(&gi->normal_map_image_index);
    return g;
}

// sample_directional_light (closest_hit_textured.glsl:77-94)
// The light's tangent frame (t, b) depends on its direction c alone, so a
// hit computes it once for all its shadow rays.
__device__ inline void light_frame(V3 c, V3& t, V3& b) {
    t = normalize(cross(c, v3(0.0f, 1.0f, 0.0f)));
    b = normalize(cross(t, c));
}
__device__ inline V3 sample_dir_light(float rx, float ry, V3 c, V3 t, V3 b, float radius) {
    float rad = sqrtf(rx);
    float angle = (ry * 2.0f) * kPi;
    float s, co;
    sincos_own(angle, &s, &co);
    float px = (rad * co) * radius, py = (rad * s) * radius;
    return normalize((c + t * px) + b * py);
}

// pbr.glsl:174-211 (brdf without the light intensity factor): returns NoL
// and the combined diffuse + specular factor.
__device__ inline void brdf_parts(V3 n, V3 view, V3 l, float pr, V3 base, float metallic, float reflectance,
                                  float& NoL_out, V3& comb) {
    V3 h = normalize(view + l);
    float rough = pr * pr;
    float NoV = clampf(dot(n, view), 1e-9f, 1.0f);
    float NoH = clampf(dot(n, h), 0.0f, 1.0f);
    float NoL = clampf(dot(n, l), 0.0f, 1.0f);
    float LoH = clampf(dot(l, h), 0.0f, 1.0f);
    float a = NoH * rough;
    float k = rough / ((1.0f - NoH * NoH) + a * a);
    float D = (k * k) * (1.0f / kPi);
    float dielectric_f0 = (0.16f * reflectance) * reflectance;
    float fd0 = dielectric_f0 * (1.0f - metallic);
    V3 f0 = v3(fd0, fd0, fd0) + base * metallic;
    float f90 = 0.5f + ((2.0f * rough) * LoH) * LoH;
    float p5 = pow5(1.0f - LoH);
    V3 F = v3(f0.x + (f90 - f0.x) * p5, f0.y + (f90 - f0.y) * p5, f0.z + (f90 - f0.z) * p5);
    float a2 = rough * rough;
    float ggxv = NoL * sqrtf((NoV * NoV) * (1.0f - a2) + a2);
    float ggxl = NoV * sqrtf((NoL * NoL) * (1.0f - a2) + a2);
    float Vis = 0.5f / (ggxv + ggxl);
    float DV = D * Vis;
    V3 spec = F * DV;
    float ls = 1.0f + (f90 - 1.0f) * pow5(1.0f - NoL);
    float vs = 1.0f + (f90 - 1.0f) * pow5(1.0f - NoV);
    float fd = (ls * vs) * (1.0f / kPi);
    V3 diffuse = base * fd;
    comb = diffuse + spec;
    NoL_out = NoL;
}

// Hit and miss stages of trace level b. ONE: the frame state holds one job,
// whose record (and TLAS header) every lane then reads as the same,
// wave-uniform data instead of looking its job up per path.
template <bool COUNT, bool ONE>
__global__ __launch_bounds__(256) void k_shade(Frame F, int level, const uint32_t* count_ptr, uint32_t fixed_count) {
    if (F.side) bounce_stream_priority();
    else frame_chain_priority();
    const uint32_t n = min(count_ptr ? *count_ptr : fixed_count, F.path_cap);
    const Queue Q = F.q[level & 1];
    const Queue NQ = F.q[(level + 1) & 1];
    uint32_t* next_count = &F.counts[level + 1];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    uint32_t n_shadow = 0, n_bounce = 0, n_textured = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint32_t i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool active = i < n;
        uint32_t want_shadow = 0;
        bool want_next = false;
        V3 sh_origin = v3(0, 0, 0), next_o = v3(0, 0, 0), next_d = v3(0, 0, 0), amb = v3(0, 0, 0);
        V3 sun = v3(0, 0, 0);
        float sun_radius = 0.0f;
        uint32_t p = 0, px = 0, py = 0, frame = 0, bn_index = 0;
        const JobDev* J = nullptr;
        if (active) {
            p = Q.path[i];
            if (p >= F.paths) {
                queue_range_error(F);
                active = false;
            }
        }
        if (active) {
            J = ONE ? &F.jobs[0] : &F.jobs[find_job(F.jobs, F.njobs, p, false)];
            uint32_t t_order, s;
            split_sample((uint32_t)(p - J->sample_base), J->spp, t_order, s);
            order_to_pixel(t_order, J->width, J->height, px, py, F.order);
            uint64_t out_index = ((uint64_t)py * J->width + px) * J->spp + s;
            frame = J->frame_index * J->spp + s;
            sun = v3(J->sun[0], J->sun[1], J->sun[2]);
            sun_radius = J->sun_radius;
            bn_index = J->bn_index;
            uint4 ha = F.hit_a[i];
            float2 hb = F.hit_b[i];
            float t = u2f(ha.x);
            uint32_t inst = ha.y, prim = ha.z, geom = ha.w;
            float4 ro = Q.o[i], rd = Q.d[i];
            V3 o = v3(ro.x, ro.y, ro.z), d = v3(rd.x, rd.y, rd.z);
            if (level == 0 && J->out_hits) {
                rtg_hit oh;
                oh.t = t;
                oh.instance = inst;
                oh.primitive = prim;
                oh.geometry = geom;
                oh.u = hb.x;
                oh.v = hb.y;
                J->out_hits[out_index] = oh;
            }
            const SceneView& S = J->scene;
            if (inst == RTG_NO_HIT) {
                // primary_ray_miss (lib.rs:40-51)
                float cs = cos_own(sun_radius);
                V3 c = dot(d, sun) > cs ? v3(1.0f, 1.0f, 1.0f) : v3(0.0f, 0.0f, 0.05f);
                F.path_rgb[p] = make_float4(c.x, c.y, c.z, 0.0f);
            } else {
                const TlasView T = tlas_view(J->tlas);
                const float4* sp = reinterpret_cast<const float4*>(T.src + inst);
                float4 r0 = gld_f4(sp), r1 = gld_f4(sp + 1), r2 = gld_f4(sp + 2);
                uint4 tail = gld_u4(sp + 3);  // custom|mask, sbt|flags, blas
                uint32_t group = tail.y & 0xffffffu;
                uint32_t custom = tail.x & 0xffffffu;
                float o2w[12] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
                V3 bc = v3((1.0f - hb.x) - hb.y, hb.x, hb.y);
                if ((group == RTG_HIT_TEXTURED || group == RTG_HIT_MIRROR) && custom >= S.model_count) {
                    // no model: colour stays 0, no continuation
                } else if (group == RTG_HIT_TEXTURED) {
                    // closest_hit_textured (closest_hit_textured.glsl:174-226)
                    const HitGeom gi = load_hit_geom(S, custom, geom, prim);
                    uint32_t i0 = gi.i0, i1 = gi.i1, i2 = gi.i2;
                    V3 pa = ld3(gi.pos, i0), pb = ld3(gi.pos, i1), pc = ld3(gi.pos, i2);
                    V3 na = ld3(gi.nrm, i0), nb = ld3(gi.nrm, i1), nc = ld3(gi.nrm, i2);
                    V2 ua = ld2(gi.uv, i0), ub = ld2(gi.uv, i1), uc = ld2(gi.uv, i2);
                    V3 pos = interp3(pa, pb, pc, bc);
                    V3 nrm = interp3(na, nb, nc, bc);
                    V2 uv = interp2(ua, ub, uc, bc);
                    // get_shadow_terminator_fix_shadow_origin (:13-39)
                    V3 off[3];
                    V3 P[3] = {pa, pb, pc}, N[3] = {na, nb, nc};
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        V3 vtp = pos - P[k];
                        float dp = fminf(0.0f, dot(vtp, N[k]));
                        off[k] = vtp - N[k] * dp;
                    }
                    V3 fixed = pos + interp3(off[0], off[1], off[2], bc);
                    sh_origin = m34_point(o2w, fixed);
                    want_shadow = J->shadow_rays;
                    // read_material_from_textures (:47-63)
                    float4 dt = sample_index(S, gi.diffuse, uv.x, uv.y);
                    float4 mr = sample_index(S, gi.mr, uv.x, uv.y);
                    V3 base = v3(dt.x, dt.y, dt.z);
                    float metallic = mr.z, roughness = mr.y;
                    // calculate_normal (:141-157)
                    float w2o[12];
                    m34_inverse(o2w, w2o);
                    V3 normal;
                    if (gi.normal_map < 0) {
                        normal = normalize(m34_tdir(w2o, nrm));
                    } else {
                        float4 mt = sample_index(S, (uint32_t)gi.normal_map, uv.x, uv.y);
                        V3 mapn = v3(mt.x * 2.0f - 1.0f, mt.y * 2.0f - 1.0f, mt.z * 2.0f - 1.0f);
                        V3 dp1 = pb - pa, dp2 = pc - pa;
                        float du1x = ub.x - ua.x, du1y = ub.y - ua.y, du2x = uc.x - ua.x, du2y = uc.y - ua.y;
                        V3 dp2perp = cross(dp2, nrm), dp1perp = cross(nrm, dp1);
                        V3 Tt = dp2perp * du1x + dp1perp * du2x;
                        V3 Bt = dp2perp * du1y + dp1perp * du2y;
                        float invmax = 1.0f / sqrtf(fmaxf(dot(Tt, Tt), dot(Bt, Bt)));
                        V3 tt = Tt * invmax, bb = Bt * invmax;
                        V3 m = normalize(mapn);
                        V3 local = (tt * m.x + bb * m.y) + nrm * m.z;
                        normal = normalize(m34_tdir(w2o, local));
                    }
                    float NoL;
                    V3 comb;
                    brdf_parts(normal, v3(-d.x, -d.y, -d.z), sun, roughness, base, metallic, 0.5f, NoL, comb);
                    F.shade_a[p] = make_float4(NoL, comb.x, comb.y, comb.z);
                    amb = v3(0.1f * base.x, 0.1f * base.y, 0.1f * base.z);
                } else if (group == RTG_HIT_MIRROR) {
                    // closest_hit_mirror (closest_hit_mirror.glsl:11-29)
                    const HitGeom gi = load_hit_geom(S, custom, geom, prim);
                    V3 na = ld3(gi.nrm, gi.i0), nb = ld3(gi.nrm, gi.i1), nc = ld3(gi.nrm, gi.i2);
                    V3 ni = interp3(na, nb, nc, bc);
                    float w2o[12];
                    m34_inverse(o2w, w2o);
                    V3 nn = normalize(m34_tdir(w2o, ni));
                    float k2 = 2.0f * dot(nn, d);
                    next_d = d - nn * k2;
                    next_o = o + d * t;
                } else if (group == RTG_HIT_PORTAL) {
                    // closest_hit_portal (lib.rs:300-312)
                    next_d = d;
                    next_o = (o + d * t) + v3(0.0f, 5.0f, 0.0f);
                }
                bool nonzero = !(next_d.x == 0.0f && next_d.y == 0.0f && next_d.z == 0.0f);
                want_next = nonzero && (uint32_t)(level + 1) < J->max_traces;
            }
        }
        // compaction: one atomic per wave for each queue; a path's shadow
        // rays take an even number of slots (a padding slot after an odd
        // count) so that the occlusion kernel's pairs never span two paths
        const uint32_t shadow_slots = want_shadow + (want_shadow & 1u);
        uint32_t sbase = wave_alloc(shadow_slots, &F.counts[4]);
        uint32_t nbase = wave_alloc(want_next ? 1u : 0u, next_count);
        if (want_next && nbase >= F.path_cap) {
            queue_range_error(F);
            want_next = false;
        }
        if (want_shadow && sbase + shadow_slots > F.shadow_cap) {
            queue_range_error(F);
            want_shadow = 0;
        }
        // a textured hit's ambient term and where its shadow rays' flags lie
        if (want_shadow) F.shade_b[p] = make_float4(amb.x, amb.y, amb.z, u2f(sbase + 1u));
        if (want_next) {
            NQ.o[nbase] = make_float4(next_o.x, next_o.y, next_o.z, 0.01f);
            NQ.d[nbase] = make_float4(next_d.x, next_d.y, next_d.z, 10000.0f);
            NQ.path[nbase] = p;
            n_bounce++;
        }
        if (want_shadow) {
            const SceneView& S = J->scene;
            // the blue-noise descriptor once, then the texels of four rays at a
            // time before their stores, so that the loads overlap
            const rtg_texture bn = load_texture(S, bn_index);
            const float anim = (float)(frame % 32u) * 0.618033988749f;
            V3 lt, lb;
            light_frame(sun, lt, lb);
            for (uint32_t k0 = 0; k0 < want_shadow; k0 += 4) {
                float n1[4], n2[4];
    #pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    // sample_blue_noise (:99-111) and animated_blue_noise (:116-120)
                    const uint32_t k = k0 + j;
                    n1[j] = n2[j] = 0.0f;
                    if (k < want_shadow) {
                        uint32_t o1x = k * 2u * 13u, o1y = k * 2u * 41u;
                        uint32_t o2x = (k * 2u + 1u) * 13u, o2y = (k * 2u + 1u) * 41u;
                        n1[j] = sample_tex(bn, S.srgb_lut, (float)(px + o1x) / 64.0f, (float)(py + o1y) / 64.0f).x;
                        n2[j] = sample_tex(bn, S.srgb_lut, (float)(px + o2x) / 64.0f, (float)(py + o2y) / 64.0f).x;
                    }
                }
    #pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint32_t k = k0 + j;
                    if (k < want_shadow) {
                        float rx = fractf(n1[j] + anim), ry = fractf(n2[j] + anim);
                        V3 sd = sample_dir_light(rx, ry, sun, lt, lb, sun_radius);
                        uint32_t slot = sbase + k;
                        F.sq.o[slot] = make_float4(sh_origin.x, sh_origin.y, sh_origin.z, 0.001f);
                        F.sq.d[slot] = make_float4(sd.x, sd.y, sd.z, 10000.0f);
                        F.sq.path[slot] = p;
                    }
                }
            }
            if (want_shadow & 1u) F.sq.path[sbase + want_shadow] = kNoPath;
            n_shadow += want_shadow;
            n_textured++;
        }
    }
    if (COUNT) {
        wave_add_counter(&F.counters[1], n_bounce);
        wave_add_counter(&F.counters[2], n_shadow);
        wave_add_counter(&F.counters[kCounterTextured], n_textured);
        if (F.side) {
            wave_add_counter(&F.counters[kCounterSideShadow], n_shadow);
            wave_add_counter(&F.counters[kCounterTextured + 1], n_textured);
            wave_add_counter(&F.counters[kCounterTextured + 2], n_bounce);
        }
    }
}

// Parts of the shadow queue an occlusion launch traces: all of it, or the
// rays of the first trace level (counts[kShadowFirst] of them, the queue's
// size when the first shade finished) and the rays appended after them.
__device__ inline void shadow_part(const Frame& F, int part, uint32_t& lo, uint32_t& hi, uint32_t*& feed) {
    hi = min(part == kPartFirst ? F.counts[kShadowFirst] : F.counts[4], F.shadow_cap);
    lo = min(part == kPartRest ? F.counts[kShadowFirst] : 0u, hi);
    feed = &F.counts[part == kPartRest ? kFeedShadowRest : kFeedShadow];
}

// The pair of shadow-queue entries 2q, 2q + 1 of [lo, hi): a path's shadow
// rays are allocated from an even slot, so a pair holds the rays 2j, 2j + 1
// of one path, or ray 2j alone when the path has an odd count (its last slot
// is padding, path kNoPath). Returns the pair's mask (3, or 1 for ray 2j
// alone) and its rays.
__device__ inline uint32_t shadow_pair(const Frame& F, uint32_t e0, uint32_t hi, uint32_t& path, float4& o,
                                       float4& d0, float4& d1) {
    path = F.sq.path[e0];
    if (path >= F.paths) {
        queue_range_error(F);
        path = 0;
    }
    o = F.sq.o[e0];
    d0 = F.sq.d[e0];
    d1 = d0;
    if (e0 + 1 >= hi) return 1u;
    const uint32_t p1 = F.sq.path[e0 + 1];
    if (p1 == kNoPath) return 1u;
    if (p1 != path) {
        // entries of different paths in one pair: a frame state driven by
        // two traces at once; its second ray is dropped (reported visible)
        queue_range_error(F);
        F.sflag[e0 + 1] = 1;
        return 1u;
    }
    d1 = F.sq.d[e0 + 1];
    return 3u;
}

// Occlusion of the shadow rays (TerminateOnFirstHit | SkipClosestHit, hit
// group offset 1, miss shader 1: closest_hit_textured.glsl:159-172);
// shadow_ray_miss (lib.rs:33-36) makes the ray count as visible. Each lane
// traces a pair of one path's shadow rays (PairTraverser); one byte per
// shadow ray, plain stores (no atomics). A pair lane holds 121 VGPRs (four
// waves per SIMD).
template <bool COUNT, bool SIDE>
__global__ __launch_bounds__(kTraceThreads) void k_trace_occlusion(Frame F, uint32_t* status, int part) {
    extern __shared__ uint2 s_stack[];
    if (SIDE) bounce_stream_priority();
    uint32_t lo, hi;
    uint32_t* feed;
    shadow_part(F, part, lo, hi, feed);
    const uint32_t npairs = (hi - lo + 1) / 2;
    const uint32_t gtid = blockIdx.x * kTraceThreads + threadIdx.x;
    Stack<false> st;
    st.init(s_stack, F.stack_lds, reinterpret_cast<uint8_t*>(F.spill), gtid, threadIdx.x, kTraceThreads);
    TravCounters cnt = {0, 0, 0, 0, 0};
    PairTraverser<COUNT> tr;
    WaveFeeder feeder;
    feeder.init(F.feed_chunk);
    bool active = false;
    uint32_t e0 = 0, started = 0, path = 0;
    const JobDev* J = nullptr;
    TlasView T;
    uint32_t c0 = 0;
    uint32_t it = 0;  // iteration parity: leaf visits on even, instance entries on odd iterations
    while (true) {
        ++it;
        uint64_t idle = __ballot(!active);
        if (want_refill(idle, feeder, F.refill_occlusion)) {
            uint32_t q = feeder.feed(idle, !active, feed, npairs);
            if (q != ~0u) {
                e0 = lo + 2 * q;
                float4 o, d0, d1;
                started = shadow_pair(F, e0, hi, path, o, d0, d1);
                J = &F.jobs[find_job(F.jobs, F.njobs, path, false)];
                T = tlas_view(J->tlas);
                tr.start(T, v3(o.x, o.y, o.z), v3(d0.x, d0.y, d0.z), v3(d1.x, d1.y, d1.z), started, o.w, d0.w, st);
                if (J->flags & RTG_TRACE_HEATMAP) c0 = (uint32_t)__builtin_readcyclecounter();
                active = true;
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            int res = tr.step(T, J->scene, st, cnt, (it & 1u) == 0u, (it & 1u) != 0u);
            if (res != kRayActive) {
                // a stack overflow leaves both rays unresolved: not visible
                const uint32_t occ = res == kRayOverflow ? started : tr.occluded(started);
                if (res == kRayOverflow) atomicOr(status, (uint32_t)kStatusStackOverflow);
                if (J->flags & RTG_TRACE_HEATMAP)
                    atomicAdd(&F.cycles[path], (uint32_t)__builtin_readcyclecounter() - c0);
                F.sflag[e0] = (occ & 1u) ? 0 : 1;
                if (started & 2u) F.sflag[e0 + 1] = (occ & 2u) ? 0 : 1;
                active = false;
            }
        }
    }
    if (COUNT) flush_counters(F.counters, cnt, SIDE ? kCountersOcclusionSide : kCountersOcclusion);
}

// The same pairs, one per lane and grid-stride round (refill_occlusion 0).
template <bool COUNT, bool SIDE>
__global__ __launch_bounds__(kTraceThreads) void k_trace_occlusion_rounds(Frame F, uint32_t* status, int part) {
    extern __shared__ uint2 s_stack[];
    if (SIDE) bounce_stream_priority();
    uint32_t lo, hi;
    uint32_t* feed;
    shadow_part(F, part, lo, hi, feed);
    const uint32_t npairs = (hi - lo + 1) / 2;
    const uint32_t gtid = blockIdx.x * kTraceThreads + threadIdx.x;
    Stack<false> st;
    st.init(s_stack, F.stack_lds, reinterpret_cast<uint8_t*>(F.spill), gtid, threadIdx.x, kTraceThreads);
    TravCounters cnt = {0, 0, 0, 0, 0};
    const uint32_t stride = gridDim.x * kTraceThreads;
    const uint32_t rounds = (npairs + stride - 1) / stride;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t q = r * stride + gtid;
        if (q < npairs) {
            const uint32_t e0 = lo + 2 * q;
            uint32_t path;
            float4 o, d0, d1;
            const uint32_t started = shadow_pair(F, e0, hi, path, o, d0, d1);
            const JobDev& J = F.jobs[find_job(F.jobs, F.njobs, path, false)];
            const TlasView T = tlas_view(J.tlas);
            PairTraverser<COUNT> tr;
            tr.start(T, v3(o.x, o.y, o.z), v3(d0.x, d0.y, d0.z), v3(d1.x, d1.y, d1.z), started, o.w, d0.w, st);
            int res;
            while ((res = tr.step(T, J.scene, st, cnt)) == kRayActive) {
            }
            const uint32_t occ = res == kRayOverflow ? started : tr.occluded(started);
            if (res == kRayOverflow) atomicOr(status, (uint32_t)kStatusStackOverflow);
            F.sflag[e0] = (occ & 1u) ? 0 : 1;
            if (started & 2u) F.sflag[e0 + 1] = (occ & 2u) ? 0 : 1;
        }
    }
    if (COUNT) flush_counters(F.counters, cnt, SIDE ? kCountersOcclusionSide : kCountersOcclusion);
}

// linear_to_srgb (lib.rs:84-92)
__device__ inline float srgb1(float x) {
    if (x <= 0.0031308f) return x * 12.92f;
    return 1.055f * pow_own(x, 1.0f / 2.4f) - 0.055f;
}

__device__ inline float smoothstep_(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}

// heatmap_temperature (shaders/ray-tracing/src/heatmap.rs:5-38)
__device__ inline V3 heatmap_temperature(float heat) {
    const float col[10][3] = {{0.0f, 2.0f, 91.0f},     {0.0f, 108.0f, 251.0f}, {0.0f, 221.0f, 221.0f},
                              {51.0f, 221.0f, 0.0f},   {255.0f, 252.0f, 0.0f}, {255.0f, 180.0f, 0.0f},
                              {255.0f, 104.0f, 0.0f},  {226.0f, 22.0f, 0.0f},  {191.0f, 0.0f, 83.0f},
                              {145.0f, 0.0f, 65.0f}};
    float h = clampf(heat, 0.0f, 1.0f) * 10.0f;
    int hi = (int)h;
    int cur = min(hi, 9), prv = max(hi - 1, 0), nxt = min(hi + 1, 9);
    float hf = floorf(h), hc = ceilf(h);
    float blur = 0.8f;
    float wc = smoothstep_(hf - blur, hf + blur, h) * (1.0f - smoothstep_(hc - blur, hc + blur, h));
    float wp = 1.0f - smoothstep_(hf - blur, hf + blur, h);
    float wn = smoothstep_(hc - blur, hc + blur, h);
    float o[3];
    for (int c = 0; c < 3; ++c) {
        float r = (wc * (col[cur][c] / 255.0f) + wp * (col[prv][c] / 255.0f)) + wn * (col[nxt][c] / 255.0f);
        o[c] = clampf(r, 0.0f, 1.0f);
    }
    return v3(o[0], o[1], o[2]);
}

__global__ __launch_bounds__(256) void k_resolve(Frame F, uint64_t total_pixels) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < total_pixels; q += stride) {
        // lanes take pixels in path order, so that they read consecutive
        // paths' state; the output pixel follows from the order
        const JobDev& J = F.jobs[find_job(F.jobs, F.njobs, q, true)];
        const uint32_t t = (uint32_t)(q - J.pixel_base);
        uint32_t px, py;
        order_to_pixel(t, J.width, J.height, px, py, F.order);
        const uint32_t pix = py * J.width + px;
        V3 acc = v3(0.0f, 0.0f, 0.0f);
        uint32_t cycles = 0;
        for (uint32_t s = 0; s < J.spp; ++s) {
            uint64_t p = J.sample_base + (uint64_t)t * J.spp + s;
            float4 b = F.shade_b[p];
            V3 c;
            const uint32_t first = f2u(b.w);
            if (first != 0) {
                float4 a = F.shade_a[p];
                uint32_t vis = 0;
                for (uint32_t k = 0; k < J.shadow_rays; ++k) vis += F.sflag[first - 1u + k];
                float sf = (float)vis / (float)J.shadow_rays;
                float in_nol = sf * a.x;
                c = v3(a.y * in_nol + b.x, a.z * in_nol + b.y, a.w * in_nol + b.z);
            } else {
                float4 r = F.path_rgb[p];
                c = v3(r.x, r.y, r.z);
            }
            acc = acc + c;
            cycles += F.cycles[p];
        }
        V3 col = J.spp == 1 ? acc : acc * (1.0f / (float)J.spp);
        if (J.flags & RTG_TRACE_HEATMAP) {
            // lib.rs:174-186: heatmap_temperature(cycles / 1e6) + colour * 1e-6
            V3 hm = heatmap_temperature((float)cycles / 1000000.0f);
            col = hm + col * 0.000001f;
        }
        if (J.out_rgb) {
            J.out_rgb[3ull * pix] = col.x;
            J.out_rgb[3ull * pix + 1] = col.y;
            J.out_rgb[3ull * pix + 2] = col.z;
        }
        if (J.out_rgba8) {
            uint32_t r = (uint32_t)(clampf(srgb1(col.x), 0.0f, 1.0f) * 255.0f + 0.5f);
            uint32_t g = (uint32_t)(clampf(srgb1(col.y), 0.0f, 1.0f) * 255.0f + 0.5f);
            uint32_t b = (uint32_t)(clampf(srgb1(col.z), 0.0f, 1.0f) * 255.0f + 0.5f);
            if (J.flags & RTG_TRACE_OUTPUT_BGRA8) {
                uint32_t t = r;
                r = b;
                b = t;
            }
            J.out_rgba8[pix] = r | (g << 8) | (b << 16) | (255u << 24);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct FrameArena {
    DeviceBuffer buf;
    DeviceBuffer spill;
    DeviceBuffer feed;  // feed counter of rtg_trace_rays
    uint64_t paths_cap = 0, shadow_cap = 0;
    uint32_t jobs_cap = 0;
    JobDev* h_jobs = nullptr;  // pinned staging
    uint32_t h_jobs_cap = 0;
    hipEvent_t staging_free = nullptr;
    // the bounce levels run on `side` beside the first level's shadow rays
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    Frame F;
    JobDev* d_jobs = nullptr;
    uint32_t* d_counts = nullptr;
    // grow to exactly the size asked (no headroom): a server slot's arena,
    // whose device bytes count against an HBM budget
    bool exact = false;
    // host calls using the arena now: a second one at the same time (two
    // threads tracing on one stream of one context) is refused
    std::atomic<int> users{0};
};

FrameArena* new_frame_arena(bool exact) {
    FrameArena* a = new FrameArena();
    a->exact = exact;
    return a;
}

void delete_frame_arena(FrameArena* a) {
    if (!a) return;
    a->buf.release();
    a->spill.release();
    a->feed.release();
    if (a->staging_free) hipEventSynchronize(a->staging_free);
    if (a->h_jobs) hipHostFree(a->h_jobs);
    if (a->staging_free) hipEventDestroy(a->staging_free);
    if (a->side) {
        hipStreamSynchronize(a->side);
        hipStreamDestroy(a->side);
    }
    if (a->fork) hipEventDestroy(a->fork);
    if (a->join) hipEventDestroy(a->join);
    delete a;
}

void arena_pointers(const FrameArena* a, const void* out[3]) {
    out[0] = a->buf.ptr;
    out[1] = a->spill.ptr;
    out[2] = a->h_jobs;
}

void free_frame_arena(rtg_ctx* ctx) {
    for (auto& kv : ctx->frames) delete_frame_arena(kv.second);
    ctx->frames.clear();
}

// Device bytes of the frame state with room for `paths` paths, `shadow`
// shadow rays and `njobs` jobs: the buffer ensure_frame carves.
size_t frame_bytes(uint64_t paths, uint64_t shadow, uint32_t njobs) {
    size_t off = 0;
    carve_size<float4>(off, paths);    // path_rgb
    carve_size<float4>(off, paths);    // shade_a
    carve_size<float4>(off, paths);    // shade_b
    carve_size<uint8_t>(off, shadow);  // sflag
    carve_size<uint32_t>(off, paths);  // cycles
    for (int q = 0; q < 2; ++q) {
        carve_size<float4>(off, paths);
        carve_size<float4>(off, paths);
        carve_size<uint32_t>(off, paths);
    }
    carve_size<uint4>(off, paths);
    carve_size<float2>(off, paths);
    carve_size<float4>(off, shadow);
    carve_size<float4>(off, shadow);
    carve_size<uint32_t>(off, shadow);
    carve_size<uint32_t>(off, kCountSlots);
    carve_size<JobDev>(off, njobs);
    return off + 4096;
}

static int trace_grid(rtg_ctx* ctx, const Tuning& tune) {
    // persistent grid: blocks_per_cu blocks of 256 lanes per CU (the LDS
    // stack takes 32 KB per block, so at most 5 fit; registers allow 4)
    return ctx->cu_count * tune.blocks_per_cu;
}
static int trace_grid(rtg_ctx* ctx) { return trace_grid(ctx, ctx->tune); }

size_t spill_bytes(rtg_ctx* ctx, const Tuning& tune) {
    return (size_t)trace_grid(ctx, tune) * kTraceThreads * kSpillLaneBytes;
}

static int ensure_spill(rtg_ctx* ctx, FrameArena* a, const Tuning& tune) {
    if (a->spill.ensure(spill_bytes(ctx, tune), !a->exact) != RTG_OK) return RTG_ERROR_OUT_OF_MEMORY;
    return a->feed.ensure(256, !a->exact);
}

size_t arena_bytes_after(const FrameArena* a, uint64_t paths, uint64_t shadow, uint32_t njobs) {
    if (a->buf.ptr && paths <= a->paths_cap && shadow <= a->shadow_cap && njobs <= a->jobs_cap) return a->buf.size;
    size_t want = frame_bytes(std::max<uint64_t>(paths, a->paths_cap), std::max<uint64_t>(shadow, a->shadow_cap),
                              std::max<uint32_t>(njobs, a->jobs_cap));
    if (!a->exact) want += want / 4;
    return std::max(want, a->buf.size);
}

size_t arena_held_bytes(const FrameArena* a) { return a->buf.size + a->spill.size + a->feed.size; }

void arena_reset(FrameArena* a) {
    a->buf.release();
    a->paths_cap = a->shadow_cap = 0;
    a->jobs_cap = 0;
}

const uint32_t* arena_status_word(const FrameArena* a) { return a->d_counts + kStatusSlot; }

// Frame state of the context's calls on stream `s`: calls on different
// streams (frames in flight) never share queues, counters or spill areas.
static FrameArena* arena(rtg_ctx* ctx, hipStream_t s) {
    FrameArena*& a = ctx->frames[reinterpret_cast<void*>(s)];
    if (!a) a = new FrameArena();
    return a;
}
FrameArena* context_arena(rtg_ctx* ctx) { return arena(ctx, nullptr); }

static int ensure_frame(rtg_ctx* ctx, FrameArena* a, uint64_t paths, uint64_t shadow, uint32_t njobs,
                        const Tuning& tune) {
    if (paths > a->paths_cap || shadow > a->shadow_cap || njobs > a->jobs_cap || a->buf.ptr == nullptr) {
        uint64_t P = std::max<uint64_t>(paths, a->paths_cap);
        uint64_t S = std::max<uint64_t>(shadow, a->shadow_cap);
        uint32_t NJ = std::max<uint32_t>(njobs, a->jobs_cap);
        if (a->buf.ensure(frame_bytes(P, S, NJ), !a->exact) != RTG_OK) return RTG_ERROR_OUT_OF_MEMORY;
        Carve c(a->buf.ptr);
        Frame& F = a->F;
        F.path_rgb = c.take<float4>(P);
        F.shade_a = c.take<float4>(P);
        F.shade_b = c.take<float4>(P);
        F.sflag = c.take<uint8_t>(S);
        F.cycles = c.take<uint32_t>(P);
        for (int q = 0; q < 2; ++q) {
            F.q[q].o = c.take<float4>(P);
            F.q[q].d = c.take<float4>(P);
            F.q[q].path = c.take<uint32_t>(P);
        }
        F.hit_a = c.take<uint4>(P);
        F.hit_b = c.take<float2>(P);
        F.sq.o = c.take<float4>(S);
        F.sq.d = c.take<float4>(S);
        F.sq.path = c.take<uint32_t>(S);
        a->d_counts = c.take<uint32_t>(kCountSlots);
        a->d_jobs = c.take<JobDev>(NJ);
        a->paths_cap = P;
        a->shadow_cap = S;
        a->jobs_cap = NJ;
    }
    if (ensure_spill(ctx, a, tune) != RTG_OK) return RTG_ERROR_OUT_OF_MEMORY;
    if (njobs > a->h_jobs_cap) {
        if (a->staging_free) hipEventSynchronize(a->staging_free);
        if (a->h_jobs) hipHostFree(a->h_jobs);
        RTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&a->h_jobs), njobs * sizeof(JobDev)));
        a->h_jobs_cap = njobs;
    }
    if (!a->staging_free) RTG_HIP(hipEventCreateWithFlags(&a->staging_free, hipEventDisableTiming));
    return RTG_OK;
}

// The bounce stream of an arena and its fork/join events, created on first
// use (a stream costs a hardware-queue share even while idle).
static int ensure_side(FrameArena* a) {
    if (!a->side) RTG_HIP(hipStreamCreateWithFlags(&a->side, hipStreamNonBlocking));
    if (!a->fork) RTG_HIP(hipEventCreateWithFlags(&a->fork, hipEventDisableTiming));
    if (!a->join) RTG_HIP(hipEventCreateWithFlags(&a->join, hipEventDisableTiming));
    return RTG_OK;
}

static int validate_job(rtg_ctx* ctx, const rtg_image_job& j, uint32_t k, bool check_tlas) {
    if (check_tlas) {
        auto it = ctx->structures.find(j.tlas);
        RTG_REQUIRE(it != ctx->structures.end() && it->second.kind == kKindTlas, RTG_ERROR_NOT_FOUND,
                    "job %u: 0x%llx is not a TLAS of this context", k, (unsigned long long)j.tlas);
    }
    RTG_REQUIRE(j.width > 0 && j.height > 0, RTG_ERROR_INVALID_ARGUMENT, "job %u: empty image %ux%u", k, j.width,
                j.height);
    RTG_REQUIRE(j.opts.spp >= 1 && j.opts.spp <= 64, RTG_ERROR_INVALID_ARGUMENT, "job %u: spp %u (1..64)", k,
                j.opts.spp);
    RTG_REQUIRE(j.opts.max_traces >= 1 && j.opts.max_traces <= 4, RTG_ERROR_INVALID_ARGUMENT,
                "job %u: max_traces %u (1..4)", k, j.opts.max_traces);
    RTG_REQUIRE(j.opts.shadow_rays >= 1 && j.opts.shadow_rays <= 16, RTG_ERROR_INVALID_ARGUMENT,
                "job %u: shadow_rays %u (1..16)", k, j.opts.shadow_rays);
    RTG_REQUIRE((j.opts.flags & ~(RTG_TRACE_HEATMAP | RTG_TRACE_OUTPUT_BGRA8)) == 0, RTG_ERROR_INVALID_ARGUMENT,
                "job %u: unknown trace flags 0x%x", k, j.opts.flags);
    RTG_REQUIRE(j.scene.model_count == 0 || j.scene.model_info, RTG_ERROR_INVALID_ARGUMENT,
                "job %u: model_count %u with a null model table", k, j.scene.model_count);
    RTG_REQUIRE(j.scene.texture_count == 0 || j.scene.textures, RTG_ERROR_INVALID_ARGUMENT,
                "job %u: texture_count %u with a null texture table", k, j.scene.texture_count);
    return RTG_OK;
}

int plan_jobs(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t njobs, FrameArena* a, bool check_tlas,
              const Tuning& tune, TracePlan& P) {
    RTG_REQUIRE(njobs > 0 && jobs, RTG_ERROR_INVALID_ARGUMENT, "trace: no jobs");
    P = TracePlan();
    P.njobs = njobs;
    for (uint32_t k = 0; k < njobs; ++k) {
        int st = validate_job(ctx, jobs[k], k, check_tlas);
        if (st != RTG_OK) return st;
        uint64_t px = (uint64_t)jobs[k].width * jobs[k].height;
        P.paths += px * jobs[k].opts.spp;
        P.pixels += px;
        P.shadow += px * jobs[k].opts.spp * shadow_slots(jobs[k].opts.shadow_rays);
        P.max_traces = std::max(P.max_traces, jobs[k].opts.max_traces);
    }
    RTG_REQUIRE(P.paths < (1ull << 31) && P.shadow < (1ull << 32), RTG_ERROR_LIMIT,
                "trace: %llu paths / %llu shadow rays exceed one launch", (unsigned long long)P.paths,
                (unsigned long long)P.shadow);
    RTG_HIP(hipSetDevice(ctx->device));
    int st = ensure_frame(ctx, a, P.paths, P.shadow, njobs, tune);
    if (st != RTG_OK) return st;
    if (P.max_traces > 1 && tune.bounce_blocks > 0) {
        st = ensure_side(a);
        if (st != RTG_OK) return st;
    }
    // the pinned job table is reused: wait until the previous upload left it
    RTG_HIP(hipEventSynchronize(a->staging_free));
    uint64_t sb = 0, pb = 0;
    for (uint32_t k = 0; k < njobs; ++k) {
        const rtg_image_job& j = jobs[k];
        JobDev& d = a->h_jobs[k];
        memset(&d, 0, sizeof(d));
        d.tlas = reinterpret_cast<const uint8_t*>(j.tlas);
        d.scene.models = reinterpret_cast<const rtg_model_info*>(j.scene.model_info);
        d.scene.textures = reinterpret_cast<const rtg_texture*>(j.scene.textures);
        d.scene.srgb_lut = ctx->d_srgb_lut;
        d.scene.model_count = j.scene.model_count;
        d.scene.texture_count = j.scene.texture_count;
        memcpy(d.view_inverse, j.uniforms.view_inverse, sizeof(d.view_inverse));
        memcpy(d.proj_inverse, j.uniforms.proj_inverse, sizeof(d.proj_inverse));
        memcpy(d.sun, j.uniforms.sun_dir, sizeof(d.sun));
        d.sun_radius = j.uniforms.sun_radius;
        d.bn_index = j.uniforms.blue_noise_texture_index;
        d.frame_index = j.uniforms.frame_index;
        d.width = j.width;
        d.height = j.height;
        d.spp = j.opts.spp;
        d.shadow_rays = j.opts.shadow_rays;
        d.max_traces = j.opts.max_traces;
        d.flags = j.opts.flags | (j.uniforms.show_heatmap ? RTG_TRACE_HEATMAP : 0u);
        d.sample_base = sb;
        d.pixel_base = pb;
        d.out_rgba8 = j.out_rgba8;
        d.out_rgb = j.out_rgb;
        d.out_hits = j.out_primary_hits;
        uint64_t px = (uint64_t)j.width * j.height;
        sb += px * j.opts.spp;
        pb += px;
    }
    return RTG_OK;
}

int upload_jobs(FrameArena* a, const TracePlan& P, hipStream_t s) {
    RTG_HIP(hipMemcpyAsync(a->d_jobs, a->h_jobs, P.njobs * sizeof(JobDev), hipMemcpyHostToDevice, s));
    RTG_HIP(hipEventRecord(a->staging_free, s));
    return RTG_OK;
}

void enqueue_trace(rtg_ctx* ctx, FrameArena* a, const TracePlan& P, const Tuning& tune, bool cnt, bool timed,
                   hipStream_t s) {
    Frame F = a->F;
    F.jobs = a->d_jobs;
    F.njobs = P.njobs;
    F.paths = P.paths;
    F.counts = a->d_counts;
    F.counters = ctx->d_counters;
    F.spill = reinterpret_cast<uint2*>(a->spill.ptr);
    F.order = (uint32_t)tune.path_order;
    F.refill_closest = tune.refill_closest;
    F.refill_occlusion = tune.refill_occlusion;
    F.feed_chunk = (uint32_t)tune.feed_chunk;
    F.stack_lds = tune.stack_lds;
    F.side = 0;
    F.path_cap = (uint32_t)std::min<uint64_t>(a->paths_cap, 0xffffffffu);
    F.shadow_cap = (uint32_t)std::min<uint64_t>(a->shadow_cap, 0xffffffffu);
    F.status = a->d_counts + kStatusSlot;
    const unsigned G = (unsigned)trace_grid(ctx, tune);
    const size_t lds = stack_lds_bytes(tune.stack_lds);  // a traversal workgroup's stacks
    const unsigned Gsmall = (unsigned)std::min<uint64_t>((P.paths + 255) / 256, (uint64_t)ctx->cu_count * 16);
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, a->d_counts, ctx->d_counters, cnt);
    {
        LaunchTimer t(ctx, RTG_KERNEL_RAYGEN, s, timed);
        hipLaunchKernelGGL(k_raygen, dim3(Gsmall), dim3(256), 0, s, F, cnt);
    }
    // trace level `level` (closest hits, then the hit and miss stages) on
    // `st` with `g` persistent blocks
    auto level_step = [&](const Frame& Fl, uint32_t level, unsigned g, unsigned gshade, hipStream_t st, bool tm) {
        const uint32_t* cp = level == 0 ? nullptr : a->d_counts + level;
        uint32_t fixed = level == 0 ? (uint32_t)P.paths : 0u;
        {
            LaunchTimer t(ctx, RTG_KERNEL_TRACE_CLOSEST, st, tm);
            const bool side = Fl.side != 0;
            const bool single = Fl.njobs == 1;
            if (tune.refill_closest == 0) {
                if (side)
                    hipLaunchKernelGGL((cnt ? k_trace_closest_rounds<true, true> : k_trace_closest_rounds<false, true>),
                                       dim3(g), dim3(kTraceThreads), lds, st, Fl, (int)level, cp, fixed, Fl.status);
                else
                    hipLaunchKernelGGL((cnt ? k_trace_closest_rounds<true, false> : k_trace_closest_rounds<false, false>),
                                       dim3(g), dim3(kTraceThreads), lds, st, Fl, (int)level, cp, fixed, Fl.status);
            } else {
                auto kernel = side ? (cnt ? (single ? k_trace_closest_one<true, true> : k_trace_closest<true, true>)
                                          : (single ? k_trace_closest_one<false, true> : k_trace_closest<false, true>))
                                   : (cnt ? (single ? k_trace_closest_one<true, false> : k_trace_closest<true, false>)
                                          : (single ? k_trace_closest_one<false, false> : k_trace_closest<false, false>));
                hipLaunchKernelGGL(kernel, dim3(g), dim3(kTraceThreads), lds, st, Fl, (int)level, cp, fixed, Fl.status);
            }
        }
        {
            LaunchTimer t(ctx, RTG_KERNEL_SHADE, st, tm);
            const bool one = Fl.njobs == 1;
            auto kernel = cnt ? (one ? k_shade<true, true> : k_shade<true, false>)
                              : (one ? k_shade<false, true> : k_shade<false, false>);
            hipLaunchKernelGGL(kernel, dim3(gshade), dim3(256), 0, st, Fl, (int)level, cp, fixed);
        }
    };
    // occlusion rays of `part` of the shadow queue; SIDE kernels count into
    // the bounce stream's counter slots
    auto occlusion = [&](const Frame& Fo, int part, unsigned g, hipStream_t st, bool tm) {
        LaunchTimer t(ctx, RTG_KERNEL_TRACE_OCCLUSION, st, tm);
        const bool side = part == kPartRest;
        if (tune.refill_occlusion == 0) {
            if (side)
                hipLaunchKernelGGL((cnt ? k_trace_occlusion_rounds<true, true> : k_trace_occlusion_rounds<false, true>),
                                   dim3(g), dim3(kTraceThreads), lds, st, Fo, Fo.status, part);
            else
                hipLaunchKernelGGL((cnt ? k_trace_occlusion_rounds<true, false> : k_trace_occlusion_rounds<false, false>),
                                   dim3(g), dim3(kTraceThreads), lds, st, Fo, Fo.status, part);
        } else if (side) {
            hipLaunchKernelGGL((cnt ? k_trace_occlusion<true, true> : k_trace_occlusion<false, true>), dim3(g),
                               dim3(kTraceThreads), lds, st, Fo, Fo.status, part);
        } else {
            hipLaunchKernelGGL((cnt ? k_trace_occlusion<true, false> : k_trace_occlusion<false, false>), dim3(g),
                               dim3(kTraceThreads), lds, st, Fo, Fo.status, part);
        }
    };
    level_step(F, 0, G, Gsmall, s, timed);
    const unsigned Gside = std::min<unsigned>((unsigned)tune.bounce_blocks, G / 2);
    if (P.max_traces > 1 && Gside > 0) {
        // The bounce levels hold few rays, so their launches are short of
        // work and bound by the latency of their longest rays. They run on
        // the side stream, on Gside persistent blocks, while the first
        // level's shadow rays take the other blocks; the shadow rays the
        // bounce levels add are traced on the side stream after them. The
        // two grids use disjoint spill areas.
        const unsigned Gmain = G - Gside;
        Frame Fs = F;
        Fs.spill = reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(F.spill) +
                                            (size_t)Gmain * kTraceThreads * kSpillLaneBytes);
        Fs.side = 1;
        hipLaunchKernelGGL(k_shadow_snapshot, dim3(1), dim3(64), 0, s, a->d_counts);
        hipEventRecord(a->fork, s);
        hipStreamWaitEvent(a->side, a->fork, 0);
        {
            // timed as one span: its kernels share the GPU with the shadow rays
            LaunchTimer t(ctx, RTG_KERNEL_BOUNCE_SIDE, a->side, timed);
            // The bounce levels' traversal and shading launches take at most
            // 128 workgroups (32 k lanes; a 1080p frame has about 46 k and
            // 3 k bounce rays): a launch ends only once each of its
            // workgroups has been dispatched, which waits for free slots while
            // the other frame's persistent kernels hold the CUs (the bounce
            // part of the frame: 1.70-1.78 instead of 1.82 ms).
            const unsigned Gbounce = std::min(Gside, 128u);
            for (uint32_t level = 1; level < P.max_traces; ++level)
                level_step(Fs, level, Gbounce, std::min(Gsmall, Gbounce), a->side, false);
            occlusion(Fs, kPartRest, Gside, a->side, false);
        }
        hipEventRecord(a->join, a->side);
        occlusion(F, kPartFirst, Gmain, s, timed);
        hipStreamWaitEvent(s, a->join, 0);
    } else {
        for (uint32_t level = 1; level < P.max_traces; ++level) level_step(F, level, G, Gsmall, s, timed);
        occlusion(F, kPartAll, G, s, timed);
    }
    {
        LaunchTimer t(ctx, RTG_KERNEL_RESOLVE, s, timed);
        unsigned Gr = (unsigned)std::min<uint64_t>((P.pixels + 255) / 256, (uint64_t)ctx->cu_count * 16);
        hipLaunchKernelGGL(k_resolve, dim3(Gr), dim3(256), 0, s, F, P.pixels);
    }
    hipLaunchKernelGGL(k_status_merge, dim3(1), dim3(64), 0, s, F.status, ctx->d_status);
}

// Marks an arena as used by this host call for its duration.
struct ArenaUse {
    FrameArena* a;
    bool ok;
    explicit ArenaUse(FrameArena* a_) : a(a_) { ok = a->users.fetch_add(1) == 0; }
    ~ArenaUse() { a->users.fetch_sub(1); }
};

int trace_jobs(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t njobs, hipStream_t s, FrameArena* a,
               bool check_tlas, const Tuning* tune) {
    if (a == nullptr) a = arena(ctx, s);
    ArenaUse use(a);
    RTG_REQUIRE(use.ok, RTG_ERROR_STATE,
                "trace: the frame state of stream %p is in use by another call (a context is used by one thread "
                "at a time)",
                (void*)s);
    const Tuning& t = tune ? *tune : ctx->tune;
    TracePlan P;
    int st = plan_jobs(ctx, jobs, njobs, a, check_tlas, t, P);
    if (st != RTG_OK) return st;
    st = upload_jobs(a, P, s);
    if (st != RTG_OK) return st;
    enqueue_trace(ctx, a, P, t, ctx->counters_enabled, true, s);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

}  // namespace rtg

using namespace rtg;

extern "C" int rtg_trace_image(rtg_ctx* ctx, const rtg_image_job* job, void* stream) {
    clear_error();
    RTG_REQUIRE(ctx && job, RTG_ERROR_INVALID_ARGUMENT, "rtg_trace_image: null argument");
    return trace_jobs(ctx, job, 1, reinterpret_cast<hipStream_t>(stream), nullptr, true);
}

extern "C" int rtg_trace_batch(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t n, void* stream) {
    clear_error();
    RTG_REQUIRE(ctx && (jobs || n == 0), RTG_ERROR_INVALID_ARGUMENT, "rtg_trace_batch: null argument");
    if (n == 0) return RTG_OK;
    return trace_jobs(ctx, jobs, n, reinterpret_cast<hipStream_t>(stream), nullptr, true);
}

// Arguments of a ray query (rtg_trace_rays, rtg_trace_rays_static).
static int check_ray_query(rtg_ctx* ctx, rtg_tlas tlas, const rtg_ray* rays, uint32_t count, uint32_t mode,
                           uint32_t ray_flags, rtg_hit* hits, uint32_t* occluded, const char* what) {
    RTG_REQUIRE(ctx, RTG_ERROR_INVALID_ARGUMENT, "%s: null context", what);
    auto it = ctx->structures.find(tlas);
    RTG_REQUIRE(it != ctx->structures.end() && it->second.kind == kKindTlas, RTG_ERROR_NOT_FOUND,
                "%s: 0x%llx is not a TLAS of this context", what, (unsigned long long)tlas);
    RTG_REQUIRE(mode <= RTG_QUERY_OCCLUSION, RTG_ERROR_INVALID_ARGUMENT, "%s: mode %u", what, mode);
    RTG_REQUIRE(count == 0 || rays, RTG_ERROR_INVALID_ARGUMENT, "%s: null rays", what);
    RTG_REQUIRE(mode != RTG_QUERY_CLOSEST || count == 0 || hits, RTG_ERROR_INVALID_ARGUMENT, "%s: null hits", what);
    RTG_REQUIRE(mode != RTG_QUERY_OCCLUSION || count == 0 || occluded, RTG_ERROR_INVALID_ARGUMENT,
                "%s: null occluded", what);
    // the 32-bit feed counter runs up to count + one chunk per wavefront
    RTG_REQUIRE(count <= RTG_MAX_QUERY_RAYS, RTG_ERROR_LIMIT, "%s: %u rays exceed %u per call", what, count,
                (uint32_t)RTG_MAX_QUERY_RAYS);
    const uint32_t allowed = RTG_RAY_FLAG_OPAQUE | RTG_RAY_FLAG_NO_OPAQUE | RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT |
                             RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER;
    RTG_REQUIRE((ray_flags & ~allowed) == 0, RTG_ERROR_INVALID_ARGUMENT, "%s: unsupported ray flags 0x%x", what,
                ray_flags & ~allowed);
    return RTG_OK;
}

// Clears the feed counter and launches the query kernel on `G` persistent
// blocks whose spill stacks start at `spill` (G * kTraceThreads lanes).
static void launch_ray_query(rtg_ctx* ctx, rtg_tlas tlas, const rtg_scene_bindings* scene, const rtg_ray* rays,
                             uint32_t count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                             uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded, uint2* spill, uint32_t* feed,
                             unsigned G, bool timed, hipStream_t s) {
    SceneView S;
    S.models = scene ? reinterpret_cast<const rtg_model_info*>(scene->model_info) : nullptr;
    S.textures = scene ? reinterpret_cast<const rtg_texture*>(scene->textures) : nullptr;
    S.srgb_lut = ctx->d_srgb_lut;
    S.model_count = scene ? scene->model_count : 0;
    S.texture_count = scene ? scene->texture_count : 0;
    const bool cnt = ctx->counters_enabled;
    uint8_t* tl = reinterpret_cast<uint8_t*>(tlas);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, feed, 1u);
    // the primary ray count of a ray query is the query size
    if (cnt) hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, ctx->d_counters, (uint64_t)count);
    LaunchTimer t(ctx, mode == RTG_QUERY_CLOSEST ? RTG_KERNEL_TRACE_CLOSEST : RTG_KERNEL_TRACE_OCCLUSION, s, timed);
    cull_mask &= 0xffu;
#define RTG_LAUNCH_RAYS(OCC, CNT)                                                                                \
    hipLaunchKernelGGL((k_trace_rays<OCC, CNT>), dim3(G), dim3(kTraceThreads), stack_lds_bytes(ctx->tune.stack_lds), \
                       s, tl, S, rays, count, ray_flags, cull_mask, sbt_offset, hits, occluded, spill, feed,         \
                       ctx->d_counters, ctx->d_status, ctx->tune.refill_rays, (uint32_t)ctx->tune.feed_chunk,       \
                       ctx->tune.stack_lds)
    if (mode == RTG_QUERY_CLOSEST) {
        if (cnt) RTG_LAUNCH_RAYS(false, true);
        else RTG_LAUNCH_RAYS(false, false);
    } else {
        if (cnt) RTG_LAUNCH_RAYS(true, true);
        else RTG_LAUNCH_RAYS(true, false);
    }
#undef RTG_LAUNCH_RAYS
}

extern "C" int rtg_trace_rays(rtg_ctx* ctx, rtg_tlas tlas, const rtg_scene_bindings* scene, const rtg_ray* rays,
                              uint32_t count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                              uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded, void* stream) {
    clear_error();
    int st = check_ray_query(ctx, tlas, rays, count, mode, ray_flags, hits, occluded, "rtg_trace_rays");
    if (st != RTG_OK) return st;
    if (count == 0) return RTG_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    RTG_HIP(hipSetDevice(ctx->device));
    FrameArena* a = arena(ctx, s);
    ArenaUse use(a);
    RTG_REQUIRE(use.ok, RTG_ERROR_STATE,
                "rtg_trace_rays: the frame state of stream %p is in use by another call (a context is used by one "
                "thread at a time)",
                (void*)s);
    if (ensure_spill(ctx, a, ctx->tune) != RTG_OK) return RTG_ERROR_OUT_OF_MEMORY;
    const unsigned G = (unsigned)std::min<uint64_t>((count + kTraceThreads - 1) / kTraceThreads, trace_grid(ctx));
    launch_ray_query(ctx, tlas, scene, rays, count, mode, ray_flags, cull_mask, sbt_offset, hits, occluded,
                     reinterpret_cast<uint2*>(a->spill.ptr), reinterpret_cast<uint32_t*>(a->feed.ptr), G, true, s);
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}

// ---------------------------------------------------------------------------
// rtg_trace_rays_static: the ray query with its scratch (the feed counter and
// the spill stacks of kStaticBlocks persistent workgroups) in a file-scope
// device array instead of the context's arenas, for callers that cannot
// allocate. Every GPU holds its own copy of `scratch` while the process
// lives; its address on a device is taken once, at rtg_create, with that
// device current, and only that address is handed to kernels. Calls that
// overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace rtg {
constexpr unsigned kStaticBlocks = RTG_STATIC_QUERY_BLOCKS;
constexpr size_t kStaticFeedBytes = 256;
constexpr size_t kStaticScratchBytes = kStaticFeedBytes + (size_t)kStaticBlocks * kTraceThreads * kSpillLaneBytes;
static_assert(kStaticScratchBytes == RTG_STATIC_SCRATCH_BYTES, "rtg.h states the static scratch size");

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

int static_scratch_init(int device) {
    RTG_REQUIRE(device >= 0 && device < kMaxDevices, RTG_ERROR_LIMIT, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return RTG_OK;
    int cur = -1;
    RTG_HIP(hipGetDevice(&cur));
    RTG_REQUIRE(cur == device, RTG_ERROR_STATE, "static scratch of device %d set up with device %d current", device,
                cur);
    void* p = nullptr;
    RTG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RTG_REQUIRE(p != nullptr, RTG_ERROR_DEVICE, "hipGetSymbolAddress(scratch) gave a null address on device %d",
                device);
    g_scratch_ptr[device] = p;
    return RTG_OK;
}

static void* static_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace rtg

extern "C" int rtg_trace_rays_static(rtg_ctx* ctx, rtg_tlas tlas, const rtg_scene_bindings* scene,
                                     const rtg_ray* rays, uint32_t count, uint32_t mode, uint32_t ray_flags,
                                     uint32_t cull_mask, uint32_t sbt_offset, rtg_hit* hits, uint32_t* occluded,
                                     void* stream) {
    clear_error();
    int st = check_ray_query(ctx, tlas, rays, count, mode, ray_flags, hits, occluded, "rtg_trace_rays_static");
    if (st != RTG_OK) return st;
    if (count == 0) return RTG_OK;
    RTG_HIP(hipSetDevice(ctx->device));
    uint8_t* base = static_cast<uint8_t*>(static_scratch(ctx->device));
    RTG_REQUIRE(base != nullptr, RTG_ERROR_STATE, "rtg_trace_rays_static: no static scratch on device %d",
                ctx->device);
    // the grid is capped by the spill stacks the array holds; the launch
    // shape never changes results
    const unsigned G = (unsigned)std::min<uint64_t>(
        std::min<uint64_t>((count + kTraceThreads - 1) / kTraceThreads, trace_grid(ctx)), kStaticBlocks);
    launch_ray_query(ctx, tlas, scene, rays, count, mode, ray_flags, cull_mask, sbt_offset, hits, occluded,
                     reinterpret_cast<uint2*>(base + kStaticFeedBytes), reinterpret_cast<uint32_t*>(base), G, false,
                     reinterpret_cast<hipStream_t>(stream));
    RTG_HIP(hipGetLastError());
    return RTG_OK;
}
