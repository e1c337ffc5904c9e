// batch.hip -- launch tuning, the autotuner and graph-captured batches.
//
// A batch is a fixed set of image jobs (different scenes, resolutions and
// sample counts) whose whole trace step is recorded once into a HIP graph by
// stream capture and then replayed with one hipGraphLaunch. The captured
// kernels are the ones rtg_trace_batch launches, with the frame state in an
// arena the batch owns; nothing in the step is a memset or a copy (counters
// and queue sizes are cleared by a kernel), so every replay starts from the
// same state. The job table is device memory read by the kernels at run time,
// so rtg_batch_update can change TLAS, uniforms and outputs between replays.
//
// The autotuner times candidate launch settings on the caller's own jobs.
// The settings never change results (traversal order and arithmetic are the
// same under every setting), only the launch shape: persistent workgroups
// per CU, rays per feed atomic, the refill threshold of the persistent
// kernels (or the non-persistent kernels) and the pixel order of paths.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "rtg_device.hpp"
#include "rtg_internal.hpp"

namespace rtg {

static int check_tuning(const rtg_tuning& t) {
    RTG_REQUIRE(t.blocks_per_cu >= 1 && t.blocks_per_cu <= 8, RTG_ERROR_INVALID_ARGUMENT,
                "tuning: blocks_per_cu %u (1..8)", t.blocks_per_cu);
    RTG_REQUIRE(t.feed_chunk >= 64 && t.feed_chunk <= 1024 && t.feed_chunk % 64 == 0, RTG_ERROR_INVALID_ARGUMENT,
                "tuning: feed_chunk %u (64..1024, multiple of 64)", t.feed_chunk);
    RTG_REQUIRE(t.refill_closest <= 64 && t.refill_occlusion <= 64, RTG_ERROR_INVALID_ARGUMENT,
                "tuning: refill thresholds %u/%u (0..64)", t.refill_closest, t.refill_occlusion);
    RTG_REQUIRE(t.path_order <= 1, RTG_ERROR_INVALID_ARGUMENT, "tuning: path_order %u (0..1)", t.path_order);
    RTG_REQUIRE(t.bounce_blocks <= 4096, RTG_ERROR_INVALID_ARGUMENT, "tuning: bounce_blocks %u (0..4096)",
                t.bounce_blocks);
    RTG_REQUIRE(t.stack_lds_entries == 0 || (t.stack_lds_entries >= (uint32_t)kStackLdsMin &&
                                             t.stack_lds_entries <= (uint32_t)kStackLdsMax &&
                                             t.stack_lds_entries % 4 == 0),
                RTG_ERROR_INVALID_ARGUMENT, "tuning: stack_lds_entries %u (0, or 8..32 in steps of 4)",
                t.stack_lds_entries);
    RTG_REQUIRE(t.reserved == 0, RTG_ERROR_INVALID_ARGUMENT, "tuning: reserved field must be zero");
    return RTG_OK;
}

// rtg_tuning leaves out the ray-query refill threshold, which is kept
static Tuning from_abi(const rtg_tuning& t, const Tuning& current) {
    Tuning u;
    u.refill_rays = current.refill_rays;
    u.blocks_per_cu = (int)t.blocks_per_cu;
    u.feed_chunk = (int)t.feed_chunk;
    u.refill_closest = (int)t.refill_closest;
    u.refill_occlusion = (int)t.refill_occlusion;
    u.path_order = (int)t.path_order;
    u.bounce_blocks = (int)t.bounce_blocks;
    u.stack_lds = t.stack_lds_entries ? (int)t.stack_lds_entries : kStackLdsDefault;
    return u;
}

static rtg_tuning to_abi(const Tuning& u) {
    rtg_tuning t;
    memset(&t, 0, sizeof(t));
    t.blocks_per_cu = (uint32_t)u.blocks_per_cu;
    t.feed_chunk = (uint32_t)u.feed_chunk;
    t.refill_closest = (uint32_t)u.refill_closest;
    t.refill_occlusion = (uint32_t)u.refill_occlusion;
    t.path_order = (uint32_t)u.path_order;
    t.bounce_blocks = (uint32_t)u.bounce_blocks;
    t.stack_lds_entries = (uint32_t)u.stack_lds;
    return t;
}

static bool operator<(const Tuning& a, const Tuning& b) {
    int x[7] = {a.blocks_per_cu, a.feed_chunk,    a.refill_closest, a.refill_occlusion,
                a.path_order,    a.bounce_blocks, a.stack_lds};
    int y[7] = {b.blocks_per_cu, b.feed_chunk,    b.refill_closest, b.refill_occlusion,
                b.path_order,    b.bounce_blocks, b.stack_lds};
    return std::lexicographical_compare(x, x + 7, y, y + 7);
}

// Owns a stream and an event pair for the autotuner's timed runs, which use
// the context's own frame state.
struct TuneRig {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FrameArena* arena = nullptr;
    ~TuneRig() {
        if (s) hipStreamSynchronize(s);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (s) hipStreamDestroy(s);
    }
};

// Median time (ms) of `repeats` runs of the jobs under `t`, after a warm-up.
static int time_setting(rtg_ctx* ctx, TuneRig& rig, const rtg_image_job* jobs, uint32_t n, uint32_t repeats,
                        const Tuning& t, double& ms) {
    TracePlan P;
    std::vector<double> v;
    for (uint32_t r = 0; r <= repeats; ++r) {
        int st = plan_jobs(ctx, jobs, n, rig.arena, true, t, P);
        if (st != RTG_OK) return st;
        st = upload_jobs(rig.arena, P, rig.s);
        if (st != RTG_OK) return st;
        RTG_HIP(hipEventRecord(rig.e0, rig.s));
        enqueue_trace(ctx, rig.arena, P, t, false, false, rig.s);
        RTG_HIP(hipGetLastError());
        RTG_HIP(hipEventRecord(rig.e1, rig.s));
        RTG_HIP(hipEventSynchronize(rig.e1));
        float e = 0.0f;
        RTG_HIP(hipEventElapsedTime(&e, rig.e0, rig.e1));
        if (r > 0) v.push_back(e);
    }
    std::sort(v.begin(), v.end());
    ms = v[v.size() / 2];
    return RTG_OK;
}

}  // namespace rtg

using namespace rtg;

extern "C" int rtg_get_tuning(rtg_ctx* ctx, rtg_tuning* out) {
    clear_error();
    RTG_REQUIRE(ctx && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_get_tuning: null argument");
    *out = to_abi(ctx->tune);
    return RTG_OK;
}

extern "C" int rtg_set_tuning(rtg_ctx* ctx, const rtg_tuning* tuning) {
    clear_error();
    RTG_REQUIRE(ctx && tuning, RTG_ERROR_INVALID_ARGUMENT, "rtg_set_tuning: null argument");
    int st = check_tuning(*tuning);
    if (st != RTG_OK) return st;
    ctx->tune = from_abi(*tuning, ctx->tune);
    return RTG_OK;
}

extern "C" int rtg_autotune(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t n, uint32_t repeats, rtg_tuning* best,
                            double* best_ms) {
    clear_error();
    RTG_REQUIRE(ctx && jobs && n > 0, RTG_ERROR_INVALID_ARGUMENT, "rtg_autotune: no jobs");
    RTG_REQUIRE(repeats >= 1 && repeats <= 64, RTG_ERROR_INVALID_ARGUMENT, "rtg_autotune: repeats %u (1..64)",
                repeats);
    RTG_HIP(hipSetDevice(ctx->device));
    // the timed runs reuse the context's frame state: let earlier work on it finish
    RTG_HIP(hipDeviceSynchronize());
    TuneRig rig;
    rig.arena = context_arena(ctx);
    RTG_HIP(hipStreamCreateWithFlags(&rig.s, hipStreamNonBlocking));
    RTG_HIP(hipEventCreate(&rig.e0));
    RTG_HIP(hipEventCreate(&rig.e1));
    // Coordinate search: each pass tries every value of one parameter with
    // the others at the best setting so far; two passes.
    const std::vector<int> axes[7] = {
        {4, 5, 6, 8},         // blocks_per_cu
        {64, 128, 256},       // feed_chunk
        {64, 48, 32, 0},      // refill_closest
        {64, 48, 32, 0},      // refill_occlusion
        {1, 0},               // path_order
        {192, 384, 128, 256, 0},  // bounce_blocks
        {16, 8, 12, 24, 32},     // stack_lds
    };
    std::map<Tuning, double> seen;
    Tuning cur = ctx->tune;
    double cur_ms = 0.0;
    int st = time_setting(ctx, rig, jobs, n, repeats, cur, cur_ms);
    if (st != RTG_OK) return st;
    seen[cur] = cur_ms;
    for (int pass = 0; pass < 2; ++pass) {
        for (int axis = 0; axis < 7; ++axis) {
            for (int value : axes[axis]) {
                Tuning t = cur;
                int* field[7] = {&t.blocks_per_cu, &t.feed_chunk, &t.refill_closest, &t.refill_occlusion,
                                 &t.path_order,    &t.bounce_blocks, &t.stack_lds};
                *field[axis] = value;
                if (seen.count(t)) continue;
                double ms = 0.0;
                st = time_setting(ctx, rig, jobs, n, repeats, t, ms);
                if (st != RTG_OK) return st;
                seen[t] = ms;
                if (ms < cur_ms) {
                    cur = t;
                    cur_ms = ms;
                }
            }
        }
    }
    ctx->tune = cur;
    if (best) *best = to_abi(cur);
    if (best_ms) *best_ms = cur_ms;
    return RTG_OK;
}

struct rtg_batch {
    rtg_ctx* ctx = nullptr;
    FrameArena* arena = nullptr;
    TracePlan plan;
    Tuning tune;
    bool counters = false;
    std::vector<uint32_t> shape;  // width, height, spp, shadow_rays, max_traces per job
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // completion of the last launch: a launch on another stream waits for it,
    // since every launch uses the batch's one frame state
    hipEvent_t done = nullptr;
    hipStream_t last_stream = nullptr;
    bool launched = false;
};

static void batch_free(rtg_batch* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    hipDeviceSynchronize();
    if (b->done) hipEventDestroy(b->done);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    delete_frame_arena(b->arena);
    delete b;
}

static std::vector<uint32_t> job_shape(const rtg_image_job* jobs, uint32_t n) {
    std::vector<uint32_t> v;
    v.reserve(5 * n);
    for (uint32_t k = 0; k < n; ++k) {
        v.push_back(jobs[k].width);
        v.push_back(jobs[k].height);
        v.push_back(jobs[k].opts.spp);
        v.push_back(jobs[k].opts.shadow_rays);
        v.push_back(jobs[k].opts.max_traces);
    }
    return v;
}

static int batch_capture(rtg_batch* b, const rtg_image_job* jobs, uint32_t n) {
    rtg_ctx* ctx = b->ctx;
    int st = plan_jobs(ctx, jobs, n, b->arena, true, b->tune, b->plan);
    if (st != RTG_OK) return st;
    hipStream_t cs = nullptr;
    RTG_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    st = upload_jobs(b->arena, b->plan, cs);
    hipError_t e = hipStreamSynchronize(cs);
    if (st == RTG_OK && e != hipSuccess) {
        set_error("rtg_batch_create: job upload: %s", hipGetErrorString(e));
        st = RTG_ERROR_DEVICE;
    }
    if (st == RTG_OK) {
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            enqueue_trace(ctx, b->arena, b->plan, b->tune, b->counters, false, cs);
            e = hipStreamEndCapture(cs, &b->graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            set_error("rtg_batch_create: graph capture: %s", hipGetErrorString(e));
            st = RTG_ERROR_DEVICE;
        }
    }
    hipStreamDestroy(cs);
    return st;
}

extern "C" int rtg_batch_create(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t n, rtg_batch** out) {
    clear_error();
    RTG_REQUIRE(ctx && out && jobs && n > 0, RTG_ERROR_INVALID_ARGUMENT, "rtg_batch_create: no jobs");
    *out = nullptr;
    RTG_HIP(hipSetDevice(ctx->device));
    rtg_batch* b = new rtg_batch();
    b->ctx = ctx;
    b->arena = new_frame_arena();
    b->tune = ctx->tune;
    b->counters = ctx->counters_enabled;
    b->shape = job_shape(jobs, n);
    int st = batch_capture(b, jobs, n);
    if (st != RTG_OK) {
        batch_free(b);
        return st;
    }
    *out = b;
    return RTG_OK;
}

extern "C" int rtg_batch_update(rtg_batch* b, const rtg_image_job* jobs, uint32_t n, void* stream) {
    clear_error();
    RTG_REQUIRE(b && jobs, RTG_ERROR_INVALID_ARGUMENT, "rtg_batch_update: null argument");
    RTG_REQUIRE(n == b->plan.njobs && job_shape(jobs, n) == b->shape, RTG_ERROR_INVALID_ARGUMENT,
                "rtg_batch_update: job count, sizes and options must equal the captured batch's");
    RTG_HIP(hipSetDevice(b->ctx->device));
    // The captured kernels hold the arena's pointers: the same shapes and
    // tuning never grow it, and this is checked.
    const void* before[3];
    const void* after[3];
    arena_pointers(b->arena, before);
    TracePlan P;
    int st = plan_jobs(b->ctx, jobs, n, b->arena, true, b->tune, P);
    if (st != RTG_OK) return st;
    arena_pointers(b->arena, after);
    RTG_REQUIRE(before[0] == after[0] && before[1] == after[1] && before[2] == after[2], RTG_ERROR_DEVICE,
                "rtg_batch_update: frame state moved");
    return upload_jobs(b->arena, P, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int rtg_batch_launch(rtg_batch* b, void* stream) {
    clear_error();
    RTG_REQUIRE(b, RTG_ERROR_INVALID_ARGUMENT, "rtg_batch_launch: null batch");
    RTG_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!b->done) RTG_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    if (b->launched && s != b->last_stream) RTG_HIP(hipStreamWaitEvent(s, b->done, 0));
    {
        LaunchTimer t(b->ctx, RTG_KERNEL_BATCH_GRAPH, s);
        RTG_HIP(hipGraphLaunch(b->exec, s));
    }
    RTG_HIP(hipEventRecord(b->done, s));
    b->last_stream = s;
    b->launched = true;
    return RTG_OK;
}

extern "C" int rtg_batch_destroy(rtg_batch* b) {
    clear_error();
    RTG_REQUIRE(b, RTG_ERROR_INVALID_ARGUMENT, "rtg_batch_destroy: null batch");
    batch_free(b);
    return RTG_OK;
}
