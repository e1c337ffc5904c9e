// context.hip -- rtg context lifetime, errors, arenas, timing and counters.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "rtg_device.hpp"
#include "rtg_internal.hpp"

namespace rtg {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}
void clear_error() { g_error[0] = 0; }

int DeviceBuffer::ensure(size_t bytes, bool slack) {
    if (bytes <= size) return RTG_OK;
    if (ptr) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            set_error("hipDeviceSynchronize: %s", hipGetErrorString(e));
            return RTG_ERROR_DEVICE;
        }
        hipFree(ptr);
        ptr = nullptr;
        size = 0;
    }
    size_t want = slack ? bytes + bytes / 4 : bytes;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return RTG_ERROR_OUT_OF_MEMORY;
    }
    size = want;
    return RTG_OK;
}

void DeviceBuffer::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    size = 0;
}

static thread_local bool t_no_timing = false;
void suppress_timing_on_this_thread() { t_no_timing = true; }

LaunchTimer::LaunchTimer(rtg_ctx* c, uint32_t k, hipStream_t s, bool on) : ctx(c), kind(k), stream(s) {
    if (!on || !ctx->timing || t_no_timing) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        a = b = nullptr;
        return;
    }
    hipEventRecord(a, stream);
}

LaunchTimer::~LaunchTimer() {
    if (!a) return;
    hipEventRecord(b, stream);
    ctx->pending.push_back(TimedLaunch{kind, a, b});
}

static int drain_timers(rtg_ctx* ctx) {
    for (auto& t : ctx->pending) {
        hipEventSynchronize(t.stop);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            ctx->kernel_ms[t.kind] += ms;
            ctx->kernel_launches[t.kind] += 1;
        }
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    ctx->pending.clear();
    return RTG_OK;
}

}  // namespace rtg

using namespace rtg;

extern "C" const char* rtg_last_error(void) { return g_error; }
extern "C" uint32_t rtg_abi_version(void) { return RTG_ABI_VERSION; }

extern "C" int rtg_create(rtg_ctx** out, int device) {
    clear_error();
    RTG_REQUIRE(out, RTG_ERROR_INVALID_ARGUMENT, "rtg_create: null output");
    *out = nullptr;
    int count = 0;
    RTG_HIP(hipGetDeviceCount(&count));
    RTG_REQUIRE(device >= 0 && device < count, RTG_ERROR_INVALID_ARGUMENT, "rtg_create: device %d of %d", device,
                count);
    RTG_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    RTG_HIP(hipGetDeviceProperties(&prop, device));
    int sst = static_scratch_init(device);
    if (sst != RTG_OK) return sst;
    rtg_ctx* ctx = new rtg_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    auto env_int = [](const char* name, int def, int lo, int hi) {
        const char* v = getenv(name);
        if (!v || !*v) return def;
        int x = atoi(v);
        return x < lo ? lo : (x > hi ? hi : x);
    };
    ctx->tune.refill_closest = env_int("RTG_REFILL_CLOSEST", ctx->tune.refill_closest, 0, 64);
    ctx->tune.refill_occlusion = env_int("RTG_REFILL_OCCLUSION", ctx->tune.refill_occlusion, 0, 64);
    ctx->tune.refill_rays = env_int("RTG_REFILL_RAYS", ctx->tune.refill_rays, 0, 64);
    ctx->tune.path_order = env_int("RTG_PATH_ORDER", ctx->tune.path_order, 0, 1);
    ctx->tune.blocks_per_cu = env_int("RTG_TRACE_BLOCKS_PER_CU", ctx->tune.blocks_per_cu, 1, 8);
    // RTG_FEED_CHUNK counts 64-ray batches taken per atomic
    ctx->tune.feed_chunk = 64 * env_int("RTG_FEED_CHUNK", ctx->tune.feed_chunk / 64, 1, 16);
    ctx->tune.bounce_blocks = env_int("RTG_BOUNCE_BLOCKS", ctx->tune.bounce_blocks, 0, 4096);
    // RTG_STACK_LDS: stack entries per lane in LDS, rounded down to a multiple of 4
    ctx->tune.stack_lds = env_int("RTG_STACK_LDS", ctx->tune.stack_lds, kStackLdsMin, kStackLdsMax) / 4 * 4;
    // IEC 61966-2-1 decode of the 256 sRGB codes, in double, rounded once.
    float lut[256];
    for (int i = 0; i < 256; ++i) {
        double c = i / 255.0;
        double l = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
        lut[i] = (float)l;
    }
    if (hipMalloc(&ctx->d_srgb_lut, sizeof(lut)) != hipSuccess || hipMalloc(&ctx->d_status, 4) != hipSuccess ||
        hipMalloc(&ctx->d_counters, sizeof(rtg_trace_counters)) != hipSuccess) {
        set_error("rtg_create: device allocation failed");
        rtg_destroy(ctx);
        return RTG_ERROR_OUT_OF_MEMORY;
    }
    hipMemcpy(ctx->d_srgb_lut, lut, sizeof(lut), hipMemcpyHostToDevice);
    hipMemset(ctx->d_status, 0, 4);
    hipMemset(ctx->d_counters, 0, sizeof(rtg_trace_counters));
    int st = refresh_registry(ctx);
    if (st != RTG_OK) {
        rtg_destroy(ctx);
        return st;
    }
    *out = ctx;
    return RTG_OK;
}

extern "C" void rtg_destroy(rtg_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    drain_timers(ctx);
    for (auto& kv : ctx->structures) hipFree(reinterpret_cast<void*>(kv.first));
    ctx->structures.clear();
    for (void* p : ctx->retired) hipFree(p);
    if (ctx->d_registry) hipFree(ctx->d_registry);
    ctx->scratch.release();
    free_frame_arena(ctx);
    if (ctx->d_srgb_lut) hipFree(ctx->d_srgb_lut);
    if (ctx->d_status) hipFree(ctx->d_status);
    if (ctx->d_counters) hipFree(ctx->d_counters);
    delete ctx;
}

extern "C" int rtg_set_counters(rtg_ctx* ctx, int enable) {
    clear_error();
    RTG_REQUIRE(ctx, RTG_ERROR_INVALID_ARGUMENT, "rtg_set_counters: null context");
    ctx->counters_enabled = enable != 0;
    return RTG_OK;
}

extern "C" int rtg_get_counters(rtg_ctx* ctx, rtg_trace_counters* out) {
    clear_error();
    RTG_REQUIRE(ctx && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_get_counters: null argument");
    RTG_HIP(hipDeviceSynchronize());
    RTG_HIP(hipMemcpy(out, ctx->d_counters, sizeof(*out), hipMemcpyDeviceToHost));
    // the device keeps the bounce stream's traversal counts apart; the
    // totals include them
    out->occ_node_visits += out->side_occ_node_visits;
    out->occ_triangle_tests += out->side_occ_triangle_tests;
    out->occ_instance_visits += out->side_occ_instance_visits;
    out->occ_anyhit_calls += out->side_occ_anyhit_calls;
    out->node_visits += out->side_node_visits;
    out->triangle_tests += out->side_triangle_tests;
    out->instance_visits += out->side_instance_visits;
    out->anyhit_calls += out->side_anyhit_calls;
    out->spill_pushes += out->side_spill_pushes;
    out->occ_spill_pushes += out->side_occ_spill_pushes;
    return RTG_OK;
}

extern "C" int rtg_get_status(rtg_ctx* ctx, uint32_t* bits, int clear) {
    clear_error();
    RTG_REQUIRE(ctx && bits, RTG_ERROR_INVALID_ARGUMENT, "rtg_get_status: null argument");
    RTG_HIP(hipSetDevice(ctx->device));
    RTG_HIP(hipDeviceSynchronize());
    RTG_HIP(hipMemcpy(bits, ctx->d_status, 4, hipMemcpyDeviceToHost));
    if (clear) RTG_HIP(hipMemset(ctx->d_status, 0, 4));
    return RTG_OK;
}

extern "C" int rtg_set_timing(rtg_ctx* ctx, int enable) {
    clear_error();
    RTG_REQUIRE(ctx, RTG_ERROR_INVALID_ARGUMENT, "rtg_set_timing: null context");
    drain_timers(ctx);
    ctx->timing = enable != 0;
    for (uint32_t k = 0; k < RTG_KERNEL_COUNT; ++k) {
        ctx->kernel_ms[k] = 0.0;
        ctx->kernel_launches[k] = 0;
    }
    return RTG_OK;
}

extern "C" int rtg_get_kernel_time(rtg_ctx* ctx, uint32_t kind, double* total_ms, uint64_t* launches) {
    clear_error();
    RTG_REQUIRE(ctx && kind < RTG_KERNEL_COUNT, RTG_ERROR_INVALID_ARGUMENT, "rtg_get_kernel_time: bad argument");
    drain_timers(ctx);
    if (total_ms) *total_ms = ctx->kernel_ms[kind];
    if (launches) *launches = ctx->kernel_launches[kind];
    return RTG_OK;
}
