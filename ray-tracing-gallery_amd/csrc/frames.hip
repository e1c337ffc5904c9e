// frames.hip -- the host frame driver: frames in flight, each with its own
// TLAS copy, instance buffers and frame state.
//
// The reference keeps two frames in flight (src/main.rs:600-668): each Frame
// owns a clone of the TLAS, an instance buffer the host writes through a
// mapping, and its own output image. Per frame it waits for that frame's
// fence, lets the scene write the new instance transforms
// (DefaultScene::write_resources, src/scene.rs:167-204), refits that frame's
// TLAS in place and records the trace (src/main.rs:916-996). Here a slot is
// one frame in flight: a TLAS copy registered with the context, a pinned host
// instance array (the mapping), a device instance array, a frame arena and an
// event that marks the end of the slot's last frame.
#include <cstring>
#include <vector>

#include "rtg_device.hpp"
#include "rtg_internal.hpp"

namespace rtg {

struct FrameSlot {
    rtg_tlas tlas = 0;
    rtg_instance* host = nullptr;  // pinned
    rtg_instance* dev = nullptr;
    FrameArena* arena = nullptr;
    hipEvent_t done = nullptr;
    bool used = false;
};

}  // namespace rtg

using namespace rtg;

struct rtg_frames {
    rtg_ctx* ctx = nullptr;
    uint32_t count = 0;  // instances of the TLAS
    std::vector<FrameSlot> slots;
};

extern "C" int rtg_tlas_clone(rtg_ctx* ctx, rtg_tlas src, rtg_tlas* out, void* stream);
extern "C" int rtg_tlas_destroy(rtg_ctx* ctx, rtg_tlas tlas);

static void frames_free(rtg_frames* f) {
    if (!f) return;
    hipSetDevice(f->ctx->device);
    for (FrameSlot& s : f->slots) {
        if (s.done) hipEventSynchronize(s.done);
    }
    hipDeviceSynchronize();
    for (FrameSlot& s : f->slots) {
        if (s.tlas) rtg_tlas_destroy(f->ctx, s.tlas);
        if (s.host) hipHostFree(s.host);
        if (s.dev) hipFree(s.dev);
        if (s.done) hipEventDestroy(s.done);
        delete_frame_arena(s.arena);
    }
    delete f;
}

extern "C" int rtg_frames_create(rtg_ctx* ctx, rtg_tlas tlas, uint32_t frames_in_flight, rtg_frames** out) {
    clear_error();
    RTG_REQUIRE(ctx && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_frames_create: null argument");
    RTG_REQUIRE(frames_in_flight >= 1 && frames_in_flight <= 8, RTG_ERROR_INVALID_ARGUMENT,
                "rtg_frames_create: %u frames in flight (1..8)", frames_in_flight);
    *out = nullptr;
    auto it = ctx->structures.find(tlas);
    RTG_REQUIRE(it != ctx->structures.end() && it->second.kind == kKindTlas, RTG_ERROR_NOT_FOUND,
                "rtg_frames_create: 0x%llx is not a TLAS of this context", (unsigned long long)tlas);
    RTG_HIP(hipSetDevice(ctx->device));
    rtg_frames* f = new rtg_frames();
    f->ctx = ctx;
    f->count = it->second.prim_count;
    f->slots.resize(frames_in_flight);
    const size_t bytes = std::max<size_t>(1, f->count) * sizeof(rtg_instance);
    // the template's source instances seed every slot's mapping
    const AsHeader* h = nullptr;
    AsHeader hd;
    hipError_t e = hipMemcpy(&hd, reinterpret_cast<const void*>(tlas), sizeof(hd), hipMemcpyDeviceToHost);
    if (e == hipSuccess) h = &hd;
    for (FrameSlot& s : f->slots) {
        int st = e == hipSuccess ? rtg_tlas_clone(ctx, tlas, &s.tlas, nullptr) : RTG_ERROR_DEVICE;
        if (st == RTG_OK && (hipHostMalloc(reinterpret_cast<void**>(&s.host), bytes) != hipSuccess ||
                             hipMalloc(&s.dev, bytes) != hipSuccess ||
                             hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess))
            st = RTG_ERROR_OUT_OF_MEMORY;
        if (st == RTG_OK && f->count &&
            hipMemcpy(s.host, reinterpret_cast<const uint8_t*>(tlas) + h->extra_off, (size_t)f->count * 64,
                      hipMemcpyDeviceToHost) != hipSuccess)
            st = RTG_ERROR_DEVICE;
        if (st != RTG_OK) {
            if (e != hipSuccess) set_error("rtg_frames_create: %s", hipGetErrorString(e));
            frames_free(f);
            return st;
        }
        s.arena = new_frame_arena();
    }
    // the clones were copied on the null stream; callers may use non-blocking streams
    RTG_HIP(hipDeviceSynchronize());
    *out = f;
    return RTG_OK;
}

extern "C" int rtg_frames_begin(rtg_frames* f, uint64_t frame, rtg_instance** host_instances, rtg_tlas* slot_tlas) {
    clear_error();
    RTG_REQUIRE(f, RTG_ERROR_INVALID_ARGUMENT, "rtg_frames_begin: null driver");
    FrameSlot& s = f->slots[frame % f->slots.size()];
    RTG_HIP(hipSetDevice(f->ctx->device));
    // ≙ waiting for the frame's fence: its last refit, trace and upload are done
    if (s.used) RTG_HIP(hipEventSynchronize(s.done));
    if (host_instances) *host_instances = s.host;
    if (slot_tlas) *slot_tlas = s.tlas;
    return RTG_OK;
}

extern "C" int rtg_frames_submit(rtg_frames* f, uint64_t frame, const rtg_image_job* job, void* stream) {
    clear_error();
    RTG_REQUIRE(f && job, RTG_ERROR_INVALID_ARGUMENT, "rtg_frames_submit: null argument");
    FrameSlot& s = f->slots[frame % f->slots.size()];
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    RTG_HIP(hipSetDevice(f->ctx->device));
    if (f->count)
        RTG_HIP(hipMemcpyAsync(s.dev, s.host, (size_t)f->count * sizeof(rtg_instance), hipMemcpyHostToDevice, st));
    int r = launch_refit(f->ctx, reinterpret_cast<uint8_t*>(s.tlas), s.dev, f->count, st);
    if (r != RTG_OK) return r;
    rtg_image_job j = *job;
    j.tlas = s.tlas;
    r = trace_jobs(f->ctx, &j, 1, st, s.arena, true);
    if (r != RTG_OK) return r;
    RTG_HIP(hipEventRecord(s.done, st));
    s.used = true;
    return RTG_OK;
}

extern "C" int rtg_frames_destroy(rtg_frames* f) {
    clear_error();
    RTG_REQUIRE(f, RTG_ERROR_INVALID_ARGUMENT, "rtg_frames_destroy: null driver");
    frames_free(f);
    return RTG_OK;
}
