// rtg_internal.hpp -- host-side state of an rtg context.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/rtg.h"

namespace rtg {

void set_error(const char* fmt, ...);
void clear_error();

#define RTG_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::rtg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                    \
            return RTG_ERROR_DEVICE;                                                       \
        }                                                                                  \
    } while (0)

#define RTG_REQUIRE(cond, code, ...)       \
    do {                                   \
        if (!(cond)) {                     \
            ::rtg::set_error(__VA_ARGS__); \
            return (code);                 \
        }                                  \
    } while (0)

struct AsRecord {
    uint32_t kind;
    uint32_t prim_count;
    uint32_t node_count;
    uint64_t size;
};

// Grow-only device buffer.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t size = 0;
    // grows to at least `bytes` (with a quarter more headroom when `slack`)
    int ensure(size_t bytes, bool slack = true);
    void release();
};

// Bump allocator over a DeviceBuffer (256-byte aligned slices).
struct Carve {
    uint8_t* base;
    size_t off = 0;
    explicit Carve(void* b) : base(static_cast<uint8_t*>(b)) {}
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};
template <class T>
inline void carve_size(size_t& off, size_t count) {
    off = (off + 255) & ~size_t(255);
    off += count * sizeof(T);
}

enum StatusBits : uint32_t {
    kStatusBadIndex = RTG_STATUS_BAD_INDEX,            // BLAS build: index > max_vertex
    kStatusBadBlas = RTG_STATUS_UNKNOWN_BLAS,          // TLAS build/refit: reference is not a live BLAS
    kStatusStackOverflow = RTG_STATUS_STACK_OVERFLOW,  // trace: traversal stack exhausted
    kStatusQueueRange = RTG_STATUS_QUEUE_RANGE,        // trace: queue size or entry out of range
};

struct TimedLaunch {
    uint32_t kind;
    hipEvent_t start, stop;
};

struct FrameArena;  // trace.hip

// Launch parameters of the traversal kernels (defaults, overridable through
// the environment: RTG_REFILL_CLOSEST, RTG_REFILL_OCCLUSION (0 selects the
// non-persistent kernels), RTG_REFILL_RAYS, RTG_PATH_ORDER,
// RTG_TRACE_BLOCKS_PER_CU, RTG_FEED_CHUNK, RTG_BOUNCE_BLOCKS).
// Measured on MI355X (1080p large scene, 4-wide BVH): refilling lanes before
// the whole wave is idle loses 13-16 % on the occlusion kernel, so the
// default refills only an idle wave, one 64-ray batch per atomic; four
// persistent workgroups per CU match the traversal kernels' occupancy
// (4 waves per SIMD at their register count) and beat 5 and 8 slightly.
struct Tuning {
    int refill_closest = 64;
    int refill_occlusion = 64;
    // idle lanes before a wavefront of rtg_trace_rays takes new rays (both
    // modes). Not part of rtg_tuning: a frame's rays start together and end
    // alike, but incoherent queries leave waves with a few long rays, and
    // refilling at 16 idle lanes measured 1,384/2,479 Mrays/s (closest /
    // occlusion; 16M rays, origins uniform in the unit cube, directions
    // uniform on the sphere, into 1M random triangles) against 1,006/1,604
    // at 64; 8 and 24 were within 4 %
    int refill_rays = 16;
    int path_order = 1;
    int blocks_per_cu = 4;
    int feed_chunk = 64;
    // persistent blocks given to the bounce stream beside the first level's
    // shadow rays (0: all levels in sequence on the caller's stream): its
    // shadow-ray launch takes them all, its bounce-level launches at most
    // 128, and the first level's shadow rays the rest of the grid. 192 of
    // 1,024 measured best on the 1080p scene with two frames in flight
    // (3,668-3,738 Mrays/s in seven runs on two boxes); 224, 256 and 320
    // often settled into a slower overlap of the two frames, 384 once in
    // seven runs, and 448 or more slowed the first level's shadow rays
    int bounce_blocks = 192;
    // stack entries per lane kept in LDS by every traversal launch (8-B
    // entries; the single-ray occlusion queries of the ray API keep twice as
    // many 4-B entries); deeper entries spill to the lane's global area
    int stack_lds = 16;
};

}  // namespace rtg

struct rtg_ctx {
    int device = 0;
    int cu_count = 0;
    std::map<uint64_t, rtg::AsRecord> structures;
    // sorted addresses of live BLASes, on the device (TLAS build/refit validate against it)
    uint64_t* d_registry = nullptr;
    uint32_t registry_count = 0;
    std::vector<void*> retired;  // old registry buffers (freed at destroy)
    rtg::DeviceBuffer scratch;   // build scratch arena
    float* d_srgb_lut = nullptr;
    uint32_t* d_status = nullptr;  // sticky kernel status bits
    uint64_t* d_counters = nullptr;  // rtg_trace_counters on the device
    bool counters_enabled = false;
    bool timing = false;
    std::vector<rtg::TimedLaunch> pending;
    double kernel_ms[RTG_KERNEL_COUNT] = {};
    uint64_t kernel_launches[RTG_KERNEL_COUNT] = {};
    // frame state of the context's own calls, one per stream they are made on
    std::map<void*, rtg::FrameArena*> frames;
    rtg::Tuning tune;
};

namespace rtg {
int refresh_registry(rtg_ctx* ctx);
// Event pair around a launch when timing is enabled.
struct LaunchTimer {
    rtg_ctx* ctx;
    uint32_t kind;
    hipStream_t stream;
    hipEvent_t a = nullptr, b = nullptr;
    // `on` false: no events (launches being captured into a graph)
    LaunchTimer(rtg_ctx* c, uint32_t k, hipStream_t s, bool on = true);
    ~LaunchTimer();
};
int check_status(rtg_ctx* ctx, hipStream_t stream, uint32_t mask, const char* what);
void free_frame_arena(rtg_ctx* ctx);
// exact: grow to exactly the sizes asked (a server slot's budgeted arena)
FrameArena* new_frame_arena(bool exact = false);
// Device bytes an arena's frame buffer has after being sized for a trace of
// `paths` paths, `shadow` shadow rays and `njobs` jobs (it only grows).
size_t arena_bytes_after(const FrameArena* a, uint64_t paths, uint64_t shadow, uint32_t njobs);
// Device bytes an arena holds (frame buffer, spill area, feed counter).
size_t arena_held_bytes(const FrameArena* a);
// Frees an idle arena's frame buffer (its next trace allocates it anew).
void arena_reset(FrameArena* a);
// The device word with the status bits of the arena's last trace.
const uint32_t* arena_status_word(const FrameArena* a);
void delete_frame_arena(FrameArena* a);
size_t frame_bytes(uint64_t paths, uint64_t shadow, uint32_t njobs);
// Shadow-queue slots a path with `shadow_rays` rays takes: an even number
// (the occlusion kernel traces the slots in pairs of one path).
inline uint64_t shadow_slots(uint32_t shadow_rays) { return (uint64_t)shadow_rays + (shadow_rays & 1u); }
// The trace step of a batch of jobs on `stream` with the frame state in `a`
// (the context's own arena if null); check_tlas validates each job's TLAS
// against the context's structures.
int trace_jobs(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t njobs, hipStream_t stream, FrameArena* a,
               bool check_tlas,
               const Tuning* tune = nullptr);
// A trace step split in its host half and its device half.
struct TracePlan {
    uint64_t paths = 0, pixels = 0, shadow = 0;
    uint32_t njobs = 0, max_traces = 0;
};
// Host half: validates the jobs, sizes the arena and fills its pinned job
// table (after the previous upload from it has left).
// The arena's spill area is sized for `tune`.
int plan_jobs(rtg_ctx* ctx, const rtg_image_job* jobs, uint32_t njobs, FrameArena* a, bool check_tlas,
              const Tuning& tune, TracePlan& P);
// The frame state of the context's own calls on the null stream (created on
// first use).
FrameArena* context_arena(rtg_ctx* ctx);
// Device pointers of an arena's buffers (a captured graph keeps them).
void arena_pointers(const FrameArena* a, const void* out[3]);
// Copies the pinned job table to the device on `stream` (asynchronous).
int upload_jobs(FrameArena* a, const TracePlan& P, hipStream_t stream);
// Device half: the kernels of the step and nothing else (no copy, memset or
// allocation), so that it can be captured into a graph. Launch parameters
// come from `tune`; `timed` records kernel events when timing is on.
void enqueue_trace(rtg_ctx* ctx, FrameArena* a, const TracePlan& P, const Tuning& tune, bool counters, bool timed,
                   hipStream_t stream);
// Device bytes of the spill area of the persistent traversal grid.
size_t spill_bytes(rtg_ctx* ctx, const Tuning& tune);
// Launches made by the calling thread are not timed (a server's worker).
void suppress_timing_on_this_thread();
// Takes the device address of rtg_trace_rays_static's scratch array on
// `device` (which must be current); once per device, at rtg_create.
int static_scratch_init(int device);
// One-kernel refit of the TLAS blob at `tlas` (no registry lookup).
int launch_refit(rtg_ctx* ctx, uint8_t* tlas, const rtg_instance* instances, uint32_t count, hipStream_t stream);
}  // namespace rtg
