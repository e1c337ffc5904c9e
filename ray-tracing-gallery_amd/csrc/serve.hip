// serve.hip -- request loop over resident scenes with continuous batching.
//
// The reference renders one frame per event-loop turn with two frames in
// flight (src/main.rs:600-668, :916-996), each with its own TLAS copy,
// refitted from that frame's instance transforms (src/scene.rs:167-204).
// The server generalises that to many requests: each request carries its own
// transforms, gets a private copy of its scene's TLAS that is refitted, and
// is traced together with the other requests of its batch. A worker thread
// keeps up to max_inflight batches on the device (one stream and one frame
// arena per slot) and forms a new batch from the pending requests, in
// arrival order, whenever a slot is free, as long as the batch fits the
// slot's share of the budget and the configured number of samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "rtg_device.hpp"
#include "rtg_internal.hpp"

namespace rtg {

struct ServerScene {
    uint64_t tlas;
    uint64_t tlas_bytes;
    uint32_t instance_count;
    rtg_scene_bindings bindings;
};

struct PendingRequest {
    rtg_request req;
    ServerScene scene;  // copied at submit: the worker never reads the scene list
    uint64_t ticket;
    double submit_s;
    uint64_t bytes;  // frame state + TLAS copy
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    FrameArena* arena = nullptr;
    uint32_t* h_status = nullptr;  // pinned copy of the status bits of the slot's last batch
    DeviceBuffer tlas_copies;
    bool busy = false;
    uint64_t bytes = 0;
    double start_s = 0.0;
    std::vector<PendingRequest> batch;
};

}  // namespace rtg

using namespace rtg;

struct rtg_server {
    rtg_ctx* ctx = nullptr;
    rtg_server_config cfg{};
    Tuning tune;                 // the context's launch setting when the server was created
    uint64_t slot_fixed = 0;     // device bytes every slot holds: its traversal spill area and feed counter
    uint64_t slot_capacity = 0;  // device bytes a slot's batch may take beyond slot_fixed
    std::vector<ServerScene> scenes;
    std::vector<Slot> slots;
    std::deque<PendingRequest> pending;
    std::vector<rtg_completion> completed;
    uint64_t next_ticket = 1;
    uint64_t submitted = 0, finished = 0;
    rtg_server_stats stats{};
    uint64_t inflight_bytes = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    std::atomic<bool> stop{false};
    std::chrono::steady_clock::time_point t0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    // completes a request that never ran (cancelled, timed out; mutex held)
    void complete_unrun(const PendingRequest& p, int32_t status, double t) {
        rtg_completion c{};
        c.ticket = p.ticket;
        c.tag = p.req.tag;
        c.status = status;
        c.batch_size = 0;
        c.submit_s = p.submit_s;
        c.start_s = t;
        c.done_s = t;
        completed.push_back(c);
        ++finished;
    }
    void run();
    bool launch(Slot& s);
    // device bytes held by all slots (frame buffers, spill areas, TLAS copies)
    uint64_t held_bytes() const {
        uint64_t b = 0;
        for (const Slot& sl : slots) b += arena_held_bytes(sl.arena) + sl.tlas_copies.size;
        return b;
    }
};

namespace {

// Device bytes of a request traced alone: its frame state and TLAS copy
// (the slot's spill area is charged apart, once per slot).
uint64_t request_bytes(const rtg_request& r, const ServerScene& sc) {
    uint64_t paths = (uint64_t)r.width * r.height * r.opts.spp;
    uint64_t bytes = frame_bytes(paths, paths * shadow_slots(r.opts.shadow_rays), 1);
    if (r.instances) bytes += (sc.tlas_bytes + 255) & ~uint64_t(255);
    return bytes;
}

// Status of a completed request from its batch's device status bits.
int32_t completion_status(hipError_t q, uint32_t bits) {
    if (q != hipSuccess) return RTG_ERROR_DEVICE;
    return (bits & (kStatusStackOverflow | kStatusQueueRange)) ? RTG_ERROR_RESULT : RTG_OK;
}

}  // namespace

// Launches the slot's batch (mutex held by the caller). Returns false on a
// launch error, which completes the batch with that status.
bool rtg_server::launch(Slot& s) {
    size_t copy_bytes = 0;
    uint64_t paths = 0, shadow = 0;
    for (auto& p : s.batch) {
        if (p.req.instances) copy_bytes += (p.scene.tlas_bytes + 255) & ~uint64_t(255);
        uint64_t rp = (uint64_t)p.req.width * p.req.height * p.req.opts.spp;
        paths += rp;
        shadow += rp * shadow_slots(p.req.opts.shadow_rays);
    }
    // Buffers only grow, each to the largest batch in each dimension; when
    // that would take the slot past its share of the budget, the idle slot's
    // buffers are freed and allocated anew for this batch alone (which was
    // admitted because it fits).
    if (arena_bytes_after(s.arena, paths, shadow, (uint32_t)s.batch.size()) +
            std::max<uint64_t>(copy_bytes, s.tlas_copies.size) >
        slot_capacity) {
        arena_reset(s.arena);
        s.tlas_copies.release();
    }
    if (copy_bytes && s.tlas_copies.ensure(copy_bytes, false) != RTG_OK) return false;
    std::vector<rtg_image_job> jobs(s.batch.size());
    size_t off = 0;
    for (size_t k = 0; k < s.batch.size(); ++k) {
        const rtg_request& r = s.batch[k].req;
        const ServerScene& sc = s.batch[k].scene;
        uint64_t tlas = sc.tlas;
        if (r.instances) {
            uint8_t* copy = static_cast<uint8_t*>(s.tlas_copies.ptr) + off;
            off += (sc.tlas_bytes + 255) & ~uint64_t(255);
            if (hipMemcpyAsync(copy, reinterpret_cast<const void*>(sc.tlas), sc.tlas_bytes, hipMemcpyDeviceToDevice,
                               s.stream) != hipSuccess)
                return false;
            if (launch_refit(ctx, copy, r.instances, sc.instance_count, s.stream) != RTG_OK) return false;
            tlas = reinterpret_cast<uint64_t>(copy);
        }
        rtg_image_job& j = jobs[k];
        memset(&j, 0, sizeof(j));
        j.tlas = tlas;
        j.scene = sc.bindings;
        j.uniforms = r.uniforms;
        j.width = r.width;
        j.height = r.height;
        j.opts = r.opts;
        j.out_rgba8 = r.out_rgba8;
        j.out_rgb = r.out_rgb;
        j.out_primary_hits = r.out_primary_hits;
    }
    if (trace_jobs(ctx, jobs.data(), (uint32_t)jobs.size(), s.stream, s.arena, false, &tune) != RTG_OK)
        return false;
    if (hipMemcpyAsync(s.h_status, arena_status_word(s.arena), sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream) !=
        hipSuccess)
        return false;
    return hipEventRecord(s.done, s.stream) == hipSuccess;
}

void rtg_server::run() {
    hipSetDevice(ctx->device);
    suppress_timing_on_this_thread();
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        bool progress = false;
        // retire finished batches
        for (Slot& s : slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            double t = now();
            for (auto& p : s.batch) {
                rtg_completion c{};
                c.ticket = p.ticket;
                c.tag = p.req.tag;
                c.status = completion_status(q, *s.h_status);
                c.batch_size = (uint32_t)s.batch.size();
                c.submit_s = p.submit_s;
                c.start_s = s.start_s;
                c.done_s = t;
                completed.push_back(c);
                ++finished;
            }
            inflight_bytes -= s.bytes;
            s.batch.clear();
            s.busy = false;
            progress = true;
        }
        // pending requests whose deadline has passed complete without running
        const double t_now = now();
        for (auto it = pending.begin(); it != pending.end();) {
            if (it->req.deadline_s > 0.0 && t_now > it->req.deadline_s) {
                complete_unrun(*it, RTG_ERROR_TIMEOUT, t_now);
                it = pending.erase(it);
                progress = true;
            } else {
                ++it;
            }
        }
        if (progress) cv_done.notify_all();
        // admit pending requests into free slots
        for (Slot& s : slots) {
            if (s.busy || pending.empty()) continue;
            uint64_t bytes = 0;
            uint64_t paths = 0, shadow = 0;
            size_t copies = 0;
            while (!pending.empty() && s.batch.size() < cfg.max_batch_jobs) {
                const PendingRequest& p = pending.front();
                const rtg_request& r = p.req;
                uint64_t rp = (uint64_t)r.width * r.height * r.opts.spp;
                uint64_t np = paths + rp, ns = shadow + rp * shadow_slots(r.opts.shadow_rays);
                size_t nc = copies + (r.instances ? ((p.scene.tlas_bytes + 255) & ~uint64_t(255)) : 0);
                uint64_t nb = frame_bytes(np, ns, (uint32_t)s.batch.size() + 1) + nc;
                if (!s.batch.empty() && nb > slot_capacity) break;
                if (!s.batch.empty() && cfg.max_batch_samples && np > cfg.max_batch_samples) break;
                paths = np;
                shadow = ns;
                copies = nc;
                bytes = nb;
                s.batch.push_back(p);
                pending.pop_front();
            }
            s.bytes = bytes;
            s.start_s = now();
            s.busy = true;
            inflight_bytes += bytes;
            stats.batches++;
            stats.max_batch = std::max<uint64_t>(stats.max_batch, s.batch.size());
            uint64_t busy = 0;
            for (Slot& o : slots) busy += o.busy;
            stats.max_inflight_seen = std::max<uint64_t>(stats.max_inflight_seen, busy);
            // the slot is owned by this thread: launch without holding the lock,
            // so that submissions never wait for kernel launches
            lk.unlock();
            bool ok = launch(s);
            lk.lock();
            stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, held_bytes());
            if (!ok) {
                // complete the batch with an error status
                hipStreamSynchronize(s.stream);
                double t = now();
                for (auto& p : s.batch) {
                    rtg_completion c{};
                    c.ticket = p.ticket;
                    c.tag = p.req.tag;
                    c.status = RTG_ERROR_DEVICE;
                    c.batch_size = (uint32_t)s.batch.size();
                    c.submit_s = p.submit_s;
                    c.start_s = s.start_s;
                    c.done_s = t;
                    completed.push_back(c);
                    ++finished;
                }
                inflight_bytes -= s.bytes;
                s.batch.clear();
                s.busy = false;
                cv_done.notify_all();
            }
            progress = true;
        }
        bool any_busy = false;
        for (Slot& s : slots) any_busy |= s.busy;
        if (stop && !any_busy && pending.empty()) break;
        if (!progress) {
            if (any_busy) {
                // batches on the device: poll their events at a fine grain
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

extern "C" int rtg_server_create(rtg_ctx* ctx, const rtg_server_config* cfg, rtg_server** out) {
    clear_error();
    RTG_REQUIRE(ctx && cfg && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_server_create: null argument");
    RTG_REQUIRE(cfg->max_batch_jobs >= 1 && cfg->max_inflight >= 1 && cfg->max_inflight <= 8,
                RTG_ERROR_INVALID_ARGUMENT, "rtg_server_create: max_batch_jobs >= 1, max_inflight in 1..8");
    RTG_REQUIRE(cfg->hbm_budget_bytes >= (64ull << 20), RTG_ERROR_INVALID_ARGUMENT,
                "rtg_server_create: budget below 64 MiB");
    RTG_HIP(hipSetDevice(ctx->device));
    // Every slot holds the spill area of the persistent traversal grid and a
    // feed counter; the rest of its share of the budget is for the frame
    // state and TLAS copies of its batches.
    const uint64_t share = cfg->hbm_budget_bytes / cfg->max_inflight;
    const uint64_t fixed = spill_bytes(ctx, ctx->tune) + 256;
    const uint64_t smallest = frame_bytes(1, 1, 1);
    RTG_REQUIRE(share >= fixed + smallest, RTG_ERROR_BUDGET,
                "rtg_server_create: a budget of %llu bytes over %u slots leaves %llu per slot, below the %llu bytes "
                "of a slot's traversal spill area and smallest frame state",
                (unsigned long long)cfg->hbm_budget_bytes, cfg->max_inflight, (unsigned long long)share,
                (unsigned long long)(fixed + smallest));
    rtg_server* s = new rtg_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->tune = ctx->tune;
    s->slot_fixed = fixed;
    s->slot_capacity = share - fixed;
    s->slots.resize(cfg->max_inflight);
    for (Slot& sl : s->slots) {
        sl.arena = new_frame_arena(true);
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&sl.h_status), sizeof(uint32_t)) != hipSuccess) {
            set_error("rtg_server_create: stream, event or pinned status word creation failed");
            rtg_server_destroy(s);
            return RTG_ERROR_DEVICE;
        }
        *sl.h_status = 0;
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return RTG_OK;
}

extern "C" int rtg_server_add_scene(rtg_server* srv, rtg_tlas tlas, const rtg_scene_bindings* scene,
                                    uint32_t* out_id) {
    clear_error();
    RTG_REQUIRE(srv && scene && out_id, RTG_ERROR_INVALID_ARGUMENT, "rtg_server_add_scene: null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->ctx->structures.find(tlas);
    RTG_REQUIRE(it != srv->ctx->structures.end() && it->second.kind == kKindTlas, RTG_ERROR_NOT_FOUND,
                "rtg_server_add_scene: 0x%llx is not a TLAS of this context", (unsigned long long)tlas);
    ServerScene sc;
    sc.tlas = tlas;
    sc.tlas_bytes = it->second.size;
    sc.instance_count = it->second.prim_count;
    sc.bindings = *scene;
    srv->scenes.push_back(sc);
    *out_id = (uint32_t)(srv->scenes.size() - 1);
    return RTG_OK;
}

extern "C" int rtg_server_submit(rtg_server* srv, const rtg_request* req, uint64_t* out_ticket) {
    clear_error();
    RTG_REQUIRE(srv && req, RTG_ERROR_INVALID_ARGUMENT, "rtg_server_submit: null argument");
    const double t_submit = srv->now();
    std::lock_guard<std::mutex> lk(srv->mu);
    RTG_REQUIRE(req->scene < srv->scenes.size(), RTG_ERROR_NOT_FOUND, "rtg_server_submit: unknown scene %u",
                req->scene);
    RTG_REQUIRE(req->width > 0 && req->height > 0 && req->opts.spp >= 1 && req->opts.spp <= 64 &&
                    req->opts.max_traces >= 1 && req->opts.max_traces <= 4 && req->opts.shadow_rays >= 1 &&
                    req->opts.shadow_rays <= 16,
                RTG_ERROR_INVALID_ARGUMENT, "rtg_server_submit: bad image size or options");
    uint64_t bytes = request_bytes(*req, srv->scenes[req->scene]);
    if (bytes > srv->slot_capacity) {
        srv->stats.rejected++;
        set_error("rtg_server_submit: request needs %llu bytes, one slot of the budget holds %llu",
                  (unsigned long long)bytes, (unsigned long long)srv->slot_capacity);
        return RTG_ERROR_BUDGET;
    }
    PendingRequest p;
    p.req = *req;
    p.scene = srv->scenes[req->scene];
    p.ticket = srv->next_ticket++;
    p.submit_s = t_submit;
    p.bytes = bytes;
    srv->pending.push_back(p);
    srv->submitted++;
    srv->stats.requests++;
    if (out_ticket) *out_ticket = p.ticket;
    srv->cv_work.notify_one();
    return RTG_OK;
}

extern "C" int rtg_server_cancel(rtg_server* srv, uint64_t ticket) {
    clear_error();
    RTG_REQUIRE(srv, RTG_ERROR_INVALID_ARGUMENT, "rtg_server_cancel: null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    for (auto it = srv->pending.begin(); it != srv->pending.end(); ++it) {
        if (it->ticket != ticket) continue;
        srv->complete_unrun(*it, RTG_ERROR_CANCELLED, srv->now());
        srv->pending.erase(it);
        srv->cv_done.notify_all();
        return RTG_OK;
    }
    set_error("rtg_server_cancel: ticket %llu is not pending (launched, completed or unknown)",
              (unsigned long long)ticket);
    return RTG_ERROR_NOT_FOUND;
}

extern "C" int rtg_server_poll(rtg_server* srv, rtg_completion* out, uint32_t capacity, uint32_t* count,
                               int wait_all) {
    clear_error();
    RTG_REQUIRE(srv && count && (out || capacity == 0), RTG_ERROR_INVALID_ARGUMENT, "rtg_server_poll: null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    if (wait_all) srv->cv_done.wait(lk, [srv] { return srv->finished == srv->submitted; });
    uint32_t n = (uint32_t)std::min<size_t>(capacity, srv->completed.size());
    for (uint32_t i = 0; i < n; ++i) out[i] = srv->completed[i];
    srv->completed.erase(srv->completed.begin(), srv->completed.begin() + n);
    *count = n;
    return RTG_OK;
}

extern "C" int rtg_server_replay(rtg_server* srv, const rtg_request* reqs, const double* arrival_s, uint32_t n,
                                 rtg_completion* out) {
    clear_error();
    RTG_REQUIRE(srv && (n == 0 || (reqs && arrival_s && out)), RTG_ERROR_INVALID_ARGUMENT,
                "rtg_server_replay: null argument");
    std::vector<uint64_t> tickets(n);
    double base = srv->now();
    for (uint32_t i = 0; i < n; ++i) {
        double at = base + arrival_s[i];
        while (srv->now() < at) {
            double left = at - srv->now();
            if (left > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((left - 100e-6) * 1e6)));
        }
        int st = rtg_server_submit(srv, &reqs[i], &tickets[i]);
        if (st != RTG_OK) return st;
    }
    std::vector<rtg_completion> all;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_done.wait(lk, [srv] { return srv->finished == srv->submitted; });
        all.swap(srv->completed);
    }
    for (uint32_t i = 0; i < n; ++i) {
        auto it = std::find_if(all.begin(), all.end(), [&](const rtg_completion& c) { return c.ticket == tickets[i]; });
        RTG_REQUIRE(it != all.end(), RTG_ERROR_STATE, "rtg_server_replay: completion of ticket %llu missing",
                    (unsigned long long)tickets[i]);
        out[i] = *it;
        out[i].submit_s = arrival_s[i];  // the scheduled arrival, not the submit call
        out[i].start_s -= base;
        out[i].done_s -= base;
    }
    return RTG_OK;
}

extern "C" double rtg_server_clock(const rtg_server* srv) { return srv ? srv->now() : 0.0; }

extern "C" int rtg_server_get_stats(const rtg_server* srv, rtg_server_stats* out) {
    clear_error();
    RTG_REQUIRE(srv && out, RTG_ERROR_INVALID_ARGUMENT, "rtg_server_get_stats: null argument");
    *out = srv->stats;
    return RTG_OK;
}

extern "C" int rtg_server_destroy(rtg_server* srv) {
    if (!srv) return RTG_OK;
    if (srv->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->stop = true;
        }
        srv->cv_work.notify_all();
        srv->worker.join();
    }
    for (Slot& s : srv->slots) {
        if (s.stream) hipStreamSynchronize(s.stream);
        delete_frame_arena(s.arena);
        s.tlas_copies.release();
        if (s.h_status) hipHostFree(s.h_status);
        if (s.done) hipEventDestroy(s.done);
        if (s.stream) hipStreamDestroy(s.stream);
    }
    delete srv;
    return RTG_OK;
}
