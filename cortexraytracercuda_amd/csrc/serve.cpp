// serve.cpp - continuous-batching request loop (crt_submit / crt_poll / crt_wait / crt_cancel).
//
// The reference serves one render per process (UI/Examples/UI/main.cpp:371-395 spawns the exe)
// with no queue, batching or residency. Here one scheduler thread owns the HIP stream:
//   - requests are admitted first-fit against the HBM budget (accumulation, tonemap and RNG
//     buffers; scenes are already resident and counted at upload);
//   - every pass packs path slots from all admitted requests (progressive passes of each request
//     interleave with the others), so requests arriving while others run share the device;
//   - a request finishes as soon as its last sample is resolved, independently of the others.
// Results do not depend on the mix: streams are keyed per (seed, pixel, sample, bounce) and each
// pixel's samples are summed in sample order whatever the pass boundaries are.
#include <algorithm>
#include <cstring>

#include "crt_kernels.h"
#include "crt_runtime.h"

namespace crt {

int Context::submit(crt_scene_id scene, const crt_render_desc* rd, crt_request_id* out) {
    if (rd->flags & CRT_RENDER_MEGAKERNEL)
        return set_error(CRT_ERR_INVALID_ARGUMENT, "the megakernel render is for blocking calls (crt_render)");
    int rc = Job::validate(*rd);
    if (rc) return rc;
    uint64_t need = Job::bytes_needed(*rd);
    DeviceScene* sc = nullptr;
    {
        std::lock_guard<std::mutex> g(gpu_mutex);
        sc = find_scene(scene);
        if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
        if (reserved + scene_bytes + need > budget)
            return set_error(CRT_ERR_OVER_BUDGET, "request can never fit the HBM budget");
        sc->refs++;
    }
    std::lock_guard<std::mutex> lk(srv_mutex);
    auto r = std::make_unique<ServeReq>();
    r->id = next_req_id++;
    r->scene = scene;
    r->sc = sc;
    r->rd = *rd;
    r->timing.submit_s = now_s();
    r->timing.status = CRT_ERR_NOT_READY;
    *out = r->id;
    srv_waiting.push_back(r.get());
    srv_all[r->id] = std::move(r);
    if (!srv_running) {
        srv_running = true;
        srv_stop = false;
        srv_thread = std::thread([this]() { server_loop(); });
    }
    srv_cv.notify_all();
    return CRT_OK;
}

void Context::stop_server() {
    {
        std::lock_guard<std::mutex> lk(srv_mutex);
        if (!srv_running) return;
        srv_stop = true;
        for (ServeReq* r : srv_waiting) r->cancel = true;
    }
    srv_cv.notify_all();
    if (srv_thread.joinable()) srv_thread.join();
    std::lock_guard<std::mutex> lk(srv_mutex);
    srv_running = false;
}

static void complete(ServeReq* r, int status, std::condition_variable& cv) {
    r->status = status;
    r->timing.status = status;
    r->timing.done_s = now_s();
    cv.notify_all();
}

void Context::server_loop() {
    uint32_t rr = 0;  // round-robin start among active requests
    for (;;) {
        std::vector<ServeReq*> admitted;
        {
            std::unique_lock<std::mutex> lk(srv_mutex);
            srv_cv.wait(lk, [&]() { return srv_stop || !srv_waiting.empty() || !srv_active.empty(); });
            // drop cancelled waiting requests
            for (auto it = srv_waiting.begin(); it != srv_waiting.end();) {
                if ((*it)->cancel) {
                    ServeReq* r = *it;
                    it = srv_waiting.erase(it);
                    r->sc->refs--;
                    complete(r, CRT_ERR_CANCELLED, srv_done_cv);
                } else {
                    ++it;
                }
            }
            if (srv_stop && srv_active.empty()) break;
            // first-fit admission against the budget
            for (auto it = srv_waiting.begin(); it != srv_waiting.end();) {
                ServeReq* r = *it;
                uint64_t need = Job::bytes_needed(r->rd);
                if (reserved + scene_bytes + request_bytes + need <= budget) {
                    request_bytes += need;
                    admitted.push_back(r);
                    it = srv_waiting.erase(it);
                } else {
                    ++it;
                }
            }
        }
        std::vector<WorkItem> items;
        std::vector<ServeReq*> in_pass;
        int err = CRT_OK;
        {
            std::lock_guard<std::mutex> g(gpu_mutex);
            for (ServeReq* r : admitted) {
                r->job = std::make_unique<Job>();
                int rc = r->job->setup(this, r->sc, r->rd);
                r->timing.admit_s = now_s();
                std::lock_guard<std::mutex> lk(srv_mutex);
                if (rc) {
                    request_bytes -= Job::bytes_needed(r->rd);
                    r->job.reset();
                    r->sc->refs--;
                    complete(r, rc, srv_done_cv);
                } else {
                    srv_active.push_back(r);
                }
            }
            std::vector<ServeReq*> act;
            {
                std::lock_guard<std::mutex> lk(srv_mutex);
                act = srv_active;
            }
            // pass packing: requests of the oldest pending rng mode, fair share first, then leftovers
            int mode = -1;
            for (ServeReq* r : act) if (!r->cancel && !r->job->scheduled_all()) { mode = (int)r->rd.rng_mode; break; }
            if (mode >= 0) {
                uint32_t cap = max_paths;
                size_t n = act.size();
                std::vector<ServeReq*> order;
                for (size_t k = 0; k < n; ++k) {
                    ServeReq* r = act[(rr + k) % n];
                    if (!r->cancel && (int)r->rd.rng_mode == mode && !r->job->scheduled_all()) order.push_back(r);
                }
                rr++;
                // one work item per request and pass (a pixel's samples resolve in order in one
                // thread, so two items of one request would race); each request gets a fair share
                uint32_t share = order.empty() ? cap : std::max<uint32_t>(cap / (uint32_t)order.size(), 1u);
                for (ServeReq* r : order) {
                    if (cap == 0 || items.size() >= kMaxPassRequests) break;
                    uint32_t lim = std::min(cap, std::max(share, r->job->npix));
                    WorkItem w;
                    if (!r->job->take(lim, w)) continue;
                    cap -= w.pix_count * w.sample_count;
                    items.push_back(w);
                    in_pass.push_back(r);
                }
                if (!items.empty()) {
                    double t = now_s();
                    for (ServeReq* r : in_pass) if (r->timing.first_pass_s == 0.0) r->timing.first_pass_s = t;
                    err = reset_counters();
                    if (!err) err = run_pass(items, false, nullptr);
                    if (!err && hipStreamSynchronize(stream) != hipSuccess) err = set_error(CRT_ERR_HIP, "pass failed");
                    crt_dev_counters c{};
                    if (!err) err = read_counters(c);
                    if (!err && c.stack_overflow) err = set_error(CRT_ERR_INTERNAL, "traversal stack overflow");
                    // attribute rays to requests in proportion to their paths in this pass
                    uint64_t total = 0;
                    for (auto& w : items) total += (uint64_t)w.pix_count * w.sample_count;
                    for (auto& w : items) {
                        for (ServeReq* r : in_pass) {
                            if (r->job.get() != w.job) continue;
                            double f = total ? (double)((uint64_t)w.pix_count * w.sample_count) / (double)total : 0.0;
                            r->timing.rays += (uint64_t)(f * (double)c.rays + 0.5);
                        }
                    }
                    for (ServeReq* r : in_pass) r->timing.passes++;
                }
            }
            // completion: everything scheduled and resolved (the pass above was synchronised)
            std::vector<ServeReq*> finished;
            for (ServeReq* r : act) {
                if (err || r->cancel || r->job->scheduled_all()) finished.push_back(r);
            }
            for (ServeReq* r : finished) {
                int st = err ? err : (r->cancel ? CRT_ERR_CANCELLED : CRT_OK);
                if (st == CRT_OK) {
                    r->rgb.resize((size_t)r->job->npix * 3);
                    r->accum.resize((size_t)r->job->npix * 4);
                    crt_image im{r->rgb.data(), r->accum.data()};
                    st = finish_job(*r->job, &im);
                    if (!st && hipStreamSynchronize(stream) != hipSuccess) st = set_error(CRT_ERR_HIP, "readback failed");
                }
                std::lock_guard<std::mutex> lk(srv_mutex);
                request_bytes -= Job::bytes_needed(r->rd);
                r->job.reset();
                srv_active.erase(std::find(srv_active.begin(), srv_active.end(), r));
                r->sc->refs--;
                complete(r, st, srv_done_cv);
            }
        }
    }
}

}  // namespace crt

using namespace crt;

extern "C" {

int crt_submit(crt_context* c, crt_scene_id scene, const crt_render_desc* rd, crt_request_id* out) {
    if (!c || !rd || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    return reinterpret_cast<Context*>(c)->submit(scene, rd, out);
}

int crt_poll(crt_context* c, crt_request_id req, int32_t* done) {
    if (!c || !done) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::lock_guard<std::mutex> lk(ctx->srv_mutex);
    auto it = ctx->srv_all.find(req);
    if (it == ctx->srv_all.end()) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown request id");
    *done = it->second->status != CRT_ERR_NOT_READY ? 1 : 0;
    return CRT_OK;
}

int crt_wait(crt_context* c, crt_request_id req, crt_image* out, crt_stats* stats) {
    if (!c) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::unique_lock<std::mutex> lk(ctx->srv_mutex);
    auto it = ctx->srv_all.find(req);
    if (it == ctx->srv_all.end()) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown request id");
    ServeReq* r = it->second.get();
    ctx->srv_done_cv.wait(lk, [&]() { return r->status != CRT_ERR_NOT_READY; });
    if (r->status != CRT_OK) {
        int st = r->status;
        return set_error(st, st == CRT_ERR_CANCELLED ? "request cancelled" : "request failed");
    }
    if (out && out->rgb8) std::memcpy(out->rgb8, r->rgb.data(), r->rgb.size());
    if (out && out->accum) std::memcpy(out->accum, r->accum.data(), r->accum.size() * sizeof(float));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rays = r->timing.rays;
        stats->host_ms = (r->timing.done_s - r->timing.submit_s) * 1e3;
    }
    return CRT_OK;
}

int crt_cancel(crt_context* c, crt_request_id req) {
    if (!c) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::lock_guard<std::mutex> lk(ctx->srv_mutex);
    auto it = ctx->srv_all.find(req);
    if (it == ctx->srv_all.end()) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown request id");
    it->second->cancel = true;
    ctx->srv_cv.notify_all();
    return CRT_OK;
}

int crt_request_get_timing(crt_context* c, crt_request_id req, crt_request_timing* out) {
    if (!c || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::lock_guard<std::mutex> lk(ctx->srv_mutex);
    auto it = ctx->srv_all.find(req);
    if (it == ctx->srv_all.end()) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown request id");
    *out = it->second->timing;
    return CRT_OK;
}

}  // extern "C"
