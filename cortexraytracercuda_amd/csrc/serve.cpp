// serve.cpp - continuous-batching request loop (crt_submit / crt_poll / crt_wait / crt_cancel).
//
// The reference serves one render per process (UI/Examples/UI/main.cpp:371-395 spawns the exe)
// with no queue, batching or residency. Here one scheduler thread owns streams and path pools of
// its own and keeps one streaming pass open per lane (two lanes, each with half the slots):
//   - requests are admitted first-fit against the HBM budget (accumulation, tonemap and RNG
//     buffers; scenes are already resident and counted at upload) and join the lane with the
//     fewest outstanding paths;
//   - an admitted request is cut into segments (pixel range x sample range); a segment becomes a
//     request-table entry of its lane's open pass and its path ids are appended behind the pass's
//     last id between two iterations, so the next regeneration stage already starts its camera
//     rays among the paths of every other request in flight;
//   - path ids are generated in order and every live path is shaded once per iteration: once the
//     generated count and the smallest surviving id (left by the shade stage) have passed a
//     segment's last id, the host enqueues its resolve (segments of one request in order), and
//     after the last one the tonemap; a completion thread reads the image back, so a request
//     completes independently of the others;
//   - when a pass has no room left (path ids or table entries) it drains and starts over.
// Results do not depend on the mix: streams are keyed per (seed, pixel, sample, bounce) and each
// pixel's samples are summed in sample order whatever the segment boundaries are.
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
        if ((rc = server_init())) return rc;
        sc = find_scene(scene);
        if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
        // the serving iterations launch the traversal kernel of the context's BVH width (their
        // captured graphs bake it in): a scene built at another width cannot join the open passes
        if (sc->host.bvh_width != bvh_width)
            return set_error(CRT_ERR_INVALID_ARGUMENT, "scene BVH width differs from the serving loop's");
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
        for (ServeReq* r : srv_active) r->cancel = true;
    }
    srv_cv.notify_all();
    if (srv_thread.joinable()) srv_thread.join();
    std::lock_guard<std::mutex> lk(srv_mutex);
    srv_running = false;
}

// Server buffers, allocated at the first submit (caller holds gpu_mutex). Counted as reserved.
int Context::server_init() {
    if (srv[0].ready) return CRT_OK;
    srv_lanes = std::max(1u, std::min<uint32_t>(sub_passes, kSrvLanes));
    // path ids of the open passes: a pass that runs out of ids drains (a tail of part-empty
    // launches) and starts over, so the serving loop takes a larger id space than a blocking pass
    const uint32_t slots_all = max_paths;
    const uint64_t cap_all = config.pass_paths ? std::max<uint32_t>(pass_capacity, slots_all)
                                               : std::max<uint64_t>(std::min<uint64_t>(1ull << 29, 32ull * slots_all), slots_all);
    for (uint32_t l = 0; l < srv_lanes; ++l) {
        int rc = server_init_lane(srv[l], std::max(1u, slots_all / srv_lanes),
                                  (uint32_t)std::max<uint64_t>(cap_all / srv_lanes, slots_all / srv_lanes), l == 0);
        if (rc) return rc;
    }
    reserved = used_bytes() + server_bytes();
    return CRT_OK;
}

int Context::server_init_lane(Server& S, uint32_t slots, uint32_t cap, bool first) {
    int rc;
    S.slots = std::max(1u, slots);
    S.cap = std::max(cap, S.slots);
    HIP_TRY(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
    if (first) HIP_TRY(hipStreamCreateWithFlags(&S.copy_s, hipStreamNonBlocking));
    const size_t n = S.slots;
    for (int k = 0; k < 2; ++k) {
        if ((rc = S.ray_o[k].alloc(n * 16, this))) return rc;
        if ((rc = S.ray_d[k].alloc(n * 16, this))) return rc;
        if ((rc = S.thr[k].alloc(n * 16, this))) return rc;
        if ((rc = S.rad[k].alloc(n * 16, this))) return rc;
        if ((rc = S.ids[k].alloc(n * 8, this))) return rc;
    }
    if ((rc = S.hit.alloc(n * 16, this))) return rc;
    if ((rc = S.hit_inst.alloc(n * 4, this))) return rc;
    if ((rc = S.out_rad.alloc((size_t)S.cap * 16, this))) return rc;
    if ((rc = S.req.alloc(sizeof(crt_dev_request) * Server::kEntries, this))) return rc;
    if ((rc = S.ctl.alloc(sizeof(uint32_t) * (2 * CRT_CTL_WORDS + 2), this))) return rc;
    S.use_graphs = config.graphs != 2u;
    S.egrid = std::min(extend_grid, std::max(1u, (S.slots + 255) / 256));
    S.sgrid = std::min(std::min(shade_grid, (uint32_t)Server::kMaxShadeGrid), std::max(1u, (S.slots + 255) / 256));
    if ((rc = S.live_min.alloc(sizeof(uint32_t) * S.sgrid, this))) return rc;
    if ((rc = S.gstack.alloc((size_t)S.egrid * 256u * crt_k_gstack_entries() * sizeof(uint32_t), this))) return rc;
    if ((rc = S.counters.alloc(sizeof(crt_dev_counters), this))) return rc;
    HIP_TRY(hipHostMalloc((void**)&S.req_host, sizeof(crt_dev_request) * Server::kEntries, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&S.open_host, sizeof(uint32_t) * 2 * Server::kOpenRing, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&S.snap, sizeof(uint32_t) * Server::kSnapWords * Server::kRing, hipHostMallocDefault));
    if (first) HIP_TRY(hipHostMalloc(&S.rays_host, sizeof(unsigned long long), hipHostMallocDefault));
    for (int k = 0; k < Server::kRing; ++k) HIP_TRY(hipEventCreateWithFlags(&S.ev[k], hipEventDisableTiming));
    if (crt_k_fill_u32(S.s, (uint32_t*)S.counters.p, sizeof(crt_dev_counters) / 4, 0u))
        return set_error(CRT_ERR_HIP, "fill launch failed");
    S.ready = true;
    server_reset_pass(S);
    return CRT_OK;
}

uint64_t Context::server_bytes() const {
    uint64_t b = 0;
    for (const Server& S : srv) b += server_lane_bytes(S);
    return b;
}

uint64_t Context::server_lane_bytes(const Server& S) const {
    uint64_t b = S.hit.bytes + S.hit_inst.bytes + S.out_rad.bytes + S.req.bytes + S.ctl.bytes + S.gstack.bytes +
                 S.counters.bytes + S.live_min.bytes;
    for (int k = 0; k < 2; ++k) b += S.ray_o[k].bytes + S.ray_d[k].bytes + S.thr[k].bytes + S.rad[k].bytes + S.ids[k].bytes;
    return b;
}

void Context::server_release() {
    for (Server& S : srv) server_release_lane(S);
}

void Context::server_release_lane(Server& S) {
    if (S.s) (void)hipStreamSynchronize(S.s);
    if (S.copy_s) (void)hipStreamSynchronize(S.copy_s);
    if (S.req_host) (void)hipHostFree(S.req_host);
    if (S.open_host) (void)hipHostFree(S.open_host);
    if (S.snap) (void)hipHostFree(S.snap);
    if (S.rays_host) (void)hipHostFree(S.rays_host);
    for (int k = 0; k < Server::kRing; ++k) if (S.ev[k]) (void)hipEventDestroy(S.ev[k]);
    for (int t = 0; t < Server::kVariants; ++t)
        for (int k = 0; k < Server::kRing; ++k)
            if (S.graph[t][k]) { (void)hipGraphExecDestroy(S.graph[t][k]); S.graph[t][k] = nullptr; }
    if (S.s) (void)hipStreamDestroy(S.s);
    if (S.copy_s) (void)hipStreamDestroy(S.copy_s);
    S.req_host = nullptr; S.open_host = nullptr; S.snap = nullptr; S.rays_host = nullptr; S.s = nullptr; S.copy_s = nullptr;
    S.ready = false;
}

// A new open pass: no entries, no path ids; control block zeroed on the stream.
void Context::server_reset_pass(Server& S) {
    S.entries.clear();
    S.total = 0;
    S.iter = 0;
    S.processed = 0;
    S.done_paths = 0;
    S.gen_seen = 0;
    S.gen_progress = 0;
    S.progress_iter = 0;
    S.full = false;
    S.two_level = false;
    S.shapes = false;
    S.mixed_rng = false;
    (void)crt_k_pass_init(S.s, (uint32_t*)S.ctl.p, 2 * CRT_CTL_WORDS + 2);
}

// True when no request is waiting or on the device and every lane has drained (caller holds
// both context locks, so no request can be submitted meanwhile).
bool Context::server_idle() {
    std::lock_guard<std::mutex> lk(srv_mutex);
    if (!srv_waiting.empty() || !srv_active.empty()) return false;
    for (uint32_t l = 0; l < (uint32_t)kSrvLanes; ++l) {
        const Server& S = srv[l];
        if (S.ready && ((uint64_t)S.total > S.done_paths || S.processed < S.iter)) return false;
    }
    return true;
}

// Drops every captured serving iteration: they bake in the traversal variant (BVH width, stack
// depth, work distribution), which the autotuner may change. Only called on an idle serving loop
// (server_idle) with both context locks held: the scheduler thread then waits for work and touches no graph.
void Context::server_drop_graphs() {
    for (Server& S : srv) {
        if (!S.ready) continue;
        (void)hipStreamSynchronize(S.s);
        for (int t = 0; t < Server::kVariants; ++t)
            for (int k = 0; k < Server::kRing; ++k)
                if (S.graph[t][k]) { (void)hipGraphExecDestroy(S.graph[t][k]); S.graph[t][k] = nullptr; }
        for (int t = 0; t < Server::kVariants; ++t) S.warm[t] = 0;
    }
}

static void complete(ServeReq* r, int status, std::condition_variable& cv) {
    r->status = status;
    r->timing.status = status;
    r->timing.done_s = now_s();
    cv.notify_all();
}

// One segment (at most seg_cap path ids, within the pass's id room) of request r becomes a
// request-table entry behind the pass's last path id. False when r has nothing left to take.
bool Context::append_segment(Server& S, ServeReq* r, uint32_t seg_cap) {
    const uint32_t room = S.cap - S.total;
    WorkItem w;
    if (!r->job->take(std::min(seg_cap, room), w)) return false;
    crt_dev_request q = r->job->req;
    q.path_base = S.total;
    q.pix_begin = w.pix_begin; q.pix_count = w.pix_count;
    q.sample_begin = w.sample_begin; q.sample_count = w.sample_count;
    const uint32_t e = (uint32_t)S.entries.size();
    S.req_host[e] = q;
    SrvEntry en;
    en.r = r;
    en.seg = r->segs++;
    en.paths = w.pix_count * w.sample_count;
    en.end = S.total + en.paths;
    S.entries.push_back(en);
    S.total += en.paths;
    if (r->sc->host.num_instances) S.two_level = true;
    if (r->sc->host.num_shapes) S.shapes = true;
    if (r->rd.rng_mode == CRT_RNG_MINSTD) S.mixed_rng = true;
    r->timing.passes++;
    if (r->timing.first_pass_s == 0.0) r->timing.first_pass_s = now_s();
    return true;
}

// Appends segments of active requests to the open pass while the backlog of ungenerated path ids
// is below two slot pools (a new request starts within a few iterations). MINSTD requests keep at
// most one segment in flight (each pixel's engine advances sample after sample).
// Order (CRT_SERVE_SMALL_THEN_OLDEST, the default): requests with at most serve_small path ids left,
// in admission order, then the others oldest first, each appending segments while the backlog
// allows. Draining the large requests in arrival order bounds the largest latency of a loop that
// is offered more work than it renders (the last request to finish is then one that arrived late);
// small requests pass them, finish within a few iterations and keep the median low.
// CRT_SERVE_ROUND_ROBIN: one segment per request and call in rotating order (every request
// progresses; the large ones finish last whenever they arrived).
int Context::server_append(Server& S, std::vector<ServeReq*>& act) {
    if (act.empty() || S.full) return CRT_OK;
    const uint32_t seg_cap = std::max<uint32_t>(S.slots / 2, 1u);
    const uint32_t n0 = (uint32_t)S.entries.size();
    const size_t na = act.size();
    std::vector<ServeReq*> order;
    order.reserve(na);
    const bool rr = serve_order == CRT_SERVE_ROUND_ROBIN;
    if (rr) {
        for (size_t t = 0; t < na; ++t) order.push_back(act[(S.rr + t) % na]);
        S.rr++;
    } else {
        for (ServeReq* r : act) if (r->job->paths_left() <= serve_small) order.push_back(r);
        for (ServeReq* r : act) if (r->job->paths_left() > serve_small) order.push_back(r);
    }
    bool stop = false;
    for (ServeReq* r : order) {
        if (stop) break;
        for (;;) {
            if (r->cancel || r->finishing || r->job->scheduled_all()) break;
            if (r->rd.rng_mode == CRT_RNG_MINSTD && r->segs_resolved < r->segs) break;
            if ((uint64_t)S.total - S.gen_seen >= 2ull * S.slots) { stop = true; break; }
            if (S.entries.size() >= Server::kEntries || S.total >= S.cap) { S.full = true; stop = true; break; }
            if (!append_segment(S, r, seg_cap)) break;
            if (rr) break;
        }
    }
    const uint32_t n1 = (uint32_t)S.entries.size();
    if (n1 == n0) return CRT_OK;
    HIP_TRY(hipMemcpyAsync((crt_dev_request*)S.req.p + n0, S.req_host + n0, sizeof(crt_dev_request) * (n1 - n0),
                           hipMemcpyHostToDevice, S.s));
    uint32_t* w2 = S.open_host + 2 * (S.appends++ % Server::kOpenRing);
    w2[0] = S.total;
    w2[1] = n1;
    HIP_TRY(hipMemcpyAsync((uint32_t*)S.ctl.p + 2 * CRT_CTL_WORDS, w2, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, S.s));
    return CRT_OK;
}

// One regenerate / closest-hit / shade iteration of the open pass, then a snapshot of the control
// block, the work counters and the shade workgroups' smallest live path ids into the pinned ring. The
// arguments depend only on the snapshot slot (which fixes the buffer parity) and on the traversal
// variant, so after a few eager iterations each (slot, variant) is captured once into a hipGraph and
// replayed: one graph launch per iteration instead of three kernel launches and three copies.
static int enqueue_iteration(Context& C, Server& S, uint64_t k) {
    uint32_t* cb = (uint32_t*)S.ctl.p;
    uint32_t* cin = cb + (k & 1) * CRT_CTL_WORDS;
    uint32_t* cout = cb + ((k + 1) & 1) * CRT_CTL_WORDS;
    const uint32_t* open = cb + 2 * CRT_CTL_WORDS;
    crt_paths P[2];
    for (int b = 0; b < 2; ++b) {
        P[b].ray_o = (float4*)S.ray_o[b].p; P[b].ray_d = (float4*)S.ray_d[b].p; P[b].thr = (float4*)S.thr[b].p;
        P[b].rad = (float4*)S.rad[b].p; P[b].ids = (uint2*)S.ids[b].p;
    }
    const crt_paths& A = P[k & 1];
    const crt_paths& B = P[(k + 1) & 1];
    const crt_dev_request* dreq = (const crt_dev_request*)S.req.p;
    const crt_dev_scene* dsc = (const crt_dev_scene*)C.scene_table.p;
    // traversal: one level, or two levels with or without the shape test; shade: the full kernel
    // once an instanced scene or one with shapes is in the pass
    const int two_level = S.two_level ? (S.shapes ? 2 : 1) : 0;
    const int full_shade = (S.two_level || S.shapes) ? 1 : 0;
    // the per-path RNG mode only once a MINSTD request is in the pass (keyed PCG otherwise)
    const int mode = S.mixed_rng ? CRT_RNG_DYNAMIC : CRT_RNG_PCG;
    if (crt_k_regen(S.s, mode, S.sgrid, dreq, 0, 0, 0, S.slots, cin, cout, open, A))
        return set_error(CRT_ERR_HIP, "regen launch failed");
    if (crt_k_extend(S.s, 0, (int)C.bvh_width, two_level, S.egrid, dsc, dreq, -1, nullptr, cin + CRT_CTL_LIVE, 0, A.ray_o,
                     A.ray_d, A.ids, (float4*)S.hit.p, (uint32_t*)S.hit_inst.p, (uint32_t*)S.gstack.p, nullptr,
                     C.extend_mode, cin + CRT_CTL_WORK, C.extend_param, (int)C.stack_lds, cin + CRT_CTL_SURV,
                     (crt_dev_counters*)S.counters.p))
        return set_error(CRT_ERR_HIP, "extend launch failed");
    if (crt_k_shade(S.s, mode, full_shade, S.sgrid, dsc, dreq, cin + CRT_CTL_LIVE, A, B, cout + CRT_CTL_SURV,
                    (const float4*)S.hit.p, (const uint32_t*)S.hit_inst.p, (float4*)S.out_rad.p, (uint32_t*)S.live_min.p,
                    (int)C.ray_sort, 0, (crt_dev_counters*)S.counters.p))
        return set_error(CRT_ERR_HIP, "shade launch failed");
    uint32_t* snap = S.snap + (size_t)(k % Server::kRing) * Server::kSnapWords;
    HIP_TRY(hipMemcpyAsync(snap, cout, sizeof(uint32_t) * CRT_CTL_WORDS, hipMemcpyDeviceToHost, S.s));
    HIP_TRY(hipMemcpyAsync(snap + CRT_CTL_WORDS, S.counters.p, sizeof(crt_dev_counters), hipMemcpyDeviceToHost, S.s));
    HIP_TRY(hipMemcpyAsync(snap + CRT_CTL_WORDS + Server::kCounterWords, S.live_min.p, sizeof(uint32_t) * S.sgrid, hipMemcpyDeviceToHost, S.s));
    return CRT_OK;
}

int Context::server_iteration(Server& S) {
    const uint64_t k = S.iter;
    const int slot = (int)(k % Server::kRing);
    const int tl = (S.two_level ? 4 : 0) + (S.shapes ? 2 : 0) + (S.mixed_rng ? 1 : 0);
    int rc;
    if (S.use_graphs && S.warm[tl] >= (uint32_t)Server::kRing) {
        hipGraphExec_t& g = S.graph[tl][slot];
        if (!g) {
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(S.s, hipStreamCaptureModeThreadLocal));
            rc = enqueue_iteration(*this, S, k);
            hipError_t e = hipStreamEndCapture(S.s, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            HIP_TRY(e);
            e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) { g = nullptr; HIP_TRY(e); }
        }
        HIP_TRY(hipGraphLaunch(g, S.s));
    } else {
        if ((rc = enqueue_iteration(*this, S, k))) return rc;
        S.warm[tl]++;
    }
    HIP_TRY(hipEventRecord(S.ev[slot], S.s));
    S.iter++;
    return CRT_OK;
}

// Reads the oldest outstanding snapshot (waits for its iteration) and enqueues the resolves of
// segments whose paths have all finished, in segment order per request; a request whose last
// segment is resolved gets its tonemap and a completion event.
int Context::server_process_snapshot(Server& S, std::vector<ServeReq*>& act) {
    const uint64_t j = S.processed;
    HIP_TRY(hipEventSynchronize(S.ev[j % Server::kRing]));
    const uint32_t* snap = S.snap + (size_t)(j % Server::kRing) * Server::kSnapWords;
    S.processed++;
    S.gen_seen = std::max(S.gen_seen, snap[CRT_CTL_GEN]);
    crt_dev_counters c;
    std::memcpy(&c, snap + CRT_CTL_WORDS, sizeof(c));
    if (c.stack_overflow) return set_error(CRT_ERR_INTERNAL, "traversal stack overflow");
    // Completion without per-path counters: path ids are generated in order, so after iteration j
    // every id below GEN (generated so far) that is not held by a surviving path has finished and
    // written its radiance. The shade stage leaves the smallest surviving id per workgroup; an
    // entry is done once GEN and that minimum have both passed its last id.
    const uint32_t gen = snap[CRT_CTL_GEN];
    const uint32_t* lm = snap + CRT_CTL_WORDS + Server::kCounterWords;
    uint32_t live = 0xffffffffu;
    for (uint32_t b = 0; b < S.sgrid; ++b) live = std::min(live, lm[b]);
    const uint32_t finished_below = std::min(gen, live);
    const size_t ne = S.entries.size();
    uint64_t done = 0;
    for (size_t e = 0; e < ne; ++e) {
        SrvEntry& en = S.entries[e];
        if (!en.done && en.end <= finished_below) en.done = true;
        if (en.done) done += en.paths;
    }
    if (gen > S.gen_progress) { S.gen_progress = gen; S.progress_iter = S.iter; }
    S.done_paths = done;
    // appended paths are generated within a few iterations (the backlog is bounded): a pass whose
    // generation stops while ids are outstanding is an error
    if (gen < S.total && S.iter - S.progress_iter > 4096)
        return set_error(CRT_ERR_INTERNAL, "serving pass did not drain");
    // resolve finished segments in order per request
    bool progress = true;
    while (progress) {
        progress = false;
        for (size_t e = 0; e < ne; ++e) {
            SrvEntry& en = S.entries[e];
            if (en.resolved || !en.done || en.r->segs_resolved != en.seg) continue;
            ServeReq* r = en.r;
            if (!r->cancel) {
                if (crt_k_resolve(S.s, (const crt_dev_request*)S.req.p + e, 1, S.req_host[e].pix_count,
                                  (const float4*)S.out_rad.p, r->job->rays_counter()))
                    return set_error(CRT_ERR_HIP, "resolve launch failed");
            }
            en.resolved = true;
            r->segs_resolved++;
            progress = true;
        }
    }
    for (ServeReq* r : act) {
        if (r->finishing || (!r->job->scheduled_all() && !r->cancel)) continue;
        if (r->segs_resolved != r->segs) continue;
        // every segment resolved (or cancelled with nothing in flight): finish on the stream
        if (!r->cancel) {
            if (crt_k_tonemap(S.s, (const float4*)r->job->accum.p, r->job->npix, r->job->rd.spp, (uint8_t*)r->job->rgb.p))
                return set_error(CRT_ERR_HIP, "tonemap launch failed");
        }
        HIP_TRY(hipEventCreateWithFlags(&r->ev_done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(r->ev_done, S.s));
        r->finishing = true;
    }
    return CRT_OK;
}

// Readback and completion of one finished request (completion thread; status != CRT_OK: no copies).
void Context::server_complete(ServeReq* r, int st) {
    Server& S = srv[0];   // owns the copy stream and the pinned ray-count word
    if (st == CRT_OK) {
        // the accumulation buffer (16 B per pixel against 3 for the image) only when asked for
        const bool want_accum = !(r->rd.flags & CRT_RENDER_IMAGE_ONLY);
        r->rgb.resize((size_t)r->job->npix * 3);
        r->accum.resize(want_accum ? (size_t)r->job->npix * 4 : 0);
        unsigned long long* rays_h = (unsigned long long*)S.rays_host;
        if (hipMemcpyAsync(rays_h, r->job->rays_counter(), sizeof(*rays_h), hipMemcpyDeviceToHost, S.copy_s) != hipSuccess ||
            hipMemcpyAsync(r->rgb.data(), r->job->rgb.p, r->rgb.size(), hipMemcpyDeviceToHost, S.copy_s) != hipSuccess ||
            (want_accum && hipMemcpyAsync(r->accum.data(), r->job->accum.p, r->accum.size() * sizeof(float),
                                          hipMemcpyDeviceToHost, S.copy_s) != hipSuccess) ||
            hipStreamSynchronize(S.copy_s) != hipSuccess)
            st = set_error(CRT_ERR_HIP, "readback failed");
        else
            r->timing.rays = *rays_h;
    }
    if (r->ev_done) { (void)hipEventDestroy(r->ev_done); r->ev_done = nullptr; }
    // the job's buffers go back to the cache (own lock); its stream work has completed
    r->job.reset();
    std::lock_guard<std::mutex> lk(srv_mutex);
    request_bytes -= Job::bytes_needed(r->rd);
    r->sc->refs--;
    complete(r, st, srv_done_cv);
    srv_cv.notify_all();   // the budget it held may admit a waiting request
}

void Context::server_copy_loop() {
    for (;;) {
        std::pair<ServeReq*, int> job;
        {
            std::unique_lock<std::mutex> lk(srv_mutex);
            srv_copy_cv.wait(lk, [&]() { return srv_copy_stop || !srv_copy_q.empty(); });
            if (srv_copy_q.empty()) return;   // stop requested and nothing left
            job = srv_copy_q.front();
            srv_copy_q.pop_front();
        }
        server_complete(job.first, job.second);
    }
}

void Context::server_loop() {
    const uint32_t L = srv_lanes;
    int err = CRT_OK;
    {
        std::lock_guard<std::mutex> lk(srv_mutex);
        srv_copy_stop = false;
    }
    std::thread copier([this]() { server_copy_loop(); });
    auto lanes_inflight = [&]() {
        for (uint32_t l = 0; l < L; ++l) if (srv[l].processed < srv[l].iter) return true;
        return false;
    };
    auto lanes_work = [&]() {
        for (uint32_t l = 0; l < L; ++l) if ((uint64_t)srv[l].total > srv[l].done_paths) return true;
        return false;
    };
    for (;;) {
        std::vector<ServeReq*> admitted, act;
        const bool inflight = lanes_inflight();
        {
            std::unique_lock<std::mutex> lk(srv_mutex);
            const bool work = lanes_work();
            srv_cv.wait(lk, [&]() { return srv_stop || !srv_waiting.empty() || !srv_active.empty() || inflight || work; });
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
            if (srv_stop && srv_active.empty() && !inflight) break;
            // first-fit admission against the budget; each request joins the lane with the fewest
            // outstanding paths
            for (auto it = srv_waiting.begin(); it != srv_waiting.end();) {
                ServeReq* r = *it;
                uint64_t need = Job::bytes_needed(r->rd);
                if (reserved + scene_bytes + request_bytes + need <= budget) {
                    request_bytes += need;
                    uint32_t best = 0;
                    for (uint32_t l = 1; l < L; ++l) if (lane_load[l] < lane_load[best]) best = l;
                    // load = the paths the request adds: its clamped sample range, not all spp
                    const uint64_t sq = (uint64_t)std::sqrt((double)r->rd.spp), ns = sq * sq;
                    const uint64_t sb = std::min<uint64_t>(r->rd.sample_begin, ns);
                    const uint64_t se = r->rd.sample_count ? std::min<uint64_t>(ns, sb + r->rd.sample_count) : ns;
                    r->lane = (int)best;
                    r->load = (uint64_t)r->rd.width * r->rd.height * (se - sb);
                    lane_load[best] += r->load;
                    admitted.push_back(r);
                    it = srv_waiting.erase(it);
                } else {
                    ++it;
                }
            }
        }
        for (ServeReq* r : admitted) {
            // no gpu_mutex: the setup touches the buffer cache (own lock), the request's scene
            // (pinned by its reference) and the lane's own stream, so a blocking render holding
            // exec_mutex does not stall admission
            r->job = std::make_unique<Job>();
            const int rc = r->job->setup(this, r->sc, r->rd, srv[r->lane].s);
            r->timing.admit_s = now_s();
            std::lock_guard<std::mutex> lk(srv_mutex);
            if (rc) {
                request_bytes -= Job::bytes_needed(r->rd);
                lane_load[r->lane] -= r->load;
                r->job.reset();
                r->sc->refs--;
                complete(r, rc, srv_done_cv);
            } else {
                srv_active.push_back(r);
            }
        }
        {
            std::lock_guard<std::mutex> lk(srv_mutex);
            act = srv_active;
        }
        bool busy = !admitted.empty(), any_step = false;
        std::vector<ServeReq*> lane_act[kSrvLanes];
        for (uint32_t l = 0; l < L && !err; ++l) {
            Server& S = srv[l];
            std::vector<ServeReq*>& act_l = lane_act[l];
            for (ServeReq* r : act) if (r->lane == (int)l) act_l.push_back(r);
            const uint64_t processed0 = S.processed;
            err = server_append(S, act_l);
            // keep up to kLag iterations ahead while paths are outstanding
            const bool outstanding = (uint64_t)S.total > S.done_paths;
            bool enqueued = false;
            if (!err && outstanding && S.iter - S.processed < (uint64_t)Server::kLag) {
                err = server_iteration(S);
                enqueued = true;
            }
            // read a snapshot when this lane is kLag ahead, or when it has nothing to enqueue and no
            // other lane does either (so one lane's wait does not hold back the other's launches)
            if (!err && S.processed < S.iter && (S.iter - S.processed >= (uint64_t)Server::kLag || (!enqueued && L == 1)))
                err = server_process_snapshot(S, act_l);
            if (!err && S.processed < S.iter && !enqueued && L > 1) {
                // poll: process the oldest snapshot only if its iteration has finished
                if (hipEventQuery(S.ev[S.processed % Server::kRing]) == hipSuccess) err = server_process_snapshot(S, act_l);
            }
            // a full pass that has drained and resolved everything starts over
            if (!err && S.processed == S.iter && (uint64_t)S.total == S.done_paths) {
                bool all_resolved = true;
                for (auto& en : S.entries) all_resolved = all_resolved && en.resolved;
                if (all_resolved && (S.full || S.total > S.cap / 2 || S.entries.size() > Server::kEntries / 2))
                    server_reset_pass(S);
            }
            busy = busy || enqueued || S.processed < S.iter;
            any_step = any_step || enqueued || S.processed != processed0;
        }
        if (!err && L > 1 && !any_step) {
            // nothing to launch and no finished iteration: wait for the oldest one in flight
            int oldest = -1;
            for (uint32_t l = 0; l < L; ++l)
                if (srv[l].processed < srv[l].iter && (oldest < 0 || srv[l].processed < srv[oldest].processed)) oldest = (int)l;
            if (oldest >= 0) err = server_process_snapshot(srv[oldest], lane_act[oldest]);
        }
        // completions: readback of finished requests whose stream work is done; with nothing else
        // to do this turn, wait for the oldest finishing request instead of polling
        std::vector<ServeReq*> finished;
        for (ServeReq* r : act) {
            if (err) { finished.push_back(r); continue; }
            if (!r->finishing) continue;
            hipError_t q = busy || !finished.empty() ? hipEventQuery(r->ev_done) : hipEventSynchronize(r->ev_done);
            if (q == hipErrorNotReady) continue;
            finished.push_back(r);
        }
        if (err && !finished.empty())
            for (uint32_t l = 0; l < L; ++l) (void)hipStreamSynchronize(srv[l].s);
        if (!finished.empty()) {
            // hand the finished requests to the completion thread (readback, release, completion)
            std::lock_guard<std::mutex> lk(srv_mutex);
            for (ServeReq* r : finished) {
                const int st = err ? err : (r->cancel ? CRT_ERR_CANCELLED : CRT_OK);
                srv_active.erase(std::find(srv_active.begin(), srv_active.end(), r));
                lane_load[r->lane] -= r->load;
                srv_copy_q.emplace_back(r, st);
            }
            srv_copy_cv.notify_all();
        }
        if (err) {
            // a failed iteration leaves the open passes unusable: wait for the streams, start over
            for (uint32_t l = 0; l < L; ++l) {
                (void)hipStreamSynchronize(srv[l].s);
                srv[l].iter = srv[l].processed = 0;
                server_reset_pass(srv[l]);
            }
            err = CRT_OK;
        }
    }
    for (uint32_t l = 0; l < L; ++l) (void)hipStreamSynchronize(srv[l].s);
    {
        std::lock_guard<std::mutex> lk(srv_mutex);
        srv_copy_stop = true;
    }
    srv_copy_cv.notify_all();
    copier.join();   // drains the queue first
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
    if (out && out->accum && (r->rd.flags & CRT_RENDER_IMAGE_ONLY))
        return set_error(CRT_ERR_INVALID_ARGUMENT, "request was submitted with CRT_RENDER_IMAGE_ONLY: no accumulation buffer");
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
