// crt_runtime.h - host runtime objects behind the C-ABI: device buffers, resident scenes, render
// jobs, the pass scheduler and the serving loop.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cortex_rt.h"
#include "crt_kernels.h"
#include "crt_layout.h"
#include "scene_build.h"

namespace crt {

int set_error(int code, const std::string& msg);

// Returns from the enclosing function with a crt status when a HIP call fails.
#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? CRT_ERR_OUT_OF_MEMORY : CRT_ERR_HIP,              \
                             std::string(#expr) + ": " + hipGetErrorString(e_));                           \
    } while (0)
double now_s();
void camera_setup(const crt_camera& c, uint32_t width, crt_dev_request& r);

struct Context;

// BVH shape used when the caller passes 0 (measured fastest for the traversal kernels with leaf
// references inline in the parent node: one primitive per leaf, four children per node)
constexpr uint32_t kDefaultLeafSize = 1, kDefaultBvhWidth = 4;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    Context* home = nullptr;   // set: the buffer came from (and returns to) the context's cache
    int alloc(size_t n, Context* ctx);
    // per-request buffers: reuse a cached allocation of the same size class (no hipMalloc / hipFree
    // per render; the previous user's work has completed before its buffers are released)
    int alloc_cached(size_t n, Context* ctx);
    void release();
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
};

struct DeviceScene {
    HostScene host;  // geometry arrays are dropped after upload; stats and small tables stay
    DevBuf geo;   // BVH nodes, then triangle, instance, sphere and quad records (HostScene offsets)
    DevBuf tri_normals, tri_uvs, sphere_map, mats, texs, texels,
        lights;
    std::vector<uint32_t> tri_gid, quad_gid;   // leaf order -> global primitive id
    std::vector<uint32_t> inst_gid;            // instance record (leaf order) -> caller's instance index
    crt_dev_scene dev{};
    uint32_t slot = 0;
    uint64_t bytes = 0;
    double upload_ms = 0.0;
    std::atomic<int> refs{0};
};

// Device buffers a streaming pass runs on: the context's pools (crt_render, crt_render_batch) or
// the fixed scratch array of crt_render_noalloc.
struct PassBufs {
    float4* ray_o[2]; float4* ray_d[2]; float4* thr[2]; float4* rad[2]; uint2* ids[2];
    float4* hit; uint32_t* hit_inst;
    float4* out_rad;                // 16 B per path id of a pass
    uint32_t* gstack;               // traversal overflow stacks of extend_grid workgroups
    uint32_t* ctl;                  // 2 * CRT_CTL_WORDS words per sub-pass
    crt_dev_counters* counters;
    crt_dev_request* req[2];        // device request tables (alternating between passes)
    uint32_t slots;                 // live path slots
    uint32_t pass_ids;              // path ids per pass
    uint32_t extend_grid;           // closest-hit workgroups
    uint32_t max_sub;               // concurrent sub-passes the buffers allow
    crt_paths paths(int k, uint32_t slot0) const {
        crt_paths p;
        p.ray_o = ray_o[k] + slot0; p.ray_d = ray_d[k] + slot0; p.thr = thr[k] + slot0; p.rad = rad[k] + slot0;
        p.ids = ids[k] + slot0;
        return p;
    }
};

struct Job;
struct WorkItem {
    Job* job = nullptr;
    uint32_t pix_begin = 0, pix_count = 0, sample_begin = 0, sample_count = 0;
};

struct Job {
    crt_render_desc rd{};
    DeviceScene* scene = nullptr;
    crt_dev_request req{};
    uint32_t npix = 0, sq = 0, nsamples = 0, s_begin = 0, s_end = 0, s_next = 0, p_next = 0;
    DevBuf accum, rgb, rng;
    uint64_t rays_in_flight = 0;
    // closest-hit queries of the job's resolved paths (served jobs): 8 bytes behind the image
    unsigned long long* rays_counter() const {
        return (unsigned long long*)((char*)rgb.p + (((size_t)npix * 3 + 7) & ~(size_t)7));
    }
    static int validate(const crt_render_desc& r);
    static uint64_t bytes_needed(const crt_render_desc& r);
    int setup(Context* ctx, DeviceScene* sc, const crt_render_desc& r, hipStream_t st = nullptr);
    // a job whose accumulation buffer the caller owns (device memory): no allocation
    void setup_external(DeviceScene* sc, const crt_render_desc& r, float* accum_dev);
    bool take(uint32_t cap, WorkItem& w);
    bool scheduled_all() const { return s_next >= s_end; }
    // path ids not yet handed to a pass
    uint64_t paths_left() const { return s_next >= s_end ? 0 : (uint64_t)(s_end - s_next) * npix - p_next; }
};

// Serving request (continuous batching)
struct ServeReq {
    crt_request_id id = 0;
    crt_scene_id scene = 0;
    DeviceScene* sc = nullptr;     // pinned by sc->refs until the request completes
    crt_render_desc rd{};
    std::unique_ptr<Job> job;
    std::vector<uint8_t> rgb;      // result (host)
    std::vector<float> accum;
    crt_request_timing timing{};
    int status = CRT_ERR_NOT_READY;
    bool cancel = false;
    uint32_t segs = 0;             // segments appended to the open pass
    uint32_t segs_resolved = 0;    // segments whose resolve has been enqueued (in segment order)
    bool finishing = false;        // tonemap + readback enqueued, waiting for ev_done
    int lane = 0;                  // open pass (server lane) that traces all of its segments
    uint64_t load = 0;             // paths it adds to its lane's load until it completes
    hipEvent_t ev_done = nullptr;
};

// One segment (work item) of a request inside the server's open pass: a request-table entry.
struct SrvEntry {
    ServeReq* r = nullptr;
    uint32_t seg = 0, paths = 0;
    uint32_t end = 0;                 // one past the entry's last path id
    bool done = false, resolved = false;
};

// The serving loop's device state: its own stream, path pool and overflow stacks, and one open
// streaming pass that grows while it runs (the host appends request-table entries and path ids
// between iterations; regeneration reads the totals from device memory).
struct Server {
    static constexpr uint32_t kEntries = 4096;   // request-table entries per pass (segments)
    static constexpr int kRing = 8, kLag = 4;    // snapshots in flight / host look-behind
    static constexpr int kOpenRing = 64;         // pinned (total, nreq) words for the appends
    static constexpr int kMaxShadeGrid = 65536;            // shade workgroups (one live-minimum word each)
    static constexpr int kCounterWords = (int)((sizeof(crt_dev_counters) + 3) / 4);
    static constexpr int kSnapWords = CRT_CTL_WORDS + kCounterWords + kMaxShadeGrid;   // control half, counters, live minima
    bool ready = false;
    hipStream_t s = nullptr, copy_s = nullptr;
    uint32_t slots = 0, cap = 0, egrid = 0, sgrid = 0;
    DevBuf ray_o[2], ray_d[2], thr[2], rad[2], ids[2], hit, hit_inst, out_rad, req, ctl, gstack, counters, live_min;
    crt_dev_request* req_host = nullptr;   // pinned mirror of the open pass's request table
    uint32_t* open_host = nullptr;         // pinned ring of (total, nreq)
    uint32_t* snap = nullptr;              // pinned ring of iteration snapshots
    void* rays_host = nullptr;             // pinned word for a finished request's ray count
    hipEvent_t ev[kRing] = {};
    // open pass (host view)
    std::vector<SrvEntry> entries;
    uint32_t total = 0;                    // path ids appended
    uint64_t iter = 0, processed = 0;      // iterations enqueued / snapshots read
    uint64_t done_paths = 0;               // paths of the entries known to be finished
    uint32_t gen_seen = 0;                 // generated path ids observed
    uint32_t gen_progress = 0;             // GEN at the last iteration that generated paths
    uint32_t appends = 0;
    bool full = false;                     // no room for another segment: drain, then reset
    bool two_level = false;                // an instanced scene has paths in the open pass
    bool shapes = false;                   // a scene with analytic shapes has paths in the open pass
    bool mixed_rng = false;                // a MINSTD request has paths in the open pass (per-path RNG mode)
    uint32_t rr = 0;                       // round-robin start among active requests
    uint64_t progress_iter = 0;            // last iteration whose snapshot showed finished paths
    bool use_graphs = true;                // replay captured iterations
    // kernel variant of an iteration: two_level * 4 + shapes * 2 + mixed_rng
    static constexpr int kVariants = 8;
    uint32_t warm[kVariants] = {};                 // eager iterations per variant (graphs are captured after kRing)
    hipGraphExec_t graph[kVariants][kRing] = {};   // [variant][snapshot slot]
};

struct Context {
    static constexpr uint32_t kMaxScenes = 1024;
    static constexpr uint32_t kMaxPassRequests = 1024;
    crt_config config{};
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, reserved = 0, scene_bytes = 0, request_bytes = 0;
    uint32_t max_paths = 0, path_capacity = 0, leaf_size = kDefaultLeafSize, bvh_width = kDefaultBvhWidth, extend_grid = 2048, shade_grid = 2048;
    int extend_mode = CRT_EXTEND_GRID_STRIDE;   // closest-hit work distribution (CRT_EXTEND_*)
    uint32_t extend_param = 2;       // lane-refill threshold or 64-ray groups per wave fetch
    uint32_t stack_lds = 16;         // LDS ring entries per lane of the traversal stack (8, 16, 32)
    uint32_t ray_sort = 0;           // shade groups each block's survivors by direction octant
    uint32_t bvh_builder = CRT_BVH_HOST;   // where crt_scene_upload builds the BVH (CRT_BVH_*)
    uint64_t batch_group = 1ull << 22;     // path ids left for a batched request to get passes of its own scene
    uint32_t serve_order = CRT_SERVE_SMALL_THEN_OLDEST;   // serving loop feed order (CRT_SERVE_*)
    uint64_t serve_small = 1ull << 22;     // path ids left below which a served request goes first
    uint32_t grid_per_cu = 16, shade_grid_per_cu = 16;
    // Lock order: exec_mutex before gpu_mutex. exec_mutex serialises the blocking calls (their
    // stream, pass buffers and `last` counters) and is held for a whole render; gpu_mutex guards
    // the scene table, the budget figures and the serving loop's set-up and is held briefly, so a
    // request can be submitted while a blocking render runs.
    std::mutex exec_mutex;
    std::mutex gpu_mutex;
    std::map<crt_scene_id, std::unique_ptr<DeviceScene>> scenes;
    bool slot_used[kMaxScenes] = {};
    crt_scene_id next_scene_id = 1;
    DevBuf scene_table, counters, ctl, gstack;
    DevBuf req_table[2];
    void* req_host[2] = {nullptr, nullptr};
    hipEvent_t req_event[2] = {nullptr, nullptr};
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_ext[2] = {nullptr, nullptr};
    void* counters_host = nullptr;
    int pass_parity = 0;
    uint64_t passes = 0, extend_launches = 0, iterations = 0;
    // streaming pass: two compact path-state buffers of max_paths slots (alternating between
    // iterations), hit records, per-path-id radiance
    DevBuf ray_o[2], ray_d[2], thr[2], rad[2], ids[2], hit, hit_inst, out_rad;
    uint32_t pass_capacity = 0;      // path ids per pass (out_rad entries)
    static constexpr int kCtlRing = 8, kCtlLag = 4;   // control snapshots in flight / host look-behind
    static constexpr int kMaxSubPasses = 4;           // concurrent sub-passes (streams) of one pass
    uint32_t sub_passes = 2;
    uint32_t* ctl_host = nullptr;    // pinned: kCtlRing snapshots per sub-pass of the control block's "out" half
    hipEvent_t ctl_ev[kCtlRing * kMaxSubPasses] = {};
    hipStream_t sub_stream[kMaxSubPasses - 1] = {};   // default priority, like `stream`
    hipEvent_t sub_done[kMaxSubPasses - 1] = {};
    hipEvent_t ev_pass_ready = nullptr;
    std::atomic<int> fail_alloc_after{-1};   // test hook (crt_debug_fail_alloc): fail the n-th next device allocation
    // released per-request device buffers, by size class (power of two), at most kCacheBytes
    static constexpr size_t kCacheBytes = 4ull << 30;
    std::mutex cache_mutex;
    std::multimap<size_t, void*> buf_cache;
    size_t cache_bytes = 0;
    void* cache_take(size_t cls);
    bool cache_put(void* p, size_t cls);
    crt_work_counters last{};        // counters of the last blocking render

    // serving
    std::mutex srv_mutex;
    std::condition_variable srv_cv, srv_done_cv;
    // finished requests waiting for their readback; a completion thread copies them back on the
    // server's copy stream so the scheduler keeps the open pass fed meanwhile
    std::deque<std::pair<ServeReq*, int>> srv_copy_q;
    std::condition_variable srv_copy_cv;
    bool srv_copy_stop = false;
    std::thread srv_thread;
    bool srv_running = false, srv_stop = false;
    std::deque<ServeReq*> srv_waiting;                 // submitted, not admitted
    std::vector<ServeReq*> srv_active;                 // admitted, on the device
    std::map<crt_request_id, std::unique_ptr<ServeReq>> srv_all;
    crt_request_id next_req_id = 1;
    // up to kSrvLanes open passes on their own streams (each with a share of the slot pool), so one
    // pass's drain tail overlaps the other's full launches; a request's segments stay in one lane
    static constexpr int kSrvLanes = 2;
    Server srv[kSrvLanes];
    uint32_t srv_lanes = 1;
    uint64_t lane_load[kSrvLanes] = {};

    int init(const crt_config* cfg);
    ~Context();
    uint64_t used_bytes() const;
    int ensure_paths(uint32_t slots, uint32_t pass_ids);
    int ensure_gstack(uint32_t blocks);
    uint64_t server_bytes() const;
    int autotune(crt_scene_id scene, const crt_scene_desc* desc, const crt_render_desc* rd, crt_tune_result* out);
    int tune_trial(crt_scene_id scene, const crt_render_desc& rd, double& mrays_s);
    int trial_scene(crt_scene_id scene, const crt_render_desc& rd, double& mrays_s, uint32_t& trials);
    crt_paths paths(int k, uint32_t slot0);
    int upload_scene(const crt_scene_desc* desc, crt_scene_id* out);
    int upload_scene_locked(const crt_scene_desc* desc, uint32_t leaf, uint32_t width, crt_scene_id* out);
    int verify_scene(const DeviceScene& sc, std::string& err);
    int release_scene(crt_scene_id id);
    int release_scene_locked(crt_scene_id id);
    DeviceScene* find_scene(crt_scene_id id);
    PassBufs pool_bufs();
    int run_pass(std::vector<WorkItem>& items, bool time_stages, double* extend_ms, const PassBufs& B);
    int run_plain(Job& j, bool time_stages, double* extend_ms);
    int finish_job(Job& j, crt_image* out);
    int reset_counters(crt_dev_counters* c);
    int read_counters(crt_dev_counters& c, const crt_dev_counters* dev);
    int render_jobs(std::vector<Job*>& jobs, bool time_stages, crt_stats* stats, const PassBufs& B);
    // no-allocation render (crt_render_noalloc): passes run on the device's scratch array
    void* scratch = nullptr;         // this device's scratch array (hipGetSymbolAddress at init)
    size_t scratch_bytes = 0;
    int scratch_bufs(PassBufs& B);
    int render_noalloc(crt_scene_id scene, const crt_render_desc* rd, float* accum_dev, uint8_t* rgb_dev,
                       crt_stats* stats);
    // serving
    int submit(crt_scene_id scene, const crt_render_desc* rd, crt_request_id* out);
    void server_loop();
    void server_copy_loop();
    void server_complete(ServeReq* r, int status);
    void stop_server();
    int server_init();
    int server_init_lane(Server& S, uint32_t slots, uint32_t cap, bool first);
    uint64_t server_lane_bytes(const Server& S) const;
    void server_release();
    void server_release_lane(Server& S);
    int server_iteration(Server& S);
    int server_append(Server& S, std::vector<ServeReq*>& act);
    bool append_segment(Server& S, ServeReq* r, uint32_t seg_cap);
    int server_process_snapshot(Server& S, std::vector<ServeReq*>& act);
    void server_reset_pass(Server& S);
    bool server_idle();
    void server_drop_graphs();
};

// Both context locks in their order (blocking calls that also change or walk the scene table).
struct ExecLock {
    std::lock_guard<std::mutex> e, g;
    explicit ExecLock(Context& c) : e(c.exec_mutex), g(c.gpu_mutex) {}
};

void fill_stats(crt_stats* st, const crt_dev_counters& c, uint64_t paths, double gpu_ms, double ext_ms, double host_ms);

// Device BVH build of a one-level scene (bvh_gpu.hip): fills the scene's allocations, its device
// record, id maps and statistics. Caller holds both context locks and has checked the budget.
int gpu_build_scene(Context& C, const crt_scene_desc* d, uint32_t leaf, uint32_t width, DeviceScene& sc, std::string& err);

}  // namespace crt
