// api.cpp - the C-ABI (include/cortex_rt.h): context, scene residency, the pass scheduler shared
// by the plain, batched and serving paths, and first-hit queries.
//
// Replaces launchGPU/renderGPU (core/src/gpu.cu:1982-2149): instead of one megakernel launch per
// render with the scene rebuilt on the device each time, scenes stay resident and every render is
// a sequence of wavefront passes whose path slots may be shared by several requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
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
#include "crt_runtime.h"
#include "crt_scratch.h"
#include "scene_build.h"

namespace crt {

static thread_local std::string g_last_error;
int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------ device allocations
int DevBuf::alloc(size_t n, Context* ctx) {
    release();
    if (n == 0) return CRT_OK;
    if (ctx && ctx->fail_alloc_after.load() >= 0) {
        // test hook: the n-th allocation from now fails (n = 0: this one); each failure re-arms it
        // at -1 so exactly one allocation fails
        int left = ctx->fail_alloc_after.fetch_sub(1);
        if (left == 0) return set_error(CRT_ERR_OUT_OF_MEMORY, "injected allocation failure");
        if (left < 0) ctx->fail_alloc_after.store(-1);
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) { p = nullptr; return set_error(CRT_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e)); }
    bytes = n;
    return CRT_OK;
}
static size_t size_class(size_t n) {
    size_t c = 4096;
    while (c < n) c <<= 1;
    return c;
}

int DevBuf::alloc_cached(size_t n, Context* ctx) {
    release();
    if (n == 0) return CRT_OK;
    const size_t cls = size_class(n);
    void* q = ctx->cache_take(cls);
    if (!q) {
        int rc = alloc(cls, ctx);
        if (rc) return rc;
    } else {
        p = q;
        bytes = cls;
    }
    home = ctx;
    return CRT_OK;
}

void DevBuf::release() {
    if (p && !(home && home->cache_put(p, bytes))) (void)hipFree(p);
    p = nullptr; bytes = 0; home = nullptr;
}

void* Context::cache_take(size_t cls) {
    std::lock_guard<std::mutex> g(cache_mutex);
    auto it = buf_cache.find(cls);
    if (it == buf_cache.end()) return nullptr;
    void* q = it->second;
    buf_cache.erase(it);
    cache_bytes -= cls;
    return q;
}

bool Context::cache_put(void* q, size_t cls) {
    std::lock_guard<std::mutex> g(cache_mutex);
    if (cache_bytes + cls > kCacheBytes) return false;
    buf_cache.emplace(cls, q);
    cache_bytes += cls;
    return true;
}

template <class T> static int upload_vec(DevBuf& b, const std::vector<T>& v, Context* ctx, hipStream_t s) {
    if (v.empty()) { b.release(); return CRT_OK; }
    int rc = b.alloc(v.size() * sizeof(T), ctx);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return CRT_OK;
}

// ------------------------------------------------------------------ camera (host, FP32)
// perspective_camera::initialize (core/src/cameras/perspective_camera.cuh:41-104) with the same
// float expression order; degrees_to_radians = deg * M_PI / 180 (math_utils.cuh:55-58). The
// orthographic camera: orthographic_camera::initialize (core/src/cameras/orthographic_camera.cuh:
// 57-118), whose viewport width follows the image's: width = render width, height =
// max(1, int(width / aspect)).
void camera_setup(const crt_camera& c, uint32_t width, crt_dev_request& r) {
    struct V { float x, y, z; };
    auto sub = [](V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto add = [](V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; };
    auto mul = [](float s, V a) { return V{s * a.x, s * a.y, s * a.z}; };
    auto crs = [](V a, V b) { return V{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; };
    auto unit = [](V v) { float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); return V{v.x / l, v.y / l, v.z / l}; };
    V from{c.lookfrom[0], c.lookfrom[1], c.lookfrom[2]}, at{c.lookat[0], c.lookat[1], c.lookat[2]}, up{c.vup[0], c.vup[1], c.vup[2]};
    float theta = c.vfov_deg * 3.14159265358979323846f / 180.0f;
    float focus = c.focus_dist, ratio = c.aspect_ratio;
    V w = unit(sub(from, at));
    V u = unit(crs(up, w));
    V v = crs(w, u);
    float hh = std::tan(theta / 2.0f);
    float hw = ratio * hh;
    V llc = sub(sub(sub(from, mul(hw * focus, u)), mul(hh * focus, v)), mul(focus, w));
    V hor = mul(2.0f * hw * focus, u);
    V ver = mul(2.0f * hh * focus, v);
    auto put = [](float* d, V a) { d[0] = a.x; d[1] = a.y; d[2] = a.z; };
    put(r.origin, from); put(r.llc, llc); put(r.horizontal, hor); put(r.vertical, ver); put(r.cu, u); put(r.cv, v);
    r.lens_radius = c.aperture / 2.0f;
    r.time0 = c.time0; r.time1 = c.time1;
    r.ortho = 0u;
    if (c.orthographic) {
        int ih = (int)((float)width / ratio);
        ih = ih < 1 ? 1 : ih;
        const float vh = 2.0f * c.ortho_height;
        const float vw = vh * ((float)width / (float)ih);
        V ollc = sub(sub(from, mul(0.5f * vw, u)), mul(0.5f * vh, v));
        put(r.llc, ollc); put(r.horizontal, mul(vw, u)); put(r.vertical, mul(vh, v));
        put(r.ortho_dir, V{-w.x, -w.y, -w.z});
        r.ortho = 1u;
    }
    (void)add;
}

// ------------------------------------------------------------------ context
int Context::init(const crt_config* cfg) {
    if (cfg) config = *cfg;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(CRT_ERR_NO_DEVICE, "no HIP device found");
    if (config.device < 0 || config.device >= n) return set_error(CRT_ERR_NO_DEVICE, "device ordinal out of range");
    device = config.device;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    size_t freeb = 0, totalb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    budget = config.hbm_budget_bytes ? config.hbm_budget_bytes : (uint64_t)(0.9 * (double)freeb);
    max_paths = config.max_paths ? config.max_paths : (1u << 26);
    leaf_size = config.leaf_size ? config.leaf_size : kDefaultLeafSize;
    bvh_width = config.bvh_width ? config.bvh_width : kDefaultBvhWidth;
    if (config.extend_kernel > CRT_EXTEND_WAVE_CHUNK) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown extend_kernel");
    extend_mode = config.extend_kernel ? (int)config.extend_kernel : CRT_EXTEND_GRID_STRIDE;
    extend_param = config.refill_min ? std::min(config.refill_min, 64u) : (extend_mode == CRT_EXTEND_PERSISTENT ? 16u : 2u);
    if (bvh_width != 2 && bvh_width != 4) return set_error(CRT_ERR_INVALID_ARGUMENT, "bvh_width must be 2 or 4");
    stack_lds = config.stack_lds ? config.stack_lds : 16u;
    if (config.ray_sort > 2) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown ray_sort");
    ray_sort = config.ray_sort == 1u ? 1u : 0u;
    if (config.bvh_builder > CRT_BVH_DEVICE) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown bvh_builder");
    bvh_builder = config.bvh_builder ? config.bvh_builder : (uint32_t)CRT_BVH_HOST;
    if (config.serve_order > CRT_SERVE_ROUND_ROBIN) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown serve_order");
    serve_order = config.serve_order ? config.serve_order : (uint32_t)CRT_SERVE_SMALL_THEN_OLDEST;
    serve_small = config.serve_small_paths ? (uint64_t)config.serve_small_paths : (1ull << 22);
    batch_group = config.batch_group_paths == 0xffffffffu ? UINT64_MAX
                : config.batch_group_paths ? (uint64_t)config.batch_group_paths : (1ull << 22);
    if (stack_lds != 8 && stack_lds != 16 && stack_lds != 32)
        return set_error(CRT_ERR_INVALID_ARGUMENT, "stack_lds must be 8, 16 or 32");
    grid_per_cu = config.grid_per_cu ? std::min(config.grid_per_cu, 256u) : 24u;
    shade_grid_per_cu = config.shade_grid_per_cu ? std::min(config.shade_grid_per_cu, 256u) : 16u;
    extend_grid = (uint32_t)num_cus * grid_per_cu;
    shade_grid = (uint32_t)num_cus * shade_grid_per_cu;
    int rc;
    if ((rc = scene_table.alloc(sizeof(crt_dev_scene) * kMaxScenes, this))) return rc;
    if ((rc = counters.alloc(sizeof(crt_dev_counters), this))) return rc;
    if ((rc = ctl.alloc(sizeof(uint32_t) * 2 * CRT_CTL_WORDS * kMaxSubPasses, this))) return rc;
    HIP_TRY(hipHostMalloc((void**)&ctl_host, sizeof(uint32_t) * CRT_CTL_WORDS * kCtlRing * kMaxSubPasses,
                          hipHostMallocDefault));
    for (int k = 0; k < kCtlRing * kMaxSubPasses; ++k) HIP_TRY(hipEventCreateWithFlags(&ctl_ev[k], hipEventDisableTiming));
    sub_passes = config.sub_passes ? std::min<uint32_t>(config.sub_passes, kMaxSubPasses) : 2u;
    for (int k = 0; k < kMaxSubPasses - 1; ++k) {
        HIP_TRY(hipStreamCreateWithFlags(&sub_stream[k], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&sub_done[k], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&ev_pass_ready, hipEventDisableTiming));
    if ((rc = ensure_gstack(extend_grid))) return rc;
    for (int k = 0; k < 2; ++k) {
        if ((rc = req_table[k].alloc(sizeof(crt_dev_request) * kMaxPassRequests, this))) return rc;
        HIP_TRY(hipHostMalloc(&req_host[k], sizeof(crt_dev_request) * kMaxPassRequests, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&req_event[k], hipEventDisableTiming));
    }
    HIP_TRY(hipHostMalloc((void**)&counters_host, sizeof(crt_dev_counters), hipHostMallocDefault));
    HIP_TRY(hipEventCreate(&ev_start));
    HIP_TRY(hipEventCreate(&ev_stop));
    HIP_TRY(hipEventCreate(&ev_ext[0]));
    HIP_TRY(hipEventCreate(&ev_ext[1]));
    // path ids per pass: every path of a 1080p 64-spp frame fits one pass by default (2 GiB of
    // per-path radiance); at least one id per slot
    uint64_t pp = config.pass_paths ? config.pass_paths : std::min<uint64_t>(1ull << 27, 16ull * max_paths);
    if ((rc = ensure_paths(max_paths, (uint32_t)std::max<uint64_t>(pp, max_paths)))) return rc;
    reserved = used_bytes();
    // the no-allocation render's scratch array: its device address on this device (current here)
    hipError_t se = (hipError_t)crt_scratch_setup(device, &scratch, &scratch_bytes);
    if (se != hipSuccess)
        return set_error(CRT_ERR_HIP, std::string("scratch array address: ") + hipGetErrorString(se));
    return CRT_OK;
}

// Carves the pass buffers of crt_render_noalloc out of the device's scratch array: one sub-pass,
// kCrtScratchSlots live paths, kCrtScratchPassIds path ids per pass, closest-hit overflow stacks for
// kCrtScratchExtendPerCu workgroups per CU (fewer if the device's CUs would not fit).
int Context::scratch_bufs(PassBufs& B) {
    if (!scratch) return set_error(CRT_ERR_INTERNAL, "no scratch array on this device");
    const uint64_t S = kCrtScratchSlots, P = kCrtScratchPassIds;
    const uint64_t per_block_stack = 256ull * crt_k_gstack_entries() * sizeof(uint32_t);
    uint64_t off = 0;
    char* base = (char*)scratch;
    auto take = [&](uint64_t bytes) -> void* {
        void* p = base + off;
        off = (off + bytes + 255u) & ~(uint64_t)255u;
        return p;
    };
    for (int k = 0; k < 2; ++k) {
        B.ray_o[k] = (float4*)take(S * 16); B.ray_d[k] = (float4*)take(S * 16); B.thr[k] = (float4*)take(S * 16);
        B.rad[k] = (float4*)take(S * 16); B.ids[k] = (uint2*)take(S * 8);
    }
    B.hit = (float4*)take(S * 16);
    B.hit_inst = (uint32_t*)take(S * 4);
    B.out_rad = (float4*)take(P * 16);
    B.ctl = (uint32_t*)take(2 * CRT_CTL_WORDS * sizeof(uint32_t));
    B.counters = (crt_dev_counters*)take(sizeof(crt_dev_counters));
    B.req[0] = (crt_dev_request*)take(sizeof(crt_dev_request));
    B.req[1] = (crt_dev_request*)take(sizeof(crt_dev_request));
    if (off >= scratch_bytes) return set_error(CRT_ERR_INTERNAL, "scratch array too small for the pass state");
    const uint64_t fit = (scratch_bytes - off) / per_block_stack;
    uint64_t grid = std::min<uint64_t>((uint64_t)kCrtScratchExtendPerCu * (uint64_t)num_cus, fit);
    if (grid == 0) return set_error(CRT_ERR_INTERNAL, "scratch array too small for the traversal stacks");
    B.gstack = (uint32_t*)take(grid * per_block_stack);
    if (off > scratch_bytes) return set_error(CRT_ERR_INTERNAL, "scratch layout exceeds the scratch array");
    B.slots = (uint32_t)S;
    B.pass_ids = (uint32_t)P;
    B.extend_grid = (uint32_t)grid;
    B.max_sub = 1;
    return CRT_OK;
}

uint64_t Context::used_bytes() const {
    uint64_t b = scene_table.bytes + counters.bytes + ctl.bytes + gstack.bytes + req_table[0].bytes + req_table[1].bytes;
    for (int k = 0; k < 2; ++k) b += ray_o[k].bytes + ray_d[k].bytes + thr[k].bytes + rad[k].bytes + ids[k].bytes;
    b += hit.bytes + hit_inst.bytes + out_rad.bytes;
    return b;
}

int Context::ensure_paths(uint32_t n, uint32_t pass_ids) {
    int rc;
    if (n > path_capacity) {
        for (int k = 0; k < 2; ++k) {
            if ((rc = ray_o[k].alloc((size_t)n * 16, this))) return rc;
            if ((rc = ray_d[k].alloc((size_t)n * 16, this))) return rc;
            if ((rc = thr[k].alloc((size_t)n * 16, this))) return rc;
            if ((rc = rad[k].alloc((size_t)n * 16, this))) return rc;
            if ((rc = ids[k].alloc((size_t)n * 8, this))) return rc;
        }
        if ((rc = hit.alloc((size_t)n * 16, this))) return rc;
        if ((rc = hit_inst.alloc((size_t)n * 4, this))) return rc;
        path_capacity = n;
    }
    if (pass_ids > pass_capacity) {
        if ((rc = out_rad.alloc((size_t)pass_ids * 16, this))) return rc;
        pass_capacity = pass_ids;
    }
    return CRT_OK;
}

// Traversal overflow stacks for `blocks` workgroups of the closest-hit launches.
int Context::ensure_gstack(uint32_t blocks) {
    const size_t need = (size_t)blocks * 256u * crt_k_gstack_entries() * sizeof(uint32_t);
    if (need <= gstack.bytes) return CRT_OK;
    if (stream) HIP_TRY(hipStreamSynchronize(stream));
    return gstack.alloc(need, this);
}

crt_paths Context::paths(int k, uint32_t slot0) {
    return pool_bufs().paths(k, slot0);
}

// The context's own pass buffers (allocated at crt_create, grown by the autotuner).
PassBufs Context::pool_bufs() {
    PassBufs B;
    for (int k = 0; k < 2; ++k) {
        B.ray_o[k] = (float4*)ray_o[k].p; B.ray_d[k] = (float4*)ray_d[k].p; B.thr[k] = (float4*)thr[k].p;
        B.rad[k] = (float4*)rad[k].p; B.ids[k] = (uint2*)ids[k].p; B.req[k] = (crt_dev_request*)req_table[k].p;
    }
    B.hit = (float4*)hit.p; B.hit_inst = (uint32_t*)hit_inst.p; B.out_rad = (float4*)out_rad.p;
    B.gstack = (uint32_t*)gstack.p; B.ctl = (uint32_t*)ctl.p; B.counters = (crt_dev_counters*)counters.p;
    B.slots = std::min(max_paths, path_capacity);
    B.pass_ids = pass_capacity;
    B.extend_grid = extend_grid;
    B.max_sub = sub_passes;
    return B;
}

Context::~Context() {
    stop_server();
    server_release();
    if (stream) (void)hipStreamSynchronize(stream);
    scenes.clear();
    for (int k = 0; k < 2; ++k) {
        if (req_host[k]) (void)hipHostFree(req_host[k]);
        if (req_event[k]) (void)hipEventDestroy(req_event[k]);
    }
    if (counters_host) (void)hipHostFree(counters_host);
    for (int k = 0; k < kMaxSubPasses - 1; ++k) if (sub_stream[k]) (void)hipStreamSynchronize(sub_stream[k]);
    if (ctl_host) (void)hipHostFree(ctl_host);
    for (int k = 0; k < kCtlRing * kMaxSubPasses; ++k) if (ctl_ev[k]) (void)hipEventDestroy(ctl_ev[k]);
    for (int k = 0; k < kMaxSubPasses - 1; ++k) {
        if (sub_done[k]) (void)hipEventDestroy(sub_done[k]);
        if (sub_stream[k]) (void)hipStreamDestroy(sub_stream[k]);
    }
    if (ev_pass_ready) (void)hipEventDestroy(ev_pass_ready);
    if (ev_start) (void)hipEventDestroy(ev_start);
    for (int k = 0; k < 2; ++k) if (ev_ext[k]) (void)hipEventDestroy(ev_ext[k]);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (stream) (void)hipStreamDestroy(stream);
    // every request has completed (stop_server) and every blocking render returned: the cache
    // holds the last per-request buffers
    for (auto& kv : buf_cache) (void)hipFree(kv.second);
    buf_cache.clear();
    cache_bytes = 0;
}

// ------------------------------------------------------------------ scenes
int Context::upload_scene(const crt_scene_desc* desc, crt_scene_id* out) {
    ExecLock g(*this);
    return upload_scene_locked(desc, leaf_size, bvh_width, out);
}

// Device bytes a device build of `d` may take at most (the scene's own allocations; the build's
// work buffers come and go within the call): one node per binary interior node, every record,
// per-triangle normals / uvs, sphere mappings and the tables.
static uint64_t device_build_bound(const crt_scene_desc* d, uint32_t width, uint64_t tables) {
    uint64_t nt = 0;
    bool nrm = false, uvs = false;
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        nt += d->meshes[m].num_triangles;
        nrm = nrm || d->meshes[m].normals; uvs = uvs || d->meshes[m].uvs;
    }
    const uint64_t nq = (uint64_t)d->num_quads + d->num_lights + d->num_shapes, ns = (uint64_t)d->num_spheres + d->num_sphere_lights,
                   n = nt + ns + nq;
    return n * (width == 4 ? 128u : 64u) + nt * 48u + ns * 32u + nq * 80u + (nrm ? nt * 48u : 0u) + (uvs ? nt * 24u : 0u) +
           ns * 24u + ((uint64_t)d->num_lights + d->num_sphere_lights) * 4u + tables + 4 * 256u;
}

// Builds and uploads a scene with the given BVH shape (caller holds both context locks): on the host (binned
// SAH, one upload), or on the device when the context asks for it and the scene is one-level with
// two primitives or more.
int Context::upload_scene_locked(const crt_scene_desc* desc, uint32_t leaf, uint32_t width, crt_scene_id* out) {
    auto sc = std::make_unique<DeviceScene>();
    std::string err;
    uint32_t slot = kMaxScenes;
    for (uint32_t k = 0; k < kMaxScenes; ++k) if (!slot_used[k]) { slot = k; break; }
    if (slot == kMaxScenes) return set_error(CRT_ERR_OVER_BUDGET, "scene table full");
    uint64_t nprims = 0;
    if (desc && desc->abi_version == CRT_ABI_VERSION) {
        nprims = (uint64_t)desc->num_lights + desc->num_sphere_lights + desc->num_spheres + desc->num_quads + desc->num_shapes;
        for (uint32_t m = 0; m < desc->num_meshes && desc->meshes; ++m) nprims += desc->meshes[m].num_triangles;
    }
    const bool on_device = bvh_builder == CRT_BVH_DEVICE && desc && desc->abi_version == CRT_ABI_VERSION &&
                           desc->num_instances == 0 && nprims >= 2 && (desc->num_meshes == 0 || desc->meshes);
    auto t0 = std::chrono::steady_clock::now();
    HostScene& H = sc->host;
    uint64_t need = 0;
    if (on_device) {
        uint64_t tables = 0;
        for (uint32_t k = 0; k < desc->num_textures && desc->textures; ++k) {
            const crt_texture& t = desc->textures[k];
            tables += 32u + (t.type == CRT_TEX_IMAGE ? (uint64_t)t.width * t.height * t.channels : 0u);
        }
        tables += ((uint64_t)desc->num_materials + desc->num_lights + 2ull * desc->num_sphere_lights + desc->num_shapes) * 32u;
        if (scene_bytes + request_bytes + reserved + device_build_bound(desc, width, tables) > budget)
            return set_error(CRT_ERR_OVER_BUDGET, "scene exceeds the HBM budget");
        int rc = gpu_build_scene(*this, desc, leaf, width, *sc, err);
        if (rc) return err.empty() ? rc : set_error(rc, err);
        need = sc->bytes;
    } else {
        unsigned hw = std::thread::hardware_concurrency();
        int rc = build_host_scene(desc, leaf, width, (int)std::min(16u, hw ? hw : 1u), H, err);
        if (rc) return set_error(rc, err);
        need = H.device_bytes();
        if (scene_bytes + request_bytes + reserved + need > budget)
            return set_error(CRT_ERR_OVER_BUDGET, "scene of " + std::to_string(need) + " bytes exceeds the HBM budget");
        t0 = std::chrono::steady_clock::now();
        {
            // nodes, triangle, instance, sphere and quad records in one allocation, in that order (the
            // traversal fetches them through 32-bit byte offsets from the node array); 256 spare bytes
            // keep the shared loads of the last record (and of a node-less scene) in bounds
            struct Part { const void* src; size_t bytes, off; };
            const Part parts[5] = {{H.nodes.data(), H.nodes.size() * sizeof(F4), 0},
                                   {H.tris.data(), H.tris.size() * sizeof(F4), (size_t)H.tri_offset()},
                                   {H.insts.data(), H.insts.size() * sizeof(F4), (size_t)H.inst_offset()},
                                   {H.spheres.data(), H.spheres.size() * sizeof(F4), (size_t)H.sphere_offset()},
                                   {H.quads.data(), H.quads.size() * sizeof(F4), (size_t)H.quad_offset()}};
            if ((rc = sc->geo.alloc((size_t)H.geometry_bytes() + 256, this))) return rc;
            for (const Part& q : parts)
                if (q.bytes) HIP_TRY(hipMemcpyAsync((char*)sc->geo.p + q.off, q.src, q.bytes, hipMemcpyHostToDevice, stream));
        }
        if ((rc = upload_vec(sc->tri_normals, H.tri_normals, this, stream))) return rc;
        if ((rc = upload_vec(sc->tri_uvs, H.tri_uvs, this, stream))) return rc;
        if ((rc = upload_vec(sc->sphere_map, H.sphere_map, this, stream))) return rc;
        if ((rc = upload_vec(sc->mats, H.mats, this, stream))) return rc;
        if ((rc = upload_vec(sc->texs, H.texs, this, stream))) return rc;
        if ((rc = upload_vec(sc->texels, H.texels, this, stream))) return rc;
        if ((rc = upload_vec(sc->lights, H.light_refs, this, stream))) return rc;
        crt_dev_scene& ds = sc->dev;
        std::memset(&ds, 0, sizeof(ds));
        char* geo = (char*)sc->geo.p;
        ds.nodes = geo; ds.tris = geo + H.tri_offset(); ds.tri_normals = sc->tri_normals.p;
        ds.tri_uvs = sc->tri_uvs.p; ds.spheres = geo + H.sphere_offset();
        ds.sphere_map = sc->sphere_map.p; ds.quads = geo + H.quad_offset(); ds.insts = geo + H.inst_offset();
        ds.mats = sc->mats.p; ds.texs = sc->texs.p;
        ds.texels = (const uint8_t*)sc->texels.p; ds.lights = (const uint32_t*)sc->lights.p;
        ds.num_lights = H.num_lights; ds.root = H.root; ds.num_quads = H.num_quads; ds.bvh_width = H.bvh_width;
        for (int k = 0; k < 3; ++k) ds.background[k] = H.background[k];
        H.num_node_records = H.nodes.size() / (H.bvh_width == 4 ? 8u : 4u);
        H.num_tri_records = H.tris.size() / 3;
    }
    HIP_TRY(hipMemcpyAsync((crt_dev_scene*)scene_table.p + slot, &sc->dev, sizeof(crt_dev_scene), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    sc->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    sc->slot = slot;
    sc->bytes = need;
    if (!on_device) {
        // leaf order -> global primitive id for first-hit queries, then drop the big host copies
        sc->tri_gid.resize(H.tris.size() / 3);
        for (size_t i = 0; i < sc->tri_gid.size(); ++i) std::memcpy(&sc->tri_gid[i], &H.tris[3 * i].w, 4);
        sc->inst_gid.resize(H.insts.size() / 4);
        for (size_t i = 0; i < sc->inst_gid.size(); ++i) std::memcpy(&sc->inst_gid[i], &H.insts[4 * i + 3].y, 4);
        sc->quad_gid.resize(H.quads.size() / 5);
        for (size_t i = 0; i < sc->quad_gid.size(); ++i) std::memcpy(&sc->quad_gid[i], &H.quads[5 * i + 3].w, 4);
        H.nodes.clear(); H.nodes.shrink_to_fit(); H.tris.clear(); H.tris.shrink_to_fit();
        H.texels.clear(); H.texels.shrink_to_fit();
    }
    slot_used[slot] = true;
    scene_bytes += need;
    crt_scene_id id = next_scene_id++;
    scenes[id] = std::move(sc);
    *out = id;
    return CRT_OK;
}

// Reads the scene's BVH and records back into a HostScene and checks its invariants (see
// crt_scene_verify). Caller holds both context locks.
int Context::verify_scene(const DeviceScene& sc, std::string& err) {
    const HostScene& H = sc.host;
    HostScene V;
    V.bvh_width = H.bvh_width;
    V.root = H.root;
    const uint32_t per = H.bvh_width == 4 ? 8u : 4u;
    const char* geo = (const char*)sc.geo.p;
    const uint64_t nq = H.num_quad_records;
    V.nodes.resize(H.num_node_records * per);
    V.tris.resize(H.num_tri_records * 3);
    V.spheres.resize((size_t)H.num_spheres * 2);
    V.quads.resize(nq * 5);
    V.insts.resize(sc.inst_gid.size() * 4);
    auto rd = [&](std::vector<F4>& v, const void* src) -> int {
        if (!v.empty()) HIP_TRY(hipMemcpyAsync(v.data(), src, v.size() * sizeof(F4), hipMemcpyDeviceToHost, stream));
        return CRT_OK;
    };
    int rc;
    if ((rc = rd(V.nodes, geo)) || (rc = rd(V.tris, sc.dev.tris)) || (rc = rd(V.spheres, sc.dev.spheres)) ||
        (rc = rd(V.quads, sc.dev.quads)) || (rc = rd(V.insts, sc.dev.insts)))
        return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    if ((rc = check_host_scene(V, err))) return rc;
    // the leaf-order records are a permutation of the scene's primitives: their global ids (kept in
    // the records and in the host maps) agree and cover the scene's id range once
    std::vector<uint32_t> ids;
    ids.reserve(V.tris.size() / 3 + V.spheres.size() / 2 + nq);
    for (size_t i = 0; i < V.tris.size() / 3; ++i) {
        uint32_t g; std::memcpy(&g, &V.tris[3 * i].w, 4);
        if (i < sc.tri_gid.size() && g != sc.tri_gid[i] && H.num_instances == 0) { err = "triangle record id differs from the id map"; return CRT_ERR_INTERNAL; }
        ids.push_back(g);
    }
    for (size_t i = 0; i < V.spheres.size() / 2; ++i) {
        uint32_t g; std::memcpy(&g, &V.spheres[2 * i + 1].x, 4);
        if (i < H.sphere_aux.size() && g != H.sphere_aux[i].x) { err = "sphere record id differs from the id map"; return CRT_ERR_INTERNAL; }
        ids.push_back(g);
    }
    for (size_t i = 0; i < nq; ++i) {
        uint32_t g; std::memcpy(&g, &V.quads[5 * i + 3].w, 4);
        if (i < sc.quad_gid.size() && g != sc.quad_gid[i]) { err = "quad record id differs from the id map"; return CRT_ERR_INTERNAL; }
        ids.push_back(g);
    }
    if (H.num_instances == 0) {
        std::sort(ids.begin(), ids.end());
        for (size_t i = 0; i < ids.size(); ++i)
            if (ids[i] != (uint32_t)i) { err = "primitive records do not cover the scene's ids once"; return CRT_ERR_INTERNAL; }
    }
    return CRT_OK;
}

int Context::release_scene(crt_scene_id id) {
    ExecLock g(*this);
    return release_scene_locked(id);
}

int Context::release_scene_locked(crt_scene_id id) {
    auto it = scenes.find(id);
    if (it == scenes.end()) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    if (it->second->refs > 0) return set_error(CRT_ERR_INVALID_ARGUMENT, "scene has requests in flight");
    (void)hipStreamSynchronize(stream);
    slot_used[it->second->slot] = false;
    scene_bytes -= it->second->bytes;
    scenes.erase(it);
    return CRT_OK;
}

DeviceScene* Context::find_scene(crt_scene_id id) {
    auto it = scenes.find(id);
    return it == scenes.end() ? nullptr : it->second.get();
}

// ------------------------------------------------------------------ jobs and passes
int Job::validate(const crt_render_desc& r) {
    if (r.width == 0 || r.height == 0) return set_error(CRT_ERR_INVALID_ARGUMENT, "width and height must be positive");
    if ((uint64_t)r.width * r.height > (1ull << 28)) return set_error(CRT_ERR_INVALID_ARGUMENT, "image too large");
    if (r.spp == 0) return set_error(CRT_ERR_INVALID_ARGUMENT, "spp must be positive");
    if (r.max_depth == 0 || r.max_depth > 255) return set_error(CRT_ERR_INVALID_ARGUMENT, "max_depth must be in [1, 255]");
    if (r.rng_mode != CRT_RNG_PCG && r.rng_mode != CRT_RNG_MINSTD) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown rng_mode");
    uint32_t sq = (uint32_t)std::sqrt((double)r.spp);
    if (sq * sq > 65536u) return set_error(CRT_ERR_INVALID_ARGUMENT, "spp too large");
    uint32_t ns = sq * sq;
    if (r.sample_begin >= ns && !(r.sample_begin == 0)) return set_error(CRT_ERR_INVALID_ARGUMENT, "sample_begin beyond spp");
    if (r.rng_mode == CRT_RNG_MINSTD && (r.sample_begin != 0 || (r.sample_count != 0 && r.sample_count != ns)))
        return set_error(CRT_ERR_INVALID_ARGUMENT, "MINSTD streams run every sample of a pixel in one call");
    return CRT_OK;
}

uint64_t Job::bytes_needed(const crt_render_desc& r) {
    uint64_t npix = (uint64_t)r.width * r.height;
    return npix * 16 + npix * 3 + 16 + (r.rng_mode == CRT_RNG_MINSTD ? npix * 4 : 0);
}

int Job::setup(Context* ctx, DeviceScene* sc, const crt_render_desc& r, hipStream_t st) {
    rd = r;
    scene = sc;
    npix = r.width * r.height;
    sq = (uint32_t)std::sqrt((double)r.spp);
    nsamples = sq * sq;
    s_begin = std::min(r.sample_begin, nsamples);
    s_end = r.sample_count ? std::min(nsamples, s_begin + r.sample_count) : nsamples;
    s_next = s_begin;
    p_next = 0;
    std::memset(&req, 0, sizeof(req));
    camera_setup(sc->host.camera, r.width, req);
    req.width = r.width; req.height = r.height; req.sqrt_spp = sq; req.spp = r.spp; req.max_depth = r.max_depth;
    req.rng_mode = r.rng_mode; req.scene = sc->slot; req.seed = r.seed;
    int rc;
    if ((rc = accum.alloc_cached((size_t)npix * 16, ctx))) return rc;
    if ((rc = rgb.alloc_cached((((size_t)npix * 3 + 7) & ~(size_t)7) + 8, ctx))) return rc;
    if (r.rng_mode == CRT_RNG_MINSTD && (rc = rng.alloc_cached((size_t)npix * 4, ctx))) return rc;
    req.accum = (float*)accum.p;
    req.rng_state = (uint32_t*)rng.p;
    if (crt_k_init_request(st ? st : ctx->stream, (float4*)accum.p, (uint32_t*)rng.p, r.width, r.height, r.seed) ||
        crt_k_fill_u32(st ? st : ctx->stream, (uint32_t*)rays_counter(), 2, 0u))
        return set_error(CRT_ERR_HIP, "init kernel launch failed");
    return CRT_OK;
}

void Job::setup_external(DeviceScene* sc, const crt_render_desc& r, float* accum_dev) {
    rd = r;
    scene = sc;
    npix = r.width * r.height;
    sq = (uint32_t)std::sqrt((double)r.spp);
    nsamples = sq * sq;
    s_begin = std::min(r.sample_begin, nsamples);
    s_end = r.sample_count ? std::min(nsamples, s_begin + r.sample_count) : nsamples;
    s_next = s_begin;
    p_next = 0;
    std::memset(&req, 0, sizeof(req));
    camera_setup(sc->host.camera, r.width, req);
    req.width = r.width; req.height = r.height; req.sqrt_spp = sq; req.spp = r.spp; req.max_depth = r.max_depth;
    req.rng_mode = r.rng_mode; req.scene = sc->slot; req.seed = r.seed;
    req.accum = accum_dev;
    req.rng_state = nullptr;
}

// Takes up to `cap` path slots of this job's remaining work (sample-major). MINSTD jobs take at
// most one sample per pass so each pixel's engine advances sample after sample.
bool Job::take(uint32_t cap, WorkItem& w) {
    if (s_next >= s_end || cap == 0) return false;
    w.job = this;
    if (p_next == 0 && cap >= npix) {
        uint32_t k = cap / npix;
        if (rd.rng_mode == CRT_RNG_MINSTD) k = 1;
        k = std::min(k, s_end - s_next);
        w.pix_begin = 0; w.pix_count = npix; w.sample_begin = s_next; w.sample_count = k;
        s_next += k;
    } else {
        uint32_t c = std::min(cap, npix - p_next);
        w.pix_begin = p_next; w.pix_count = c; w.sample_begin = s_next; w.sample_count = 1;
        p_next += c;
        if (p_next == npix) { p_next = 0; s_next += 1; }
    }
    return true;
}

// Runs one fused pass over the given work items (all with the same rng mode). Asynchronous on
// ctx->stream; the caller synchronizes before reading results.
int Context::run_pass(std::vector<WorkItem>& items, bool time_stages, double* extend_ms, const PassBufs& B) {
    if (items.empty()) return CRT_OK;
    int buf = pass_parity;
    pass_parity ^= 1;
    HIP_TRY(hipEventSynchronize(req_event[buf]));
    crt_dev_request* host = (crt_dev_request*)req_host[buf];
    uint32_t base = 0, max_pix = 0, max_depth = 1;
    int mode = (int)items[0].job->rd.rng_mode;
    for (size_t k = 0; k < items.size(); ++k) {
        WorkItem& w = items[k];
        crt_dev_request r = w.job->req;
        r.path_base = base;
        r.pix_begin = w.pix_begin; r.pix_count = w.pix_count;
        r.sample_begin = w.sample_begin; r.sample_count = w.sample_count;
        host[k] = r;
        base += w.pix_count * w.sample_count;
        max_pix = std::max(max_pix, w.pix_count);
        max_depth = std::max(max_depth, w.job->rd.max_depth);
        w.job->rays_in_flight += (uint64_t)w.pix_count * w.sample_count;
    }
    uint32_t npaths = base;
    int32_t uniform_scene = (int32_t)items[0].job->req.scene;
    int two_level = 0, full_shade = 0, any_shapes = 0;
    const uint32_t width = items[0].job->scene->host.bvh_width;   // the traversal kernel follows the scenes' BVH
    for (auto& w : items) {
        if ((int32_t)w.job->req.scene != uniform_scene) uniform_scene = -1;
        if (w.job->scene->host.num_instances) two_level = 1;
        if (w.job->scene->host.num_instances || w.job->scene->host.num_shapes) full_shade = 1;
        if (w.job->scene->host.num_shapes) any_shapes = 1;
        if (w.job->scene->host.bvh_width != width)
            return set_error(CRT_ERR_INVALID_ARGUMENT, "scenes of different BVH widths in one pass");
    }
    if (npaths > B.pass_ids) return set_error(CRT_ERR_INTERNAL, "pass exceeds the pass capacity");
    crt_dev_request* dreq = B.req[buf];
    HIP_TRY(hipMemcpyAsync(dreq, host, sizeof(crt_dev_request) * items.size(), hipMemcpyHostToDevice, stream));
    const crt_dev_scene* dsc = (const crt_dev_scene*)scene_table.p;
    int nreq = (int)items.size();
    // The pass runs as K independent sub-passes on K streams (contiguous path-id ranges, disjoint
    // slot ranges, own control blocks and overflow stacks), so the tail of one sub-pass's launch
    // overlaps full launches of the others. Timed-stage runs use one stream.
    const uint32_t nslots_all = std::min(B.slots, npaths);
    uint32_t K = time_stages ? 1u : std::min<uint32_t>(std::min(sub_passes, B.max_sub), std::max<uint32_t>(1u, nslots_all / 65536u));
    K = std::max(1u, K);
    struct Sub { uint32_t lo, hi, slot0, nslots, egrid, sgrid; hipStream_t s; uint64_t k; bool done; };
    Sub subs[kMaxSubPasses];
    uint32_t* cb = B.ctl;
    if (crt_k_pass_init(stream, cb, 2u * CRT_CTL_WORDS * K)) return set_error(CRT_ERR_HIP, "pass init launch failed");
    HIP_TRY(hipEventRecord(ev_pass_ready, stream));
    const uint32_t egrid_all = B.extend_grid / K;
    for (uint32_t q = 0; q < K; ++q) {
        Sub& u = subs[q];
        u.lo = (uint32_t)((uint64_t)npaths * q / K);
        u.hi = (uint32_t)((uint64_t)npaths * (q + 1) / K);
        u.slot0 = (uint32_t)((uint64_t)nslots_all * q / K);
        u.nslots = std::min<uint32_t>((uint32_t)((uint64_t)nslots_all * (q + 1) / K) - u.slot0, u.hi - u.lo);
        u.egrid = std::min(egrid_all, std::max(1u, (u.nslots + 255) / 256));
        u.sgrid = std::min(shade_grid, std::max(1u, (u.nslots + 255) / 256));
        u.s = q == 0 ? stream : sub_stream[q - 1];
        u.k = 0;
        u.done = u.lo == u.hi;
        if (q > 0) HIP_TRY(hipStreamWaitEvent(u.s, ev_pass_ready, 0));
    }
    uint32_t remaining = 0;
    for (uint32_t q = 0; q < K; ++q) remaining += subs[q].done ? 0u : 1u;
    while (remaining) {
        for (uint32_t q = 0; q < K; ++q) {
            Sub& u = subs[q];
            if (u.done) continue;
            const uint32_t total = u.hi - u.lo;
            uint32_t* ring = ctl_host + (size_t)q * kCtlRing * CRT_CTL_WORDS;
            hipEvent_t* evs = ctl_ev + (size_t)q * kCtlRing;
            if (u.k >= (uint64_t)kCtlLag) {
                // a sub-pass has drained once an iteration ends with every path generated and none alive
                uint64_t j = u.k - kCtlLag;
                HIP_TRY(hipEventSynchronize(evs[j % kCtlRing]));
                const uint32_t* h = ring + (j % kCtlRing) * CRT_CTL_WORDS;
                if (h[CRT_CTL_GEN] >= total && h[CRT_CTL_SURV] == 0) {
                    u.done = true;
                    remaining--;
                    if (q > 0) {
                        HIP_TRY(hipEventRecord(sub_done[q - 1], u.s));
                        HIP_TRY(hipStreamWaitEvent(stream, sub_done[q - 1], 0));
                    }
                    continue;
                }
            }
            // every path finishes within max_depth iterations of entering; a slot hosts at most
            // ceil(total / nslots) paths in turn
            const uint64_t max_iter = (uint64_t)max_depth * ((total + u.nslots - 1) / u.nslots + 1) + 2;
            if (u.k > max_iter) return set_error(CRT_ERR_INTERNAL, "pass did not drain");
            const uint64_t k = u.k;
            uint32_t* cin = cb + (q * 2 + (k & 1)) * CRT_CTL_WORDS;
            uint32_t* cout = cb + (q * 2 + ((k + 1) & 1)) * CRT_CTL_WORDS;
            crt_paths A = B.paths((int)(k & 1), u.slot0), Bn = B.paths((int)((k + 1) & 1), u.slot0);
            float4* hq = B.hit + u.slot0;
            uint32_t* hiq = B.hit_inst + u.slot0;
            uint32_t* gq = B.gstack + (size_t)q * egrid_all * 256u * crt_k_gstack_entries();
            if (crt_k_regen(u.s, mode, u.sgrid, dreq, nreq, u.lo, total, u.nslots, cin, cout, nullptr, A))
                return set_error(CRT_ERR_HIP, "regen launch failed");
            if (time_stages) HIP_TRY(hipEventRecord(ev_ext[0], u.s));
            if (crt_k_extend(u.s, config.count_work ? 1 : 0, (int)width, two_level ? 1 + any_shapes : 0, u.egrid, dsc, dreq, uniform_scene,
                             nullptr, cin + CRT_CTL_LIVE, 0, A.ray_o, A.ray_d, A.ids, hq, hiq, gq, nullptr, extend_mode,
                             cin + CRT_CTL_WORK, extend_param, (int)stack_lds, cin + CRT_CTL_SURV, B.counters))
                return set_error(CRT_ERR_HIP, "extend launch failed");
            if (time_stages) {
                HIP_TRY(hipEventRecord(ev_ext[1], u.s));
                HIP_TRY(hipEventSynchronize(ev_ext[1]));
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ev_ext[0], ev_ext[1]));
                if (extend_ms) *extend_ms += ms;
                extend_launches++;
            }
            if (crt_k_shade(u.s, mode, full_shade, u.sgrid, dsc, dreq, cin + CRT_CTL_LIVE, A, Bn, cout + CRT_CTL_SURV, hq, hiq,
                            B.out_rad, nullptr, (int)ray_sort, nreq == 1 ? 1 : 0, B.counters))
                return set_error(CRT_ERR_HIP, "shade launch failed");
            HIP_TRY(hipMemcpyAsync(ring + (k % kCtlRing) * CRT_CTL_WORDS, cout, sizeof(uint32_t) * CRT_CTL_WORDS,
                                   hipMemcpyDeviceToHost, u.s));
            HIP_TRY(hipEventRecord(evs[k % kCtlRing], u.s));
            u.k++;
            iterations++;
        }
    }
    if (crt_k_resolve(stream, dreq, nreq, max_pix, B.out_rad, nullptr))
        return set_error(CRT_ERR_HIP, "resolve launch failed");
    HIP_TRY(hipEventRecord(req_event[buf], stream));
    passes++;
    return CRT_OK;
}

// One megakernel launch over the whole request (every pixel, every remaining sample in order).
int Context::run_plain(Job& j, bool time_stages, double* extend_ms) {
    if (j.scheduled_all()) return CRT_OK;
    int buf = pass_parity;
    pass_parity ^= 1;
    HIP_TRY(hipEventSynchronize(req_event[buf]));
    crt_dev_request* host = (crt_dev_request*)req_host[buf];
    crt_dev_request r = j.req;
    r.path_base = 0;
    r.pix_begin = 0; r.pix_count = j.npix;
    r.sample_begin = j.s_next; r.sample_count = j.s_end - j.s_next;
    host[0] = r;
    j.s_next = j.s_end;
    crt_dev_request* dreq = (crt_dev_request*)req_table[buf].p;
    HIP_TRY(hipMemcpyAsync(dreq, host, sizeof(crt_dev_request), hipMemcpyHostToDevice, stream));
    // the baseline keeps its own launch shape (one 256-thread workgroup per 256 pixels, at most 32
    // per CU), independent of the tuned closest-hit grid
    const uint32_t grid = std::min((uint32_t)num_cus * 32u, std::max(1u, (j.npix + 255) / 256));
    int rc = ensure_gstack(grid);
    if (rc) return rc;
    if (time_stages) HIP_TRY(hipEventRecord(ev_ext[0], stream));
    if (crt_k_plain(stream, (int)j.rd.rng_mode, (int)j.scene->host.bvh_width, j.scene->host.num_instances ? (j.scene->host.num_shapes ? 2 : 1) : 0, grid,
                    (const crt_dev_scene*)scene_table.p, dreq, (uint32_t*)gstack.p, (crt_dev_counters*)counters.p))
        return set_error(CRT_ERR_HIP, "megakernel launch failed");
    if (time_stages) {
        HIP_TRY(hipEventRecord(ev_ext[1], stream));
        HIP_TRY(hipEventSynchronize(ev_ext[1]));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev_ext[0], ev_ext[1]));
        if (extend_ms) *extend_ms += ms;
        extend_launches++;
    }
    HIP_TRY(hipEventRecord(req_event[buf], stream));
    passes++;
    return CRT_OK;
}

int Context::finish_job(Job& j, crt_image* out) {
    if (crt_k_tonemap(stream, (const float4*)j.accum.p, j.npix, j.rd.spp, (uint8_t*)j.rgb.p))
        return set_error(CRT_ERR_HIP, "tonemap launch failed");
    if (out && out->rgb8) HIP_TRY(hipMemcpyAsync(out->rgb8, j.rgb.p, (size_t)j.npix * 3, hipMemcpyDeviceToHost, stream));
    if (out && out->accum) HIP_TRY(hipMemcpyAsync(out->accum, j.accum.p, (size_t)j.npix * 16, hipMemcpyDeviceToHost, stream));
    return CRT_OK;
}

int Context::reset_counters(crt_dev_counters* c) {
    if (crt_k_fill_u32(stream, (uint32_t*)c, sizeof(crt_dev_counters) / 4, 0u))
        return set_error(CRT_ERR_HIP, "fill launch failed");
    return CRT_OK;
}

int Context::read_counters(crt_dev_counters& c, const crt_dev_counters* dev) {
    HIP_TRY(hipMemcpyAsync(counters_host, dev, sizeof(crt_dev_counters), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    c = *(crt_dev_counters*)counters_host;
    return CRT_OK;
}

void fill_stats(crt_stats* st, const crt_dev_counters& c, uint64_t paths, double gpu_ms, double ext_ms, double host_ms) {
    if (!st) return;
    st->paths = paths;
    // every closest-hit query is a ray, pass-through re-traces included (the reference issues a
    // second world hit for them, core/src/cameras/camera.cuh:180-183)
    st->rays = c.rays + c.retraces;
    st->retraces = c.retraces;
    st->light_pdf_tests = c.light_tests;
    st->node_visits = c.node_visits;
    st->prim_tests = c.tri_tests + c.sphere_tests + c.quad_tests;
    st->gpu_ms = gpu_ms;
    st->extend_ms = ext_ms;
    st->host_ms = host_ms;
}

// Plain and batched rendering: all items are admitted at once, passes are packed greedily.
int Context::render_jobs(std::vector<Job*>& jobs, bool time_stages, crt_stats* stats, const PassBufs& B) {
    double h0 = now_s();
    int rc;
    if ((rc = reset_counters(B.counters))) return rc;
    HIP_TRY(hipEventRecord(ev_start, stream));
    double ext_ms = 0.0;
    uint64_t paths = 0;
    uint64_t launches0 = extend_launches, passes0 = passes, iter0 = iterations;
    for (Job* j : jobs) {
        if (!(j->rd.flags & CRT_RENDER_MEGAKERNEL)) continue;
        paths += (uint64_t)j->npix * (j->s_end - j->s_next);
        if ((rc = run_plain(*j, time_stages, &ext_ms))) return rc;
    }
    // Passes of one scene run the single-scene traversal kernel and keep one tree in the caches.
    // A request with at least batch_group path ids left leads a pass of its own scene (with every
    // other request on that scene); the smaller remainders share passes in request order.
    auto mine = [](const Job* j, int mode) {
        return (int)j->rd.rng_mode == mode && !(j->rd.flags & CRT_RENDER_MEGAKERNEL) && !j->scheduled_all();
    };
    for (int mode = 0; mode < 2; ++mode) {
        for (;;) {
            std::vector<WorkItem> items;
            uint32_t cap = B.pass_ids;
            const DeviceScene* lead = nullptr;
            for (Job* j : jobs)
                if (mine(j, mode) && j->paths_left() >= batch_group) { lead = j->scene; break; }
            for (Job* j : jobs) {
                if (!mine(j, mode)) continue;
                if (lead ? j->scene != lead : j->paths_left() >= batch_group) continue;
                // at most one work item per request and pass: a pixel's samples are resolved by one
                // thread in sample order, so two items of one request never race on a pixel
                if (cap > 0 && items.size() < kMaxPassRequests) {
                    WorkItem w;
                    if (j->take(cap, w)) {
                        cap -= w.pix_count * w.sample_count;
                        paths += (uint64_t)w.pix_count * w.sample_count;
                        items.push_back(w);
                    }
                }
            }
            if (items.empty()) break;
            if ((rc = run_pass(items, time_stages, &ext_ms, B))) return rc;
        }
    }
    HIP_TRY(hipEventRecord(ev_stop, stream));
    HIP_TRY(hipEventSynchronize(ev_stop));
    float gms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&gms, ev_start, ev_stop));
    crt_dev_counters c;
    if ((rc = read_counters(c, B.counters))) return rc;
    if (c.stack_overflow) return set_error(CRT_ERR_INTERNAL, "traversal stack overflow");
    fill_stats(stats, c, paths, gms, ext_ms, (now_s() - h0) * 1e3);
    last.rays = c.rays + c.retraces; last.retraces = c.retraces; last.light_pdf_tests = c.light_tests;
    last.node_visits = c.node_visits; last.leaf_visits = c.leaf_visits; last.tri_tests = c.tri_tests;
    last.sphere_tests = c.sphere_tests; last.quad_tests = c.quad_tests; last.inst_visits = c.inst_visits;
    last.extend_launches = (time_stages ? extend_launches - launches0 : 0);
    last.passes = passes - passes0;
    last.iterations = iterations - iter0;
    last.exact_queries = c.exact;
    return CRT_OK;
}

// True when `p` is device memory of `device` (hipMalloc'd or managed), as the no-allocation render
// needs for its caller-owned outputs.
static bool device_pointer_on(const void* p, int device) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
    if (a.isManaged) return true;
    return a.type == hipMemoryTypeDevice && a.device == device;
}

// crt_render on the device's scratch array: no device allocation and no allocating library call.
// The caller owns the outputs (device memory); the scene is resident. Keyed PCG streams only (the
// reference stream would need per-pixel engine state).
int Context::render_noalloc(crt_scene_id scene, const crt_render_desc* rd, float* accum_dev, uint8_t* rgb_dev,
                            crt_stats* stats) {
    std::lock_guard<std::mutex> ex(exec_mutex);
    int rc = Job::validate(*rd);
    if (rc) return rc;
    if (rd->rng_mode != CRT_RNG_PCG)
        return set_error(CRT_ERR_INVALID_ARGUMENT, "crt_render_noalloc renders keyed PCG streams only");
    if (rd->flags & CRT_RENDER_MEGAKERNEL)
        return set_error(CRT_ERR_INVALID_ARGUMENT, "crt_render_noalloc runs the wavefront passes only");
    if (!accum_dev) return set_error(CRT_ERR_INVALID_ARGUMENT, "accumulation buffer is null");
    if (!device_pointer_on(accum_dev, device) || (rgb_dev && !device_pointer_on(rgb_dev, device)))
        return set_error(CRT_ERR_INVALID_ARGUMENT, "outputs must be device memory of the context's device");
    DeviceScene* sc = nullptr;
    {
        std::lock_guard<std::mutex> g(gpu_mutex);
        sc = find_scene(scene);
    }
    if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    PassBufs B;
    if ((rc = scratch_bufs(B))) return rc;
    Job j;
    j.setup_external(sc, *rd, accum_dev);
    if (crt_k_init_request(stream, (float4*)accum_dev, nullptr, rd->width, rd->height, rd->seed))
        return set_error(CRT_ERR_HIP, "init kernel launch failed");
    std::vector<Job*> jobs{&j};
    if ((rc = render_jobs(jobs, (rd->flags & CRT_RENDER_TIME_STAGES) != 0, stats, B))) return rc;
    if (rgb_dev && crt_k_tonemap(stream, (const float4*)accum_dev, j.npix, rd->spp, rgb_dev))
        return set_error(CRT_ERR_HIP, "tonemap launch failed");
    HIP_TRY(hipStreamSynchronize(stream));
    return CRT_OK;
}

}  // namespace crt

using namespace crt;

extern "C" {

const char* crt_last_error(void) { return g_last_error.c_str(); }

int crt_render_noalloc(crt_context* c, crt_scene_id s, const crt_render_desc* rd, float* accum_dev, uint8_t* rgb_dev,
                       crt_stats* stats) {
    if (!c || !rd) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    return reinterpret_cast<Context*>(c)->render_noalloc(s, rd, accum_dev, rgb_dev, stats);
}

int crt_device_count(int32_t* count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return CRT_OK;
}

int crt_create(const crt_config* cfg, crt_context** out) {
    if (!out) return set_error(CRT_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    auto ctx = std::make_unique<Context>();
    int rc = ctx->init(cfg);
    if (rc) return rc;
    *out = reinterpret_cast<crt_context*>(ctx.release());
    return CRT_OK;
}

int crt_destroy(crt_context* c) {
    delete reinterpret_cast<Context*>(c);
    return CRT_OK;
}

int crt_scene_upload(crt_context* c, const crt_scene_desc* d, crt_scene_id* out) {
    if (!c || !d || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    return reinterpret_cast<Context*>(c)->upload_scene(d, out);
}

int crt_scene_release(crt_context* c, crt_scene_id s) {
    if (!c) return set_error(CRT_ERR_INVALID_ARGUMENT, "null context");
    return reinterpret_cast<Context*>(c)->release_scene(s);
}

int crt_scene_get_info(crt_context* c, crt_scene_id s, crt_scene_info* info) {
    if (!c || !info) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::lock_guard<std::mutex> g(ctx->gpu_mutex);
    DeviceScene* sc = ctx->find_scene(s);
    if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    info->num_nodes = sc->host.stats.nodes;
    info->num_leaves = sc->host.stats.leaves;
    info->num_prims = sc->host.stats.prims;
    info->device_bytes = sc->bytes;
    info->max_depth = sc->host.stats.depth;
    info->num_instances = sc->host.num_instances;
    info->bvh_width = sc->host.bvh_width;
    info->node_bytes = sc->host.bvh_width == 4 ? 128u : 64u;
    info->leaf_size = sc->host.leaf_size;
    info->builder = sc->host.device_built ? (uint32_t)CRT_BVH_DEVICE : (uint32_t)CRT_BVH_HOST;
    info->build_ms = sc->host.stats.build_ms;
    info->upload_ms = sc->upload_ms;
    info->sah_cost = sc->host.stats.sah_cost;
    return CRT_OK;
}

int crt_render(crt_context* c, crt_scene_id s, const crt_render_desc* rd, crt_image* out, crt_stats* stats) {
    if (!c || !rd) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    crt_batch_item item;
    item.scene = s; item.rd = *rd;
    if (out) item.out = *out; else { item.out.rgb8 = nullptr; item.out.accum = nullptr; }
    return crt_render_batch(c, &item, 1, stats);
}

int crt_render_batch(crt_context* c, crt_batch_item* items, uint32_t n, crt_stats* stats) {
    if (!c || (!items && n)) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    // the whole render holds exec_mutex (the context stream and pass buffers); the scene table and
    // the budget are read under gpu_mutex, so the serving loop keeps admitting meanwhile. Scenes
    // stay resident: only release_scene frees one, and it takes exec_mutex.
    std::lock_guard<std::mutex> ex(ctx->exec_mutex);
    std::vector<DeviceScene*> scs(n);
    {
        std::lock_guard<std::mutex> g(ctx->gpu_mutex);
        uint64_t need = 0;
        for (uint32_t k = 0; k < n; ++k) {
            if (!(scs[k] = ctx->find_scene(items[k].scene))) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
            int rc = Job::validate(items[k].rd);
            if (rc) return rc;
            need += Job::bytes_needed(items[k].rd);
        }
        if (ctx->scene_bytes + ctx->request_bytes + ctx->reserved + need > ctx->budget)
            return set_error(CRT_ERR_OVER_BUDGET, "requests exceed the HBM budget");
    }
    std::vector<std::unique_ptr<Job>> jobs;
    std::vector<Job*> jp;
    for (uint32_t k = 0; k < n; ++k) {
        auto j = std::make_unique<Job>();
        int rc = j->setup(ctx, scs[k], items[k].rd);
        if (rc) return rc;
        jp.push_back(j.get());
        jobs.push_back(std::move(j));
    }
    bool time_stages = (n > 0) && (items[0].rd.flags & CRT_RENDER_TIME_STAGES);
    int rc = ctx->render_jobs(jp, time_stages, stats, ctx->pool_bufs());
    if (rc) return rc;
    for (uint32_t k = 0; k < n; ++k) {
        if ((rc = ctx->finish_job(*jobs[k], &items[k].out))) return rc;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return set_error(CRT_ERR_HIP, "stream synchronize failed");
    return CRT_OK;
}

int crt_trace_rays(crt_context* c, crt_scene_id s, const float* rays, uint64_t n, int32_t flags, int32_t* prim,
                   int32_t* inst, float* t, float* uv) {
    if (!c || (!rays && n)) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    if (n > (1ull << 30)) return set_error(CRT_ERR_INVALID_ARGUMENT, "too many rays");
    Context* ctx = reinterpret_cast<Context*>(c);
    ExecLock g(*ctx);
    DeviceScene* sc = ctx->find_scene(s);
    if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    if (n == 0) return CRT_OK;
    DevBuf dr, dh, di;
    int rc;
    const bool camera = (flags & CRT_TRACE_CAMERA) != 0;
    if ((rc = dh.alloc(n * 16, ctx))) return rc;
    if ((rc = di.alloc(n * 4, ctx))) return rc;
    if ((rc = ctx->reset_counters((crt_dev_counters*)ctx->counters.p))) return rc;
    uint32_t grid = std::min(ctx->extend_grid, (uint32_t)std::max<uint64_t>(1, (n + 255) / 256));
    const int two_level = sc->host.num_instances ? (sc->host.num_shapes ? 2 : 1) : 0;
    if (flags & CRT_TRACE_STAGE) {
        // the render's closest-hit stage: path-state rays (origin, time 0; direction, depth 1 and the
        // camera bit), one uniform scene, the stage's control words
        DevBuf dro, drd, dids, dctl;
        if ((rc = dro.alloc(n * 16, ctx)) || (rc = drd.alloc(n * 16, ctx)) || (rc = dids.alloc(n * 8, ctx)) ||
            (rc = dctl.alloc(CRT_CTL_WORDS * 4, ctx)))
            return rc;
        std::vector<float> po(n * 4), pd(n * 4);
        const uint32_t state = 1u | (camera ? 0x10000u : 0u);   // depth 1 | camera bit of the path state
        for (uint64_t k = 0; k < n; ++k) {
            for (int a = 0; a < 3; ++a) { po[4 * k + a] = rays[8 * k + a]; pd[4 * k + a] = rays[8 * k + 4 + a]; }
            po[4 * k + 3] = 0.0f;
            std::memcpy(&pd[4 * k + 3], &state, 4);
        }
        HIP_TRY(hipMemcpyAsync(dro.p, po.data(), n * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(drd.p, pd.data(), n * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemsetAsync(dids.p, 0, n * 8, ctx->stream));
        HIP_TRY(hipMemsetAsync(dctl.p, 0, CRT_CTL_WORDS * 4, ctx->stream));
        uint32_t* ctl = (uint32_t*)dctl.p;
        if (crt_k_extend(ctx->stream, ctx->config.count_work ? 1 : 0, (int)sc->host.bvh_width, two_level, grid,
                         (const crt_dev_scene*)ctx->scene_table.p, nullptr, (int32_t)sc->slot, nullptr, nullptr,
                         (uint32_t)n, (const float4*)dro.p, (const float4*)drd.p, (const uint2*)dids.p, (float4*)dh.p,
                         (uint32_t*)di.p, (uint32_t*)ctx->gstack.p, nullptr, ctx->extend_mode, ctl + CRT_CTL_WORK,
                         ctx->extend_param, (int)ctx->stack_lds, nullptr, (crt_dev_counters*)ctx->counters.p))
            return set_error(CRT_ERR_HIP, "extend launch failed");
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    } else {
        if ((rc = dr.alloc(n * 32, ctx))) return rc;
        HIP_TRY(hipMemcpyAsync(dr.p, rays, n * 32, hipMemcpyHostToDevice, ctx->stream));
        if (crt_k_trace_rays(ctx->stream, (int)sc->host.bvh_width, two_level, grid,
                             (const crt_dev_scene*)ctx->scene_table.p + sc->slot, (const float4*)dr.p,
                             (uint32_t)n, camera ? 1 : 0, (float4*)dh.p, (uint32_t*)di.p, (uint32_t*)ctx->gstack.p,
                             (crt_dev_counters*)ctx->counters.p))
            return set_error(CRT_ERR_HIP, "trace launch failed");
    }
    std::vector<float> hh(n * 4);
    std::vector<uint32_t> hi(n);
    HIP_TRY(hipMemcpyAsync(hh.data(), dh.p, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(hi.data(), di.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    crt_dev_counters cnt;
    if ((rc = ctx->read_counters(cnt, (const crt_dev_counters*)ctx->counters.p))) return rc;
    if (cnt.stack_overflow) return set_error(CRT_ERR_INTERNAL, "traversal stack overflow");
    // map leaf-ordered hits back to global ids: copies of the id words are kept on the host
    const HostScene& H = sc->host;
    for (uint64_t k = 0; k < n; ++k) {
        uint32_t p;
        std::memcpy(&p, &hh[4 * k + 1], 4);
        int32_t gid = -1;
        if (p != CRT_MISS) {
            uint32_t type = CRT_HIT_TYPE(p), idx = CRT_HIT_INDEX(p);
            if (type == CRT_PRIM_TRI) gid = (int32_t)sc->tri_gid[idx];
            else if (type == CRT_PRIM_SPHERE) gid = (int32_t)H.sphere_aux[idx].x;
            else gid = (int32_t)sc->quad_gid[idx];
        }
        if (prim) prim[k] = gid;
        const uint32_t iw = hi[k] & CRT_INST_MASK;
        if (inst) inst[k] = (p == CRT_MISS || iw == CRT_NO_INST) ? -1 : (int32_t)sc->inst_gid[iw];
        if (t) t[k] = p == CRT_MISS ? 0.0f : hh[4 * k];
        if (uv) { uv[2 * k] = p == CRT_MISS ? 0.0f : hh[4 * k + 2]; uv[2 * k + 1] = p == CRT_MISS ? 0.0f : hh[4 * k + 3]; }
    }
    return CRT_OK;
}

int crt_trace_primary(crt_context* c, crt_scene_id s, uint32_t W, uint32_t H, int32_t* prim, float* t) {
    if (!c) return set_error(CRT_ERR_INVALID_ARGUMENT, "null context");
    Context* ctx = reinterpret_cast<Context*>(c);
    crt_dev_request r;
    {
        std::lock_guard<std::mutex> g(ctx->gpu_mutex);
        DeviceScene* sc = ctx->find_scene(s);
        if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
        camera_setup(sc->host.camera, W, r);
    }
    std::vector<float> rays((size_t)W * H * 8);
    for (uint32_t jj = 0; jj < H; ++jj) {
        uint32_t j = H - 1 - jj;
        for (uint32_t i = 0; i < W; ++i) {
            float u = ((float)i + 0.5f) / (float)W, v = ((float)j + 0.5f) / (float)H;
            float* q = &rays[8 * ((size_t)jj * W + i)];
            for (int k = 0; k < 3; ++k) {
                if (r.ortho) {
                    q[k] = r.llc[k] + u * r.horizontal[k] + v * r.vertical[k];
                    q[4 + k] = r.ortho_dir[k];
                } else {
                    q[k] = r.origin[k];
                    q[4 + k] = r.llc[k] + u * r.horizontal[k] + v * r.vertical[k] - r.origin[k];
                }
            }
            q[3] = 0.00001f; q[7] = 3.402823466e+38f;
        }
    }
    std::vector<float> uv((size_t)W * H * 2);
    return crt_trace_rays(c, s, rays.data(), (uint64_t)W * H, 1, prim, nullptr, t, uv.data());
}

int crt_scene_verify(crt_context* c, crt_scene_id s) {
    if (!c) return set_error(CRT_ERR_INVALID_ARGUMENT, "null context");
    Context* ctx = reinterpret_cast<Context*>(c);
    ExecLock g(*ctx);
    DeviceScene* sc = ctx->find_scene(s);
    if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    std::string err;
    int rc = ctx->verify_scene(*sc, err);
    return rc && !err.empty() ? set_error(rc, err) : rc;
}

int crt_scene_build_check(const crt_scene_desc* d, uint32_t leaf_size, uint32_t bvh_width, crt_scene_info* info) {
    HostScene H;
    std::string err;
    unsigned hw = std::thread::hardware_concurrency();
    int rc = build_host_scene(d, leaf_size ? leaf_size : kDefaultLeafSize, bvh_width ? bvh_width : kDefaultBvhWidth, (int)std::min(16u, hw ? hw : 1u),
                              H, err);
    if (rc) return set_error(rc, err);
    rc = check_host_scene(H, err);
    if (rc) return set_error(rc, err);
    if (info) {
        info->num_nodes = H.stats.nodes; info->num_leaves = H.stats.leaves; info->num_prims = H.stats.prims;
        info->device_bytes = H.device_bytes(); info->max_depth = H.stats.depth; info->num_instances = H.num_instances;
        info->bvh_width = H.bvh_width; info->node_bytes = H.bvh_width == 4 ? 128u : 64u; info->leaf_size = H.leaf_size;
        info->builder = CRT_BVH_HOST;
        info->build_ms = H.stats.build_ms; info->upload_ms = 0.0; info->sah_cost = H.stats.sah_cost;
    }
    return CRT_OK;
}

int crt_last_counters(crt_context* c, crt_work_counters* out) {
    if (!c || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::lock_guard<std::mutex> g(ctx->exec_mutex);
    *out = ctx->last;
    return CRT_OK;
}

int crt_memory_get_info(crt_context* c, crt_memory_info* out) {
    if (!c || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    Context* ctx = reinterpret_cast<Context*>(c);
    std::lock_guard<std::mutex> g(ctx->gpu_mutex);
    std::lock_guard<std::mutex> lk(ctx->srv_mutex);
    out->budget = ctx->budget; out->reserved = ctx->reserved; out->scenes = ctx->scene_bytes;
    out->requests = ctx->request_bytes;
    return CRT_OK;
}

int crt_debug_fail_alloc(crt_context* c, int32_t n) {
    if (!c) return set_error(CRT_ERR_INVALID_ARGUMENT, "null context");
    reinterpret_cast<Context*>(c)->fail_alloc_after.store(n < 0 ? -1 : n);
    return CRT_OK;
}

int crt_debug_valu_probe(crt_context* c, int32_t kind, uint32_t waves_per_simd, uint32_t iters, double* ms,
                         uint64_t* wave_insts) {
    if (!c || !ms || !wave_insts) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    if (kind < 0 || kind > 15 || waves_per_simd < 1 || waves_per_simd > 8 || iters < 1 || iters > (1u << 24))
        return set_error(CRT_ERR_INVALID_ARGUMENT, "probe kind 0-15, 1-8 waves per SIMD, 1-2^24 rounds");
    Context* ctx = reinterpret_cast<Context*>(c);
    ExecLock g(*ctx);
    // four SIMDs per CU, four waves per 256-lane workgroup: waves_per_simd workgroups per CU
    const uint32_t blocks = ctx->num_cus * waves_per_simd;
    DevBuf out;
    int rc;
    if ((rc = out.alloc((size_t)blocks * 256 * sizeof(float), ctx))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) { (void)hipEventDestroy(e0); HIP_TRY(e); }
    rc = crt_k_valu_probe(ctx->stream, kind, blocks, iters, (float*)out.p);
    if (!rc) { (void)hipEventRecord(e0, ctx->stream); rc = crt_k_valu_probe(ctx->stream, kind, blocks, iters, (float*)out.p); }
    (void)hipEventRecord(e1, ctx->stream);
    float t = 0.0f;
    hipError_t se = hipEventSynchronize(e1);
    if (!rc && se == hipSuccess) se = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return set_error(CRT_ERR_HIP, "probe launch failed");
    HIP_TRY(se);
    *ms = t;
    *wave_insts = (uint64_t)blocks * 4u * iters * 16u;
    return CRT_OK;
}

int crt_struct_sizes(uint64_t* sizes, uint32_t n) {
    uint64_t v[] = {sizeof(crt_camera), sizeof(crt_material), sizeof(crt_texture), sizeof(crt_uvmap), sizeof(crt_sphere),
                    sizeof(crt_quad), sizeof(crt_area_light), sizeof(crt_mesh), sizeof(crt_instance), sizeof(crt_scene_desc),
                    sizeof(crt_render_desc), sizeof(crt_image), sizeof(crt_stats), sizeof(crt_config), sizeof(crt_scene_info),
                    sizeof(crt_batch_item), sizeof(crt_request_timing), sizeof(crt_work_counters), sizeof(crt_memory_info), sizeof(crt_obj_mesh),
                    sizeof(crt_scenefile_info), sizeof(crt_tune_result), sizeof(crt_sphere_light), sizeof(crt_transform)};
    uint32_t m = (uint32_t)(sizeof(v) / sizeof(v[0]));
    for (uint32_t k = 0; k < n && k < m; ++k) sizes[k] = v[k];
    return (int)m;
}

}  // extern "C"
