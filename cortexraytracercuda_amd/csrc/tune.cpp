// tune.cpp - in-library autotuner for the launch and memory shape of the streaming passes.
//
// The reference fixes its launch (16x16-thread blocks over the image, core/src/gpu.cu:2072-2080) and
// has nothing to tune. Here the throughput of a render depends on how many paths are in flight per
// launch (live path slots, sub-passes), on the launch shape of the closest-hit and shade stages, on
// the closest-hit work distribution and traversal stack depth, and on the BVH (width, leaf size).
// The best values differ between scenes and image sizes, so crt_autotune measures: it times trial
// renders of the caller's request (full image, a sample range sized to fill the largest slot pool
// twice) for each candidate of one tunable at a time, keeping the others at their best so far.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "crt_kernels.h"
#include "crt_runtime.h"

namespace crt {

namespace {

struct Shape {
    uint32_t slots, grid, sgrid, sub, mode, param, lds;
};

Shape current_shape(const Context& c) {
    return Shape{c.max_paths, c.grid_per_cu, c.shade_grid_per_cu, c.sub_passes, (uint32_t)c.extend_mode, c.extend_param,
                 c.stack_lds};
}

}  // namespace

// One timed trial render of `rd` on `scene` with the context's current shape (caller holds
// both context locks). Throughput = closest-hit queries / GPU time of the passes.
int Context::tune_trial(crt_scene_id scene, const crt_render_desc& rd, double& mrays_s) {
    DeviceScene* sc = find_scene(scene);
    if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    Job j;
    int rc = j.setup(this, sc, rd);
    if (rc) return rc;
    std::vector<Job*> jobs{&j};
    crt_stats st{};
    if ((rc = render_jobs(jobs, false, &st, pool_bufs()))) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    mrays_s = st.gpu_ms > 0.0 ? (double)st.rays / (st.gpu_ms * 1e-3) / 1e6 : 0.0;
    return CRT_OK;
}

int Context::autotune(crt_scene_id scene, const crt_scene_desc* desc, const crt_render_desc* rd, crt_tune_result* out) {
    ExecLock g(*this);
    int rc = Job::validate(*rd);
    if (rc) return rc;
    DeviceScene* sc = find_scene(scene);
    if (!sc) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown scene id");
    if (sc->refs > 0) return set_error(CRT_ERR_INVALID_ARGUMENT, "scene has requests in flight");
    // the serving loop's captured iterations bake in the settings tuned here: tune only while it
    // is idle, and drop its graphs afterwards (they are captured again from the new settings)
    if (!server_idle()) return set_error(CRT_ERR_INVALID_ARGUMENT, "serving requests in flight");
    struct DropGraphs {
        Context* c;
        ~DropGraphs() { c->server_drop_graphs(); }
    } drop_graphs{this};
    if (rd->flags & CRT_RENDER_MEGAKERNEL) return set_error(CRT_ERR_INVALID_ARGUMENT, "the megakernel render has no tunables");
    std::memset(out, 0, sizeof(*out));

    // trial request: the full image, enough samples to fill the largest slot pool twice
    const uint64_t max_slots = 1ull << 26;
    crt_render_desc tr = *rd;
    const uint32_t sq = (uint32_t)std::sqrt((double)rd->spp), ns = sq * sq;
    const uint64_t npix = (uint64_t)rd->width * rd->height;
    uint32_t want = (uint32_t)std::min<uint64_t>(ns, std::max<uint64_t>(1, (2 * max_slots + npix - 1) / npix));
    if (rd->rng_mode == CRT_RNG_MINSTD) want = ns;   // the reference stream renders all samples in one call
    tr.sample_begin = 0;
    tr.sample_count = want;
    tr.flags = 0;
    if (Job::bytes_needed(tr) + scene_bytes + request_bytes + reserved > budget)
        return set_error(CRT_ERR_OVER_BUDGET, "trial render exceeds the HBM budget");

    double best = 0.0;
    uint32_t trials = 0;
    auto trial = [&](double& v) -> int {
        int r = tune_trial(scene, tr, v);
        trials++;
        return r;
    };
    double warm;
    if ((rc = trial(warm))) return rc;   // first use of the kernels and buffers
    if ((rc = trial(best))) return rc;
    out->initial_mrays_s = best;

    // BVH shape: rebuild the scene at each width / leaf size, keep the fastest in place
    if (desc) {
        bool other_width = false;
        for (auto& kv : scenes) other_width = other_width || (kv.first != scene);
        const uint32_t shapes[4][2] = {{4, 1}, {4, 2}, {4, 4}, {2, 2}};
        crt_scene_id best_id = scene;
        uint32_t best_w = sc->host.bvh_width, best_l = leaf_size;
        for (auto& sh : shapes) {
            // the traversal kernel of a pass follows its scenes' width: with other scenes resident
            // only the context's width is tried
            if (other_width && sh[0] != bvh_width) continue;
            if (sh[0] == sc->host.bvh_width && sh[1] == leaf_size) continue;
            crt_scene_id tmp;
            if ((rc = upload_scene_locked(desc, sh[1], sh[0], &tmp))) return rc;
            double v;
            if ((rc = trial_scene(tmp, tr, v, trials))) { release_scene_locked(tmp); return rc; }
            if (v > best) {
                if (best_id != scene) release_scene_locked(best_id);
                best = v; best_id = tmp; best_w = sh[0]; best_l = sh[1];
            } else {
                release_scene_locked(tmp);
            }
        }
        if (best_id != scene) {
            // the winner takes the caller's scene id; the old BVH is released
            std::swap(scenes[scene], scenes[best_id]);
            if ((rc = release_scene_locked(best_id))) return rc;
            bvh_width = best_w;
            leaf_size = best_l;
        }
        sc = find_scene(scene);
    }

    // launch and memory shape, one tunable at a time
    Shape bs = current_shape(*this);
    auto apply = [&](const Shape& s) -> int {
        if (s.slots > path_capacity) {
            const uint64_t extra = (uint64_t)(s.slots - path_capacity) * 164ull;
            if (reserved + scene_bytes + request_bytes + Job::bytes_needed(tr) + extra > budget)
                return set_error(CRT_ERR_OVER_BUDGET, "slot pool exceeds the HBM budget");
            HIP_TRY(hipStreamSynchronize(stream));
            int r = ensure_paths(s.slots, pass_capacity);
            if (r) return r;
        }
        max_paths = s.slots;
        grid_per_cu = s.grid;
        shade_grid_per_cu = s.sgrid;
        extend_grid = (uint32_t)num_cus * s.grid;
        shade_grid = (uint32_t)num_cus * s.sgrid;
        int r = ensure_gstack(extend_grid);
        if (r) return r;
        sub_passes = s.sub;
        extend_mode = (int)s.mode;
        extend_param = s.param;
        stack_lds = s.lds;
        reserved = used_bytes() + server_bytes();
        return CRT_OK;
    };
    auto sweep = [&](auto field_set, std::initializer_list<uint32_t> values) -> int {
        Shape base = bs;
        for (uint32_t v : values) {
            Shape s = base;
            field_set(s, v);
            if (std::memcmp(&s, &base, sizeof(s)) == 0) continue;
            int r = apply(s);
            if (r == CRT_ERR_OVER_BUDGET) { set_error(CRT_OK, ""); continue; }
            if (r) return r;
            double m;
            if ((r = trial(m))) return r;
            if (m > best) { best = m; bs = s; }
        }
        return apply(bs);
    };
    if ((rc = sweep([](Shape& s, uint32_t v) { s.slots = v; }, {1u << 23, 1u << 24, 1u << 25, 1u << 26}))) return rc;
    if ((rc = sweep([](Shape& s, uint32_t v) { s.grid = v; }, {12u, 24u, 32u}))) return rc;
    if ((rc = sweep([](Shape& s, uint32_t v) { s.sgrid = v; }, {4u, 8u, 16u}))) return rc;
    if ((rc = sweep([](Shape& s, uint32_t v) { s.sub = v; }, {1u, 2u, 3u, 4u}))) return rc;
    // closest-hit work distribution: grid-stride (coherent waves) or per-lane refill from the
    // wave's share of the queue after 16 / 32 idle lanes (fewer idle lanes on long, uneven rays)
    if ((rc = sweep(
             [](Shape& s, uint32_t v) {
                 s.mode = v == 0 ? (uint32_t)CRT_EXTEND_GRID_STRIDE : (uint32_t)CRT_EXTEND_PERSISTENT;
                 s.param = v == 0 ? 2u : v;
             },
             {0u, 16u, 32u})))
        return rc;
    if ((rc = sweep([](Shape& s, uint32_t v) { s.lds = v; }, {8u, 16u, 32u}))) return rc;

    out->max_paths = max_paths;
    out->grid_per_cu = grid_per_cu;
    out->shade_grid_per_cu = shade_grid_per_cu;
    out->sub_passes = sub_passes;
    out->extend_kernel = (uint32_t)extend_mode;
    out->refill_min = extend_param;
    out->stack_lds = stack_lds;
    out->bvh_width = sc->host.bvh_width;
    out->leaf_size = leaf_size;
    out->trials = trials;
    out->best_mrays_s = best;
    return CRT_OK;
}

// Trial on a temporary scene (two renders: the first warms its BVH into the caches as the
// timed renders of the other shapes were).
int Context::trial_scene(crt_scene_id scene, const crt_render_desc& rd, double& mrays_s, uint32_t& trials) {
    double warm;
    int rc = tune_trial(scene, rd, warm);
    if (rc) return rc;
    rc = tune_trial(scene, rd, mrays_s);
    trials += 2;
    return rc;
}

}  // namespace crt

using namespace crt;

extern "C" int crt_autotune(crt_context* c, crt_scene_id scene, const crt_scene_desc* desc, const crt_render_desc* rd,
                            crt_tune_result* out) {
    if (!c || !rd || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    return reinterpret_cast<Context*>(c)->autotune(scene, desc, rd, out);
}
