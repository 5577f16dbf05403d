"""Benchmark of the render hot path on one MI355X.

Default workload: the single large scene (icosphere subdivided to 1,310,720 triangles displaced by
3-octave value noise, plus 500 random spheres), 1920x1080, 64 spp, max depth 10, FP32, keyed PCG
streams, seed 1. One step = one complete render of that image (every sample of every pixel).

Printed JSON line (rank 0):
  value        Mrays/s over the timed steps (closest-hit queries / wall time of the renders)
  roofline     traversal kernel (k_extend) against its binding resource, VALU issue: VALU wave
               instructions per ray (rocprofv3 SQ counters, profiles/) x rays per launch / HIP-event
               launch time, against 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction (dense
               FMA-class issue; the measured probe rates of profiles/valu_issue_probe_mi355x.json
               beside it); with the measured DRAM traffic (FETCH_SIZE / WRITE_SIZE) and the
               algorithmic bytes
  cpu_baseline reference-faithful CPU renderer (oracle/, FP32 minstd) on the host cores, 4 spp,
               plus a one-thread figure and the host CPU model
  megakernel   the plain path on the same workload: one kernel launch renders the image
  batched      the batched-launch workload (64 scenes of 10k-2M triangles, 256^2-1024^2, 1-64 spp):
               one crt_render_batch vs one megakernel launch per request
  serving      the serving workload (Poisson arrivals, seeded scenes of 10k-2M triangles uploaded
               and built on first use, 256^2-2048^2, 1-256 spp, HBM budget): continuous batching vs
               the plain one-launch-per-request server, p50/p99 latency and aggregate Mrays/s of each
  instanced    the instanced workload (4,096 instances, 3840x2160, 16 progressive 1-spp passes) with
               its own roofline
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue capacity of the MI355X: 256 CUs x 4 SIMDs, each issuing one wave64 VALU instruction per
# 4 cycles (16 lanes) at the 2.4 GHz peak engine clock
VALU_PEAK_GINST = 1024 * 2.4 / 4.0
# Algorithmic bytes of one closest-hit query in k_extend (declared layout, crt_layout.h):
#   ray 32 (ray_o + ray_d) + ids 8 + queue index 4 + hit record 16 + hit instance 4 = 64 B of state
#   + 64 B per BVH2 / 128 B per BVH4 interior node visit (child boxes + refs; a leaf's type, count
#   and first primitive are packed into its reference in the parent, so a leaf visit reads nothing),
#   + 48 B per triangle, 16 B per sphere, 80 B per quad/light, 64 B per instance record.
B_STATE = 64
B_LEAF, B_TRI, B_SPHERE, B_QUAD, B_INST = 0, 48, 16, 80, 64   # node bytes come from the scene (64 or 128)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def instanced_workload():
    from cortexraytracercuda_amd import scenes
    scene = scenes.instanced_scene(seed=1, n_instances=4096, n_textures=4, tex_size=2048, nu=50, nv=50)
    cfg = dict(width=3840, height=2160, spp=16, max_depth=12)
    workload = ("instanced scene: 4,096 instances of a 5,000-triangle displaced torus (20,480,000 triangles), "
                "four 2048^2 RGBA8 textures, 3840x2160, 16 progressive 1-spp passes, depth 12")
    return scene, cfg, workload


def large_workload():
    from cortexraytracercuda_amd import scenes
    return scenes.large_scene(seed=1, levels=8, n_spheres=500), dict(width=1920, height=1080, spp=64, max_depth=10)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, cfg, spp=4, rows_1t=48):
    """The oracle's reference-faithful renderer on the host cores, on a bounded sample of the
    workload: every pixel at `spp` samples on all threads (about 25 core-seconds on the bench host at
    spp=4), then `rows_1t` rows from the middle of the image on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    acc, st = O.render(scene, cfg["width"], cfg["height"], spp, cfg["max_depth"], seed=2580, rng="minstd", precision=32,
                       threads=threads, ref_style=True)
    r0 = max(0, cfg["height"] // 2 - rows_1t // 2)
    _, st1 = O.render(scene, cfg["width"], cfg["height"], spp, cfg["max_depth"], seed=2580, rng="minstd", precision=32,
                      threads=1, ref_style=True, rows=(r0, min(cfg["height"], r0 + rows_1t)))
    return {"value": st["rays"] / st["seconds"] / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": "%dx%d at %d spp (all pixels), depth %d, reference-faithful FP32 minstd renderer with the "
                      "reference's median-split BVH; %d rays in %.1f s on %d threads (BVH build excluded)" %
                      (cfg["width"], cfg["height"], spp, cfg["max_depth"], st["rays"], st["seconds"], threads),
            "single_thread": {"value": st1["rays"] / st1["seconds"] / 1e6, "unit": "Mrays/s", "cores": 1,
                              "sample": "rows %d-%d of the same image at %d spp: %d rays in %.1f s" %
                                        (r0, min(cfg["height"], r0 + rows_1t) - 1, spp, st1["rays"], st1["seconds"])}}


def _plain_and_batched(r, jobs):
    """Times the same requests three ways: the plain path (one megakernel launch per request), the
    wavefront pipeline one request at a time, and one fused crt_render_batch over all of them. The
    timed calls return each request's image (RGB8), as a renderer does; the accumulation buffers
    used for the equality checks come from untimed repeats of the same calls."""
    j0 = jobs[0]
    r.render(j0["sid"], 64, 64, 1, 8, megakernel=True)
    r.render(j0["sid"], 64, 64, 1, 8)
    r.render_batch([dict(j0, width=64, height=64)])
    out = {}
    modes = (("plain", dict(megakernel=True)), ("wavefront_per_request", {}))
    for name, kw in modes:
        t0 = time.perf_counter()
        rays = 0
        for j in jobs:
            _, _, st = r.render(j["sid"], j["width"], j["height"], j["spp"], j["max_depth"], seed=j["seed"],
                                want_accum=False, **kw)
            rays += st["rays"]
        dt = time.perf_counter() - t0
        out[name + "_mrays_s"] = rays / dt / 1e6
        out[name + "_s"] = dt
    t0 = time.perf_counter()
    _, st = r.render_batch(jobs, want_accum=False)
    dt = time.perf_counter() - t0
    out["batched_mrays_s"] = st["rays"] / dt / 1e6
    out["batched_s"] = dt
    out["batched_gpu_ms"] = st["gpu_ms"]
    out["speedup"] = out["batched_mrays_s"] / out["plain_mrays_s"]
    out["speedup_vs_wavefront_per_request"] = out["batched_mrays_s"] / out["wavefront_per_request_mrays_s"]
    # batched accumulations equal the per-request wavefront ones bit for bit; the megakernel runs the
    # same estimator on the same streams (pixels agreeing to 1e-4 relative)
    accs = {name: [r.render(j["sid"], j["width"], j["height"], j["spp"], j["max_depth"], seed=j["seed"], **kw)[1]
                   for j in jobs] for name, kw in modes}
    outs, _ = r.render_batch(jobs)
    out["batched_identical_to_alone"] = bool(all(np.array_equal(a, b[1]) for a, b in zip(accs["wavefront_per_request"], outs)))
    agree = [np.all(np.abs(a - b[1]) <= 1e-4 * np.maximum(np.abs(b[1]), 1.0), axis=-1).mean() for a, b in zip(accs["plain"], outs)]
    out["plain_vs_batched_pixels_agree_min"] = float(min(agree))
    return out


def batched_vs_plain(seed=7, n=24):
    """Many small requests on resident scenes (10k-200k triangles, 256^2-512^2, 1-16 spp)."""
    from cortexraytracercuda_amd import Renderer, scenes
    rng = np.random.default_rng(seed)
    r = Renderer()
    jobs = []
    pool = {}
    for k in range(n):
        tris = int(math.exp(rng.uniform(math.log(1e4), math.log(2e5))))
        key = round(math.log(tris), 1)
        if key not in pool:
            pool[key] = r.upload(scenes.sphere_mesh_scene(tris, seed=k))
        res = int(rng.choice([256, 384, 512]))
        spp = int(rng.choice([1, 4, 9, 16]))
        jobs.append(dict(sid=pool[key], width=res, height=res, spp=spp, max_depth=8, seed=1000 + k))
    out = dict(requests=n, **_plain_and_batched(r, jobs))
    r.close()
    return out


def _poisson_requests(rng, pool, n, rate_hz, res_choices, spp_choices, max_depth=8, seed0=5000):
    reqs = []
    for k in range(n):
        res = int(rng.choice(res_choices))
        reqs.append(dict(sid=pool[int(rng.integers(0, len(pool)))], width=res, height=res,
                         spp=int(rng.choice(spp_choices)), max_depth=max_depth, seed=seed0 + k))
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    return reqs, arrivals


def _serve_batched(r, reqs, arrivals):
    """Continuous batching (crt_submit): requests join the fused passes as they arrive."""
    t0 = time.perf_counter()
    ids = []
    for q, a in zip(reqs, arrivals):
        while time.perf_counter() < t0 + a:
            time.sleep(0.0005)
        ids.append(r.submit(q["sid"], q["width"], q["height"], q["spp"], q["max_depth"], seed=q["seed"], image_only=True))
    lat, rays, imgs = [], 0, []
    for q, i in zip(reqs, ids):
        rgb, _, _ = r.wait(i, q["width"], q["height"], want_accum=False)
        tm = r.timing(i)
        lat.append(tm["done_s"] - tm["submit_s"])
        rays += tm["rays"]
        imgs.append(rgb)
    wall = time.perf_counter() - t0
    return np.array(lat), rays, wall, imgs


def _serve_plain(r, reqs, arrivals):
    """The plain server: requests in arrival order, each rendered by one megakernel launch; a
    request waits until the ones before it are done. Latency = completion - arrival."""
    t0 = time.perf_counter()
    lat, rays, imgs = [], 0, []
    for q, a in zip(reqs, arrivals):
        while time.perf_counter() < t0 + a:
            time.sleep(0.0005)
        rgb, _, st = r.render(q["sid"], q["width"], q["height"], q["spp"], q["max_depth"], seed=q["seed"],
                              want_accum=False, megakernel=True)
        lat.append(time.perf_counter() - (t0 + a))
        rays += st["rays"]
        imgs.append(rgb)
    wall = time.perf_counter() - t0
    return np.array(lat), rays, wall, imgs


def _serving_compare(r, reqs, arrivals):
    """Both servers on the same Poisson arrivals (latency percentiles, aggregate rate at that
    offered load), then on the same requests all arriving at once (sustained capacity: the
    aggregate Mrays/s each server reaches when it is never idle)."""
    out = {}
    lat_b, rays_b, wall_b, img_b = _serve_batched(r, reqs, arrivals)
    lat_p, rays_p, wall_p, img_p = _serve_plain(r, reqs, arrivals)
    for name, lat, rays, wall in (("batched", lat_b, rays_b, wall_b), ("plain", lat_p, rays_p, wall_p)):
        out[name] = {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "aggregate_mrays_s": rays / wall / 1e6, "wall_s": wall}
    zero = np.zeros(len(reqs))
    _, rays_bs, wall_bs, _ = _serve_batched(r, reqs, zero)
    _, rays_ps, wall_ps, _ = _serve_plain(r, reqs, zero)
    out["batched"]["saturated_mrays_s"] = rays_bs / wall_bs / 1e6
    out["plain"]["saturated_mrays_s"] = rays_ps / wall_ps / 1e6
    out["speedup"] = out["batched"]["saturated_mrays_s"] / out["plain"]["saturated_mrays_s"]
    out["speedup_at_offered_load"] = out["batched"]["aggregate_mrays_s"] / out["plain"]["aggregate_mrays_s"]
    # equal image quality: 8-bit pixels of the two servers agree (same estimator and streams)
    out["rgb_equal_fraction"] = float(np.mean([np.mean(np.all(a == b, axis=-1)) for a, b in zip(img_b, img_p)]))
    return out


def serving(seed=11, n=48, rate_hz=60.0):
    """Poisson arrivals into the continuous-batching loop and into the plain server."""
    from cortexraytracercuda_amd import Renderer, scenes
    rng = np.random.default_rng(seed)
    r = Renderer()
    pool = [r.upload(scenes.sphere_mesh_scene(int(t), seed=i)) for i, t in enumerate([1e4, 5e4, 2e5, 1e6])]
    reqs, arrivals = _poisson_requests(rng, pool, n, rate_hz, (256, 384, 512, 768), (1, 4, 16, 64))
    r.render(pool[0], 64, 64, 1, 8)
    r.render(pool[0], 64, 64, 1, 8, megakernel=True)
    out = dict(requests=n, rate_hz=rate_hz, **_serving_compare(r, reqs, arrivals))
    r.close()
    return out


def _profile_json(kind, workload_key, inf):
    """A committed counter summary (profiles/k_extend_<kind>_<workload>.json) for the BVH shape
    it was taken on, or None."""
    if not workload_key:
        return None
    path = os.path.join(ROOT, "profiles", "k_extend_%s_%s.json" % (kind, workload_key))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if (t["bvh_width"], t["leaf_size"]) != (inf["bvh_width"], inf["leaf_size"]):
        return None
    return t


def traversal_roofline(scene, cfg, args, timed_renderer, sid, progressive=False, workload_key=""):
    """k_extend roofline for one render of the workload. The launches are timed with HIP events on
    their stream. The binding resource is VALU issue: VALU wave instructions per ray (SQ_INSTS_VALU
    of the same kernel on the same workload, profiles/) times the rays of a launch over its time,
    against VALU_PEAK_GINST. Beside it: DRAM bytes per launch (FETCH_SIZE / WRITE_SIZE, profiles/)
    and the algorithmic bytes (node and primitive records of every visit, from the counting variant
    of the kernel on the same deterministic workload), which count cache hits as well and exceed
    the HBM peak."""
    from cortexraytracercuda_amd import Renderer
    rc = Renderer(count_work=True, bvh_width=args.bvh_width, max_paths=args.max_paths, extend_kernel=args.extend_kernel,
                  refill_min=args.refill_min, grid_per_cu=args.grid_per_cu, shade_grid_per_cu=args.shade_grid_per_cu,
                  sub_passes=args.sub_passes, leaf_size=args.leaf_size, stack_lds=getattr(args, "stack_lds", 0))
    sidc = rc.upload(scene)
    inf = rc.info(sidc)
    node_b = inf["node_bytes"]
    cnt = dict.fromkeys(["rays", "node_visits", "leaf_visits", "tri_tests", "sphere_tests", "quad_tests", "inst_visits"], 0)
    ext_ms = 0.0
    launches = 0
    ranges = [(k, 1) for k in range(cfg["spp"])] if progressive else [(0, 0)]
    for sb, sc in ranges:
        rc.render(sidc, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, sample_begin=sb,
                  sample_count=sc, want_accum=False)
        c = rc.counters()
        for k in cnt:
            cnt[k] += c[k]
        _, _, stt = timed_renderer.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1,
                                          sample_begin=sb, sample_count=sc, want_accum=False, time_stages=True)
        ext_ms += stt["extend_ms"]
        launches += timed_renderer.extend_launches()
    rc.close()
    alg = (B_STATE * cnt["rays"] + node_b * cnt["node_visits"] + B_LEAF * cnt["leaf_visits"] + B_TRI * cnt["tri_tests"] +
           B_SPHERE * cnt["sphere_tests"] + B_QUAD * cnt["quad_tests"] + B_INST * cnt["inst_visits"])
    launches = max(launches, 1)
    launch_s = ext_ms / launches * 1e-3
    rays_per_launch = cnt["rays"] / launches
    algorithmic_gbs = alg / (ext_ms * 1e-3) / 1e9
    # measured DRAM bytes (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy) per ray of this
    # workload's closest-hit kernel, times the rays of one launch; null for workloads not profiled
    tr = _profile_json("traffic", workload_key, inf)
    traffic = None if tr is None else tr["hbm_bytes_per_ray"] * rays_per_launch
    dram = None if traffic is None else traffic / launch_s / 1e9
    sq = _profile_json("sq", workload_key, inf)
    probe_path = os.path.join(ROOT, "profiles", "valu_issue_probe_mi355x.json")
    probe = json.load(open(probe_path))["rates_ginst_s"] if os.path.exists(probe_path) else {}
    roof = {"kernel": "k_extend", "launches": launches, "avg_launch_ms": ext_ms / launches,
            "rays_per_launch": rays_per_launch,
            "traffic": traffic, "traffic_unit": "bytes per launch",
            "dram_gbs": dram, "dram_frac": None if dram is None else dram / HBM_PEAK_GBS,
            # the algorithmic byte model prices every node and primitive fetch, most of which hit
            # L2 or the MALL: a rate above the HBM peak, not a fraction of it
            "algorithmic_gbs": algorithmic_gbs,
            "bytes_per_launch": alg / launches, "bytes_per_ray": alg / max(cnt["rays"], 1),
            "node_bytes": node_b, "node_visits_per_ray": cnt["node_visits"] / max(cnt["rays"], 1),
            "prim_tests_per_ray": (cnt["tri_tests"] + cnt["sphere_tests"] + cnt["quad_tests"]) / max(cnt["rays"], 1)}
    if sq is not None:
        valu = sq["valu_insts_per_ray"] * rays_per_launch / launch_s / 1e9
        roof.update({"bound": "valu", "achieved": valu, "peak": VALU_PEAK_GINST, "unit": "G wave64 VALU instructions/s",
                     "frac": valu / VALU_PEAK_GINST, "valu_insts_per_ray": sq["valu_insts_per_ray"],
                     "lanes_active_per_valu": sq.get("lanes_active_per_valu"),
                     "probe_dense_fma_ginst": probe.get("v_fma_f32_8wps"), "probe_mov_ginst": probe.get("v_mov_b32_8wps")})
    elif dram is not None:
        roof.update({"bound": "hbm", "achieved": dram, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dram / HBM_PEAK_GBS})
    else:
        roof.update({"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINST, "unit": "G wave64 VALU instructions/s",
                     "frac": None, "note": "no counter profile committed for this workload and BVH shape"})
    # a fraction of the binding resource cannot exceed it: a larger value means the committed
    # counter profile is stale for this build. The VALU peak prices every wave64 instruction at 4
    # cycles; moves issue in 2 (the probe reaches 1030 G/s with v_mov_b32), so a stream rich in
    # them can pass the dense peak, never twice it
    assert roof["frac"] is None or roof["frac"] <= (2.0 if roof["bound"] == "valu" else 1.0), roof
    return roof


def _megakernel_bench(r, sid, cfg, ranges):
    """The plain path on the same workload: one megakernel launch per render (per pass when progressive)."""
    r.render(sid, 64, 64, 1, cfg["max_depth"], megakernel=True)
    t1 = time.perf_counter()
    mk_rays = 0
    for sb, sc in ranges:
        _, _, stm = r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, sample_begin=sb,
                             sample_count=sc, want_accum=False, megakernel=True)
        mk_rays += stm["rays"]
    mk_wall = time.perf_counter() - t1
    log("megakernel: %.1f ms, %d rays" % (mk_wall * 1e3, mk_rays))
    return {"mrays_s": mk_rays / mk_wall / 1e6, "ms_per_step": mk_wall * 1e3}


def _progressive_served(r, sid, cfg, ranges, args):
    """The progressive passes submitted to the serving loop together (each pass a request of its
    own sample range), so one pass's drain overlaps the next passes' full launches; wall time from
    the first submit to the last image; every pass image is available as soon as its pass is done.
    Then the same passes fused into one crt_render_batch: that is not progressive (no pass image
    before the whole batch is done), it is reported apart. Every pass image of both equals the
    blocking render of the same sample range (checked untimed)."""
    W, H = cfg["width"], cfg["height"]

    def step():
        rids = [r.submit(sid, W, H, cfg["spp"], cfg["max_depth"], seed=1, image_only=True, sample_begin=sb,
                         sample_count=sc) for sb, sc in ranges]
        rays, imgs = 0, []
        for rid in rids:
            rgb, _, st = r.wait(rid, W, H, want_accum=False)
            rays += st["rays"]
            imgs.append(rgb)
        return rays, imgs

    for _ in range(args.warmup):
        step()
    rays = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rays += step()[0]
    wall = time.perf_counter() - t0
    _, imgs = step()
    refs = [r.render(sid, W, H, cfg["spp"], cfg["max_depth"], seed=1, sample_begin=sb, sample_count=sc,
                     want_accum=False)[0] for sb, sc in ranges]
    served = {"mrays_s": rays / wall / 1e6, "ms_per_step": wall / args.steps * 1e3, "passes": len(ranges),
              "every_pass_identical_to_blocking": bool(all(np.array_equal(a, b) for a, b in zip(imgs, refs)))}
    # the same passes fused into one batched launch (one request per pass): all images at the end
    jobs = [dict(sid=sid, width=W, height=H, spp=cfg["spp"], max_depth=cfg["max_depth"], seed=1, sample_begin=sb,
                 sample_count=sc) for sb, sc in ranges]
    for _ in range(args.warmup):
        r.render_batch(jobs, want_accum=False)
    rays = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs, st = r.render_batch(jobs, want_accum=False)
        rays += st["rays"]
    wall = time.perf_counter() - t0
    served["fused_batch_not_progressive"] = {
        "mrays_s": rays / wall / 1e6, "ms_per_step": wall / args.steps * 1e3,
        "every_pass_identical_to_blocking": bool(all(np.array_equal(o[0], b) for o, b in zip(outs, refs)))}
    return served


def _device_bvh_bench(scene, cfg, args, r_host, sid_host):
    """The same request on a tree built on the device (crt_config.bvh_builder = device): build time
    (a first build warms the sort's and scan's kernels up), one render, and its image against the
    host-built tree's (closest hits do not depend on the tree, so the images are identical)."""
    from cortexraytracercuda_amd import Renderer
    r = Renderer(bvh_builder="device", bvh_width=args.bvh_width, max_paths=args.max_paths,
                 extend_kernel=args.extend_kernel, refill_min=args.refill_min, grid_per_cu=args.grid_per_cu,
                 shade_grid_per_cu=args.shade_grid_per_cu, sub_passes=args.sub_passes, leaf_size=args.leaf_size)
    r.release(r.upload(scene))
    t0 = time.perf_counter()
    sid = r.upload(scene)
    wall_ms = (time.perf_counter() - t0) * 1e3
    info = r.info(sid)
    r.verify(sid)
    rgb_d, acc_d, st_d = r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1)
    r.close()
    rgb_h, acc_h, st_h = r_host.render(sid_host, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1)
    out = {"builder": info["builder"], "build_ms": info["build_ms"], "upload_call_ms": wall_ms, "nodes": info["num_nodes"],
           "depth": info["max_depth"], "sah_cost": info["sah_cost"], "device_bytes": info["device_bytes"],
           "render_gpu_ms": st_d["gpu_ms"], "host_tree_render_gpu_ms": st_h["gpu_ms"],
           "image_identical_to_host_tree": bool(np.array_equal(acc_d, acc_h, equal_nan=True) and np.array_equal(rgb_d, rgb_h))}
    log("device BVH: %s" % out)
    return out


def valu_probe(iters=100000):
    """Wave64 VALU instructions per second of independent v_fma_f32, v_add_u32 and v_mul_f32 chains
    on every SIMD (2, 4 and 8 waves per SIMD), and of v_fma_f32 with part of each wave active (16
    lanes in one group or spread over all four, 32, 48 or 63 lanes): the issue rate the closest-hit
    kernel's VALU count is priced against, and how it depends on the active lanes."""
    import ctypes as C
    from cortexraytracercuda_amd import Renderer
    out = {}
    with Renderer() as r:
        for kind, name in ((0, "v_fma_f32"), (1, "v_add_u32"), (2, "v_mul_f32"), (3, "v_fma_f32_lanes0_15"),
                           (4, "v_fma_f32_every4th_lane"), (5, "v_fma_f32_lanes0_31"), (6, "v_fma_f32_lanes0_62"),
                           (7, "v_fma_f32_lanes0_47"), (8, "v_fma_f32_no_lane"), (9, "v_cndmask_b32"),
                           (10, "v_max_f32"), (11, "v_cmp_lt_f32"), (12, "v_mov_b32"), (13, "v_cndmask_b32_e64_sgpr"),
                           (14, "v_fma_f32_exec_lanes0_15"), (15, "v_fma_f32_exec_lanes0_31")):
            for wps in (2, 4, 8):
                ms, n = C.c_double(), C.c_uint64()
                rc = r.lib.crt_debug_valu_probe(r.h, kind, wps, iters, C.byref(ms), C.byref(n))
                if rc:
                    raise RuntimeError("valu probe failed: %d" % rc)
                out["%s_%dwps" % (name, wps)] = {"ms": ms.value, "wave_insts": n.value,
                                                 "ginst_s": n.value / (ms.value * 1e-3) / 1e9}
    return {"valu_probe": out}


def one_render_probe(scene, cfg, args, progressive):
    """One render of the workload with the library's default shape, in the form the roofline times
    (stages timed with HIP events: one sub-pass, the launches of traversal_roofline; no counting
    variant): the rays and kernel launches a counter run divides its per-dispatch sums by."""
    from cortexraytracercuda_amd import Renderer
    r = Renderer(bvh_width=args.bvh_width, max_paths=args.max_paths, extend_kernel=args.extend_kernel,
                 refill_min=args.refill_min, grid_per_cu=args.grid_per_cu, shade_grid_per_cu=args.shade_grid_per_cu,
                 sub_passes=args.sub_passes, leaf_size=args.leaf_size)
    sid = r.upload(scene)
    info = r.info(sid)
    ranges = [(k, 1) for k in range(cfg["spp"])] if progressive else [(0, 0)]
    rays = launches = exact = 0
    for sb, sc in ranges:
        _, _, st = r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, sample_begin=sb,
                            sample_count=sc, want_accum=False, time_stages=True)
        c = r.counters()
        rays += st["rays"]
        launches += c["extend_launches"]
        exact += c["exact_queries"]
    r.close()
    return {"one_render": True, "config": cfg, "passes": len(ranges), "rays": rays, "extend_launches": launches,
            "exact_queries": exact, "bvh_width": info["bvh_width"], "leaf_size": info["leaf_size"]}


def single_render_bench(scene, cfg, args, workload, progressive=False, workload_key=""):
    from cortexraytracercuda_amd import Renderer
    r = Renderer(bvh_width=args.bvh_width, max_paths=args.max_paths, extend_kernel=args.extend_kernel,
                 refill_min=args.refill_min, grid_per_cu=args.grid_per_cu, shade_grid_per_cu=args.shade_grid_per_cu,
                  sub_passes=args.sub_passes, leaf_size=args.leaf_size)
    t0 = time.perf_counter()
    sid = r.upload(scene)
    info = r.info(sid)
    log("upload+build %.2f s %s" % (time.perf_counter() - t0, info))
    tune = None
    if args.autotune:
        t0 = time.perf_counter()
        tune = r.autotune(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, scene=scene)
        tune["seconds"] = time.perf_counter() - t0
        info = r.info(sid)
        log("autotune %s" % tune)
    ranges = [(k, 1) for k in range(cfg["spp"])] if progressive else [(0, 0)]

    def one_step():
        rays = 0
        gpu = 0.0
        for sb, sc in ranges:
            _, _, st = r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, sample_begin=sb,
                                sample_count=sc, want_accum=False)
            rays += st["rays"]
            gpu += st["gpu_ms"]
        return rays, gpu, st

    for _ in range(args.warmup):
        one_step()
        log("warmup done")
    rays = 0
    gpu_ms = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        rr, gg, st = one_step()
        rays += rr
        gpu_ms += gg
        log("step %d: %.1f ms GPU, %d rays" % (k, gg, rr))
    wall = time.perf_counter() - t0
    value = rays / wall / 1e6
    # closest-hit queries of the last render that the float triangle test left to the exact re-trace
    exact_frac = None
    if not progressive:
        exact_last = r.counters()["exact_queries"]
        exact_frac = exact_last / max(rr, 1)
        log("exact re-traces in the last step: %d of %d queries" % (exact_last, rr))
    served = None
    if progressive:
        served = _progressive_served(r, sid, cfg, ranges, args)
    mk = None
    if not args.no_megakernel:
        mk = _megakernel_bench(r, sid, cfg, ranges)
    dev_bvh = None
    if not progressive and not args.autotune:
        dev_bvh = _device_bvh_bench(scene, cfg, args, r, sid)
    if tune is not None:
        # the counting and event-timed renders use the tuned shape too
        args = argparse.Namespace(**vars(args))
        args.max_paths, args.grid_per_cu, args.shade_grid_per_cu = tune["max_paths"], tune["grid_per_cu"], tune["shade_grid_per_cu"]
        args.sub_passes, args.extend_kernel, args.refill_min = tune["sub_passes"], tune["extend_kernel"], tune["refill_min"]
        args.stack_lds, args.bvh_width, args.leaf_size = tune["stack_lds"], tune["bvh_width"], tune["leaf_size"]
    roof = traversal_roofline(scene, cfg, args, r, sid, progressive, workload_key)

    r.close()
    return {"metric": "Mrays/s on 1 MI355X + p50/p99 request latency, batched vs plain, equal RMSE",
            "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": dict(workload=workload, **cfg),
            "rays_per_step": rays // args.steps, "segments_per_s": value * 1e6, "gpu_ms_per_step": gpu_ms / args.steps,
            "exact_retrace_fraction": exact_frac,
            "bvh": {"nodes": info["num_nodes"], "width": info["bvh_width"], "depth": info["max_depth"],
                    "build_ms": info["build_ms"], "upload_ms": info["upload_ms"], "device_bytes": info["device_bytes"],
                    "builder": info["builder"]},
            "bvh_device": dev_bvh, "roofline": roof, "autotune": tune, "progressive_served": served,
            "megakernel": mk and dict(mk, wavefront_speedup=value / mk["mrays_s"])}


def batch_workload(seed=4, n=64, tri_range=(1e4, 2e6), res_choices=(256, 384, 512, 768, 1024), spp_range=(1, 64)):
    """Config 4: n scenes of uneven size (log-uniform triangle counts), resolutions 256^2-1024^2,
    1-64 spp: one crt_render per request vs one crt_render_batch over all of them."""
    from cortexraytracercuda_amd import Renderer, scenes
    rng = np.random.default_rng(seed)
    r = Renderer()
    jobs = []
    tg = time.perf_counter()
    tris_total = 0
    for k in range(n):
        tris = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        sid = r.upload(scenes.sphere_mesh_scene(tris, seed=k))
        tris_total += tris
        res = int(rng.choice(res_choices))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        jobs.append(dict(sid=sid, width=res, height=res, spp=spp, max_depth=8, seed=1000 + k))
    log("batch workload: %d scenes, %d triangles, built in %.1f s" % (n, tris_total, time.perf_counter() - tg))
    out = dict(requests=n, triangles_total=tris_total, **_plain_and_batched(r, jobs))
    r.close()
    return out


def _serve_with_uploads(r, reqs, arrivals, catalogue, plain):
    """One server over Poisson arrivals where each request names a seeded scene of the catalogue:
    the first request of a scene has it uploaded and its BVH built (crt_scene_upload) before it is
    submitted, later requests find it resident. Latency = completion - arrival, upload and build
    included. plain: one megakernel render per request in arrival order (one thread); batched: an
    arrival thread submits requests whose scene is resident and hands the others to an upload
    thread, which uploads in arrival order and submits behind each upload."""
    import queue
    import threading
    resident, builds = {}, []
    lock = threading.Lock()
    t0 = time.perf_counter()
    done_at = [0.0] * len(reqs)
    rays = [0] * len(reqs)

    def ensure(key):
        with lock:
            if key in resident:
                return resident[key]
        tb = time.perf_counter()
        sid = r.upload(catalogue[key])
        builds.append(time.perf_counter() - tb)
        with lock:
            resident[key] = sid
        return sid

    if plain:
        for k, (q, a) in enumerate(zip(reqs, arrivals)):
            while time.perf_counter() < t0 + a:
                time.sleep(0.0005)
            sid = ensure(q["scene"])
            _, _, st = r.render(sid, q["width"], q["height"], q["spp"], q["max_depth"], seed=q["seed"], want_accum=False,
                                megakernel=True)
            done_at[k] = time.perf_counter() - t0
            rays[k] = st["rays"]
    else:
        ids = [None] * len(reqs)
        py_submit = [0.0] * len(reqs)
        upq = queue.Queue()

        def submit(k, sid):
            q = reqs[k]
            py_submit[k] = time.perf_counter()
            ids[k] = r.submit(sid, q["width"], q["height"], q["spp"], q["max_depth"], seed=q["seed"], image_only=True)

        def uploader():
            while True:
                k = upq.get()
                if k is None:
                    return
                submit(k, ensure(reqs[k]["scene"]))

        th = threading.Thread(target=uploader)
        th.start()
        pending = set()
        for k, (q, a) in enumerate(zip(reqs, arrivals)):
            while time.perf_counter() < t0 + a:
                time.sleep(0.0005)
            with lock:
                sid = resident.get(q["scene"])
            if sid is not None and q["scene"] not in pending:
                submit(k, sid)
            else:
                # behind the scene's upload (requests of one scene keep their order)
                pending.add(q["scene"])
                upq.put(k)
        upq.put(None)
        th.join()
        tms = []
        for k, q in enumerate(reqs):
            r.wait(ids[k], q["width"], q["height"], want_accum=False)
            tms.append(r.timing(ids[k]))
            rays[k] = tms[-1]["rays"]
        # the library stamps submit / done on its steady clock: map it onto this process's clock by
        # the submit stamps (taken just before each crt_submit here)
        offset = float(np.median([tm["submit_s"] - ps for tm, ps in zip(tms, py_submit)]))
        for k, tm in enumerate(tms):
            done_at[k] = tm["done_s"] - offset - t0
    wall = max(done_at)
    lat = np.array([d - a for d, a in zip(done_at, arrivals)])
    return lat, int(sum(rays)), wall, builds


def serving_workload(seed=5, n=96, rate_hz=20.0, n_scenes=24, tri_range=(1e4, 2e6), res_choices=(256, 512, 1024, 2048),
                     spp_choices=(1, 4, 16, 64, 256), budget_gb=200.0):
    """Config 5: Poisson arrivals at rate_hz; each request names one of n_scenes seeded scenes
    (log-uniform 10k-2M triangles; uploaded and built at its first request, then resident), a
    resolution 256^2-2048^2 and 1-256 spp; admission against an HBM budget; latency = completion -
    arrival, scene upload and BVH build included. The same requests then arrive all at once on
    resident scenes: the sustained capacity of each server."""
    from cortexraytracercuda_amd import Renderer, scenes
    rng = np.random.default_rng(seed)
    tris = [int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))) for _ in range(n_scenes)]
    tg = time.perf_counter()
    catalogue = [scenes.sphere_mesh_scene(t, seed=100 + k) for k, t in enumerate(tris)]
    log("serving workload: %d scenes, %d triangles, generated in %.1f s" % (n_scenes, sum(tris), time.perf_counter() - tg))
    reqs = []
    for k in range(n):
        res = int(rng.choice(res_choices))
        reqs.append(dict(scene=int(rng.integers(0, n_scenes)), width=res, height=res, spp=int(rng.choice(spp_choices)),
                         max_depth=8, seed=5000 + k))
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    out = dict(requests=n, rate_hz=rate_hz, budget_gb=budget_gb, scenes=n_scenes, triangles=sum(tris))
    for name, plain, builder in (("batched", False, "host"), ("batched_device_bvh", False, "device"),
                                 ("plain", True, "host")):
        r = Renderer(hbm_budget=int(budget_gb * 1e9), bvh_builder=builder)
        warm = r.upload(scenes.box_scene())
        r.render(warm, 64, 64, 1, 8)
        r.render(warm, 64, 64, 1, 8, megakernel=True)
        lat, rays, wall, builds = _serve_with_uploads(r, reqs, arrivals, catalogue, plain)
        out[name] = {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "aggregate_mrays_s": rays / wall / 1e6, "wall_s": wall,
                     "scene_builds": len(builds), "build_ms_max": max(builds) * 1e3 if builds else 0.0,
                     "build_ms_total": sum(builds) * 1e3}
        log("serving %s: %s" % (name, out[name]))
        # sustained capacity on resident scenes: every request at t = 0
        sids = [r.upload(c) for c in catalogue]
        resident_reqs = [dict(q, sid=sids[q["scene"]]) for q in reqs]
        zero = np.zeros(n)
        if plain:
            _, rays_s, wall_s, _ = _serve_plain(r, resident_reqs, zero)
        else:
            _, rays_s, wall_s, _ = _serve_batched(r, resident_reqs, zero)
        out[name]["saturated_mrays_s"] = rays_s / wall_s / 1e6
        r.close()
    out["speedup"] = out["batched"]["saturated_mrays_s"] / out["plain"]["saturated_mrays_s"]
    out["speedup_at_offered_load"] = out["batched"]["aggregate_mrays_s"] / out["plain"]["aggregate_mrays_s"]
    out["p99_speedup_device_bvh"] = out["plain"]["p99_ms"] / out["batched_device_bvh"]["p99_ms"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="large", choices=["large", "box", "instanced", "batch", "serving"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller image for a fast check")
    ap.add_argument("--bvh-width", type=int, default=0, help="2 or 4 children per node (0: library default)")
    ap.add_argument("--max-paths", type=int, default=0, help="path slots per pass (0: library default)")
    ap.add_argument("--extend-kernel", type=int, default=0, help="1 grid-stride, 2 persistent refill (0: default)")
    ap.add_argument("--refill-min", type=int, default=0, help="persistent stage: idle lanes before a refill")
    ap.add_argument("--grid-per-cu", type=int, default=0, help="workgroups per CU of the closest-hit launches (0: default)")
    ap.add_argument("--shade-grid-per-cu", type=int, default=0, help="workgroups per CU of regenerate/shade (0: default)")
    ap.add_argument("--sub-passes", type=int, default=0, help="concurrent sub-passes per pass (0: default)")
    ap.add_argument("--leaf-size", type=int, default=0, help="BVH max primitives per leaf (0: default)")
    ap.add_argument("--no-megakernel", action="store_true", help="skip the megakernel comparison render")
    ap.add_argument("--autotune", action="store_true",
                    help="run crt_autotune on the workload before timing (default: the library's default shape)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the VALU issue-rate probe (calibrates the closest-hit kernel's VALU roofline)")
    ap.add_argument("--one-render", action="store_true",
                    help="only upload the workload's scene and render it once (every pass of a progressive "
                         "workload), for counter runs: every closest-hit dispatch then belongs to that render")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 1 << 28                                  # 2^28 float32 = 1 GiB
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 3}), flush=True)
        return
    if args.valu_probe:
        print(json.dumps(valu_probe()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X, and this path is single-GPU "
              "(multi-GPU sharding is a design note only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from cortexraytracercuda_amd import scenes
    if args.config == "batch":
        print(json.dumps({"config": "batch", "batched": batch_workload()}), flush=True)
        return
    if args.config == "serving":
        print(json.dumps({"config": "serving", "serving": serving_workload()}), flush=True)
        return
    t_gen = time.perf_counter()
    progressive = False
    if args.config == "box":
        scene, cfg = scenes.box_scene(), dict(width=256, height=256, spp=16, max_depth=8)
        workload = "small box scene: 12 wall triangles, area light, 3 spheres, 256x256, 16 spp, depth 8"
    elif args.config == "instanced":
        scene, cfg, workload = instanced_workload()
        progressive = True
    else:
        scene, cfg = large_workload()
        workload = ("large scene: icosphere subdivided to 1,310,720 triangles displaced by 3-octave value noise "
                    "(amplitude 0.05) + 500 random spheres (60/30/10 Lambertian/metal/dielectric), 1920x1080, "
                    "64 spp, depth 10, keyed PCG seed 1")
    if args.quick:
        cfg = dict(cfg, width=cfg["width"] // 4, height=cfg["height"] // 4, spp=min(cfg["spp"], 16))
    log("scene generated in %.1f s" % (time.perf_counter() - t_gen))
    if args.one_render:
        print(json.dumps(one_render_probe(scene, cfg, args, progressive)), flush=True)
        return
    key = "" if (args.quick or args.bvh_width or args.leaf_size) else {
        "large": "large_1080p_64spp", "box": "box_256_16spp", "instanced": "instanced_4k_16spp"}[args.config]
    out = single_render_bench(scene, cfg, args, workload, progressive, key)
    if args.config == "large" and not args.no_extras:
        # the batched-launch, serving and instanced workloads at their configured sizes
        try:
            out["batched"] = dict(workload="64 scenes of log-uniform 10k-2M triangles, 256^2-1024^2, 1-64 spp, depth 8: "
                                           "one crt_render_batch vs one megakernel launch per request",
                                  **batch_workload())
            log("batched: %s" % out["batched"])
        except Exception as e:  # reported, never hidden
            out["batched"] = {"error": repr(e)}
        try:
            out["serving"] = dict(workload="96 Poisson requests at 20 Hz over 24 seeded scenes of log-uniform 10k-2M "
                                           "triangles (uploaded and built at first use, inside the latency), "
                                           "256^2-2048^2, 1-256 spp, depth 8, 200 GB HBM budget",
                                  **serving_workload())
            log("serving: %s" % out["serving"])
        except Exception as e:
            out["serving"] = {"error": repr(e)}
        try:
            iscene, icfg, iwork = instanced_workload()
            iargs = argparse.Namespace(**vars(args))
            iargs.steps, iargs.warmup = min(args.steps, 3), min(args.warmup, 1)
            ins = single_render_bench(iscene, icfg, iargs, iwork, True, "instanced_4k_16spp")
            out["instanced"] = {k: ins[k] for k in ("value", "unit", "ms_per_step", "rays_per_step", "gpu_ms_per_step",
                                                     "config", "bvh", "roofline", "progressive_served", "megakernel")}
            log("instanced: %s" % out["instanced"])
        except Exception as e:
            out["instanced"] = {"error": repr(e)}
    if not args.no_cpu and args.config in ("large", "box"):
        out["cpu_baseline"] = cpu_baseline(scene, cfg)
        log("cpu baseline: %s" % out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
