"""Benchmark of the render hot path on one MI355X.

Default workload: the single large scene (icosphere subdivided to 1,310,720 triangles displaced by
3-octave value noise, plus 500 random spheres), 1920x1080, 64 spp, max depth 10, FP32, keyed PCG
streams, seed 1. One step = one complete render of that image (every sample of every pixel).

Printed JSON line (rank 0):
  value        Mrays/s over the timed steps (closest-hit queries / wall time of the renders)
  roofline     traversal kernel (k_extend): algorithmic bytes per launch / HIP-event time
  cpu_baseline reference-faithful CPU renderer (oracle/, FP32 minstd) on the host cores, 1 spp
  batched      fused passes over many small requests vs one render per launch (same images)
  serving      Poisson request loop: p50/p99 latency and aggregate Mrays/s
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
# Algorithmic bytes of one closest-hit query in k_extend (declared layout, crt_layout.h):
#   ray 32 (ray_o + ray_d) + ids 8 + queue index 4 + hit record 16 + hit instance 4 = 64 B of state
#   + 64 B per interior node visit (two child boxes + refs), 8 B per leaf record,
#   + 48 B per triangle, 16 B per sphere, 80 B per quad/light, 64 B per instance record.
B_STATE = 64
B_NODE, B_LEAF, B_TRI, B_SPHERE, B_QUAD, B_INST = 64, 8, 48, 16, 80, 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def large_workload():
    from cortexraytracercuda_amd import scenes
    return scenes.large_scene(seed=1, levels=8, n_spheres=500), dict(width=1920, height=1080, spp=64, max_depth=10)


def cpu_baseline(scene, cfg):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    acc, st = O.render(scene, cfg["width"], cfg["height"], 1, cfg["max_depth"], seed=2580, rng="minstd", precision=32,
                       threads=threads, ref_style=True)
    return {"value": st["rays"] / st["seconds"] / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%dx%d at 1 spp (all pixels), depth %d, reference-faithful FP32 minstd renderer with the "
                      "reference's median-split BVH; %d rays in %.1f s" % (cfg["width"], cfg["height"], cfg["max_depth"],
                                                                         st["rays"], st["seconds"])}


def batched_vs_plain(seed=7, n=24):
    """Many small requests: one crt_render per request vs crt_render_batch (fused passes)."""
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
    # warm up both paths
    r.render(jobs[0]["sid"], 64, 64, 1, 8)
    r.render_batch([dict(jobs[0], width=64, height=64)])
    t0 = time.perf_counter()
    plain_rays = 0
    plain_imgs = []
    for j in jobs:
        rgb, _, st = r.render(j["sid"], j["width"], j["height"], j["spp"], j["max_depth"], seed=j["seed"], want_accum=False)
        plain_rays += st["rays"]
        plain_imgs.append(rgb)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    outs, st = r.render_batch(jobs, want_accum=False)
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a, b[0]) for a, b in zip(plain_imgs, outs))
    r.close()
    return {"requests": n, "plain_mrays_s": plain_rays / t_plain / 1e6, "batched_mrays_s": st["rays"] / t_batch / 1e6,
            "speedup": (st["rays"] / t_batch) / (plain_rays / t_plain), "identical_images": bool(same),
            "plain_s": t_plain, "batched_s": t_batch}


def serving(seed=11, n=48, rate_hz=60.0):
    """Poisson arrivals at rate_hz into the continuous-batching loop; latency = done - submit."""
    from cortexraytracercuda_amd import Renderer, scenes
    rng = np.random.default_rng(seed)
    r = Renderer()
    pool = [r.upload(scenes.sphere_mesh_scene(int(t), seed=i)) for i, t in enumerate([1e4, 5e4, 2e5, 1e6])]
    reqs = []
    for k in range(n):
        res = int(rng.choice([256, 384, 512, 768]))
        reqs.append(dict(sid=pool[int(rng.integers(0, len(pool)))], width=res, height=res,
                         spp=int(rng.choice([1, 4, 16, 64])), max_depth=8, seed=5000 + k))
    gaps = rng.exponential(1.0 / rate_hz, n)
    r.render(pool[0], 64, 64, 1, 8)
    t0 = time.perf_counter()
    ids = []
    for k, q in enumerate(reqs):
        target = t0 + gaps[:k + 1].sum()
        while time.perf_counter() < target:
            time.sleep(0.0005)
        ids.append(r.submit(q["sid"], q["width"], q["height"], q["spp"], q["max_depth"], seed=q["seed"]))
    lat = []
    rays = 0
    for q, i in zip(reqs, ids):
        r.wait(i, q["width"], q["height"], want_accum=False)
        tm = r.timing(i)
        lat.append(tm["done_s"] - tm["submit_s"])
        rays += tm["rays"]
    wall = time.perf_counter() - t0
    r.close()
    lat = np.array(lat)
    return {"requests": n, "rate_hz": rate_hz, "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "aggregate_mrays_s": rays / wall / 1e6}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller image for a fast check")
    ap.add_argument("--bvh-width", type=int, default=0, help="2 or 4 children per node (0: library default)")
    ap.add_argument("--max-paths", type=int, default=0, help="path slots per pass (0: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X, and this path is single-GPU "
              "(multi-GPU sharding is a design note only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from cortexraytracercuda_amd import Renderer
    t_gen = time.perf_counter()
    scene, cfg = large_workload()
    if args.quick:
        cfg = dict(width=480, height=270, spp=16, max_depth=10)
    log("scene generated in %.1f s" % (time.perf_counter() - t_gen))
    r = Renderer(bvh_width=args.bvh_width, max_paths=args.max_paths)
    t0 = time.perf_counter()
    sid = r.upload(scene)
    info = r.info(sid)
    log("upload+build %.2f s %s" % (time.perf_counter() - t0, info))
    for _ in range(args.warmup):
        r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, want_accum=False)
        log("warmup done")
    rays = 0
    gpu_ms = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        rgb, _, st = r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, want_accum=False)
        rays += st["rays"]
        gpu_ms += st["gpu_ms"]
        log("step %d: %.1f ms, %d rays" % (k, st["host_ms"], st["rays"]))
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = rays / wall / 1e6
    # roofline of the traversal kernel, measured outside the timed region
    rc = Renderer(count_work=True, bvh_width=args.bvh_width, max_paths=args.max_paths)
    sidc = rc.upload(scene)
    _, _, stc = rc.render(sidc, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, want_accum=False)
    cnt = rc.counters()
    rc.close()
    _, _, stt = r.render(sid, cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], seed=1, want_accum=False,
                         time_stages=True)
    launches = r.extend_launches()
    r.close()
    alg_bytes = (B_STATE * cnt["rays"] + B_NODE * cnt["node_visits"] + B_LEAF * cnt["leaf_visits"] +
                 B_TRI * cnt["tri_tests"] + B_SPHERE * cnt["sphere_tests"] + B_QUAD * cnt["quad_tests"] +
                 B_INST * cnt["inst_visits"])
    achieved = alg_bytes / (stt["extend_ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "kernel": "k_extend", "launches": launches,
            "avg_launch_ms": stt["extend_ms"] / max(launches, 1), "bytes_per_launch": alg_bytes / max(launches, 1),
            "bytes_per_ray": alg_bytes / max(cnt["rays"], 1), "node_visits_per_ray": cnt["node_visits"] / max(cnt["rays"], 1)}
    out = {"metric": "Mrays/s on 1 MI355X + p50/p99 request latency, batched vs plain, equal RMSE",
           "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "large scene: displaced icosphere 1,310,720 tris + 500 spheres, %dx%d, %d spp, "
                                  "depth %d, keyed PCG seed 1" % (cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"]),
                      "triangles": 1310720, "spheres": 500, "width": cfg["width"], "height": cfg["height"],
                      "spp": cfg["spp"], "max_depth": cfg["max_depth"]},
           "rays_per_step": rays // args.steps, "paths_per_step": st["paths"],
           "segments_per_s": value * 1e6, "gpu_ms_per_step": gpu_ms / args.steps,
           "bvh": {"nodes": info["num_nodes"], "depth": info["max_depth"], "build_ms": info["build_ms"],
                   "upload_ms": info["upload_ms"]},
           "roofline": roof}
    if not args.no_extras:
        try:
            out["batched"] = batched_vs_plain()
            log("batched: %s" % out["batched"])
        except Exception as e:  # reported, never hidden
            out["batched"] = {"error": repr(e)}
        try:
            out["serving"] = serving()
            log("serving: %s" % out["serving"])
        except Exception as e:
            out["serving"] = {"error": repr(e)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, cfg)
        log("cpu baseline: %s" % out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
