"""Python front end of the C-ABI (ctypes over the in-tree libcortex_rt.so).

Mirrors the reference's render interface: `Renderer.render` is the blocking render of one scene
(launchGPU, core/src/gpu.cu:2127-2149), `launch_gpu` keeps launchGPU's argument list and writes the
PNG, and the batched / serving calls expose the fused-pass scheduler.
"""
import ctypes as C
import struct
import zlib

import numpy as np

from . import _abi as A

_RNG = {"pcg": A.CRT_RNG_PCG, "minstd": A.CRT_RNG_MINSTD}


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def stats_dict(st):
    return {k: getattr(st, k) for k, _ in A.crt_stats._fields_}


def render_desc(width, height, spp, max_depth, seed=1, rng="pcg", sample_begin=0, sample_count=0, flags=0):
    rd = A.crt_render_desc()
    rd.width, rd.height, rd.spp, rd.max_depth = int(width), int(height), int(spp), int(max_depth)
    rd.seed = int(seed)
    rd.rng_mode = _RNG[rng] if isinstance(rng, str) else int(rng)
    rd.sample_begin, rd.sample_count, rd.flags = int(sample_begin), int(sample_count), int(flags)
    return rd


class Renderer:
    """One device context: resident scenes, path-state buffers and the serving loop."""

    def __init__(self, device=0, hbm_budget=0, max_paths=0, count_work=False, leaf_size=0, bvh_width=0,
                 extend_kernel=0, refill_min=0, pass_paths=0, grid_per_cu=0, shade_grid_per_cu=0, sub_passes=0,
                 graphs=0, stack_lds=0, ray_sort=0, bvh_builder="host", batch_group_paths=0, serve_order="small_then_oldest",
                 serve_small_paths=0):
        """bvh_builder: "host" (binned SAH on the CPU) or "device" (linear BVH built on the GPU).
        batch_group_paths: render_batch gives a request with at least this many path ids left passes
        of its own scene (0: the library default, 0xffffffff: every pass mixes requests in order).
        serve_order: "small_then_oldest" (requests with at most serve_small_paths path ids left first,
        then the others oldest first) or "round_robin" (one segment per request in turn)."""
        self.lib = A.load()
        cfg = A.crt_config()
        cfg.device = device
        cfg.hbm_budget_bytes = int(hbm_budget)
        cfg.max_paths = int(max_paths)
        cfg.count_work = 1 if count_work else 0
        cfg.leaf_size = int(leaf_size)
        cfg.bvh_width = int(bvh_width)
        cfg.extend_kernel = int(extend_kernel)
        cfg.refill_min = int(refill_min)
        cfg.pass_paths = int(pass_paths)
        cfg.grid_per_cu = int(grid_per_cu)
        cfg.shade_grid_per_cu = int(shade_grid_per_cu)
        cfg.sub_passes = int(sub_passes)
        cfg.graphs = int(graphs)
        cfg.stack_lds = int(stack_lds)
        cfg.ray_sort = int(ray_sort)
        builders = {"host": A.CRT_BVH_HOST, "device": A.CRT_BVH_DEVICE}
        if bvh_builder not in builders:
            raise ValueError("bvh_builder must be 'host' or 'device'")
        cfg.bvh_builder = builders[bvh_builder]
        cfg.batch_group_paths = int(batch_group_paths)
        orders = {"small_then_oldest": A.CRT_SERVE_SMALL_THEN_OLDEST, "round_robin": A.CRT_SERVE_ROUND_ROBIN}
        if serve_order not in orders:
            raise ValueError("serve_order must be 'small_then_oldest' or 'round_robin'")
        cfg.serve_order = orders[serve_order]
        cfg.serve_small_paths = int(serve_small_paths)
        h = C.c_void_p()
        A.check(self.lib.crt_create(C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.crt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- scenes
    def upload(self, scene):
        desc, keep = scene.to_desc()
        sid = C.c_uint32()
        A.check(self.lib.crt_scene_upload(self.h, C.byref(desc), C.byref(sid)))
        del keep
        return sid.value

    def upload_scene_file(self, path):
        """Loads a reference `.scene` file with the library's reader and uploads it.
        Returns (scene id, file info dict)."""
        desc, info = read_scene_file(path)
        sid = C.c_uint32()
        A.check(self.lib.crt_scene_upload(self.h, C.byref(desc._desc), C.byref(sid)))
        desc.close()
        return sid.value, info

    def release(self, sid):
        A.check(self.lib.crt_scene_release(self.h, sid))

    def info(self, sid):
        inf = A.crt_scene_info()
        A.check(self.lib.crt_scene_get_info(self.h, sid, C.byref(inf)))
        d = {k: getattr(inf, k) for k, _ in A.crt_scene_info._fields_}
        d["builder"] = {A.CRT_BVH_HOST: "host", A.CRT_BVH_DEVICE: "device"}.get(d["builder"], d["builder"])
        return d

    def verify(self, sid):
        """Reads the scene's BVH back from the device and checks its invariants (raises CrtError)."""
        A.check(self.lib.crt_scene_verify(self.h, sid))

    # ---------------------------------------------------------------- rendering
    def render(self, sid, width, height, spp, max_depth, seed=1, rng="pcg", sample_begin=0, sample_count=0,
               want_accum=True, time_stages=False, megakernel=False):
        """Blocking render. megakernel=True renders the request with one kernel launch (one thread
        per pixel running its samples and bounces in turn) instead of the wavefront passes."""
        flags = (A.CRT_RF_TIME_STAGES if time_stages else 0) | (A.CRT_RF_MEGAKERNEL if megakernel else 0)
        rd = render_desc(width, height, spp, max_depth, seed, rng, sample_begin, sample_count, flags)
        rgb = np.zeros((height, width, 3), np.uint8)
        acc = np.zeros((height, width, 4), np.float32) if want_accum else None
        img = A.crt_image(rgb.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(acc) if acc is not None else None)
        st = A.crt_stats()
        A.check(self.lib.crt_render(self.h, sid, C.byref(rd), C.byref(img), C.byref(st)))
        return rgb, acc, stats_dict(st)

    def autotune(self, sid, width, height, spp, max_depth, seed=1, rng="pcg", scene=None):
        """Times trial renders of this request over the launch/memory shape (and, with `scene`, the
        BVH shape, rebuilt in place) and keeps the fastest as the context's setting."""
        rd = render_desc(width, height, spp, max_depth, seed, rng)
        res = A.crt_tune_result()
        keep = None
        dp = None
        if scene is not None:
            desc, keep = scene.to_desc()
            dp = C.byref(desc)
        A.check(self.lib.crt_autotune(self.h, sid, dp, C.byref(rd), C.byref(res)))
        del keep
        return {k: getattr(res, k) for k, _ in A.crt_tune_result._fields_}

    def counters(self):
        """Work counters of the last blocking render (traversal counts need count_work=True)."""
        c = A.crt_work_counters()
        A.check(self.lib.crt_last_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in A.crt_work_counters._fields_}

    def memory(self):
        m = A.crt_memory_info()
        A.check(self.lib.crt_memory_get_info(self.h, C.byref(m)))
        return {k: getattr(m, k) for k, _ in A.crt_memory_info._fields_}

    def extend_launches(self):
        return self.counters()["extend_launches"]

    def render_batch(self, jobs, want_accum=True, time_stages=False):
        """jobs: list of dicts (sid, width, height, spp, max_depth, seed, rng; optional sample_begin,
        sample_count for one pass of a progressive render). Returns ([(rgb, accum)], stats) with every
        request rendered in fused passes."""
        n = len(jobs)
        items = (A.crt_batch_item * n)()
        outs = []
        for k, j in enumerate(jobs):
            W, H = j["width"], j["height"]
            rgb = np.zeros((H, W, 3), np.uint8)
            acc = np.zeros((H, W, 4), np.float32) if want_accum else None
            items[k].scene = j["sid"]
            items[k].rd = render_desc(W, H, j["spp"], j["max_depth"], j.get("seed", 1), j.get("rng", "pcg"),
                                      sample_begin=j.get("sample_begin", 0), sample_count=j.get("sample_count", 0),
                                      flags=A.CRT_RF_TIME_STAGES if time_stages else 0)
            items[k].out = A.crt_image(rgb.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(acc) if acc is not None else None)
            outs.append((rgb, acc))
        st = A.crt_stats()
        A.check(self.lib.crt_render_batch(self.h, items, n, C.byref(st)))
        return outs, stats_dict(st)

    def trace_rays(self, sid, rays, camera=False, stage=False):
        """Closest hits of n x 8 rays (origin, tmin, direction, tmax). stage=True answers them with the
        render's closest-hit stage (estimator interval, pass-through of one-sided lights seen from
        behind) instead of the query kernel."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        prim = np.zeros(n, np.int32)
        inst = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        flags = (A.CRT_TRACE_CAMERA if camera else 0) | (A.CRT_TRACE_STAGE if stage else 0)
        A.check(self.lib.crt_trace_rays(self.h, sid, _fp(rays), n, flags, _ip(prim), _ip(inst), _fp(t), _fp(uv)))
        return prim, inst, t, uv

    def trace_primary(self, sid, width, height):
        prim = np.zeros((height, width), np.int32)
        t = np.zeros((height, width), np.float32)
        A.check(self.lib.crt_trace_primary(self.h, sid, width, height, _ip(prim), _fp(t)))
        return prim, t

    # ---------------------------------------------------------------- serving
    def submit(self, sid, width, height, spp, max_depth, seed=1, rng="pcg", image_only=False, sample_begin=0,
               sample_count=0):
        """Queues a request in the serving loop. image_only: only the RGB8 image is copied back
        (wait(..., want_accum=False) then). sample_begin / sample_count: a sample range of the
        request, as for render() (one pass of a progressive render)."""
        rd = render_desc(width, height, spp, max_depth, seed, rng, sample_begin=sample_begin, sample_count=sample_count,
                         flags=A.CRT_RF_IMAGE_ONLY if image_only else 0)
        rid = C.c_uint64()
        A.check(self.lib.crt_submit(self.h, sid, C.byref(rd), C.byref(rid)))
        return rid.value

    def poll(self, rid):
        d = C.c_int32()
        A.check(self.lib.crt_poll(self.h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid, width, height, want_accum=True):
        rgb = np.zeros((height, width, 3), np.uint8)
        acc = np.zeros((height, width, 4), np.float32) if want_accum else None
        img = A.crt_image(rgb.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(acc) if acc is not None else None)
        st = A.crt_stats()
        A.check(self.lib.crt_wait(self.h, rid, C.byref(img), C.byref(st)))
        return rgb, acc, stats_dict(st)

    def cancel(self, rid):
        A.check(self.lib.crt_cancel(self.h, rid))

    def timing(self, rid):
        t = A.crt_request_timing()
        A.check(self.lib.crt_request_get_timing(self.h, rid, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.crt_request_timing._fields_}


class SceneFile:
    """A `.scene` file parsed by the library (crt_scenefile_load); owns the descriptor's memory."""

    def __init__(self, path):
        self.lib = A.load()
        self.h = C.c_void_p()
        A.check(self.lib.crt_scenefile_load(str(path).encode(), C.byref(self.h)))
        self._desc = A.crt_scene_desc()
        self._info = A.crt_scenefile_info()
        A.check(self.lib.crt_scenefile_desc(self.h, C.byref(self._desc), C.byref(self._info)))

    def to_desc(self):
        return self._desc, [self]

    def info(self):
        d = {k: getattr(self._info, k) for k, _ in A.crt_scenefile_info._fields_}
        d["output_path"] = d["output_path"].decode()
        return d

    def close(self):
        if self.h:
            self.lib.crt_scenefile_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_scene_file(path):
    f = SceneFile(path)
    return f, f.info()


def write_png(path, rgb):
    """RGB8 PNG writer (stbi_write_png replacement, core/src/gpu.cu:2110); rows top-down."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xffffffff)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def pixel_lines(rgb):
    """The reference's stdout pixel protocol 'p x y r g b' (core/src/gpu.cu:801): x left to right,
    y counted from the top row of the written image as H - j - 1 with j from the bottom."""
    h, w, _ = rgb.shape
    out = []
    for jj in range(h):
        for i in range(w):
            r, g, b = rgb[jj, i]
            out.append("p %d %d %d %d %d" % (i, jj, r, g, b))
    return out


def launch_gpu(scene, width, height, spp, max_depth, tx=16, ty=16, filepath=None, quiet=True, seed=2580,
               rng="minstd", renderer=None):
    """launchGPU(sceneConfig, width, height, spp, max_depth, tx, ty, filepath, quietMode)
    (core/src/renderers/gpu_cuda_renderer.h:17). tx/ty are accepted and ignored: the tile shape
    is the pass scheduler's. Defaults reproduce the reference's RNG stream (minstd, seed 2580)."""
    own = renderer is None
    r = renderer or Renderer()
    try:
        sid = r.upload(scene)
        rgb, _, st = r.render(sid, width, height, spp, max_depth, seed=seed, rng=rng, want_accum=False)
        r.release(sid)
    finally:
        if own:
            r.close()
    if filepath:
        write_png(filepath, rgb)
    if not quiet:
        for line in pixel_lines(rgb):
            print(line)
    return rgb, st
