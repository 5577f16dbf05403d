"""ctypes front end of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py, never by the product package."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class oracle_opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("accel", C.c_int32), ("threads", C.c_int32), ("row_begin", C.c_uint32),
                ("row_end", C.c_uint32), ("pixel_stride", C.c_uint32), ("ref_style", C.c_int32)]


class oracle_stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays", C.c_uint64), ("retraces", C.c_uint64), ("light_pdf_tests", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64), ("seconds", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, os.path.join(HERE, "liboracle.so")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.c_void_p)


def render(scene, width, height, spp, max_depth, seed=1, rng="pcg", precision=64, accel=0, threads=1, rows=None,
           pixel_stride=1, ref_style=None):
    """CPU render. rng='minstd' + precision=32 + ref_style=True is the reference-faithful renderer;
    rng='pcg' + precision=64 is the FP64 oracle. Returns (accum HxWx4 float32, stats dict)."""
    from cortexraytracercuda_amd import render_desc
    desc, keep = scene.to_desc()
    rd = render_desc(width, height, spp, max_depth, seed, rng)
    o = oracle_opts()
    o.precision = precision
    o.accel = accel
    o.threads = threads
    o.row_begin, o.row_end = (rows if rows else (0, 0))
    o.pixel_stride = pixel_stride
    if ref_style is None:
        ref_style = rng == "minstd"
    o.ref_style = 1 if ref_style else 0
    acc = np.zeros((height, width, 4), np.float32)
    st = oracle_stats()
    rc = lib().oracle_render(C.byref(desc), C.byref(rd), C.byref(o), _fp(acc), C.byref(st))
    if rc:
        raise RuntimeError("oracle_render failed: %d" % rc)
    del keep
    return acc, {k: getattr(st, k) for k, _ in oracle_stats._fields_}


def trace(scene, rays, camera=False, precision=64, accel=1, ref_style=False, normals=False):
    """Closest hits: (prim, inst, t, uv), and with normals=True also the hit record's normal (the
    face the ray meets, as set_face_normal leaves it)."""
    desc, keep = scene.to_desc()
    rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 8)
    n = rays.shape[0]
    prim = np.zeros(n, np.int32)
    inst = np.zeros(n, np.int32)
    t = np.zeros(n, np.float64)
    uv = np.zeros((n, 2), np.float64)
    nrm = np.zeros((n, 3), np.float64)
    f = lib().oracle_trace
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_void_p]
    rc = f(C.cast(C.pointer(desc), C.c_void_p), _fp(rays), n, 1 if camera else 0, precision, accel,
           1 if ref_style else 0, _fp(prim), _fp(inst), _fp(t), _fp(uv), _fp(nrm) if normals else None)
    if rc:
        raise RuntimeError("oracle_trace failed: %d" % rc)
    del keep
    if normals:
        return prim, inst, t, uv, nrm
    return prim, inst, t, uv


def primary_rays(scene, width, height):
    desc, keep = scene.to_desc()
    rays = np.zeros((height * width, 8), np.float64)
    f = lib().oracle_primary_rays
    f.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    f(C.cast(C.pointer(desc), C.c_void_p), width, height, _fp(rays))
    del keep
    return rays


def tonemap(accum, spp):
    a = np.ascontiguousarray(accum, dtype=np.float32).reshape(-1, 4)
    out = np.zeros((a.shape[0], 3), np.uint8)
    f = lib().oracle_tonemap
    f.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]
    f(_fp(a), a.shape[0], spp, _fp(out))
    return out.reshape(tuple(accum.shape[:-1]) + (3,))


def minstd(seed, n):
    out = np.zeros(n, np.uint32)
    u = np.zeros(n, np.float32)
    f = lib().oracle_minstd_stream
    f.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    f(seed, n, _fp(out), _fp(u))
    return out, u


def minstd_uniform_int(seed, n, lo, hi):
    out = np.zeros(n, np.int32)
    f = lib().oracle_minstd_uniform_int
    f.argtypes = [C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, C.c_void_p]
    f(seed, n, lo, hi, _fp(out))
    return out


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    f = lib().oracle_pcg32_stream
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p]
    f(initstate, initseq, n, _fp(out))
    return out


def keyed(seed, pixel, sample, bounce, n):
    out = np.zeros(n, np.uint32)
    f = lib().oracle_keyed_stream
    f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    f(seed, pixel, sample, bounce, n, _fp(out))
    return out


def light_pdf(scene, origins, dirs):
    desc, keep = scene.to_desc()
    o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(o.shape[0])
    f = lib().oracle_light_pdf
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    f(C.cast(C.pointer(desc), C.c_void_p), _fp(o), _fp(v), o.shape[0], _fp(out))
    del keep
    return out


def light_sample(scene, origins, seed=1):
    """lights_random (FP64) from each origin: the direction towards a sample of light 0."""
    desc, keep = scene.to_desc()
    o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(o)
    f = lib().oracle_light_sample
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]
    f(C.cast(C.pointer(desc), C.c_void_p), _fp(o), o.shape[0], seed, _fp(out))
    del keep
    return out


def ref_bvh_info(scene):
    desc, keep = scene.to_desc()
    nodes = C.c_uint64()
    depth = C.c_uint32()
    f = lib().oracle_ref_bvh_info
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    f(C.cast(C.pointer(desc), C.c_void_p), C.byref(nodes), C.byref(depth))
    del keep
    return nodes.value, depth.value


def camera_basis(scene, width=0):
    """setup_camera in FP64 for an image `width` columns wide (the orthographic viewport follows the
    image): dict of origin, llc, horizontal, vertical, u, v, w, lens_radius."""
    desc, keep = scene.to_desc()
    out = np.zeros(22)
    f = lib().oracle_camera_basis
    f.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    f(C.cast(C.pointer(desc), C.c_void_p), width, _fp(out))
    del keep
    names = ["origin", "llc", "horizontal", "vertical", "u", "v", "w"]
    d = {k: out[3 * i:3 * i + 3].copy() for i, k in enumerate(names)}
    d["lens_radius"] = float(out[21])
    return d


def camera_rays(scene, uv, seed=1, width=0):
    """camera_ray at image-plane points (u, v) in FP64 for an image `width` columns wide:
    (origins, times, directions)."""
    desc, keep = scene.to_desc()
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((uv.shape[0], 8))
    f = lib().oracle_camera_rays
    f.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p]
    f(C.cast(C.pointer(desc), C.c_void_p), width, uv.shape[0], _fp(uv), seed, _fp(out))
    del keep
    return out[:, 0:3], out[:, 3], out[:, 4:7]


def metal(dirs, normals, fuzz, seed=1):
    d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    n = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(d)
    f = lib().oracle_metal
    f.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_double, C.c_uint64, C.c_void_p]
    f(d.shape[0], _fp(d), _fp(n), fuzz, seed, _fp(out))
    return out


def dielectric(dirs, normals, front, ior, draws):
    """Returns (directions, reflected flags, draw-consumed flags)."""
    d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    n = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
    fr = np.ascontiguousarray(np.broadcast_to(np.asarray(front, dtype=np.int32), (d.shape[0],)))
    dr = np.ascontiguousarray(np.broadcast_to(np.asarray(draws, dtype=np.float64), (d.shape[0],)))
    out = np.zeros_like(d)
    refl = np.zeros(d.shape[0], np.int32)
    used = np.zeros(d.shape[0], np.int32)
    f = lib().oracle_dielectric
    f.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    f(d.shape[0], _fp(d), _fp(n), _fp(fr), ior, _fp(dr), _fp(out), _fp(refl), _fp(used))
    return out, refl.astype(bool), used.astype(bool)


def sphere_uv(points, mapping=(1, 1, 0, 0, 1, 1), precision=64):
    """get_sphere_uv; mapping = (scale_u, scale_v, offset_u, offset_v, repeat_u, repeat_v)."""
    from cortexraytracercuda_amd import _abi as A
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    m = A.crt_uvmap(*[float(x) for x in mapping])
    out = np.zeros((p.shape[0], 2))
    f = lib().oracle_sphere_uv
    f.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    f(p.shape[0], _fp(p), C.cast(C.pointer(m), C.c_void_p), precision, _fp(out))
    return out


def tex_value(scene, tex, u, v, points):
    """tex_value of texture index `tex` at texture coordinates (u, v) and hit points."""
    desc, keep = scene.to_desc()
    u = np.asarray(u, dtype=np.float64).reshape(-1)
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(np.column_stack([u, v, p]))
    out = np.zeros((q.shape[0], 3))
    f = lib().oracle_tex_value
    f.argtypes = [C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p]
    rc = f(C.cast(C.pointer(desc), C.c_void_p), tex, q.shape[0], _fp(q), _fp(out))
    del keep
    if rc:
        raise RuntimeError("oracle_tex_value failed: %d" % rc)
    return out
