"""GPU test of crt_render_noalloc, the variant of crt_render that allocates nothing (its passes run in
a static scratch array), called through the C-ABI: two calls in a row, the second with a smaller
image, each equal bit for bit to crt_render's output for the same request. The caller-owned
outputs are device memory allocated here with hipMalloc."""
import ctypes as C

import numpy as np
import pytest

from cortexraytracercuda_amd import Renderer, scenes, _abi as A
from cortexraytracercuda_amd.renderer import render_desc, stats_dict

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), n) == 0
        return p

    def free(self, p):
        assert self.lib.hipFree(p) == 0

    def to_host(self, p, arr):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(C.c_void_p), p, arr.nbytes, 2) == 0   # device to host
        return arr


def _noalloc(hip, r, sid, W, H, spp, depth, seed):
    acc_d = hip.malloc(W * H * 16)
    rgb_d = hip.malloc(W * H * 3)
    try:
        rd = render_desc(W, H, spp, depth, seed)
        st = A.crt_stats()
        A.check(r.lib.crt_render_noalloc(r.h, sid, C.byref(rd), acc_d, rgb_d, C.byref(st)))
        acc = hip.to_host(acc_d, np.zeros((H, W, 4), np.float32))
        rgb = hip.to_host(rgb_d, np.zeros((H, W, 3), np.uint8))
    finally:
        hip.free(acc_d)
        hip.free(rgb_d)
    return rgb, acc, stats_dict(st)


def test_noalloc_equals_render_on_two_calls():
    hip = _Hip()
    with Renderer() as r:
        sid = r.upload(scenes.sphere_mesh_scene(30000, seed=4))
        # the first request needs two passes of the scratch array's 1M path ids (160 x 120 x 64 paths),
        # the second is smaller; the scratch keeps the first call's contents when the second starts
        for W, H, spp, depth, seed in ((160, 120, 64, 8, 7), (64, 48, 4, 6, 9)):
            print("crt_render_noalloc %dx%d %d spp depth %d seed %d" % (W, H, spp, depth, seed), flush=True)
            rgb0, acc0, st0 = r.render(sid, W, H, spp, depth, seed=seed)
            rgb1, acc1, st1 = _noalloc(hip, r, sid, W, H, spp, depth, seed)
            assert np.array_equal(acc0, acc1, equal_nan=True)
            assert np.array_equal(rgb0, rgb1)
            assert st0["rays"] == st1["rays"] and st0["paths"] == st1["paths"]


def test_noalloc_rejects_what_it_cannot_render():
    hip = _Hip()
    with Renderer() as r:
        sid = r.upload(scenes.box_scene())
        acc_d = hip.malloc(32 * 32 * 16)
        try:
            st = A.crt_stats()
            rd = render_desc(32, 32, 4, 4, 2580, rng="minstd")
            assert r.lib.crt_render_noalloc(r.h, sid, C.byref(rd), acc_d, None, C.byref(st)) == A.CRT_ERR_INVALID_ARGUMENT
            host = np.zeros((32, 32, 4), np.float32)
            rd = render_desc(32, 32, 4, 4, 1)
            assert r.lib.crt_render_noalloc(r.h, sid, C.byref(rd), host.ctypes.data_as(C.c_void_p), None,
                                            C.byref(st)) == A.CRT_ERR_INVALID_ARGUMENT
            assert r.lib.crt_render_noalloc(r.h, sid, C.byref(rd), None, None, C.byref(st)) == A.CRT_ERR_INVALID_ARGUMENT
            # still usable after the refusals
            A.check(r.lib.crt_render_noalloc(r.h, sid, C.byref(rd), acc_d, None, C.byref(st)))
            acc = hip.to_host(acc_d, np.zeros((32, 32, 4), np.float32))
            assert np.all(acc[..., 3] == 4)
        finally:
            hip.free(acc_d)
