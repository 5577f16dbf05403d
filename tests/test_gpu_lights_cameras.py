"""GPU parity of the omni and spot lights (core/src/lights/omni_light.cuh, spot_light.cuh,
core/src/materials/diffuse_spot_light.cuh) and of the orthographic camera
(core/src/cameras/orthographic_camera.cuh) against the FP64 oracle: first hits, images, the spot's
pass-through, device-built trees."""
import os

import numpy as np
import pytest

import oracle_py as O
from cortexraytracercuda_amd import Renderer, scenes

from test_gpu_parity import _first_hit_parity, _same_rays

pytestmark = pytest.mark.gpu
NT = os.cpu_count() or 4


def _lit_box(invisible):
    """The box scene lit by an omni light, a spot light aimed down at the floor and its area light
    (omni first and spot last in the world light order; the omni light is the sampled one)."""
    s = scenes.box_scene()
    s.omni_light((0.55, 0.45, -0.3), 0.12, rgb=(1.0, 0.8, 0.6), intensity=6.0, invisible=invisible)
    s.spot_light((-0.5, 0.7, -0.5), (0.2, -1.0, 0.1), 35.0, 4.0, 0.08, rgb=(0.6, 0.8, 1.0), intensity=20.0,
                 invisible=invisible)
    return s


def _compare(g, o, label, min_close, max_rmse):
    diff = np.abs(g - o)
    close = np.all(diff <= 1e-3 * np.maximum(o, 1.0), axis=-1)
    rmse = float(np.sqrt(np.mean((g - o) ** 2)))
    print("%s: pixels within 1e-3 %.5f rmse %.6f" % (label, close.mean(), rmse), flush=True)
    assert close.mean() >= min_close, (label, close.mean())
    assert rmse <= max_rmse, (label, rmse)


@pytest.mark.parametrize("invisible", [False, True])
def test_sphere_lights_first_hits(renderer, invisible):
    s = _lit_box(invisible)
    W = H = 192
    r32, r64 = _same_rays(s, W, H)
    sid = renderer.upload(s)
    for camera in (True, False):
        prim, _, t, _ = renderer.trace_rays(sid, r32, camera=camera)
        oprim, _, ot, _ = O.trace(s, r64, camera=camera, precision=64, accel=1)
        _first_hit_parity(prim, t, oprim, ot, "lit box camera=%s invisible=%s" % (camera, invisible), min_agree=1.0)
        lights_seen = np.isin(oprim, [0, 2]).sum()   # omni = id 0, spot = id 2 (after the area light)
        if camera and invisible:
            assert lights_seen == 0
        else:
            assert lights_seen > 0
    renderer.release(sid)


@pytest.mark.parametrize("invisible", [False, True])
def test_sphere_lights_render_vs_fp64(renderer, invisible):
    s = _lit_box(invisible)
    W = H = 64
    spp, depth = 16, 8
    sid = renderer.upload(s)
    rgb, acc, st = renderer.render(sid, W, H, spp, depth, seed=3)
    oacc, ost = O.render(s, W, H, spp, depth, seed=3, rng="pcg", precision=64, threads=NT)
    assert np.all(acc[..., 3] == spp)
    _compare(acc[..., :3] / spp, oacc[..., :3] / spp, "lit box invisible=%s" % invisible, 0.97, 0.05)
    renderer.release(sid)


def test_spot_light_back_face_pass_through():
    """Camera rays from inside a visible spot light meet its back face: white (colour(1,1,1,0))
    is added and the re-trace continues behind the light; the re-traces are counted as rays."""
    s = scenes.box_scene()
    s.spot_light((0.0, 0.0, 0.9), (0.0, 0.0, 1.0), 30.0, 2.0, 0.25, rgb=(1, 1, 1), intensity=5.0, invisible=False)
    W = H = 32
    with Renderer() as r:
        sid = r.upload(s)
        rgb, acc, st = r.render(sid, W, H, 4, 6, seed=9)
    oacc, ost = O.render(s, W, H, 4, 6, seed=9, rng="pcg", precision=64, threads=NT)
    assert st["retraces"] >= W * H * 4 and ost["retraces"] >= W * H * 4
    # the same paths but for the few whose float decisions differ from the FP64 oracle's at some bounce
    assert abs(st["rays"] - ost["rays"]) <= 0.01 * ost["rays"]
    _compare(acc[..., :3] / 4, oacc[..., :3] / 4, "inside a spot light", 0.97, 0.05)


@pytest.mark.parametrize("builder", ["host", "device"])
def test_sphere_lights_device_built_equal_host_built(builder):
    s = _lit_box(False)
    with Renderer(bvh_builder=builder) as r:
        sid = r.upload(s)
        assert r.info(sid)["builder"] == builder
        r.verify(sid)
        out = r.render(sid, 48, 48, 9, 8, seed=4)
    with Renderer() as r:
        sid = r.upload(s)
        ref = r.render(sid, 48, 48, 9, 8, seed=4)
    assert np.array_equal(out[1], ref[1], equal_nan=True)


def test_orthographic_camera_first_hits_and_render():
    s = scenes.box_scene()
    s.set_camera((0.3, 0.2, 0.95), (0.1, -0.2, -1.0), vfov=70.0, aspect_ratio=1.0, aperture=0.0,
                 orthographic=True, ortho_height=0.8)
    W = H = 128
    r32, r64 = _same_rays(s, W, H)
    assert np.ptp(r64[:, 0]) > 0.5 and np.allclose(r64[:, 4:7], r64[0, 4:7])   # parallel rays
    with Renderer() as r:
        sid = r.upload(s)
        prim, _, t, _ = r.trace_rays(sid, r32, camera=True)
        pprim, _ = r.trace_primary(sid, W, H)
        rgb, acc, st = r.render(sid, 48, 48, 16, 8, seed=6)
    oprim, _, ot, _ = O.trace(s, r64, camera=True, precision=64, accel=1)
    # the viewport reaches past the box walls: some rays start just outside and hit a wall at small t
    _first_hit_parity(prim, t, oprim, ot, "orthographic box", min_agree=1.0, t_floor=0.05)
    assert (pprim.reshape(-1) == prim).mean() >= 0.999
    oacc, _ = O.render(s, 48, 48, 16, 8, seed=6, rng="pcg", precision=64, threads=NT)
    _compare(acc[..., :3] / 16, oacc[..., :3] / 16, "orthographic box", 0.97, 0.05)
