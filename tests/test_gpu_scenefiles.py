"""Scene files the reference ships (data/scenes/all_lights_types.scene: a quad light and seven
spheres, inactive spot and omni lights; data/scenes/simple_spot_light.scene: a spot light, two
spheres and a box), read by crt_scenefile_load and rendered on the GPU: first hits and a
reduced-resolution image at the file's own depth against the FP64 oracle on the same descriptor.
The files are kept byte for byte under tests/golden/scenes as input data (they name no mesh or
texture file)."""
import os

import numpy as np
import pytest

import oracle_py as O
from cortexraytracercuda_amd import SceneFile

from test_gpu_parity import _first_hit_parity, _same_rays

pytestmark = pytest.mark.gpu
NT = os.cpu_count() or 4
HERE = os.path.dirname(os.path.abspath(__file__))
SHIPPED = ["shipped_all_lights_types.scene", "shipped_simple_spot_light.scene"]


@pytest.mark.parametrize("name", SHIPPED)
def test_shipped_scene_first_hits_and_image_against_fp64(renderer, name):
    f = SceneFile(os.path.join(HERE, "golden", "scenes", name))
    info = f.info()
    assert info["unsupported"] == 0 and info["skipped"] == 0
    W = 128
    H = int(round(W * info["height"] / info["width"]))
    r32, r64 = _same_rays(f, W, H)
    sid = renderer.upload(f)
    prim, _, t, _ = renderer.trace_rays(sid, r32, camera=True)
    oprim, _, ot, _ = O.trace(f, r64, camera=True, precision=64, accel=1)
    _first_hit_parity(prim, t, oprim, ot, name, min_agree=1.0)
    assert (oprim >= 0).mean() > 0.3
    spp, depth = 16, info["max_depth"]
    rgb, acc, st = renderer.render(sid, W, H, spp, depth, seed=7)
    oacc, ost = O.render(f, W, H, spp, depth, seed=7, rng="pcg", precision=64, threads=NT)
    renderer.release(sid)
    g, o = acc[..., :3] / spp, oacc[..., :3] / spp
    close = np.all(np.abs(g - o) <= 1e-3 * np.maximum(o, 1.0), axis=-1)
    rmse = float(np.sqrt(np.mean((g - o) ** 2)))
    scale = max(1.0, float(np.abs(o).mean()))
    print("%s %dx%d %d spp depth %d: pixels within 1e-3 %.5f rmse %.5f (mean %.3f) rays %d vs %d"
          % (name, W, H, spp, depth, close.mean(), rmse, np.abs(o).mean(), st["rays"], ost["rays"]), flush=True)
    assert np.all(acc[..., 3] == spp)
    assert float(o.max()) > 0.1       # the lights reach the image
    assert close.mean() >= 0.97
    assert rmse <= 0.1 * scale
    assert abs(st["rays"] - ost["rays"]) <= 0.01 * ost["rays"]
