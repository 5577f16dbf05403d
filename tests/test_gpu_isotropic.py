"""GPU parity of the isotropic material (core/src/materials/isotropic.cuh with sphere_pdf,
core/src/pdf/sphere_pdf.cuh) and the Oren-Nayar material (core/src/materials/oren_nayar.cuh)
against the FP64 oracle: the box scene with a solid and a textured isotropic ball and an Oren-Nayar
ball, on the keyed streams and on the reference's minstd stream."""
import os

import numpy as np
import pytest

import oracle_py as O
from cortexraytracercuda_amd import scenes

pytestmark = pytest.mark.gpu
NT = os.cpu_count() or 4


def _iso_box():
    s = scenes.box_scene()
    s.checker_texture("iso_check", 0.1, (0.9, 0.6, 0.2), (0.2, 0.3, 0.8))
    s.isotropic("iso_solid", (0.7, 0.7, 0.7))
    s.isotropic("iso_tex", texture="iso_check")
    s.oren_nayar("rough", (0.8, 0.5, 0.3), roughness=0.6, albedo_temp=0.5)
    s.spheres[0]["material"] = s.material_names["iso_solid"]
    s.spheres[1]["material"] = s.material_names["iso_tex"]
    s.sphere((0.45, -0.75, 0.2), 0.25, "rough")
    return s


@pytest.mark.parametrize("rng", ["pcg", "minstd"])
def test_isotropic_render_against_fp64(renderer, rng):
    s = _iso_box()
    W = H = 64
    spp, depth = 16, 8
    sid = renderer.upload(s)
    _, acc, st = renderer.render(sid, W, H, spp, depth, seed=12, rng=rng)
    renderer.release(sid)
    oacc, _ = O.render(s, W, H, spp, depth, seed=12, rng=rng, precision=64, threads=NT)
    acc32, ost32 = O.render(s, W, H, spp, depth, seed=12, rng=rng, precision=32, threads=NT)
    g, o, o32 = acc[..., :3] / spp, oacc[..., :3] / spp, acc32[..., :3] / spp
    close = np.all(np.abs(g - o) <= 1e-3 * np.maximum(o, 1.0), axis=-1)
    close32 = np.all(np.abs(o32 - o) <= 1e-3 * np.maximum(o, 1.0), axis=-1)
    rmse, rmse32 = float(np.sqrt(np.mean((g - o) ** 2))), float(np.sqrt(np.mean((o32 - o) ** 2)))
    print("isotropic box (%s): pixels within 1e-3 %.5f (fp32 oracle %.5f) rmse %.6f (fp32 oracle %.6f) rays %d vs %d"
          % (rng, close.mean(), close32.mean(), rmse, rmse32, st["rays"], ost32["rays"]), flush=True)
    assert np.all(acc[..., 3] == spp)
    # keyed streams: one float decision changes one path, and the GPU matches the FP32 oracle's
    # distance to oracle64 pixel for pixel. The minstd stream runs on through a pixel's samples, so
    # one decision taken the other way changes every later sample of that pixel, and an isotropic
    # surface sends half its rays back into the surface it sits on, where FP32 self-hits just past
    # the 1e-5 floor are decided by the last ulp of the hit point (measured: 0.80 of pixels within
    # 1e-3 on the GPU, 0.94 for the FP32 oracle, whose rounding differs). There the image is held
    # to the RMSE bound and the ray count below.
    if rng == "pcg":
        assert close.mean() >= 0.97 and close.mean() >= close32.mean() - 0.01
        assert rmse <= 1.5 * rmse32 + 1e-3
        assert abs(st["rays"] - ost32["rays"]) <= 0.002 * ost32["rays"]
    else:
        # with the minstd stream the FP32 oracle is the reference-faithful renderer (its sphere
        # quadratic with cancellation decides the self-hits differently from the GPU's robust test,
        # measured 0.4 % more rays), so the yardstick is oracle64 on another seed: the GPU image is
        # no farther from oracle64 than that, and its mean radiance agrees
        ob, _ = O.render(s, W, H, spp, depth, seed=13, rng=rng, precision=64, threads=NT)
        rmse_seed = float(np.sqrt(np.mean((ob[..., :3] / spp - o) ** 2)))
        print("minstd: rmse %.6f, oracle64 seed-to-seed %.6f" % (rmse, rmse_seed), flush=True)
        assert rmse <= rmse_seed
        assert abs(float(g.mean()) - float(o.mean())) <= 0.03 * float(o.mean())
