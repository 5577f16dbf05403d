"""GPU parity of the device BVH builder (linear BVH: Morton codes, radix sort, radix tree, bottom-up
boxes, collapse to the BVH width). The tree read back from the device satisfies the structural
invariants; closest hits through it equal the FP64 oracle's; and images rendered through it equal,
bit for bit, the images rendered through the host's binned-SAH tree (closest hits do not depend on
the tree under the deterministic tie rule)."""
import numpy as np
import pytest

import oracle_py as O
from cortexraytracercuda_amd import Renderer, scenes
from cortexraytracercuda_amd.scene import Scene

from test_gpu_parity import _first_hit_parity, _same_rays

pytestmark = pytest.mark.gpu


def _mixed_scene():
    """Every one-level primitive kind and side array: area lights, spheres (with uv mappings), quads,
    a textured mesh with per-vertex normals and uvs, per-triangle materials, a one-sided emitter."""
    s = scenes.sphere_mesh_scene(20000, seed=3)
    tex = s.image_texture("noise", scenes.checker_noise_texture(64, 5))
    tm = s.lambertian("textured", (1.0, 1.0, 1.0), texture=tex)
    pos, nrm, uv, f = scenes.torus(0.25, 0.08, 24, 16)
    pos = pos + np.array([0.55, 0.3, -0.6], np.float32)
    s.mesh(pos.astype(np.float32), f.astype(np.uint32), tm, normals=nrm.astype(np.float32), uvs=uv.astype(np.float32))
    s.quad((-0.9, -0.95, -0.95), (0.5, 0.0, 0.0), (0.0, 0.4, 0.0), "metal_ball")
    s.sphere((-0.4, 0.4, -0.7), 0.15, tm, mapping=(2.0, 1.0, 0.25, 0.0, 1.0, 1.0))
    em = s.diffuse_light("panel", (3.0, 2.0, 1.0), intensity=1.0, one_sided=True)
    s.quad((0.6, -0.2, -0.99), (0.0, 0.3, 0.0), (0.3, 0.0, 0.0), em)
    return s


@pytest.mark.parametrize("width", [4, 2])
def test_device_tree_invariants_and_first_hits_on_the_large_scene(width):
    """The large scene at its full size (1,310,720 triangles + 500 spheres + an area light)."""
    s = scenes.large_scene(levels=8, n_spheres=500)
    W, H = 320, 180
    r32, r64 = _same_rays(s, W, H)
    with Renderer(bvh_builder="device", bvh_width=width) as r:
        sid = r.upload(s)
        info = r.info(sid)
        print(info, flush=True)
        assert info["builder"] == "device" and info["bvh_width"] == width
        assert info["num_prims"] == 1310720 + 500 + 1
        r.verify(sid)
        prim, _, t, _ = r.trace_rays(sid, r32, camera=True)
        sprim, _, st, _ = r.trace_rays(sid, r32, camera=True, stage=True)
    oprim, _, ot, _ = O.trace(s, r64, camera=True, precision=64, accel=0)
    _first_hit_parity(prim, t, oprim, ot, "large device-built")
    _first_hit_parity(sprim, st, oprim, ot, "large device-built stage")


@pytest.mark.parametrize("width", [4, 2])
def test_device_built_images_equal_host_built(width):
    s = _mixed_scene()
    imgs = {}
    for builder in ("host", "device"):
        with Renderer(bvh_builder=builder, bvh_width=width) as r:
            sid = r.upload(s)
            assert r.info(sid)["builder"] == builder
            r.verify(sid)
            imgs[builder] = r.render(sid, 96, 96, 16, 8, seed=11)
    (rgb0, acc0, st0), (rgb1, acc1, st1) = imgs["host"], imgs["device"]
    assert np.array_equal(acc0, acc1, equal_nan=True)
    assert np.array_equal(rgb0, rgb1)
    assert st0["rays"] == st1["rays"]


def test_device_built_box_first_hits_exact():
    """The box scene's diagonal rays through shared triangle edges, on a device-built tree."""
    box = scenes.box_scene()
    r32, r64 = _same_rays(box, 256, 256)
    with Renderer(bvh_builder="device") as r:
        sid = r.upload(box)
        r.verify(sid)
        prim, _, t, _ = r.trace_rays(sid, r32, camera=True)
        sprim, _, st, _ = r.trace_rays(sid, r32, camera=True, stage=True)
    oprim, _, ot, _ = O.trace(box, r64, camera=True, precision=64, accel=1)
    _first_hit_parity(prim, t, oprim, ot, "box device-built", min_agree=1.0)
    _first_hit_parity(sprim, st, oprim, ot, "box device-built stage", min_agree=1.0)


def test_small_and_instanced_scenes_fall_back_to_the_host_build():
    one = Scene().set_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    one.sphere((0.0, 0.0, 0.0), 1.0, one.lambertian("m", (0.5, 0.5, 0.5)))
    inst = scenes.instanced_scene(n_instances=8, tex_size=16, nu=8, nv=8)
    with Renderer(bvh_builder="device") as r:
        for sc in (one, inst):
            sid = r.upload(sc)
            assert r.info(sid)["builder"] == "host"
            r.verify(sid)


def test_verify_accepts_host_built_scenes():
    with Renderer() as r:
        for sc in (scenes.box_scene(), scenes.instanced_scene(n_instances=16, tex_size=16, nu=8, nv=8)):
            sid = r.upload(sc)
            r.verify(sid)
