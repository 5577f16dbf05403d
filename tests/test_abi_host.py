"""CPU tests of the C-ABI library: it loads, exports every declared symbol, agrees with the ctypes
mirror on struct layout, and its host-side scene flattening / BVH build is valid (no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cortexraytracercuda_amd import Scene, _abi as A, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "cortex_rt.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(crt_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = A.load()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(A.EXPORTED) == names


def test_struct_layout_matches_ctypes():
    lib = A.load()
    n = len(A.STRUCTS_IN_ABI_ORDER)
    sizes = (C.c_uint64 * n)()
    m = lib.crt_struct_sizes(sizes, n)
    assert m == n
    for k, s in enumerate(A.STRUCTS_IN_ABI_ORDER):
        assert sizes[k] == C.sizeof(s), s.__name__


def _check(scene, leaf=4, width=4):
    desc, keep = scene.to_desc()
    info = A.crt_scene_info()
    rc = A.load().crt_scene_build_check(C.byref(desc), leaf, width, C.byref(info))
    return rc, {k: getattr(info, k) for k, _ in A.crt_scene_info._fields_}


def test_bvh_invariants_box_and_large():
    rc, info = _check(scenes.box_scene())
    assert rc == 0, A.load().crt_last_error()
    assert info["num_prims"] == 12 + 3 + 1
    rc, info = _check(scenes.large_scene(levels=5, n_spheres=500))
    assert rc == 0, A.load().crt_last_error()
    assert info["num_prims"] == 20 * 4 ** 5 + 500 + 1
    assert info["max_depth"] <= 64


@pytest.mark.parametrize("leaf", [1, 2, 4, 8])
@pytest.mark.parametrize("width", [2, 4])
def test_bvh_invariants_leaf_sizes(leaf, width):
    rc, _ = _check(scenes.large_scene(levels=3, n_spheres=100), leaf, width)
    assert rc == 0, A.load().crt_last_error()


def test_leaf_size_reported_and_bounded():
    """Leaf references carry 1-8 primitives: 0 selects the default (1), larger sizes are capped at 8.
    With one primitive per leaf every primitive has its own leaf."""
    s = scenes.large_scene(levels=3, n_spheres=100)
    for asked, built in ((0, 1), (1, 1), (3, 3), (8, 8), (20, 8)):
        rc, info = _check(s, asked, 4)
        assert rc == 0, A.load().crt_last_error()
        assert info["leaf_size"] == built
        assert info["num_prims"] / built <= info["num_leaves"] <= info["num_prims"]
    rc, info = _check(s, 1, 4)
    assert info["num_leaves"] == info["num_prims"]


@pytest.mark.parametrize("width", [2, 4])
def test_bvh_invariants_small_and_instanced(width):
    for s in (scenes.box_scene(), Scene(), scenes.instanced_scene(n_instances=20, n_textures=1, tex_size=8, nu=6, nv=4)):
        rc, _ = _check(s, 4, width)
        assert rc == 0, A.load().crt_last_error()


def test_bvh_invariants_instanced():
    s = scenes.instanced_scene(n_instances=64, n_textures=2, tex_size=32, nu=10, nv=8)
    rc, info = _check(s)
    assert rc == 0, A.load().crt_last_error()
    assert info["num_instances"] == 64


def test_sah_beats_median_depth():
    # binned SAH on a clustered scene: fewer expected node visits than a balanced split's depth
    rc, info = _check(scenes.large_scene(levels=4, n_spheres=200))
    assert rc == 0
    assert info["sah_cost"] > 0


def test_invalid_scenes_are_rejected():
    lib = A.load()
    s = Scene()
    s.lambertian("m", (0.5, 0.5, 0.5))
    s.sphere((0, 0, 0), 1.0, 0)
    s.spheres[0]["material"] = 7          # dangling material index
    rc, _ = _check(s)
    assert rc == A.CRT_ERR_INVALID_SCENE
    assert b"material" in lib.crt_last_error()
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]]), m)   # index out of range
    rc, _ = _check(s)
    assert rc == A.CRT_ERR_INVALID_SCENE
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5), texture=3)                 # missing texture
    s.sphere((0, 0, 0), 1.0, m)
    rc, _ = _check(s)
    assert rc == A.CRT_ERR_INVALID_SCENE
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.sphere((0, float("nan"), 0), 1.0, m)                             # NaN geometry
    rc, _ = _check(s)
    assert rc == A.CRT_ERR_INVALID_SCENE
    desc, keep = Scene().to_desc()
    desc.abi_version = 99
    assert lib.crt_scene_build_check(C.byref(desc), 4, 4, None) == A.CRT_ERR_INVALID_ARGUMENT
    desc, keep = Scene().to_desc()
    assert lib.crt_scene_build_check(C.byref(desc), 4, 3, None) == A.CRT_ERR_INVALID_ARGUMENT


def test_empty_scene_builds():
    rc, info = _check(Scene())
    assert rc == 0
    assert info["num_prims"] == 0


def test_no_device_is_reported_not_crashed():
    lib = A.load()
    n = C.c_int32(-1)
    lib.crt_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = lib.crt_create(None, C.byref(h))
    assert rc == A.CRT_ERR_NO_DEVICE
    assert not h.value


def _tiny_scenes():
    out = {"empty": Scene()}
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.sphere((0.0, 0.0, 0.0), 1.0, m)
    out["one_sphere"] = s
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.mesh([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 1.0, 0.0]], [[0, 1, 2]], m)
    out["one_triangle"] = s
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.quad((-1.0, -1.0, 0.0), (2.0, 0.0, 0.0), (0.0, 2.0, 0.0), m)
    out["one_quad"] = s
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    rng = np.random.default_rng(5)
    for _ in range(37):
        s.sphere(tuple(rng.uniform(-1.5, 1.5, 3)), float(rng.uniform(0.05, 0.3)), m)
    out["spheres_only"] = s
    return out


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("leaf", [1, 4])
def test_bvh_invariants_tiny_scenes(width, leaf):
    # one-primitive and empty trees (the root wraps the leaf, or holds only empty child slots)
    for name, s in _tiny_scenes().items():
        rc, info = _check(s, leaf=leaf, width=width)
        assert rc == 0, (name, A.load().crt_last_error())
        assert info["num_prims"] == s.num_world_prims, name
        assert info["num_nodes"] >= 1, name
