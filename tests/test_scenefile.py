"""The `.scene` reader (crt_scenefile_load) and the command-line renderer's argument handling,
on CPU: syntax, key mapping and reference semantics, checked against a hand-built Scene."""
import os
import glob

import numpy as np
import pytest

import oracle_py as O
from cortexraytracercuda_amd import SceneFile, _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
BOX = os.path.join(HERE, "golden", "scenes", "box_c1.scene")


def _arr(ptr, n):
    return [ptr[k] for k in range(n)]


def test_box_scene_file_mapping():
    f = SceneFile(BOX)
    d, _ = f.to_desc()
    info = f.info()
    assert (info["width"], info["height"], info["spp"], info["max_depth"]) == (256, 256, 16, 8)
    assert abs(info["aspect_ratio"] - 1.0) < 1e-7
    assert info["skipped"] == 1            # the sphere whose material does not exist
    assert info["unsupported"] == 1        # the cone
    assert d.num_lights == 1 and d.num_quads == 6 and d.num_spheres == 3 and d.num_meshes == 2
    L = d.lights[0]
    assert list(L.center) == pytest.approx([0, 0.99, 0]) and L.invisible == 1 and list(L.rgb) == [15, 15, 15]
    mats = _arr(d.materials, d.num_materials)
    red = mats[1]
    assert red.type == A.CRT_MAT_LAMBERTIAN and list(red.albedo) == pytest.approx([166 / 255, 13 / 255, 13 / 255])
    metal = [m for m in mats if m.type == A.CRT_MAT_METAL][0]
    assert metal.fuzz == pytest.approx(0.1)
    glass = [m for m in mats if m.type == A.CRT_MAT_DIELECTRIC][0]
    assert glass.ior == pytest.approx(1.5)
    floor = mats[3]
    tex = d.textures[floor.texture]
    assert tex.type == A.CRT_TEX_CHECKER and tex.scale == pytest.approx(0.25)
    assert list(d.textures[tex.odd].color) == pytest.approx([0.2, 0.2, 0.2])
    # box: centre +- size/2, 12 triangles; OBJ: 4 triangles, flat (use_smoothing = false)
    box = d.meshes[0]
    pos = np.ctypeslib.as_array(box.positions, (24,)).reshape(8, 3)
    assert pos.min(0) == pytest.approx([-0.65, -1.0, 0.15]) and pos.max(0) == pytest.approx([-0.35, -0.7, 0.45])
    assert box.num_triangles == 12
    obj = d.meshes[1]
    assert obj.num_triangles == 4 and not obj.normals
    assert d.camera.vfov_deg == pytest.approx(70.0) and list(d.camera.lookfrom) == pytest.approx([0, 0, 0.9])


def test_scene_file_renders_on_the_oracle():
    # the descriptor from the file is a complete scene: the CPU oracle renders it
    f = SceneFile(BOX)
    acc, st = O.render(f, 32, 32, 4, 6, seed=1, rng="pcg", precision=64, threads=os.cpu_count() or 2)
    assert st["rays"] > 32 * 32 * 4
    assert np.isfinite(acc).all() and acc[..., :3].mean() > 0


@pytest.mark.parametrize("text,err", [
    ("image: { width = 10 ", "expected"),
    ("image = { width = 1e; };", "bad number"),
    ("camera: { lookFrom = { x = 1.0 y = 2.0; }; };", ""),
    ('materials: { lambertian: ( { name = "a"; texture = "nope"; } ); };', "not found"),
])
def test_malformed_scene_files(tmp_path, text, err):
    p = tmp_path / "bad.scene"
    p.write_text(text)
    if "y = 2.0" in text:
        SceneFile(p)      # a missing separator between settings is accepted, like libconfig
        return
    with pytest.raises(A.CrtError) as e:
        SceneFile(p)
    assert err in str(e.value)


def test_missing_file():
    with pytest.raises(A.CrtError) as e:
        SceneFile("/nonexistent/x.scene")
    assert e.value.code == A.CRT_ERR_IO


@pytest.mark.skipif(not os.path.isdir("/root/reference/data/scenes"), reason="reference scene files absent")
def test_reference_scene_files_parse():
    # every shipped scene parses; missing mesh/texture files are the only accepted failures
    for path in sorted(glob.glob("/root/reference/data/scenes/*.scene")):
        try:
            SceneFile(path)
        except A.CrtError as e:
            assert e.code == A.CRT_ERR_IO and ("cannot load mesh" in str(e) or "cannot open" in str(e)), (path, str(e))
