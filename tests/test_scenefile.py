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
    assert info["unsupported"] == 1        # the torus
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


def _rot(axis, deg):
    a = np.radians(deg)
    c, s = np.cos(a), np.sin(a)
    i, j = (axis + 1) % 3, (axis + 2) % 3
    R = np.eye(3)
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    return R


def test_primitive_transforms(tmp_path):
    # translate, rotate (degrees), scale as the reference's wrappers see them: a world ray is divided
    # by the scale, rotated by Rx(-x) Ry(-y) Rz(-z) and offset by -T, so the surface it meets is
    # S Rz(z) Ry(y) Rx(x) (p + T) (scale.cuh:50-56, rotate.cuh:104-119, translate.cuh:38-43)
    text = """
    materials: { lambertian: ( { name = "m"; color = { r = 0.5; g = 0.5; b = 0.5; }; } ); };
    primitives: {
      spheres: ( { name = "s"; position = { x = 1.0; y = 0.0; z = 0.0; }; radius = 0.5; material = "m";
                   transform = { translate = { x = 0.0; y = 2.0; z = 0.0; }; rotate = { x = 0.0; y = 0.0; z = 90.0; };
                                 scale = { x = 2.0; y = 2.0; z = 2.0; }; }; } );
      quads: ( { name = "q"; position = { x = 0.0; y = 0.0; z = 0.0; }; u = { x = 1.0; y = 0.0; z = 0.0; };
                 v = { x = 0.0; y = 1.0; z = 0.0; }; material = "m";
                 transform = { translate = { x = 1.0; y = 0.0; z = 0.0; }; rotate = { x = 30.0; y = 45.0; z = 0.0; };
                               scale = { x = 1.0; y = 2.0; z = 3.0; }; }; } );
      boxes: ( { name = "b"; position = { x = 0.0; y = 0.0; z = 0.0; }; size = { x = 2.0; y = 2.0; z = 2.0; };
                 material = "m"; transform = { translate = { x = 5.0; y = 0.0; z = 0.0; }; }; } );
    };
    """
    p = tmp_path / "xf.scene"
    p.write_text(text)
    f = SceneFile(p)
    d, _ = f.to_desc()
    assert f.info()["unsupported"] == 0
    s = d.spheres[0]
    want_c = 2.0 * (_rot(2, 90.0) @ (np.array([1.0, 0.0, 0.0]) + [0.0, 2.0, 0.0]))
    assert list(s.center) == pytest.approx(list(want_c), abs=1e-5) and s.radius == pytest.approx(1.0)
    q = d.quads[0]
    L = np.diag([1.0, 2.0, 3.0]) @ _rot(2, 0.0) @ _rot(1, 45.0) @ _rot(0, 30.0)
    assert list(q.corner) == pytest.approx(list(L @ np.array([1.0, 0.0, 0.0])), abs=1e-5)
    assert list(q.u) == pytest.approx(list(L @ [1.0, 0.0, 0.0]), abs=1e-5)
    assert list(q.v) == pytest.approx(list(L @ [0.0, 1.0, 0.0]), abs=1e-5)
    box = np.ctypeslib.as_array(d.meshes[0].positions, (24,)).reshape(8, 3)
    assert box.min(0) == pytest.approx([4.0, -1.0, -1.0]) and box.max(0) == pytest.approx([6.0, 1.0, 1.0])
    # a non-uniform scale of a sphere (an ellipsoid) is not representable: counted as unsupported
    p.write_text(text.replace("scale = { x = 2.0; y = 2.0; z = 2.0; }", "scale = { x = 2.0; y = 1.0; z = 2.0; }"))
    assert SceneFile(p).info()["unsupported"] == 1


def _scene_texture(tmp_path, image_name):
    p = tmp_path / "tex.scene"
    p.write_text('textures: { image: ( { name = "t"; filepath = "%s"; } ); };\n'
                 'materials: { lambertian: ( { name = "m"; texture = "t"; } ); };\n' % image_name)
    f = SceneFile(p)
    d, _ = f.to_desc()
    t = d.textures[0]
    assert t.type == A.CRT_TEX_IMAGE
    px = np.ctypeslib.as_array(t.pixels, (t.height * t.width * t.channels,)).reshape(t.height, t.width, t.channels).copy()
    return px


def _synthetic_rgb(h, w, seed=0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([(x * 255 // max(w - 1, 1)), (y * 255 // max(h - 1, 1)), ((x + y) * 3) % 256], axis=-1).astype(np.float64)
    img += rng.normal(0, 6, img.shape)
    img[h // 3:h // 2, w // 4:w // 2] = [250, 20, 30]     # a sharp-edged patch
    return np.clip(img, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("opts", [dict(subsampling=0, quality=95), dict(subsampling=2, quality=90),
                                  dict(subsampling=1, quality=85), dict(subsampling=0, quality=92, progressive=True),
                                  dict(subsampling=2, quality=80, progressive=True), dict(subsampling=2, quality=90, restart_marker_blocks=3),
                                  dict(subsampling=2, quality=85, progressive=True, restart_marker_blocks=5)])
def test_jpeg_textures_decode_like_libjpeg(tmp_path, opts):
    # JPEGs written by Pillow (libjpeg): baseline and progressive, 4:4:4 / 4:2:2 / 4:2:0, restart
    # intervals; the library's decoder agrees with Pillow's (libjpeg's integer IDCT, triangle-filter
    # chroma upsampling) within the rounding of the IDCT and of the upsampling
    PIL = pytest.importorskip("PIL.Image")
    img = _synthetic_rgb(37, 61)
    path = tmp_path / "t.jpg"
    kw = dict(opts)
    rst = kw.pop("restart_marker_blocks", None)
    if rst:
        kw["restart_marker_blocks"] = rst
    PIL.fromarray(img).save(path, "JPEG", **kw)
    ref = np.asarray(PIL.open(path).convert("RGB")).astype(int)
    got = _scene_texture(tmp_path, "t.jpg").astype(int)
    assert got.shape == ref.shape
    diff = np.abs(got - ref)
    print(opts, "mean |diff| %.3f max %d" % (diff.mean(), diff.max()))
    if opts["subsampling"] == 0:
        assert diff.max() <= 3 and diff.mean() < 0.1
    else:
        assert diff.max() <= 6 and diff.mean() < 0.6


def test_jpeg_grayscale_texture(tmp_path):
    PIL = pytest.importorskip("PIL.Image")
    img = _synthetic_rgb(40, 24)[..., 0]
    PIL.fromarray(img, "L").save(tmp_path / "g.jpg", "JPEG", quality=93)
    ref = np.asarray(PIL.open(tmp_path / "g.jpg").convert("RGB")).astype(int)
    got = _scene_texture(tmp_path, "g.jpg").astype(int)
    assert got.shape == ref.shape and np.abs(got - ref).max() <= 3


@pytest.mark.skipif(not os.path.isdir("/root/reference/data/textures"), reason="reference textures absent")
def test_shipped_jpeg_textures_decode(tmp_path):
    # every JPEG texture the shipped scenes can name (baseline and progressive) decodes close to
    # Pillow's result
    PIL = pytest.importorskip("PIL.Image")
    for path in sorted(glob.glob("/root/reference/data/textures/*.jpg")):
        link = tmp_path / os.path.basename(path)
        if not link.exists():
            os.symlink(path, link)
        got = _scene_texture(tmp_path, os.path.basename(path)).astype(int)
        ref = np.asarray(PIL.open(path).convert("RGB")).astype(int)
        assert got.shape == ref.shape, path
        diff = np.abs(got - ref)
        assert diff.mean() < 1.0 and np.percentile(diff, 99.9) <= 8, (path, diff.mean())


def test_camera_defaults_and_orthographic_camera(tmp_path):
    # scene_builder.cpp:49 defaults (look from (0, 0, 10) at the origin, focus 100, fov 70) and the
    # orthographic keys of scene_loader.cpp:361-366 (test_cameras.scene uses them)
    p = tmp_path / "cam.scene"
    p.write_text('camera: { aspectRatio = "16:9"; };')
    d, _ = SceneFile(str(p)).to_desc()
    c = d.camera
    assert list(c.lookfrom) == [0, 0, 10] and list(c.lookat) == [0, 0, 0] and list(c.vup) == [0, 1, 0]
    assert c.vfov_deg == 70 and c.focus_dist == 100 and c.aperture == 0 and c.orthographic == 0
    assert c.aspect_ratio == pytest.approx(16 / 9)
    p.write_text('camera: { aspectRatio = "16:9"; lookFrom = { x = 0.0; y = 2.0; z = 9.0; }; '
                 'lookAt = { x = 0.0; y = 0.6; z = 0.0; }; orthographic = true; orthoHeight = 2.0; };')
    f = SceneFile(str(p))
    d, _ = f.to_desc()
    assert d.camera.orthographic == 1 and d.camera.ortho_height == 2.0
    assert f.info()["unsupported"] == 0
    # every camera ray of the orthographic camera is parallel to lookat - lookfrom
    b = O.camera_basis(f, width=512)
    assert np.allclose(np.linalg.norm(b["vertical"]), 4.0)


SHIPPED = ["all_lights_types.scene", "simple_spot_light.scene", "extended_primitives.scene"]


@pytest.mark.parametrize("name", SHIPPED)
def test_shipped_scene_fixtures(name):
    # the copies the GPU tests render are the reference's files byte for byte and read in full:
    # all_lights_types has its quad light only (spot and omni are active = false), simple_spot_light
    # a spot light between four boxes (ground, roof, two walls) of 12 triangles each,
    # extended_primitives a cone, a cylinder and a disk with two boxes, a sphere and an omni light
    # (its torus is counted unsupported)
    path = os.path.join(HERE, "golden", "scenes", "shipped_" + name)
    src = os.path.join("/root/reference/data/scenes", name)
    if os.path.exists(src):
        with open(src, "rb") as a, open(path, "rb") as b:
            assert a.read() == b.read()
    f = SceneFile(path)
    extended = name == "extended_primitives.scene"
    assert f.info()["unsupported"] == (1 if extended else 0) and f.info()["skipped"] == 0
    d, _ = f.to_desc()
    if extended:
        assert (d.num_spheres, d.num_lights, d.num_sphere_lights, d.num_shapes, d.num_meshes) == (1, 0, 1, 3, 2)
        kinds = [d.shapes[k].kind for k in range(3)]
        assert kinds == [A.CRT_SHAPE_CONE, A.CRT_SHAPE_CYLINDER, A.CRT_SHAPE_DISK]   # load_scene's order
        cyl = d.shapes[1]
        assert list(cyl.center) == pytest.approx([-2.0, 0.0, 1.0]) and (cyl.radius, cyl.height) == pytest.approx((0.4, 0.6))
        assert (cyl.mapping.repeat_u, cyl.mapping.repeat_v) == pytest.approx((1.8, 0.2)) and cyl.transform.has == 0
    elif name == "all_lights_types.scene":
        assert (d.num_spheres, d.num_lights, d.num_sphere_lights) == (7, 1, 0)
    else:
        assert (d.num_spheres, d.num_lights, d.num_sphere_lights) == (2, 0, 1)
        assert sum(d.meshes[k].num_triangles for k in range(d.num_meshes)) == 4 * 12
