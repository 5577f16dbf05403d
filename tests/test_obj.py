"""OBJ ingestion (crt_obj_load): round trips OBJ text -> indexed mesh -> triangles, relative
indices, polygon fans, missing attributes, malformed files (host only, no GPU)."""
import numpy as np
import pytest

from cortexraytracercuda_amd import Scene, _abi as A


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_round_trip_triangles_and_quads(tmp_path):
    rng = np.random.default_rng(1)
    v = rng.normal(size=(6, 3)).astype(np.float32)
    lines = ["# test", "o thing"] + ["v %.9g %.9g %.9g" % tuple(x) for x in v]
    lines += ["vt 0.1 0.2", "vt 0.3 0.4", "vt 0.5 0.6", "vt 0.7 0.8"]
    lines += ["vn 0 0 1", "vn 0 1 0"]
    lines += ["f 1/1/1 2/2/1 3/3/1", "f 3/3/2 4/4/2 5/1/2 6/2/2"]   # a triangle and a quad-ish pentagon? (4+ corners)
    p = _write(tmp_path, "\n".join(lines) + "\n")
    s = Scene()
    m = s.lambertian("m")
    k = s.load_obj(p, m)
    me = s.meshes[k]
    tri = me["positions"][me["indices"]]                # (T, 3, 3)
    expect = np.array([[v[0], v[1], v[2]], [v[2], v[3], v[4]], [v[2], v[4], v[5]]], np.float32)
    assert np.array_equal(tri, expect)
    uvs = me["uvs"][me["indices"]]
    assert np.allclose(uvs[0], [[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]])
    assert np.allclose(me["normals"][me["indices"]][1:], [0, 1, 0])
    # the shared corner (3/3/*) with different normals is two vertices; identical corners are one
    assert me["positions"].shape[0] == 7


def test_relative_indices_and_missing_attributes(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf -4 -3 -2\nf 2//1 4//1 3//1\nvn 0 0 1\n"
    # the second face refers to a normal declared later: out of range at parse time -> error
    p = _write(tmp_path, text)
    with pytest.raises(A.CrtError) as e:
        Scene().load_obj(p, 0)
    assert e.value.code == A.CRT_ERR_IO
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf -4 -3 -2\nf 2 4 3\n"
    p = _write(tmp_path, text)
    s = Scene()
    s.lambertian("m")
    k = s.load_obj(p, 0)
    me = s.meshes[k]
    assert me["normals"] is None and me["uvs"] is None
    assert me["indices"].tolist() == [[0, 1, 2], [1, 3, 2]]


def test_malformed_and_missing_files(tmp_path):
    with pytest.raises(A.CrtError) as e:
        Scene().load_obj(tmp_path / "nope.obj", 0)
    assert e.value.code == A.CRT_ERR_IO
    for bad in ["v 0 0\n", "v 0 0 0\nf 1 2\n", "v 0 0 0\nf 1 2 9\n", "v 0 0 0\nf a b c\n"]:
        with pytest.raises(A.CrtError):
            Scene().load_obj(_write(tmp_path, bad), 0)


def test_transform_and_flat_option(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\n"
    p = _write(tmp_path, text)
    s = Scene()
    s.lambertian("m")
    xf = np.array([[0, -1, 0, 5], [1, 0, 0, 6], [0, 0, 1, 7]], np.float32)   # 90 degrees about z + translation
    k = s.load_obj(p, 0, transform=xf)
    assert np.allclose(s.meshes[k]["positions"], [[5, 6, 7], [5, 7, 7], [4, 6, 7]])
    assert np.allclose(s.meshes[k]["normals"], [[0, 0, 1]] * 3)
    k2 = s.load_obj(p, 0, use_smoothing=False)
    assert s.meshes[k2]["normals"] is None
