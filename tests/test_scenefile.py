"""Scene files (the reference's YAML schema) against the reference's own
fixtures (src/lightmetrica-test/test_scene3.cpp: SimpleLoad :71-98,
SimpleLoadWithAssets :141-212, Transform :216-278, SensorNode :282-300) and
the closed forms of the transform nodes; OBJ meshes as tinyobjloader splits,
de-duplicates and fans them."""
import numpy as np
import pytest

from lmx import scenefile, scenes

yaml = pytest.importorskip("yaml")


def _nodes(text):
    return scenefile.parse_nodes(yaml.safe_load(text))


def test_simple_load_ids_and_order():
    prims, s = _nodes("""
sensor: n1
nodes:
  - id: n1
  - id: n2
    child:
      - id: n2_1
      - id: n2_2
        child:
          - id: n2_2_1
          - id: n2_2_2
""")
    assert [p["id"] for p in prims] == ["n1", "n2", "n2_1", "n2_2", "n2_2_1", "n2_2_2"]
    assert s == 0


def test_simple_load_with_assets_binds_ids():
    prims, _ = _nodes("""
sensor: n1
nodes:
  - id: n1
    sensor: sensor_1
    mesh: mesh_1
    bsdf: bsdf_1
  - id: n2
    light: light_1
    mesh: mesh_2
    bsdf: bsdf_1
""")
    assert (prims[0]["sensor"], prims[0]["mesh"], prims[0]["bsdf"], prims[0]["light"]) == ("sensor_1", "mesh_1", "bsdf_1", None)
    assert (prims[1]["light"], prims[1]["mesh"], prims[1]["bsdf"], prims[1]["sensor"]) == ("light_1", "mesh_2", "bsdf_1", None)


def _colmajor(*v):
    return np.array(v, np.float64).reshape(4, 4).T


def test_transform_fixture():
    prims, _ = _nodes("""
sensor: n1
nodes:
  - id: n1
    transform:
      matrix: >
        1 0 0 0
        0 1 0 0
        0 0 1 0
        0 0 0 1
  - id: n2
    transform:
      translate: 0 0 0
      scale: 1 1 1
      rotate:
        axis: 0 1 0
        angle: 0
  - id: n3
    transform:
      matrix: >
        1 0 0 1
        0 1 0 1
        0 0 1 1
        0 0 0 1
    child:
      - id: n4
        transform:
          matrix: >
            2 0 0 0
            0 2 0 0
            0 0 2 0
            0 0 0 1
""")
    T = {p["id"]: p["transform"] for p in prims}
    np.testing.assert_allclose(T["n1"], np.eye(4))
    np.testing.assert_allclose(T["n2"], np.eye(4), atol=1e-12)
    np.testing.assert_allclose(T["n3"], _colmajor(1, 0, 0, 1, 0, 1, 0, 1, 0, 0, 1, 1, 0, 0, 0, 1))
    np.testing.assert_allclose(T["n4"], _colmajor(2, 0, 0, 2, 0, 2, 0, 2, 0, 0, 2, 2, 0, 0, 0, 1))


def test_sensor_node_selection_and_errors():
    _, s = _nodes("sensor: n2\nnodes:\n  - id: n1\n  - id: n2\n")
    assert s == 1
    with pytest.raises(scenefile.SceneFileError, match="Missing primitive ID"):
        _nodes("sensor: n3\nnodes:\n  - id: n1\n")
    with pytest.raises(scenefile.SceneFileError, match="Missing 'sensor'"):
        _nodes("nodes:\n  - id: n1\n")
    with pytest.raises(scenefile.SceneFileError, match="Missing 'nodes'"):
        _nodes("sensor: n1\n")
    with pytest.raises(scenefile.SceneFileError, match="same time"):
        _nodes("sensor: n1\nnodes:\n  - id: n1\n    light: a\n    sensor: b\n")


def test_transform_closed_forms():
    # T R S: (1, 0, 0) -> scale (2, 0, 0) -> 90 degrees about z (0, 2, 0) -> + (1, 2, 3)
    m = scenefile.parse_transform(yaml.safe_load("translate: 1 2 3\nrotate:\n  axis: 0 0 1\n  angle: 90\nscale: 2 1 1\n"))
    np.testing.assert_allclose(m @ [1, 0, 0, 1], [1, 4, 3, 1], atol=1e-12)
    # a right-handed 90 degree turn about y takes x to -z
    r = scenefile.parse_transform(yaml.safe_load("rotate:\n  axis: 0 2 0\n  angle: 90\n"))
    np.testing.assert_allclose(r @ [1, 0, 0, 1], [0, 0, -1, 1], atol=1e-12)
    # the matrix text is column-major: numbers 13-15 translate
    t = scenefile.parse_transform(yaml.safe_load("matrix: 1 0 0 0 0 1 0 0 0 0 1 0 5 6 7 1\n"))
    np.testing.assert_allclose(t @ [0, 0, 0, 1], [5, 6, 7, 1])
    # lookat: columns vx, vy, vz, eye as the pinhole camera frame
    la = scenefile.parse_transform(yaml.safe_load("lookat:\n  eye: 0 1 4\n  center: 0 0 0\n  up: 0 1 0\n"))
    pos, vx, vy, vz = scenes.camera_frame(dict(eye=(0, 1, 4), center=(0, 0, 0), up=(0, 1, 0)))
    np.testing.assert_allclose(la[:3, 3], pos)
    np.testing.assert_allclose(np.stack([la[:3, 0], la[:3, 1], la[:3, 2]]), np.stack([vx, vy, vz]), atol=1e-6)
    with pytest.raises(scenefile.SceneFileError):
        scenefile.parse_transform({"matrix": "1 2 3"})


def test_obj_shapes_dedup_and_fans(tmp_path):
    (tmp_path / "m.obj").write_text("""# quad with texcoords and normals, then a group with a negative-index triangle
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
f 1/1/1 2/2/1 3/3/1 4/4/1
g second
v 2 0 0
f -1/1/1 -4/2/1 -3/3/1
""")
    m = scenefile.load_obj(tmp_path / "m.obj")
    # shape 1: 4 vertices, fan (0,1,2), (0,2,3); shape 2: 3 new vertices appended
    np.testing.assert_array_equal(m["faces"], [[0, 1, 2], [0, 2, 3], [4, 5, 6]])
    np.testing.assert_array_equal(m["positions"], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                                                   [2, 0, 0], [1, 0, 0], [1, 1, 0]])
    np.testing.assert_array_equal(m["normals"], [[0, 0, 1]] * 7)
    np.testing.assert_array_equal(m["texcoords"][:4], [[0, 0], [1, 0], [1, 1], [0, 1]])
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\ng b\nf 1 2 3\n")
    with pytest.raises(scenefile.SceneFileError, match="Inconsistency"):
        scenefile.load_obj(tmp_path / "bad.obj")


def test_scene_file_root_and_version():
    with pytest.raises(scenefile.SceneFileError, match="lightmetrica"):
        scenefile.load_scene("a: 1\n")
    with pytest.raises(scenefile.SceneFileError, match="version"):
        scenefile.load_scene("lightmetrica:\n  version: 1.0.0\n")
