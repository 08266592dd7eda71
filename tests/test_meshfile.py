"""trianglemesh::raw parsing against the reference's own fixture
(src/lightmetrica-test/test_trianglemesh.cpp:41-123: the YAML and the
expected arrays), its validation rules, and a round trip of a generated mesh
through dump_raw_mesh -> load_raw_mesh (float32 values exact)."""
import numpy as np
import pytest

from lmx import meshfile, scenes

FIXTURE = """
positions: >
  0 0 0
  1 0 0
  1 1 0
  0 1 0
normals: >
  0 0 1
  0 0 1
  0 0 1
  0 0 1
texcoords: >
  0 0
  1 0
  1 1
  0 1
faces: >
  0 1 2
  0 2 3
"""


def test_reference_fixture():
    m = meshfile.load_raw_mesh(FIXTURE)
    np.testing.assert_array_equal(m["positions"], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(m["normals"], [[0, 0, 1]] * 4)
    np.testing.assert_array_equal(m["texcoords"], [[0, 0], [1, 0], [1, 1], [0, 1]])
    np.testing.assert_array_equal(m["faces"], [[0, 1, 2], [0, 2, 3]])
    assert m["positions"].dtype == np.float32 and m["faces"].dtype == np.uint32


@pytest.mark.parametrize("bad,msg", [
    ({"positions": "0 0 0 1"}, "positions"),
    ({"positions": "0 0 0", "normals": "0 0"}, "normals"),
    ({"positions": "0 0 0", "texcoords": "0"}, "texcoords"),
    ({"positions": "0 0 0", "faces": "0 0"}, "faces"),
    ({"faces": "0 1 2"}, "positions"),
])
def test_validation(bad, msg):
    with pytest.raises(ValueError, match=msg):
        meshfile.load_raw_mesh(bad)


def test_lists_stop_at_the_first_non_number():
    m = meshfile.load_raw_mesh({"positions": "1 2 3 x 4 5 6", "faces": "0 0 0 -1 2 2"})
    np.testing.assert_array_equal(m["positions"], [[1, 2, 3]])
    np.testing.assert_array_equal(m["faces"], [[0, 0, 0]])
    assert meshfile.load_raw_mesh({"positions": "1 2 3"})["faces"].shape == (0, 3)


def test_round_trip_generated_mesh():
    v, f = scenes.icosphere(3)
    v = (v * np.float32(1.37)).astype(np.float32)
    m = meshfile.load_raw_mesh(meshfile.dump_raw_mesh(v, f, normals=v))
    np.testing.assert_array_equal(m["positions"], v)
    np.testing.assert_array_equal(m["normals"], v)
    np.testing.assert_array_equal(m["faces"], f)
