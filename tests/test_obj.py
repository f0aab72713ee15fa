"""OBJ loader (rtx_obj_parse / rtx_obj_load; host code, no GPU): known answers written by hand from the
Wavefront OBJ format (vertex statements, 1-based and negative face references, v/vt/vn corner forms, polygon
fan triangulation, o / g groups)."""
import numpy as np
import pytest

import pyrtx

CUBE = """# unit cube, six quads
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
vn 0 0 -1
vt 0 0
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 2/1 3/1 7/1 6/1
f 3//1 4//1 8//1 7//1
f 4/1/1 1/1/1 5/1/1 8/1/1
"""


def test_cube_quads_fan_triangulated():
    v, i, f = pyrtx.load_obj(text=CUBE)
    assert v.shape == (8, 3) and v[6].tolist() == [1, 1, 1]
    assert i.shape == (12, 3)
    # quad (a, b, c, d) -> (a, b, c), (a, c, d), zero-based
    assert i[:2].tolist() == [[0, 3, 2], [0, 2, 1]]
    assert i[6:8].tolist() == [[1, 2, 6], [1, 6, 5]]
    assert i[10:].tolist() == [[3, 0, 4], [3, 4, 7]]
    assert f.tolist() == [0, 12]


def test_groups_negative_indices_and_comments():
    text = """mtllib x.mtl
o first
v 0 0 0
v 1 0 0 1.0
v 0 1 0
f -3 -2 -1   # the last three vertices
g empty
g second
v 2 0 0
v 2 1 0
f 2 4 5 3
s off
usemtl m
f 1/1/1 2/2/2 3/3/3
"""
    v, i, f = pyrtx.load_obj(text=text)
    assert v.shape == (5, 3) and v[1].tolist() == [1, 0, 0]
    assert i.tolist() == [[0, 1, 2], [1, 3, 4], [1, 4, 2], [0, 1, 2]]
    assert f.tolist() == [0, 1, 4]  # the empty group is dropped


@pytest.mark.parametrize("bad", ["v 0 0\n", "v 0 0 0\nv 1 0 0\nf 1 2\n", "v 0 0 0\nf 1 2 3\n",
                                 "v 0 0 0\nv 1 1 1\nv 2 2 2\nf 1 2 x\n", "v 0 0 0\nf 0 1 1\n"])
def test_malformed_input_is_a_format_error(bad):
    with pytest.raises(pyrtx.RtxError) as e:
        pyrtx.load_obj(text=bad)
    assert e.value.code == pyrtx.RTX_ERROR_FORMAT


def test_load_from_file_and_missing_file(tmp_path):
    p = tmp_path / "cube.obj"
    p.write_text(CUBE)
    v, i, f = pyrtx.load_obj(path=str(p))
    assert i.shape == (12, 3)
    with pytest.raises(pyrtx.RtxError):
        pyrtx.load_obj(path=str(tmp_path / "missing.obj"))
    v, i, f = pyrtx.load_obj(text="# nothing\n")
    assert v.shape == (0, 3) and i.shape == (0, 3) and f.tolist() == [0]
