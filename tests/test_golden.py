"""CPU tests against the committed R0 fixtures (tests/golden/make_r0_fixture.py wrote them from the reference's
scene literals in vulkan-raytraced-triangle/main.cpp and the closed-form picture those literals imply).

Pins (1) the scene this repository builds for R0 to the reference's numbers, and (2) the oracle's film to the
closed form; the GPU film is compared with the same fixture in test_gpu_parity.py."""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import oracle_scene
from pyrtx import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_r0():
    with open(os.path.join(GOLDEN, "r0_scene.json")) as f:
        sc = json.load(f)
    z = np.load(os.path.join(GOLDEN, "r0_film.npz"), allow_pickle=False)
    return sc, int(z["stride"]), z["rgba"], z["tie"]


@pytest.fixture(autouse=True)
def _reset():
    O.reset()
    yield
    O.reset()


def test_r0_scene_uses_the_reference_literals():
    g, _, _, _ = load_r0()
    sc = S.scene_r0()
    geoms = sc.blas[0]
    V = np.array(g["vertices"]["value"], np.float32)
    idx = np.array(g["indices"]["value"], np.uint32).reshape(-1, 3)
    assert len(geoms) == len(g["geometry_transforms"]["value"]) == 2  # two identical geometries, main.cpp:743
    for geo, T in zip(geoms, g["geometry_transforms"]["value"]):
        assert np.array_equal(geo.vertices, V)
        assert np.array_equal(geo.indices.reshape(-1, 3), idx)
        assert np.array_equal(np.asarray(geo.transform, np.float32).reshape(3, 4), np.array(T, np.float32))
    assert np.array_equal(sc.transforms.reshape(-1, 3, 4), np.array(g["instance_transforms"]["value"], np.float32))
    inst = g["instances"]
    assert sc.custom_index.tolist() == inst["custom_index"]
    assert sc.mask.tolist() == inst["mask"]
    assert sc.sbt_offset.tolist() == inst["sbt_record_offset"]
    assert all(f == "VK_GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT_KHR" for f in inst["flags"])
    assert sc.flags.tolist() == [S.INSTANCE_TRIANGLE_FACING_CULL_DISABLE] * 2
    assert list(sc.camera) == g["camera"]["position"] + [g["camera"]["y_fov_degree"]]
    assert list(sc.miss_color) == g["miss_color"]["value"]
    assert (sc.width, sc.height) == (g["film"]["width"], g["film"]["height"])
    cols = np.array(g["hit_colors"]["value"], np.float32)
    assert np.array_equal(sc.records["color"], cols)
    m = g["closest_hit_match"]["value"]
    assert m == {"gl_PrimitiveID": 1, "gl_InstanceID": 1, "gl_InstanceCustomIndexEXT": 100, "gl_GeometryIndexEXT": 1}
    assert np.all(sc.records["program"] == S.PROGRAM_MATCH_BARYCENTRIC)
    assert g["trace"]["tmin"] == 0.0 and g["trace"]["tmax"] == 100.0


def test_oracle_film_matches_r0_golden():
    g, stride, rgba, tie = load_r0()
    sc = S.scene_r0()
    t = oracle_scene(sc)
    film, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color,
                            tmin=g["trace"]["tmin"], tmax=g["trace"]["tmax"])
    f = film[::stride, ::stride]
    assert f.shape == rgba.shape
    keep = ~tie
    assert keep.sum() > 14000
    assert np.abs(f[keep] - rgba[keep]).max() < 2e-6
    # the oracle calls every fixture tie a tie or agrees with the fixture there too
    s = st[::stride, ::stride]
    assert np.all((s[tie] == 2) | (np.abs(f[tie] - rgba[tie]).max(axis=-1) < 2e-6))
