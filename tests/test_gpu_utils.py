"""GPU tests of the steps next to the hot path: film conversion, acceleration-structure compaction / clone, and
OBJ meshes built into a BLAS (parity with the FP64 oracle on the same arrays)."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import compare_closest, oracle_scene, small_mixed_scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import pyrtx
    c = pyrtx.Context(0)
    yield c
    c.sync()
    c.close()


def _unorm8(x):
    """UNORM quantisation restated in float32: trunc(clamp(x) * 255 + 0.5); NaN and negatives -> 0."""
    x = np.asarray(x, np.float32)
    q = (np.clip(np.nan_to_num(x, nan=0.0), 0, 1).astype(np.float32) * np.float32(255) + np.float32(0.5))
    return np.where(x >= 1, 255, q.astype(np.uint32)).astype(np.uint8)


def _srgb(x):
    x = np.clip(np.nan_to_num(np.asarray(x, np.float32), nan=0.0), 0, 1).astype(np.float32)
    return np.where(x <= np.float32(0.0031308), np.float32(12.92) * x,
                    np.float32(1.055) * np.power(x, np.float32(1 / 2.4)) - np.float32(0.055)).astype(np.float32)


def test_film_to_rgba8(ctx):
    import pyrtx
    rng = np.random.default_rng(5)
    f = rng.uniform(-0.2, 1.2, (37, 53, 4)).astype(np.float32)
    f[0, 0] = [np.nan, 1.0, 0.0, 0.5]
    f[0, 1] = [1 / 255 * 0.5, 254.5 / 255, 0.0, 1.0]
    d = torch.from_numpy(f).cuda()
    got = ctx.film_to_rgba8(d, 0).cpu().numpy()
    assert np.array_equal(got, _unorm8(f))
    ref = ctx.film_to_rgba8(d, pyrtx.RTX_FILM_REFERENCE_BYTES).cpu().numpy()
    assert np.array_equal(ref[..., [2, 1, 0, 3]], _unorm8(f))  # bytes B, G, R, A
    op = ctx.film_to_rgba8(d, pyrtx.RTX_FILM_OPAQUE_BIT).cpu().numpy()
    assert np.all(op[..., 3] == 255) and np.array_equal(op[..., :3], _unorm8(f)[..., :3])
    s = ctx.film_to_rgba8(d, pyrtx.RTX_FILM_SRGB_BIT).cpu().numpy().astype(int)
    e = _unorm8(np.concatenate([_srgb(f[..., :3]), f[..., 3:]], -1)).astype(int)
    assert np.abs(s - e).max() <= 1 and (s != e).mean() < 1e-3


def test_reference_scene_bytes(ctx):
    """The R0 film as the reference's swapchain would receive it: B, G, R, alpha 0."""
    import pyrtx
    from pyrtx import scenes as S
    sc = S.scene_r0()
    blases, tlas = ctx.upload_scene(sc)
    film = ctx.trace_rays(tlas, sc.records, sc.camera, 1200, 800)
    b = ctx.film_to_rgba8(film, pyrtx.RTX_FILM_REFERENCE_BYTES).cpu().numpy()
    fl = film.cpu().numpy()
    assert np.all(b[..., 3] == 0)
    miss = np.all(fl[..., :3] == np.float32([0.0, 0.0, 0.2]), -1)
    assert miss.any() and np.all(b[miss][:, 0] == 51) and np.all(b[miss][:, 2] == 0)  # 0.2 * 255 = 51 in byte 0


def test_compaction_and_clone_keep_traversal(ctx):
    import pyrtx
    sc = small_mixed_scene(seed=9, instances=16, freq=12, cells=16)
    blases, tlas = ctx.upload_scene(sc)
    rays = torch.from_numpy(np.concatenate([__import__("pyrtx").scenes.camera_rays(sc.camera, 96, 64),
                                            __import__("pyrtx").scenes.incoherent_rays(sc, 8000, seed=2)])).cuda()
    ref = ctx.trace_ray_buffer(tlas, rays).clone()
    ctx.sync()
    small = [ctx.copy_as(b) for b in blases]
    for b, c in zip(blases, small):
        assert ctx.compacted_size(b) < b.storage.numel()
        assert c.storage.numel() == ctx.compacted_size(b)
    rec = pyrtx.instance_records(sc, [c.address for c in small])
    t2 = ctx.build_tlas(rec)
    t3 = ctx.copy_as(t2)
    t4 = ctx.copy_as(t3, pyrtx.RTX_COPY_MODE_CLONE)
    assert ctx.compacted_size(t2) == t3.storage.numel() < t2.storage.numel()
    for t in (t2, t3, t4):
        got = ctx.trace_ray_buffer(t, rays)
        ctx.sync()
        assert torch.equal(got, ref)
    info = ctx.as_info(t3)
    assert info.primitive_count == 16


OBJ = """o box
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 2 3 7 6
f 3 4 8 7
f 4 1 5 8
g floor
v -2 -0.5 -2
v 2 -0.5 -2
v 2 -0.5 2
v -2 -0.5 2
f 9 10 11 12
"""


def test_obj_mesh_blas_matches_oracle(ctx):
    import pyrtx
    from pyrtx import scenes as S
    _say("OBJ -> BLAS")
    v, i, first = pyrtx.load_obj(text=OBJ)
    assert first.tolist() == [0, 12, 14]
    blas = ctx.build_blas_obj(v, i, first)
    assert ctx.as_info(blas).geometry_count == 2 and ctx.as_info(blas).primitive_count == 14
    sc = S.Scene(blas=[[S.Geometry(v, i[first[g]:first[g + 1]]) for g in range(2)]],
                 transforms=S.translation()[None], blas_of_instance=np.array([0]),
                 custom_index=np.zeros(1, np.uint32), mask=np.full(1, 0xFF, np.uint32),
                 sbt_offset=np.zeros(1, np.uint32), flags=np.zeros(1, np.uint32), camera=(0.4, 0.7, 4.0, 60.0),
                 width=96, height=64, meta={"name": "obj"})
    tlas = ctx.build_tlas(pyrtx.instance_records(sc, [blas.address]))
    rays = np.concatenate([S.camera_rays(sc.camera, 96, 64), S.incoherent_rays(sc, 5000, seed=3)])
    h = ctx.trace_ray_buffer(tlas, torch.from_numpy(rays).cuda())
    ctx.sync()
    g = pyrtx.hits_to_numpy(h)
    tuv, ids, st = O.trace(oracle_scene(sc), rays)
    rep = compare_closest(g, tuv, ids, st)
    _say(f"  {rep}")
    assert rep["robust_hits"] > 500 and rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
    assert set(np.unique(g[4][g[5] != 0xFFFFFFFF]).tolist()) == {0, 1}  # both geometries are hit


def test_blas_update_refits_in_place(ctx):
    """A BLAS built with ALLOW_UPDATE, updated with displaced vertices (same topology), traces like the oracle on
    the new vertices; wrong counts or a non-updatable BLAS are rejected."""
    import pyrtx
    from pyrtx import scenes as S
    _say("BLAS update")
    a = S.icosphere(12, radius=1.0, seed=1, amplitude=0.05)
    rng = np.random.default_rng(3)
    vb = (a.vertices * np.float32(1.3) + rng.normal(0, 0.02, a.vertices.shape).astype(np.float32)
          + np.float32([0.2, -0.1, 0.0])).astype(np.float32)
    b = S.Geometry(vb, a.indices)

    def dev(g):
        return [{"vertices": torch.from_numpy(g.vertices).cuda(),
                 "indices": torch.from_numpy(g.indices.astype(np.uint32).view(np.int32)).cuda(), "transform": None}]

    flags = pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | pyrtx.RTX_BUILD_ALLOW_UPDATE_BIT
    blas = ctx.build_blas(dev(a), flags)
    assert blas.storage.numel() > ctx.build_blas(dev(a)).storage.numel()
    ctx.update_blas(blas, dev(b))
    sc = S.Scene(blas=[[b]], transforms=S.translation()[None], blas_of_instance=np.array([0]),
                 custom_index=np.zeros(1, np.uint32), mask=np.full(1, 0xFF, np.uint32),
                 sbt_offset=np.zeros(1, np.uint32), flags=np.zeros(1, np.uint32), camera=(0.0, 0.0, 4.0, 60.0),
                 width=96, height=64, meta={"name": "update"})
    tlas = ctx.build_tlas(pyrtx.instance_records(sc, [blas.address]))
    info = ctx.as_info(blas)
    assert np.allclose(info.bounds[:3], vb.min(0)) and np.allclose(info.bounds[3:], vb.max(0))
    rays = np.concatenate([S.camera_rays(sc.camera, 96, 64), S.incoherent_rays(sc, 5000, seed=8)])
    h = ctx.trace_ray_buffer(tlas, torch.from_numpy(rays).cuda())
    ctx.sync()
    g = pyrtx.hits_to_numpy(h)
    tuv, ids, st = O.trace(oracle_scene(sc), rays)
    rep = compare_closest(g, tuv, ids, st)
    _say(f"  {rep}")
    assert rep["robust_hits"] > 500 and rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
    with pytest.raises(pyrtx.RtxError):
        ctx.update_blas(blas, dev(S.icosphere(6, seed=1)))            # other triangle count
    with pytest.raises(pyrtx.RtxError):
        ctx.update_blas(ctx.build_blas(dev(a)), dev(b))               # built without ALLOW_UPDATE
