"""GPU parity: librtx (HIP, gfx950) against the FP64 oracle on the same seeded inputs.

Bar: hit instance / geometry / primitive IDs, miss/occluded decisions, Morton codes, sort order and the
binary-radix-tree topology are bit-exact on every ray the oracle does not flag as an epsilon-tie; t is within
1e-5 relative and the barycentrics within 1e-5 absolute of the FP64 values."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import assert_ties_ok, compare_closest, oracle_scene, small_mixed_scene
import pyrtx
from pyrtx import scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import pyrtx
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a GPU")
    c = pyrtx.Context(0)
    yield c
    c.sync()
    c.close()


@pytest.fixture(autouse=True)
def _reset_oracle():
    O.reset()
    yield
    O.reset()


def _trace_closest(ctx, tlas, rays):
    import pyrtx
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    h = ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_CLOSEST_HIT)
    ctx.sync()
    return pyrtx.hits_to_numpy(h)


def _trace_any(ctx, tlas, rays):
    import pyrtx
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    o = ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_ANY_HIT)
    ctx.sync()
    return o.cpu().numpy().astype(np.uint32)


# ---- build steps ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 5, 4095, 4096, 4097, 100003])
@pytest.mark.parametrize("kb", [4, 8])
def test_radix_sort_matches_stable_sort(ctx, n, kb):
    _say(f"sort n={n} key_bytes={kb}")
    rng = np.random.default_rng(n + kb)
    if kb == 4:
        keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = keys[: n // 3] & 0xFF00  # many duplicates
    else:
        keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        keys[: n // 3] = keys[: n // 3] & 0xFFFF000000000000
    sk, order = ctx.debug_sort(keys)
    ok, oo = O.sort_pairs(keys)
    assert np.array_equal(sk, ok)
    assert np.array_equal(order, oo)


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("n", [1, 2, 3, 1000, 50000])
def test_morton_sort_karras_bit_exact(ctx, bits, n):
    _say(f"lbvh bits={bits} n={n}")
    rng = np.random.default_rng(bits * 7 + n)
    lo = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    ext = rng.uniform(0, 0.3, (n, 3)).astype(np.float32)
    if n > 100:
        lo[10:60] = lo[10]  # identical boxes -> equal codes -> index tie-breaks
        ext[10:60] = ext[10]
    boxes = np.concatenate([lo, lo + ext], 1).astype(np.float32)
    codes, order, left, right = ctx.debug_lbvh(torch.from_numpy(boxes).cuda(), bits)
    cen = ((boxes[:, :3] + boxes[:, 3:]) * np.float32(0.5)).astype(np.float32)
    oc = O.morton(cen, bits)
    ok, oo = O.sort_pairs(oc)
    assert np.array_equal(codes, ok)
    assert np.array_equal(order, oo)
    if n > 1:
        ol, orr = O.karras(ok, 32 if bits == 30 else 64)
        assert np.array_equal(left, ol)
        assert np.array_equal(right, orr)


def _random_boxes(n, seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    ext = (rng.uniform(0, 0.3, (n, 3)) ** 2).astype(np.float32)
    if n > 100:
        lo[10:60] = lo[10]  # identical boxes: equal codes, equal areas, ties in the treelet DP
        ext[10:60] = ext[10]
    return np.concatenate([lo, lo + ext], 1).astype(np.float32)


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("n,rounds", [(2, 1), (7, 1), (1000, 0), (1000, 1), (50000, 2), (200000, 3)])
def test_treelet_refinement_bit_exact(ctx, bits, n, rounds):
    _say(f"treelet bits={bits} n={n} rounds={rounds}")
    boxes = _random_boxes(n, bits + n + rounds)
    db = torch.from_numpy(boxes).cuda()
    _, order, kl, kr = ctx.debug_lbvh(db, bits)  # the Karras tree, bit-exact with the oracle's (test above)
    o2, L, R, cost, hgt = ctx.debug_treelet(db, rounds, bits)
    assert np.array_equal(o2, order)
    el, er, ec, eh, _ = O.treelet_refine(kl, kr, boxes[order], rounds)
    assert np.array_equal(L, el) and np.array_equal(R, er)
    assert np.array_equal(cost.view(np.uint32), ec.view(np.uint32))
    assert np.array_equal(hgt, eh)
    if rounds == 0:
        assert np.array_equal(L, kl) and np.array_equal(R, kr)


def test_treelet_refinement_on_mesh_and_reported_cost(ctx):
    _say("treelet on terrain + icosphere BLAS")
    for g in (S.terrain(96, seed=5), S.icosphere(30, radius=1.0, seed=2, amplitude=0.05)):
        tri = g.vertices[g.indices]
        boxes = np.concatenate([tri.min(1), tri.max(1)], 1).astype(np.float32)
        db = torch.from_numpy(boxes).cuda()
        _, order, kl, kr = ctx.debug_lbvh(db, 30)
        sah = {}
        for rounds in (0, 2):
            o2, L, R, cost, hgt = ctx.debug_treelet(db, rounds, 30)
            el, er, ec, eh, esah = O.treelet_refine(kl, kr, boxes[order], rounds)
            assert np.array_equal(L, el) and np.array_equal(R, er)
            assert np.array_equal(cost.view(np.uint32), ec.view(np.uint32))
            sah[rounds] = esah
        assert sah[2] < 0.95 * sah[0]
        # the BLAS build runs the same passes: its reported SAH cost is the oracle's
        geom = [{"vertices": torch.from_numpy(g.vertices.astype(np.float32)).cuda(),
                 "indices": torch.from_numpy(g.indices.astype(np.uint32).view(np.int32)).cuda(), "transform": None}]
        for flags, rounds in ((pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT, 2),
                              (pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | pyrtx.RTX_BUILD_NO_TREELET_BIT, 0)):
            b = ctx.build_blas(geom, flags)
            info = ctx.as_info(b)
            assert abs(info.sah_cost - sah[rounds]) <= 1e-12 * sah[rounds], (info.sah_cost, sah[rounds])


# ---- reference scenes -------------------------------------------------------------------------------------------
def test_r0_film_matches_oracle(ctx):
    _say("R0 film 1200x800")
    sc = S.scene_r0()
    blases, tlas = ctx.upload_scene(sc)
    film = ctx.trace_rays(tlas, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    ctx.sync()
    f = film.cpu().numpy()
    t = oracle_scene(sc)
    of, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    robust = st != 2
    assert np.all(f[..., 3] == 0.0)
    diff = np.abs(f - of)[robust]
    assert diff.max() <= 1e-5, diff.max()
    # record colours and the miss colour are bit-exact
    rec_or_miss = robust & ~((of[..., 0] + of[..., 1] + of[..., 2] > 0.999) & (of[..., 0] + of[..., 1] + of[..., 2] < 1.001))
    assert np.array_equal(f[rec_or_miss], of[rec_or_miss])
    assert (st == 2).sum() < 500


def test_r0_film_matches_golden(ctx):
    """The HIP film against the committed closed-form R0 samples (tests/golden/r0_film.npz), ties excluded."""
    from test_golden import load_r0
    _say("R0 film vs golden samples")
    g, stride, rgba, tie = load_r0()
    sc = S.scene_r0()
    blases, tlas = ctx.upload_scene(sc)
    film = ctx.trace_rays(tlas, sc.records, sc.camera, sc.width, sc.height, sc.miss_color,
                          tmin=g["trace"]["tmin"], tmax=g["trace"]["tmax"])
    ctx.sync()
    f = film.cpu().numpy()[::stride, ::stride]
    keep = ~tie
    assert np.all(f[..., 3] == 0.0)
    assert np.abs(f[keep] - rgba[keep]).max() <= 1e-5


@pytest.mark.parametrize("z", [-2.5, 2.5])
def test_c1_film_matches_oracle(ctx, z):
    _say(f"C1 camera z={z}")
    sc = S.scene_c1(z)
    blases, tlas = ctx.upload_scene(sc)
    film = ctx.trace_rays(tlas, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    ctx.sync()
    f = film.cpu().numpy()
    t = oracle_scene(sc)
    of, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    robust = st != 2
    assert np.abs(f - of)[robust].max() <= 1e-5
    if z < 0:
        assert np.all(st == 0)


def test_raygen_rays_match_oracle_transcription():
    # the numpy generator used for ray buffers and the oracle's raygen agree bit for bit
    cam = (0.0, 0.0, 150.0, 60.0)
    a = S.camera_rays(cam, 640, 360, 0.0, 100.0)
    b = O.raygen(cam, 640, 360, 0.0, 100.0)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---- ray buffers ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [3, 11])
def test_mixed_scene_closest_and_any_hit(ctx, seed):
    _say(f"mixed scene seed={seed}")
    sc = small_mixed_scene(seed=seed, instances=24, freq=14, cells=20)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    coh = S.camera_rays(sc.camera, 160, 120)
    inc = S.incoherent_rays(sc, 20000, seed=seed)
    for name, rays in (("coherent", coh), ("incoherent", inc)):
        g = _trace_closest(ctx, tlas, rays)
        tuv, ids, st = O.trace(t, rays)
        rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
        _say(f"  {name}: {rep}")
        assert rep["robust_hits"] > 100
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
        assert rep["t_bad"] == 0 and rep["uv_bad"] == 0
        assert rep["ties"] < 0.01 * len(rays)
        assert_ties_ok(rep)
        occ = _trace_any(ctx, tlas, rays)
        _, _, sta = O.trace(t, rays, any_hit=True)
        assert np.all(occ[sta == 1] == 1) and np.all(occ[sta == 0] == 0)


def test_trace_results_independent_of_launch_config(ctx):
    _say("trace configs")
    sc = small_mixed_scene(seed=5, instances=24, freq=14, cells=20)
    blases, tlas = ctx.upload_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=5)])
    base_cfg = ctx.config()
    ref = _trace_closest(ctx, tlas, rays)
    ref_any = _trace_any(ctx, tlas, rays)
    try:
        for cfg in ({"blocks_per_cu": 4}, {"blocks_per_cu": 4, "lds_stack_depth": 4},
                    {"block_size": 64, "blocks_per_cu": 16, "ray_chunk": 1024},
                    {"block_size": 128, "blocks_per_cu": 6, "refill_lanes": 64, "refill_interval": 1},
                    {"block_size": 256, "blocks_per_cu": 1, "ray_chunk": 64, "refill_lanes": 8,
                     "refill_interval": 64}, {"leaf_lanes": 16}, {"leaf_lanes": 64, "block_size": 64}):
            ctx.set_config(**cfg)
            got = _trace_closest(ctx, tlas, rays)
            assert all(np.array_equal(a, b) for a, b in zip(got, ref)), cfg
            assert np.array_equal(_trace_any(ctx, tlas, rays), ref_any), cfg
            ctx.set_config(**base_cfg)
    finally:
        ctx.set_config(**base_cfg)


def _batch_scene(seed=21, instances=30):
    """Uneven BLASes for a batched build: 1 and 2 triangles, two icospheres, a terrain, a transformed two-geometry
    BLAS, plus (built, never instanced) an empty one."""
    from pyrtx import scenes as S
    rng = np.random.default_rng(seed)
    one = S.Geometry(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    two = S.Geometry(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32),
                     np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    s1 = S.icosphere(6, radius=1.0, seed=seed, amplitude=0.05)
    s2 = S.icosphere(15, radius=1.0, seed=seed + 1, amplitude=0.08)
    ter = S.terrain(20, size=2.0, height=0.3, seed=seed + 2)
    xf = np.zeros((3, 4), np.float32)
    xf[:, :3] = np.eye(3)
    xf[:, 3] = (0.5, 0.0, 0.0)
    pair = [S.Geometry(s1.vertices, s1.indices), S.Geometry(two.vertices, two.indices, transform=xf)]
    blas = [[one], [two], [s1], [s2], [ter], pair]
    R = S.random_rotation(rng, instances)
    sc = rng.uniform(0.5, 2.0, instances)
    pos = rng.uniform(-6, 6, (instances, 3))
    tx = np.zeros((instances, 3, 4))
    tx[:, :, :3] = R * sc[:, None, None]
    tx[:, :, 3] = pos
    return S.Scene(blas=blas, transforms=tx.astype(np.float32), blas_of_instance=np.arange(instances) % len(blas),
                   custom_index=np.arange(instances, dtype=np.uint32), mask=np.full(instances, 0xFF, np.uint32),
                   sbt_offset=np.zeros(instances, np.uint32), flags=np.zeros(instances, np.uint32),
                   camera=(0.0, 0.0, 20.0, 60.0), width=96, height=64, meta={"name": "batch"})


def test_batched_blas_build(ctx):
    _say("batched BLAS build")
    sc = _batch_scene()
    blases_b, tlas_b = ctx.upload_scene(sc, batched=True)
    blases_s, tlas_s = ctx.upload_scene(sc)
    for bb, bs in zip(blases_b, blases_s):
        ib, is_ = ctx.as_info(bb), ctx.as_info(bs)
        assert ib.primitive_count == is_.primitive_count and ib.geometry_count == is_.geometry_count
        assert list(ib.bounds) == list(is_.bounds)
        assert abs(ib.sah_cost - is_.sah_cost) <= 0.05 * is_.sah_cost + 1e-9
    # an empty BLAS in a batch
    import pyrtx
    e = ctx.build_blas_batch([[], [{"vertices": torch.zeros((3, 3), device="cuda"),
                                    "indices": torch.tensor([[0, 1, 2]], dtype=torch.int32, device="cuda")}]])
    assert ctx.as_info(e[0]).primitive_count == 0 and ctx.as_info(e[0]).slot_count == 0
    assert ctx.as_info(e[1]).primitive_count == 1
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=4)])
    g = _trace_closest(ctx, tlas_b, rays)
    tuv, ids, st = O.trace(t, rays)
    rep = compare_closest(g, tuv, ids, st)
    _say(f"  batched build vs oracle: {rep}")
    assert rep["robust_hits"] > 100
    assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0 and rep["uv_bad"] == 0
    occ = _trace_any(ctx, tlas_b, rays)
    _, _, sta = O.trace(t, rays, any_hit=True)
    assert np.all(occ[sta == 1] == 1) and np.all(occ[sta == 0] == 0)


def test_query_scene_reduced_parity(ctx):
    _say("query scene 1/200 scale")
    sc = S.query_scene(scale_tris=0.005)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    coh = S.camera_rays(sc.camera, 256, 256)
    inc = S.incoherent_rays(sc, 30000)
    for name, rays in (("coherent", coh), ("incoherent", inc)):
        g = _trace_closest(ctx, tlas, rays)
        tuv, ids, st = O.trace(t, rays)
        rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
        _say(f"  {name}: {rep}")
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
        assert rep["t_bad"] == 0 and rep["uv_bad"] == 0
        assert_ties_ok(rep)


def test_query_scene_full_size_sample(ctx):
    """The 19,994,340-triangle scene at full size: both 16,777,216-ray sets traced in full-size launches
    (closest and any hit), every 256th ray compared with the FP64 oracle on the same scene."""
    import pyrtx
    _say("query scene full size (20M triangles, 2 x 16M rays)")
    sc = S.query_scene()
    assert sc.unique_triangle_count == 19_994_340
    blases, tlas = ctx.upload_scene(sc)
    ctx.sync()
    t = oracle_scene(sc)
    W = H = 4096
    for name, rays in (("coherent", S.camera_rays(sc.camera, W, H)),
                       ("incoherent", S.incoherent_rays(sc, W * H, seed=1234))):
        r = torch.from_numpy(rays).cuda()
        h = ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_CLOSEST_HIT)
        o = ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_ANY_HIT)
        ctx.sync()
        idx = torch.arange(0, W * H, 256, device=r.device)
        g = pyrtx.hits_to_numpy(h[idx])
        occ = o[idx].cpu().numpy().astype(np.uint32)
        sample = rays[::256]
        tuv, ids, st = O.trace(t, sample)
        rep = compare_closest(g, tuv, ids, st, tlas=t, rays=sample)
        _say(f"  {name}: {rep}")
        assert rep["robust_hits"] > 1000
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
        assert rep["t_bad"] == 0 and rep["uv_bad"] == 0
        assert rep["ties"] < 0.01 * len(sample)
        assert_ties_ok(rep)
        _, _, sta = O.trace(t, sample, any_hit=True)
        assert np.all(occ[sta == 1] == 1) and np.all(occ[sta == 0] == 0)
        del r, h, o


def test_cull_mask_and_instance_mask(ctx):
    _say("cull mask")
    sc = small_mixed_scene(seed=5, instances=12, freq=8, cells=8)
    sc.mask = np.where(np.arange(12) % 2 == 0, 0x01, 0x02).astype(np.uint32)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    rays = S.incoherent_rays(sc, 5000, seed=2)
    import pyrtx
    r = torch.from_numpy(rays).cuda()
    for cull in (0x01, 0x02, 0x03, 0x00):
        h = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_CLOSEST_HIT, cull_mask=cull))
        ctx.sync()
        tuv, ids, st = O.trace(t, rays, cull_mask=cull)
        rep = compare_closest(h, tuv, ids, st)
        _say(f"  cull {cull:#x}: {rep}, gpu hits {int((h[5] != 0xFFFFFFFF).sum())}")
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
        if cull == 0:
            assert np.all(h[5] == 0xFFFFFFFF)


def test_empty_and_single_triangle(ctx):
    _say("degenerate sizes")
    import pyrtx
    # one triangle, one instance
    sc = S.scene_c1(2.5)
    blases, tlas = ctx.upload_scene(sc)
    info = ctx.as_info(blases[0])
    assert info.primitive_count == 1 and info.node_count == 1 and info.slot_count == 2
    # TLAS with zero instances: every ray misses
    empty = ctx.build_tlas(np.zeros(0, pyrtx.INSTANCE_DTYPE))
    rays = torch.from_numpy(S.camera_rays(sc.camera, 16, 16)).cuda()
    h = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(empty, rays))
    ctx.sync()
    assert np.all(h[5] == 0xFFFFFFFF)


def test_error_paths(ctx):
    _say("error paths")
    import pyrtx
    v = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    i = torch.zeros((1, 3), dtype=torch.int32, device="cuda")
    geoms = [{"vertices": v, "indices": i, "transform": None}]
    as_size, scratch_size = ctx.blas_build_sizes(geoms)
    arr = ctx._geometry_structs(geoms)
    import ctypes
    dst = torch.empty(as_size, dtype=torch.uint8, device="cuda")
    small = torch.empty(max(256, scratch_size // 2), dtype=torch.uint8, device="cuda")
    addr = ctypes.c_uint64()
    rc = ctx.lib.rtx_cmd_build_blas(ctx.handle, ctypes.cast(arr, ctypes.c_void_p), 1, 0, dst.data_ptr(), as_size,
                                    small.data_ptr(), small.numel(), ctypes.byref(addr))
    assert rc == -2  # RTX_ERROR_OUT_OF_DEVICE_MEMORY
    rc = ctx.lib.rtx_cmd_trace_ray_buffer(ctx.handle, 0, None, 0, 0, 0, 0xFF, None, None)
    assert rc == -1002
    with pytest.raises(pyrtx.RtxError):
        ctx.set_config(block_size=100)


def test_cpp_host_sample_matches_oracle(tmp_path):
    """The C++ transcription of main.cpp's flow on the C ABI renders the reference picture."""
    import os
    import subprocess
    from conftest import PKG
    exe = os.path.join(PKG, "host", "raytraced_triangle")
    assert os.path.exists(exe), "build the sample first (make -C build-up-phase_amd)"
    out = tmp_path / "r0.ppm"
    _say("C++ host sample")
    subprocess.run([exe, str(out)], check=True, timeout=120)
    data = out.read_bytes()
    assert data.startswith(b"P6\n1200 800\n255\n")
    img = np.frombuffer(data[len(b"P6\n1200 800\n255\n"):], np.uint8).reshape(800, 1200, 3)
    sc = S.scene_r0()
    t = oracle_scene(sc)
    of, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    ref = np.clip(of[..., :3], 0, 1) * 255.0
    robust = st != 2
    # 8-bit rounding of identical values: at most one code apart (half-way cases)
    assert np.abs(img.astype(np.float64) - ref)[robust].max() <= 0.5 + 1e-3


# ---- ray and instance flags -------------------------------------------------------------------------------------
def test_ray_flags_match_oracle(ctx):
    import pyrtx
    _say("ray flags")
    sc = small_mixed_scene(seed=13, instances=24, freq=12, cells=16)
    # instance flags: none / facing culls disabled / flipped facing, cycling
    sc.flags = np.array([(0, pyrtx.RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT,
                          pyrtx.RTX_INSTANCE_TRIANGLE_FLIP_FACING_BIT)[i % 3] for i in range(24)], np.uint32)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=6)])
    dr = torch.from_numpy(rays).cuda()
    for rf in (pyrtx.RTX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES, pyrtx.RTX_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES,
               pyrtx.RTX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES | pyrtx.RTX_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES):
        h = ctx.trace_ray_buffer(tlas, dr, ray_flags=rf)
        ctx.sync()
        g = pyrtx.hits_to_numpy(h)
        tuv, ids, st = O.trace(t, rays, ray_flags=rf)
        rep = compare_closest(g, tuv, ids, st)
        _say(f"  flags {rf:#x}: {rep}")
        assert rep["robust_hits"] > 200 and rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
        occ = ctx.trace_ray_buffer(tlas, dr, pyrtx.RTX_QUERY_ANY_HIT, ray_flags=rf).cpu().numpy()
        _, _, sta = O.trace(t, rays, any_hit=True, ray_flags=rf)
        assert np.all(occ[sta == 1] == 1) and np.all(occ[sta == 0] == 0)
    # nothing to hit: every triangle is opaque
    for rf in (pyrtx.RTX_RAY_FLAG_SKIP_TRIANGLES, pyrtx.RTX_RAY_FLAG_CULL_OPAQUE):
        g = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, dr, ray_flags=rf))
        assert np.all(g[5] == 0xFFFFFFFF)
    # first hit: hits exactly where something is hit, never closer than the closest
    g = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, dr, ray_flags=pyrtx.RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT))
    tuv, ids, st = O.trace(t, rays)
    robust = st != 2
    assert np.array_equal((g[5] != 0xFFFFFFFF)[robust], (st == 1)[robust])
    hit = robust & (st == 1)
    assert np.all(g[0][hit] >= tuv[hit, 0] * (1 - 1e-5))


def _cutout_mixed_scene(pyrtx):
    """small_mixed_scene with non-opaque icospheres, instance flags cycling none / FORCE_OPAQUE / FORCE_NO_OPAQUE
    and one hit record per instance: record colours, cutout any-hit programs of 2, 3 or 5 cells per edge."""
    sc = small_mixed_scene(seed=17, instances=24, freq=12, cells=16)
    sc.blas[0][0].opaque = False
    n = len(sc.blas_of_instance)
    sc.flags = np.array([(0, pyrtx.RTX_INSTANCE_FORCE_OPAQUE_BIT, pyrtx.RTX_INSTANCE_FORCE_NO_OPAQUE_BIT)[i % 3]
                         for i in range(n)], np.uint32)
    sc.sbt_offset = np.arange(n, dtype=np.uint32)
    rng = np.random.default_rng(3)
    cells = [(2, 3, 5)[i % 3] for i in range(n)]
    sc.records = S.hit_records([pyrtx.RTX_PROGRAM_RECORD_COLOR] * n, rng.uniform(0.1, 0.9, (n, 3)),
                               any_hit=[S.ANY_HIT_CUTOUT | (c << 8) for c in cells])
    return sc


def test_opacity_flags_match_oracle(ctx):
    """Geometry opacity, instance FORCE flags and the ray's OPAQUE / NO_OPAQUE / CULL_*OPAQUE on ray buffers (no
    hit-record table, so no any-hit program runs)."""
    import pyrtx
    _say("opacity flags")
    sc = _cutout_mixed_scene(pyrtx)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=8)])
    dr = torch.from_numpy(rays).cuda()
    F = pyrtx
    for rf in (0, F.RTX_RAY_FLAG_OPAQUE, F.RTX_RAY_FLAG_NO_OPAQUE, F.RTX_RAY_FLAG_CULL_OPAQUE,
               F.RTX_RAY_FLAG_CULL_NO_OPAQUE, F.RTX_RAY_FLAG_NO_OPAQUE | F.RTX_RAY_FLAG_CULL_OPAQUE,
               F.RTX_RAY_FLAG_OPAQUE | F.RTX_RAY_FLAG_CULL_NO_OPAQUE):
        g = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, dr, ray_flags=rf))
        tuv, ids, st = O.trace(t, rays, ray_flags=rf)
        rep = compare_closest(g, tuv, ids, st)
        _say(f"  flags {rf:#x}: {rep}")
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
        occ = ctx.trace_ray_buffer(tlas, dr, pyrtx.RTX_QUERY_ANY_HIT, ray_flags=rf).cpu().numpy()
        _, _, sta = O.trace(t, rays, any_hit=True, ray_flags=rf)
        assert np.all(occ[sta == 1] == 1) and np.all(occ[sta == 0] == 0)
        if rf in (F.RTX_RAY_FLAG_CULL_OPAQUE, F.RTX_RAY_FLAG_CULL_NO_OPAQUE):
            assert rep["robust_hits"] > 100  # both opacities are present
    for rf in (F.RTX_RAY_FLAG_CULL_OPAQUE | F.RTX_RAY_FLAG_CULL_NO_OPAQUE,
               F.RTX_RAY_FLAG_OPAQUE | F.RTX_RAY_FLAG_CULL_OPAQUE):
        g = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, dr, ray_flags=rf))
        assert np.all(g[5] == 0xFFFFFFFF)


def test_any_hit_programs_match_oracle(ctx):
    """Cutout any-hit programs on the film (rtx_cmd_trace_rays with the hit-record table): colours bit-exact
    against the oracle off epsilon-ties (edges, cell boundaries within 2^-16 of the GPU's FP32 barycentrics)."""
    import pyrtx
    _say("any-hit programs")
    sc = _cutout_mixed_scene(pyrtx)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    W, H = 192, 128
    F = pyrtx
    holes = None
    for rf in (0, F.RTX_RAY_FLAG_NO_OPAQUE, F.RTX_RAY_FLAG_OPAQUE, F.RTX_RAY_FLAG_CULL_OPAQUE):
        film = ctx.trace_rays(tlas, sc.records, sc.camera, W, H, sc.miss_color, ray_flags=rf).cpu().numpy()
        of, st = O.trace_rays(t, sc.records, sc.camera, W, H, sc.miss_color, ray_flags=rf)
        robust = st != 2
        bad = robust & np.any(film != of, axis=-1)
        _say(f"  flags {rf:#x}: robust {robust.mean():.4f}, mismatched robust pixels {int(bad.sum())}")
        assert bad.sum() == 0
        assert robust.mean() > 0.9
        if rf == 0:
            holes = film
        if rf == F.RTX_RAY_FLAG_OPAQUE:
            # the cutouts must have changed something: with OPAQUE no program runs
            assert np.any(np.any(film != holes, axis=-1) & robust)


def _procedural_scene(pyrtx, seed=21, instances=12):
    """Instances of one BLAS holding a geometry of 3000 random boxes (AABB primitives) and a terrain (triangles)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-2, 2, (3000, 3))
    h = rng.uniform(0.03, 0.25, (3000, 3))
    boxes = np.concatenate([c - h, c + h], 1).astype(np.float32)
    ter = S.terrain(12, size=4.0, height=0.3, seed=seed)
    R = S.random_rotation(rng, instances)
    xf = np.zeros((instances, 3, 4))
    xf[:, :, :3] = R * rng.uniform(0.5, 1.5, instances)[:, None, None]
    xf[:, :, 3] = rng.uniform(-6, 6, (instances, 3))
    sc = S.Scene(blas=[[S.AabbGeometry(boxes), ter]], transforms=xf.astype(np.float32),
                 blas_of_instance=np.zeros(instances, int), custom_index=np.arange(instances, dtype=np.uint32),
                 mask=np.full(instances, 0xFF, np.uint32), sbt_offset=2 * np.arange(instances, dtype=np.uint32),
                 flags=np.zeros(instances, np.uint32), camera=(0.0, 0.0, 20.0, 60.0), width=192, height=128,
                 meta={"name": "procedural"})
    progs = []
    for i in range(instances):  # geometry 0 (boxes): sphere or box program; geometry 1 (terrain): triangles
        progs += [S.hit_group((pyrtx.RTX_PROGRAM_BARYCENTRIC, pyrtx.RTX_PROGRAM_RECORD_COLOR)[i % 2],
                              (S.INTERSECTION_SPHERE, S.INTERSECTION_BOX)[(i // 2) % 2]), pyrtx.RTX_PROGRAM_RECORD_COLOR]
    sc.records = S.hit_records(progs, rng.uniform(0.1, 0.9, (2 * instances, 3)))
    return sc


def test_procedural_geometry_matches_oracle(ctx):
    """AABB geometry: ray buffers (no hit-record table: box program) and films (sphere / box intersection programs
    from the records) against the oracle; SKIP_AABBS / SKIP_TRIANGLES."""
    import pyrtx
    _say("procedural geometry")
    sc = _procedural_scene(pyrtx)
    blases, tlas = ctx.upload_scene(sc)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=9)])
    dr = torch.from_numpy(rays).cuda()
    for rf in (0, pyrtx.RTX_RAY_FLAG_SKIP_AABBS, pyrtx.RTX_RAY_FLAG_SKIP_TRIANGLES):
        g = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, dr, ray_flags=rf))
        tuv, ids, st = O.trace(t, rays, ray_flags=rf)
        rep = compare_closest(g, tuv, ids, st)
        _say(f"  flags {rf:#x}: {rep}")
        assert rep["robust_hits"] > 500
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0 and rep["uv_bad"] == 0
        occ = ctx.trace_ray_buffer(tlas, dr, pyrtx.RTX_QUERY_ANY_HIT, ray_flags=rf).cpu().numpy()
        _, _, sta = O.trace(t, rays, any_hit=True, ray_flags=rf)
        assert np.all(occ[sta == 1] == 1) and np.all(occ[sta == 0] == 0)
    W, H = sc.width, sc.height
    film = ctx.trace_rays(tlas, sc.records, sc.camera, W, H, sc.miss_color).cpu().numpy()
    of, st = O.trace_rays(t, sc.records, sc.camera, W, H, sc.miss_color)
    robust = st != 2
    close = np.all(np.abs(film - of) <= 1e-5, axis=-1)
    _say(f"  film: robust {robust.mean():.4f}, robust pixels off by > 1e-5: {int((robust & ~close).sum())}")
    assert robust.mean() > 0.95 and np.all(close[robust])


def test_skip_closest_hit_program(ctx):
    import pyrtx
    sc = S.scene_r0()
    blases, tlas = ctx.upload_scene(sc)
    plain = ctx.trace_rays(tlas, sc.records, sc.camera, 300, 200).cpu().numpy()
    skip = ctx.trace_rays(tlas, sc.records, sc.camera, 300, 200,
                          ray_flags=pyrtx.RTX_RAY_FLAG_OPAQUE | pyrtx.RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER).cpu().numpy()
    miss = np.all(plain[..., :3] == np.float32([0, 0, 0.2]), -1)
    assert miss.any() and (~miss).any()
    assert np.array_equal(skip[miss], plain[miss]) and np.all(skip[~miss] == 0)


def _children_per_node(storage):
    """Walk a built structure from its root (64-B node slots after the 256-B header; internal-child mask byte 9,
    present-child mask byte 10, first-child slot at bytes 12-15): the number of children of every node."""
    raw = storage.cpu().numpy()
    slots = raw[256:].reshape(-1, 64)
    counts, todo = [], [0]
    while todo:
        s = todo.pop()
        imask, vmask = int(slots[s, 9]), int(slots[s, 10])
        base = int(slots[s, 12:16].view(np.uint32)[0])
        counts.append(bin(vmask).count("1"))
        for k in range(8):
            if vmask >> k & 1 and imask >> k & 1:
                todo.append(base + bin(vmask & ((1 << k) - 1)).count("1"))
    return np.array(counts)


def test_node_width_cap(ctx):
    """RTX_BUILD_MAX_WIDTH(w) caps the SAH wide collapse at w children per node (same node format): every node has
    at most w children, the trees get deeper, and traces stay exact against the oracle."""
    import pyrtx
    _say("node width cap")
    sc = small_mixed_scene(seed=8, instances=20, freq=14, cells=20)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=8)])
    tuv, ids, st = O.trace(t, rays)
    depth = {}
    for w in (2, 4, 8):
        flags = pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | (pyrtx.RTX_BUILD_MAX_WIDTH(w) if w < 8 else 0)
        blases, tlas = ctx.upload_scene(sc, flags=flags)
        for b in blases:
            c = _children_per_node(b.storage)
            assert c.max() <= w and c.min() >= 1
        depth[w] = ctx.as_info(tlas).max_depth
        g = _trace_closest(ctx, tlas, rays)
        rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
        _say(f"  width {w}: depth {depth[w]}, children per BLAS node {c.mean():.2f}, {rep['robust_hits']} hits")
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0
        assert_ties_ok(rep)
    assert depth[2] > depth[4] > depth[8]


def test_prefer_fast_build_skips_treelets(ctx):
    """PREFER_FAST_BUILD builds without treelet refinement (each BLAS reports the SAH cost of a NO_TREELET build),
    RTX_BUILD_TREELET_ROUNDS still asks for rounds under it (the cost of a PREFER_FAST_TRACE build, lower), and
    every variant traces exactly against the oracle."""
    import pyrtx
    _say("PREFER_FAST_BUILD and treelet rounds")
    sc = small_mixed_scene(seed=9, instances=12, freq=24, cells=24)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 96, 72), S.incoherent_rays(sc, 12000, seed=9)])
    tuv, ids, st = O.trace(t, rays)
    fast_trace = pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT
    fast_build = pyrtx.RTX_BUILD_PREFER_FAST_BUILD_BIT
    cost = {}
    for name, flags in (("fast_build", fast_build), ("no_treelet", fast_trace | pyrtx.RTX_BUILD_NO_TREELET_BIT),
                        ("fast_build_2_rounds", fast_build | pyrtx.RTX_BUILD_TREELET_ROUNDS(2)),
                        ("fast_trace", fast_trace)):
        blases, tlas = ctx.upload_scene(sc, flags=flags)
        ctx.sync()
        cost[name] = [ctx.as_info(b).sah_cost for b in blases]
        g = _trace_closest(ctx, tlas, rays)
        rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0, name
        assert_ties_ok(rep)
    _say(f"  SAH cost per BLAS: fast build {cost['fast_build']}, fast trace {cost['fast_trace']}")
    assert cost["fast_build"] == cost["no_treelet"]
    assert cost["fast_build_2_rounds"] == cost["fast_trace"]
    assert all(a < b for a, b in zip(cost["fast_trace"], cost["fast_build"]))


def test_inactive_and_degenerate_geometry_and_rays(ctx):
    """Triangles with a NaN or infinite vertex component (inactive: never hit, primitive indices kept), a zero-area
    triangle, and degenerate rays (empty and single-point [tmin, tmax], origins on the plane, rays in the plane,
    axis-parallel rays) give the oracle's results, closest and any hit, through a transformed instance."""
    import pyrtx
    from test_oracle import degenerate_geometry
    _say("inactive / degenerate triangles and rays")
    v, idx = degenerate_geometry()
    dead = v.copy()
    dead[:, 0] = np.nan  # a BLAS whose triangles are all inactive, instanced beside the first
    rng = np.random.default_rng(21)
    xf = np.stack([S.translation(1.0, -2.0, 3.0), S.translation(1.0, 0.5, 3.0)]).astype(np.float32)
    sc = S.Scene(blas=[[S.Geometry(v, idx)], [S.Geometry(dead, idx)]], transforms=xf, blas_of_instance=np.array([0, 1]),
                 custom_index=np.zeros(2, np.uint32), mask=np.full(2, 0xFF, np.uint32),
                 sbt_offset=np.zeros(2, np.uint32), flags=np.zeros(2, np.uint32), meta={"name": "degenerate"})
    n = 4000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = rng.uniform(-2, 16, n) + 1.0
    rays[:, 1] = rng.uniform(-1, 5.5, n) - 2.0
    rays[:, 2] = 3.0 + rng.choice([1.0, -1.0, 0.0], n)
    rays[:, 6] = np.where(rays[:, 2] > 3.0, -1.0, 1.0)
    rays[:, 3] = rng.choice([0.0, 0.5, 1.0, 2.0], n)
    rays[:, 7] = rng.choice([0.5, 1.0, 10.0, np.inf], n)
    k = n // 4  # a quarter of the rays travel in the triangles' plane along +x or +y
    rays[:k, 2] = 3.0
    rays[:k, 4:7] = [[1, 0, 0], [0, 1, 0]][0]
    rays[k // 2:k, 4:7] = [0, 1, 0]
    # and 1,000 ordinary rays straight down onto the boxes of the two active triangles (world x in [1, 3] and
    # [13, 15], y in [-2, 0]; each triangle fills half its box)
    aim = np.zeros((1000, 8), np.float32)
    aim[:, 0] = rng.uniform(1.0, 3.0, 1000) + 12.0 * rng.integers(0, 2, 1000)
    aim[:, 1] = rng.uniform(-2.0, 0.0, 1000)
    aim[:, 2] = 5.0
    aim[:, 6] = -1.0
    aim[:, 7] = 10.0
    rays = np.concatenate([rays, aim])
    t = oracle_scene(sc)
    tuv, ids, st = O.trace(t, rays)
    occl_o = O.trace(t, rays, any_hit=True)[2]
    blases, tlas = ctx.upload_scene(sc)
    assert ctx.as_info(blases[0]).primitive_count == 5 and ctx.as_info(blases[1]).primitive_count == 5
    g = _trace_closest(ctx, tlas, rays)
    rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
    _say(f"  {rep}")
    assert rep["robust_hits"] > 400
    assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0
    assert_ties_ok(rep)
    hit = g[5] != 0xFFFFFFFF
    assert set(int(p) for p in g[3][hit]) <= {0, 4}  # the inactive and degenerate triangles are never hit
    assert set(int(i) for i in g[5][hit]) <= {0}  # nor anything of the all-inactive BLAS
    r = torch.from_numpy(rays).cuda()
    occ = ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_ANY_HIT).cpu().numpy()
    ctx.sync()
    robust = occl_o != 2
    assert np.array_equal(occ[robust] != 0, occl_o[robust] == 1)


def test_scene_far_from_the_origin(ctx):
    """The mixed scene moved 10^5 units away from the world origin, with the camera and the incoherent rays moved
    along: the TLAS is built and traversed in FP32 world coordinates of size 10^5, the BLASes in their own rebased
    frames, and the results still match the FP64 oracle (ties checked)."""
    _say("scene far from the origin")
    sc = small_mixed_scene(seed=11, instances=16, freq=16, cells=16)
    off = np.array([1.0e5, -3.0e4, 2.0e5], np.float32)
    sc.transforms = sc.transforms.copy()
    sc.transforms[:, :, 3] += off
    cam = list(sc.camera)
    cam[:3] = [c + o for c, o in zip(cam[:3], off)]
    sc.camera = tuple(cam)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=11)])
    tuv, ids, st = O.trace(t, rays)
    blases, tlas = ctx.upload_scene(sc)
    g = _trace_closest(ctx, tlas, rays)
    rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
    _say(f"  {rep}")
    assert rep["robust_hits"] > 2000
    assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0
    assert_ties_ok(rep)


def test_inactive_instances_keep_instance_ids(ctx):
    """Instances with a null BLAS reference (Vulkan's inactive instances) or a singular transform are never hit
    and keep their place in the instance numbering: the other instances report the oracle's instance IDs, the
    oracle scene holding mask-0 instances (never hit) in the same places."""
    import pyrtx
    _say("inactive instances")
    sc = small_mixed_scene(seed=12, instances=16, freq=14, cells=16)
    blases, _ = ctx.upload_scene(sc)
    rec = pyrtx.instance_records(sc, [b.address for b in blases])
    rec["addr"][3] = 0  # null reference
    rec["xf"][9] = 0.0  # singular transform
    tlas = ctx.build_tlas(rec)
    masked = S.Scene(**{**sc.__dict__, "mask": sc.mask.copy()})
    masked.mask[[3, 9]] = 0
    t = oracle_scene(masked)
    rays = np.concatenate([S.camera_rays(sc.camera, 128, 96), S.incoherent_rays(sc, 20000, seed=12)])
    tuv, ids, st = O.trace(t, rays)
    g = _trace_closest(ctx, tlas, rays)
    rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
    _say(f"  {rep}")
    assert rep["robust_hits"] > 1000
    assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0
    assert_ties_ok(rep)
    hit = g[5] != 0xFFFFFFFF
    assert not np.isin(g[5][hit], [3, 9]).any()
    assert len(set(int(i) for i in g[5][hit])) >= 8


def test_morton63_builds_trace_exactly(ctx):
    """Structures built with 63-bit Morton keys trace like the oracle, their reported SAH cost is within a few
    percent of a 30-bit build's (the same LBVH method on a finer grid), and a batched build, which keys BLAS id and
    a 30-bit code into one word, rejects the flag."""
    import pyrtx
    _say("63-bit Morton builds")
    sc = small_mixed_scene(seed=13, instances=14, freq=20, cells=20)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 96, 72), S.incoherent_rays(sc, 15000, seed=13)])
    tuv, ids, st = O.trace(t, rays)
    cost = {}
    for name, flags in (("30", pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT),
                        ("63", pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | pyrtx.RTX_BUILD_MORTON63_BIT)):
        blases, tlas = ctx.upload_scene(sc, flags=flags)
        ctx.sync()
        cost[name] = [ctx.as_info(b).sah_cost for b in blases]
        g = _trace_closest(ctx, tlas, rays)
        rep = compare_closest(g, tuv, ids, st, tlas=t, rays=rays)
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0, name
        assert_ties_ok(rep)
    _say(f"  SAH cost per BLAS: 30-bit {cost['30']}, 63-bit {cost['63']}")
    for a, b in zip(cost["30"], cost["63"]):
        assert abs(a - b) <= 0.05 * a
    with pytest.raises(pyrtx.RtxError) as e:
        ctx.upload_scene(sc, flags=pyrtx.RTX_BUILD_MORTON63_BIT, batched=True)
    assert e.value.code == -1000  # RTX_ERROR_INVALID_ARGUMENT
