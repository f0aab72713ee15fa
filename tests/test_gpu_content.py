"""GPU tests of the per-structure content bits (non-opaque geometry, AABB primitives) that choose the traversal
variant and the integrator form per call, and of the traversal stack sized from the traced TLAS's depth.

A structure records at build time what rays through it can meet; a trace or render picks the opaque-triangle
traversal (and whole-path launches) or the one that resolves opacity and runs any-hit / intersection programs from
the traced TLAS alone, so building a cutout scene changes nothing for other scenes in the process."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import compare_closest, oracle_scene
from pyrtx import scenes as S

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


@pytest.fixture(autouse=True)
def _reset_oracle():
    O.reset()
    yield
    O.reset()


def _cutout_cornell(level=3, W=48, H=32):
    sc, lights = S.cornell_scene(level, width=W, height=H)
    sc.blas[1][0].opaque = False
    sc.records["any_hit"][4] = S.ANY_HIT_CUTOUT | (8 << 8)
    return sc, lights


def test_content_bits_per_structure(ctx):
    import pyrtx
    _say("content bits of BLASes and TLASes")
    opaque, _ = S.cornell_scene(2, width=16, height=16)
    b_op, t_op = ctx.upload_scene(opaque)
    assert all(ctx.as_info(b).content_flags == 0 for b in b_op)
    assert ctx.as_info(t_op).content_flags == 0
    cut, _ = _cutout_cornell(2)
    b_cut, t_cut = ctx.upload_scene(cut)
    assert [ctx.as_info(b).content_flags for b in b_cut] == [0, pyrtx.RTX_AS_CONTENT_NON_OPAQUE_BIT]
    assert ctx.as_info(t_cut).content_flags == pyrtx.RTX_AS_CONTENT_NON_OPAQUE_BIT
    proc, _ = S.cornell_scene(2, width=16, height=16)
    proc.blas[1] = [S.AabbGeometry(np.array([[0.3, 0.0, 0.1, 0.9, 0.6, 0.7]], np.float32))]
    b_pr, t_pr = ctx.upload_scene(proc)
    assert ctx.as_info(b_pr[1]).content_flags == pyrtx.RTX_AS_CONTENT_AABBS_BIT
    assert ctx.as_info(t_pr).content_flags == pyrtx.RTX_AS_CONTENT_AABBS_BIT
    # an instance with FORCE_NO_OPAQUE makes an opaque BLAS's triangles non-opaque
    rec = pyrtx.instance_records(opaque, [b.address for b in b_op])
    rec["sf"][1] |= np.uint32(pyrtx.RTX_INSTANCE_FORCE_NO_OPAQUE_BIT << 24)
    t_force = ctx.build_tlas(rec)
    assert ctx.as_info(t_force).content_flags == pyrtx.RTX_AS_CONTENT_NON_OPAQUE_BIT
    # batched builds: only the BLAS holding the non-opaque geometry carries the bit
    sc, views, lights = S.batch_scenes(count=4, seed=3, tri_lo=200, tri_hi=2000)
    sc.blas[2][0].opaque = False
    blases, tlas = ctx.upload_scene(sc, batched=True)
    assert [ctx.as_info(b).content_flags for b in blases] == [0, 0, 1, 0]
    assert ctx.as_info(tlas).content_flags == 1


def test_opaque_scene_renders_whole_paths_after_a_cutout_scene(ctx, monkeypatch):
    """A process that built and rendered a cutout scene still renders an opaque scene with whole-path launches (with
    the film the per-vertex wavefront gives), and the cutout scene keeps the wavefront with any-hit programs."""
    import pyrtx
    _say("integrator form per TLAS")
    monkeypatch.delenv("RTX_PATH_WAVEFRONT", raising=False)
    W, H = 64, 48
    cut, lights = _cutout_cornell(3, W, H)
    b_cut, t_cut = ctx.upload_scene(cut)  # the BLASes must outlive every trace of the TLAS
    f_cut = ctx.render(t_cut, cut.records, cut.camera, lights, W, H, spp=2).clone()
    assert ctx.render_form() == pyrtx.RTX_RENDER_FORM_WAVEFRONT
    opaque, lights2 = S.cornell_scene(3, width=W, height=H)
    b_op, t_op = ctx.upload_scene(opaque)
    f_whole = ctx.render(t_op, opaque.records, opaque.camera, lights2, W, H, spp=2).clone()
    assert ctx.render_form() == pyrtx.RTX_RENDER_FORM_WHOLE_PATH
    monkeypatch.setenv("RTX_PATH_WAVEFRONT", "1")
    f_wave = ctx.render(t_op, opaque.records, opaque.camera, lights2, W, H, spp=2).clone()
    assert ctx.render_form() == pyrtx.RTX_RENDER_FORM_WAVEFRONT
    monkeypatch.delenv("RTX_PATH_WAVEFRONT")
    ctx.sync()
    assert torch.equal(f_whole, f_wave)
    assert float((f_whole - f_cut).abs().max()) > 0.05  # the holes matter
    # and the cutout scene again: wavefront, same film as before
    f_cut2 = ctx.render(t_cut, cut.records, cut.camera, lights, W, H, spp=2)
    assert ctx.render_form() == pyrtx.RTX_RENDER_FORM_WAVEFRONT
    ctx.sync()
    assert torch.equal(f_cut, f_cut2)


def test_graph_follows_tlas_rebuilt_with_new_content(ctx):
    """A recorded batched-render graph bakes in the traversal variant and the integrator form.  Rebuilding the TLAS
    in place (same address, same record table) with FORCE_NO_OPAQUE instances and cutout any-hit programs must
    record a new graph: the replayed film equals the direct call's, not the opaque scene's."""
    import pyrtx
    _say("batched graph after an in-place TLAS rebuild")
    sc, views, lights = S.batch_scenes(count=4, seed=3, tri_lo=200, tri_hi=5000)
    for k, v in enumerate(views):
        v["width"], v["height"] = [(40, 30), (64, 48), (33, 17), (20, 20)][k]
    blases, tlas = ctx.upload_scene(sc, batched=True)
    rec = torch.from_numpy(np.ascontiguousarray(sc.records).view(np.uint8).reshape(-1).copy()).cuda()
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).copy()).cuda()
    g1 = ctx.render_batch(tlas, rec, views, lt, spp=2, graph=True).clone()
    g1b = ctx.render_batch(tlas, rec, views, lt, spp=2, graph=True).clone()
    ctx.sync()
    assert torch.equal(g1, g1b)
    # cutout programs on every record, every instance non-opaque, TLAS rebuilt into its own storage
    r2 = sc.records.copy()
    r2["any_hit"][:] = S.ANY_HIT_CUTOUT | (6 << 8)
    rec.copy_(torch.from_numpy(np.ascontiguousarray(r2).view(np.uint8).reshape(-1).copy()).cuda())
    inst = pyrtx.instance_records(sc, [b.address for b in blases])
    inst["sf"] |= np.uint32(pyrtx.RTX_INSTANCE_FORCE_NO_OPAQUE_BIT << 24)
    t2 = ctx.build_tlas(inst, into=tlas)
    assert t2.address == tlas.address
    assert ctx.as_info(t2).content_flags == pyrtx.RTX_AS_CONTENT_NON_OPAQUE_BIT
    g2 = ctx.render_batch(t2, rec, views, lt, spp=2, graph=True).clone()
    assert ctx.render_form() == pyrtx.RTX_RENDER_FORM_WAVEFRONT
    d2 = ctx.render_batch(t2, rec, views, lt, spp=2, graph=False).clone()
    ctx.sync()
    assert torch.equal(g2, d2)
    assert float((g2 - g1).abs().max()) > 0.05


def _comb_scene(n=3000, seed=11):
    """A deliberately deep BVH: small triangles at geometrically spaced distances along a diagonal (each one
    splits off at its own Morton level), in several rotated instances."""
    rng = np.random.default_rng(seed)
    d = 1.004 ** np.arange(n)                                  # 1 .. ~1.6e5
    c = np.stack([d, d * 0.7, d * 0.4], 1) + rng.normal(0, 0.01, (n, 3)) * d[:, None]
    tri = c[:, None, :] + rng.normal(0, 0.05, (n, 3, 3)) * d[:, None, None]
    g = S.Geometry(tri.reshape(-1, 3).astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3))
    k = 6
    R = S.random_rotation(rng, k)
    xf = np.zeros((k, 3, 4))
    xf[:, :, :3] = R * 1e-4
    xf[:, :, 3] = rng.uniform(-3, 3, (k, 3))
    return S.Scene(blas=[[g]], transforms=xf.astype(np.float32), blas_of_instance=np.zeros(k, np.int64),
                   custom_index=np.arange(k, dtype=np.uint32), mask=np.full(k, 0xFF, np.uint32),
                   sbt_offset=np.zeros(k, np.uint32), flags=np.zeros(k, np.uint32), camera=(0.0, 0.0, 30.0, 60.0),
                   width=64, height=64, meta={"name": "comb"})


def test_stack_sized_from_depth_and_overflow_reported(ctx):
    """The derived stack (2 entries per BVH level + 4) is enough for every ray (no error, hits equal the oracle's,
    also with most of it in the global spill area); a stack limit of one entry makes rays run out of stack: they
    end early and rtx_sync reports RTX_ERROR_STACK_OVERFLOW, after which the context traces normally again."""
    import pyrtx
    from test_gpu_parity import _trace_closest
    _say("stack depth from the TLAS, forced overflow")
    sc = _comb_scene()
    blases, tlas = ctx.upload_scene(sc)
    depth = ctx.as_info(tlas).max_depth
    _say(f"  TLAS + BLAS depth {depth} (BLAS {ctx.as_info(blases[0]).max_depth})")
    assert depth >= 4
    rays = np.concatenate([S.camera_rays(sc.camera, 96, 96), S.incoherent_rays(sc, 20000, seed=3)])
    t = oracle_scene(sc)
    o_tuv, o_ids, o_st = O.trace(t, rays)
    base = ctx.config()
    try:
        for lds in (12, 2):
            ctx.set_config(lds_stack_depth=lds)
            got = _trace_closest(ctx, tlas, rays)
            ctx.sync()
            rep = compare_closest(got, o_tuv, o_ids, o_st)
            assert rep["robust_hits"] > 500
            assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0 and rep["t_bad"] == 0, rep
        ctx.set_config(stack_limit=1)
        with pytest.raises(pyrtx.RtxError) as e:
            _trace_closest(ctx, tlas, rays)  # synchronises
        assert e.value.code == -1001
        ctx.set_config(stack_limit=pyrtx.RTX_TRACE_STACK_LIMIT_DERIVED)
        assert ctx.config()["stack_limit"] == 0
        got = _trace_closest(ctx, tlas, rays)
        ctx.sync()
        rep = compare_closest(got, o_tuv, o_ids, o_st)
        assert rep["id_mismatch"] == 0 and rep["miss_mismatch"] == 0
    finally:
        ctx.set_config(**dict(base, stack_limit=pyrtx.RTX_TRACE_STACK_LIMIT_DERIVED))
