"""GPU tests of batching: several scenes of uneven size built in one batched BLAS build and rendered by one
wavefront (rtx_cmd_render_batch), directly and through its recorded HIP graph."""
import sys

import numpy as np
import pytest

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


def _small_batch(count=6):
    from pyrtx import scenes as S
    sc, views, lights = S.batch_scenes(count=count, seed=3, tri_lo=200, tri_hi=20000)
    for k, v in enumerate(views):  # small films of different shapes
        v["width"], v["height"] = [(40, 30), (64, 48), (33, 17)][k % 3]
    return sc, views, lights


def test_batched_render_equals_single_scene_renders(ctx):
    """Each view of a batched render equals rendering its scene alone (same instance, camera, lights, random
    numbers): the clip box keeps every ray inside its scene."""
    import pyrtx
    _say("batched render vs single-scene renders")
    sc, views, lights = _small_batch()
    blases, tlas = ctx.upload_scene(sc, batched=True)
    films = ctx.render_batch(tlas, sc.records, views, lights, spp=8, max_depth=6)
    ctx.sync()
    f = films.cpu().numpy()
    inst = pyrtx.instance_records(sc, [b.address for b in blases])
    base, worst, exact = 0, 0.0, 0
    for k, v in enumerate(views):
        t1 = ctx.build_tlas(inst[k:k + 1])
        one = ctx.render(t1, sc.records, v["camera"], lights[v["light_first"]:v["light_first"] + v["light_count"]],
                         v["width"], v["height"], spp=8, max_depth=6)
        ctx.sync()
        o = one.cpu().numpy().reshape(-1, 4)
        g = f[base:base + v["width"] * v["height"]]
        base += v["width"] * v["height"]
        worst = max(worst, float(np.abs(g - o).max()))
        exact += int(np.array_equal(g, o))
        assert np.isfinite(g).all() and g[:, :3].mean() > 0.01
    _say(f"  {exact} of {len(views)} views bit-identical, max |diff| {worst:.3g}")
    assert worst <= 1e-5


def test_batch_graph_replays_match_direct_calls(ctx):
    """The recorded graph replays (with other GPU work in between) give the direct call's films, for any
    sample_offset."""
    _say("batched render graph replays")
    sc, views, lights = _small_batch(5)
    blases, tlas = ctx.upload_scene(sc, batched=True)
    direct = {}
    for so in (0, 16):
        direct[so] = ctx.render_batch(tlas, sc.records, views, lights, spp=4, sample_offset=so).clone()
    ctx.sync()
    for rep, so in enumerate((0, 16, 0, 16, 0)):
        got = ctx.render_batch(tlas, sc.records, views, lights, spp=4, sample_offset=so, graph=True).clone()
        ctx.sync()
        assert torch.equal(got, direct[so]), (rep, so)
        # unrelated work between replays: torch kernels, a direct batched render, a single render
        x = torch.randn(1 << 20, device="cuda").cumsum(0)
        ctx.render_batch(tlas, sc.records, views, lights, spp=1, sample_offset=99)
        ctx.sync()
        assert torch.isfinite(x).all()
    assert not torch.equal(direct[0], direct[16])


def test_batched_whole_paths_match_wavefront(ctx, monkeypatch):
    """A batched render traces whole paths in one launch per sample; RTX_PATH_WAVEFRONT=1 runs the per-vertex
    wavefront instead: bit-identical films."""
    _say("batched render: whole paths vs wavefront")
    sc, views, lights = _small_batch(4)
    blases, tlas = ctx.upload_scene(sc, batched=True)
    monkeypatch.delenv("RTX_PATH_WAVEFRONT", raising=False)
    whole = ctx.render_batch(tlas, sc.records, views, lights, spp=2, max_depth=6).clone()
    monkeypatch.setenv("RTX_PATH_WAVEFRONT", "1")
    wave = ctx.render_batch(tlas, sc.records, views, lights, spp=2, max_depth=6).clone()
    monkeypatch.delenv("RTX_PATH_WAVEFRONT")
    ctx.sync()
    assert float(whole[:, :3].mean()) > 0.01
    assert torch.equal(whole, wave)


def _relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def test_batched_views_match_fp64_oracle(ctx):
    """The batched-scenes workload's scene kind at oracle-affordable films: 8 uneven scenes (seed 1: 2,991 to 1,372,489
    sphere triangles) in one batched build and one batched render (films 256^2 / 512^2 / 1024x768 scaled by 1/8),
    each view against the FP64 oracle's estimator on its scene alone with the same camera, lights and random
    numbers (the clip box keeps a batched view's rays inside its scene)."""
    import oracle as O
    from helpers import oracle_scene
    from pyrtx import scenes as S
    _say("batched views vs the FP64 oracle (8 uneven scenes, one above 1M triangles)")
    sc, views, lights = S.batch_scenes(count=8, seed=1)
    assert max(v["triangles"] for v in views) > 1_000_000
    for v in views:
        v["width"], v["height"] = v["width"] // 8, v["height"] // 8
    spp, depth = 4, 6
    blases, tlas = ctx.upload_scene(sc, batched=True)
    films = ctx.render_batch(tlas, sc.records, views, lights, spp=spp, max_depth=depth)
    ctx.sync()
    f = films.cpu().numpy()
    base, worst_err, worst_close = 0, 0.0, 1.0
    try:
        for k, v in enumerate(views):
            w, h = v["width"], v["height"]
            one = S.Scene(blas=[sc.blas[k]], transforms=sc.transforms[k:k + 1], blas_of_instance=np.array([0]),
                          custom_index=sc.custom_index[k:k + 1], mask=sc.mask[k:k + 1],
                          sbt_offset=sc.sbt_offset[k:k + 1], flags=sc.flags[k:k + 1], records=sc.records,
                          camera=v["camera"], width=w, height=h, meta={"name": f"batch view {k}"})
            O.reset()
            t = oracle_scene(one)
            lt = lights[v["light_first"]:v["light_first"] + v["light_count"]]
            o = O.render(t, sc.records, v["camera"], lt, w, h, spp, max_depth=depth) / spp
            g = f[base:base + w * h, :3].astype(np.float64).reshape(h, w, 3)
            base += w * h
            err = _relmse(g, o)
            close = float((np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o))).all(-1).mean())
            _say(f"  view {k}: {v['triangles']} triangles, {w}x{h}: relMSE {err:.3g}, pixels within 1e-3 {close:.5f}")
            worst_err, worst_close = max(worst_err, err), min(worst_close, close)
            assert np.isfinite(g).all() and g.mean() > 0.01
    finally:
        O.reset()
    # measured: relMSE 1.5e-15 .. 3.8e-11 per view, every pixel within 1e-3
    assert worst_err < 4e-10
    assert worst_close >= 0.9999


def test_full_64_scene_batch(ctx):
    """The bench's 64-scene batch itself (64 scenes, seed 7: 15.5M triangles, 23.3M pixels) for one sample:
    every view finite, non-negative and lit, and the views of the largest and the smallest scene equal to rendering
    those scenes alone."""
    import pyrtx
    from pyrtx import scenes as S
    _say("64-scene batch, one sample")
    sc, views, lights = S.batch_scenes(count=64, seed=7)
    blases, tlas = ctx.upload_scene(sc, batched=True)
    films = ctx.render_batch(tlas, sc.records, views, lights, spp=1, max_depth=6)
    ctx.sync()
    f = films.cpu().numpy()
    assert np.isfinite(f).all() and (f[:, :3] >= 0).all()
    starts = np.cumsum([0] + [v["width"] * v["height"] for v in views])
    means = [float(f[starts[k]:starts[k + 1], :3].mean()) for k in range(len(views))]
    _say(f"  {len(views)} views, {starts[-1]} pixels, mean radiance per view {min(means):.3f} .. {max(means):.3f}")
    assert min(means) > 0.01
    inst = pyrtx.instance_records(sc, [b.address for b in blases])
    tri = [v["triangles"] for v in views]
    for k in (int(np.argmax(tri)), int(np.argmin(tri))):
        v = views[k]
        t1 = ctx.build_tlas(inst[k:k + 1])
        one = ctx.render(t1, sc.records, v["camera"], lights[v["light_first"]:v["light_first"] + v["light_count"]],
                         v["width"], v["height"], spp=1, max_depth=6)
        ctx.sync()
        d = float(np.abs(one.cpu().numpy().reshape(-1, 4) - f[starts[k]:starts[k + 1]]).max())
        _say(f"  view {k} ({tri[k]} triangles) vs its scene alone: max |diff| {d:.3g}")
        assert d <= 1e-5
