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
