"""GPU tests of the neural-radiance-cache kernels against the FP64 oracle MLP.

The oracle emulates the GPU's BF16 rounding points (weights, input features, activations, layer gradients) and
accumulates in FP64, so the remaining differences are FP32 accumulation order and rare one-ulp BF16 flips; the
oracle's own backprop is pinned by finite differences in tests/test_oracle.py."""
import sys

import numpy as np
import pytest

import oracle as O

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


LO, HI = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)


def random_samples(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(0, 1, (n, 3))
    nrm = rng.standard_normal((n, 3))
    s[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.standard_normal((n, 3))
    s[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 9:12] = rng.uniform(0, 1, (n, 3))
    # a smooth radiance field to learn
    s[:, 12:15] = (0.5 + 0.4 * np.sin(3 * s[:, 0:3]) * s[:, 9:12]).astype(np.float32)
    return s


def test_gradient_matches_oracle(ctx):
    import pyrtx
    _say("nrc gradient")
    nrc = pyrtx.Nrc(ctx, LO, HI, seed=3, batch_size=1024)
    smp = random_samples(1000, 1)
    g, loss = nrc.gradient(torch.from_numpy(smp).cuda())
    p = nrc.params()
    _, og, oloss = O.nrc(p, smp, np.array(LO, np.float32), np.array(HI, np.float32), bf16=True)
    rel = np.linalg.norm(g - og) / np.linalg.norm(og)
    _say(f"  |g - g_oracle| / |g_oracle| = {rel:.3g}, loss {loss:.6f} vs {oloss:.6f}")
    assert rel < 1e-3  # measured ~2e-5
    assert abs(loss - oloss) / oloss < 1e-4
    # per layer
    offs = [0, 6208, 10368, 14528, 18688, 22848, 23888]
    for l in range(6):
        a, b = g[offs[l]:offs[l + 1]], og[offs[l]:offs[l + 1]]
        _say(f"  layer {l}: {np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30):.3g}")
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b) + 1e-9, l
    nrc.close()


def test_inference_matches_oracle(ctx):
    import pyrtx
    _say("nrc inference")
    nrc = pyrtx.Nrc(ctx, LO, HI, seed=5, batch_size=1024)
    smp = random_samples(777, 2)
    q = torch.from_numpy(np.ascontiguousarray(smp[:, :12])).cuda()
    y = nrc.infer(q).cpu().numpy()
    oy, _, _ = O.nrc(nrc.params(), smp, np.array(LO, np.float32), np.array(HI, np.float32), bf16=True,
                     want_grad=False)
    oy = np.maximum(oy, 0.0)
    err = np.abs(y - oy).max()
    scale = np.abs(oy).max()
    _say(f"  max |y - y_oracle| = {err:.3g} (scale {scale:.3g})")
    assert err <= 1e-4 * max(scale, 1.0)
    nrc.close()


def test_inference_encoding_bit_exact(ctx):
    """Every input feature of the GPU encoding, read through selector weights (a network that copies feature f to
    its outputs: out0 = max(0, f), out1 = max(0, -f)), equals the oracle's (closed-form-pinned) value exactly, at
    positions whose normalised coordinates are 0, 1/4, 1/2 and 1."""
    import pyrtx
    from test_oracle_pins import _selector
    _say("nrc encoding, feature by feature")
    lo = np.array([-2.0, -2.0, -2.0], np.float32)
    hi = np.array([2.0, 2.0, 2.0], np.float32)
    smp = np.zeros((12, 16), np.float32)
    for a in range(3):
        for i, x in enumerate((-2.0, -1.0, 0.0, 2.0)):
            smp[4 * a + i, a] = x
    smp[:, 3:12] = np.arange(1, 10, dtype=np.float32) * np.float32(0.0625) - np.float32(0.3)
    q = torch.from_numpy(np.ascontiguousarray(smp[:, :12])).cuda()
    nrc = pyrtx.Nrc(ctx, tuple(lo), tuple(hi), seed=1, batch_size=64)
    try:
        for f in range(96):
            p = _selector(f)
            nrc.set_params(p)
            y = nrc.infer(q).cpu().numpy()
            oy, _, _ = O.nrc(p, smp, lo, hi, bf16=True, want_grad=False)
            assert np.array_equal(y[:, 0], np.maximum(oy[:, 0], 0.0).astype(np.float32)), f
            assert np.array_equal(y[:, 1], np.maximum(oy[:, 1], 0.0).astype(np.float32)), f
    finally:
        nrc.close()


def test_render_nrc_frame_answers_queries_with_the_updated_cache(ctx):
    """Order within an NRC frame: the training paths' Adam steps come first, then the pixel paths' cache queries
    are answered with the updated parameters.  The frame's film equals a frame without training rendered with the
    parameters after the update (same sample index), and differs from one with the parameters before it."""
    import pyrtx
    from pyrtx import scenes as S
    _say("nrc frame order")
    W, H = 64, 48
    sc, lights = S.cornell_scene(3, width=W, height=H)
    blases, tlas = ctx.upload_scene(sc)
    lo, hi = S.scene_bounds(sc)
    nrc = pyrtx.Nrc(ctx, lo, hi, seed=1, batch_size=4096, cache_vertex=1)
    other = pyrtx.Nrc(ctx, lo, hi, seed=2, batch_size=4096, cache_vertex=1)
    try:
        for fr in range(2):
            ctx.render_nrc(tlas, sc.records, sc.camera, lights, nrc, W, H, spp=1, train_paths=8192, sample_offset=fr)
        before = nrc.params()
        f1 = ctx.render_nrc(tlas, sc.records, sc.camera, lights, nrc, W, H, spp=1, train_paths=8192,
                            sample_offset=5).clone()
        after = nrc.params()
        assert not np.array_equal(before, after)
        other.set_params(after)
        f_after = ctx.render_nrc(tlas, sc.records, sc.camera, lights, other, W, H, spp=1, train_paths=0,
                                 sample_offset=5).clone()
        other.set_params(before)
        f_before = ctx.render_nrc(tlas, sc.records, sc.camera, lights, other, W, H, spp=1, train_paths=0,
                                  sample_offset=5).clone()
        ctx.sync()
    finally:
        nrc.close()
        other.close()
    assert torch.equal(f1, f_after)
    assert not torch.equal(f1, f_before)


@pytest.mark.parametrize("wavefront", [False, True])
def test_train_only_frames_train_like_rendered_frames(ctx, monkeypatch, wavefront):
    """spp = 0 makes an NRC frame that only trains: the same training paths and Adam steps as a rendered frame of
    the same sample index, so the cache ends bit-identical, while the film it is given stays untouched.  Rendered
    frames after the training-only ones then equal those after rendered ones."""
    import pyrtx
    from pyrtx import scenes as S
    _say(f"nrc train-only frames (wavefront {wavefront})")
    if wavefront:
        monkeypatch.setenv("RTX_PATH_WAVEFRONT", "1")
    else:
        monkeypatch.delenv("RTX_PATH_WAVEFRONT", raising=False)
    W, H = 64, 48
    sc, lights = S.cornell_scene(3, width=W, height=H)
    blases, tlas = ctx.upload_scene(sc)
    lo, hi = S.scene_bounds(sc)
    rendered = pyrtx.Nrc(ctx, lo, hi, seed=4, batch_size=4096, cache_vertex=1)
    trained = pyrtx.Nrc(ctx, lo, hi, seed=4, batch_size=4096, cache_vertex=1)
    try:
        sentinel = torch.full((H, W, 4), 7.0, dtype=torch.float32, device="cuda")
        for fr in range(2):
            ctx.render_nrc(tlas, sc.records, sc.camera, lights, rendered, W, H, spp=1, train_paths=8192,
                           sample_offset=fr)
            ctx.render_nrc(tlas, sc.records, sc.camera, lights, trained, W, H, spp=0, train_paths=8192,
                           sample_offset=fr, film=sentinel)
        assert ctx.render_form() == (pyrtx.RTX_RENDER_FORM_WAVEFRONT if wavefront else
                                     pyrtx.RTX_RENDER_FORM_WHOLE_PATH)
        ctx.sync()
        assert rendered.status()[0] > 0
        assert rendered.status() == trained.status()
        assert np.array_equal(rendered.params(), trained.params())
        assert bool((sentinel == 7.0).all())
        f_r = ctx.render_nrc(tlas, sc.records, sc.camera, lights, rendered, W, H, spp=1, train_paths=8192,
                             sample_offset=2).clone()
        f_t = ctx.render_nrc(tlas, sc.records, sc.camera, lights, trained, W, H, spp=1, train_paths=8192,
                             sample_offset=2).clone()
        ctx.sync()
        assert torch.equal(f_r, f_t)
        assert np.array_equal(rendered.params(), trained.params())
        with pytest.raises(pyrtx.RtxError):
            ctx.render_nrc(tlas, sc.records, sc.camera, lights, trained, W, H, spp=0, train_paths=0)
    finally:
        rendered.close()
        trained.close()


def test_render_nrc_graph_replays_match_direct_frames(ctx):
    """NRC frames recorded into a HIP graph (both streams, no host synchronisation inside the frame) and replayed:
    film, parameters and step count equal direct frames from an identical cache, frame after frame, with other GPU
    work between the replays; the first two frames only train (spp 0, a graph of their own)."""
    import pyrtx
    from pyrtx import scenes as S
    _say("nrc frame graph replays")
    W, H = 80, 60
    sc, lights = S.cornell_scene(3, width=W, height=H)
    blases, tlas = ctx.upload_scene(sc)
    rec = torch.from_numpy(np.ascontiguousarray(sc.records).view(np.uint8).reshape(-1).copy()).cuda()
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).copy()).cuda()
    lo, hi = S.scene_bounds(sc)
    direct = pyrtx.Nrc(ctx, lo, hi, seed=1, batch_size=2048, cache_vertex=1)
    graph = pyrtx.Nrc(ctx, lo, hi, seed=1, batch_size=2048, cache_vertex=1)
    fd = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    fg = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    try:
        for fr in range(6):
            spp = 0 if fr < 2 else 2
            ctx.render_nrc(tlas, rec, sc.camera, lt, direct, W, H, spp=spp, train_paths=4096, sample_offset=fr,
                           film=fd)
            ctx.render_nrc(tlas, rec, sc.camera, lt, graph, W, H, spp=spp, train_paths=4096, sample_offset=fr,
                           film=fg, graph=True)
            ctx.sync()
            assert torch.equal(fd, fg), fr
            assert np.array_equal(direct.params(), graph.params()), fr
            assert direct.status() == graph.status(), fr
            x = torch.randn(1 << 20, device="cuda").cumsum(0)  # unrelated work between replays
            ctx.render(tlas, rec, sc.camera, lt, W, H, spp=1, sample_offset=100 + fr)
            ctx.sync()
            assert torch.isfinite(x).all()
            if fr < 2:
                assert not bool(fg.any()), fr  # training-only frames leave the film alone
        steps, _ = graph.status()
        assert steps >= 6
        assert float(fg[..., :3].mean()) > 0.01
    finally:
        direct.close()
        graph.close()


def test_adam_step_matches_restatement(ctx):
    import pyrtx
    _say("nrc adam")
    nrc = pyrtx.Nrc(ctx, LO, HI, seed=9, batch_size=512)
    smp = random_samples(512, 3)
    dev = torch.from_numpy(smp).cuda()
    p0 = nrc.params()
    g, _ = nrc.gradient(dev)
    nrc.train(dev)
    p1 = nrc.params()
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    pe = p0.copy()
    O.adam(pe, m, v, g.astype(np.float64), 1e-3, 0.9, 0.99, 1e-8, 1)
    assert np.abs(p1 - pe).max() < 1e-6
    assert nrc.status()[0] == 1
    nrc.close()


def test_adam_writes_the_packed_weights(ctx):
    """Adam writes each updated weight's BF16 fragment slots itself: inference after training must equal, bit for
    bit, inference after re-packing the same parameters from scratch; the step's loss equals the oracle's."""
    import pyrtx
    _say("nrc adam fragments")
    nrc = pyrtx.Nrc(ctx, LO, HI, seed=4, batch_size=2048)
    train = torch.from_numpy(random_samples(2048 * 3, 12)).cuda()
    for _ in range(2):
        nrc.train(train)
    q = torch.from_numpy(np.ascontiguousarray(random_samples(3001, 13)[:, :12])).cuda()
    y_fused = nrc.infer(q).cpu().numpy()
    p = nrc.params()
    nrc.set_params(p)
    y_repacked = nrc.infer(q).cpu().numpy()
    assert np.array_equal(y_fused, y_repacked)
    # the last batch's loss (summed per block, then over blocks) matches the oracle's loss for those samples
    last = train[2048 * 2:].cpu().numpy()
    _, loss = nrc.gradient(train[2048 * 2:])
    _, _, oloss = O.nrc(p, last, np.array(LO, np.float32), np.array(HI, np.float32), bf16=True)
    assert abs(loss - oloss) / oloss < 1e-3
    nrc.close()


def test_training_learns_a_smooth_field(ctx):
    import pyrtx
    _say("nrc training")
    nrc = pyrtx.Nrc(ctx, LO, HI, seed=1, batch_size=4096)
    train = torch.from_numpy(random_samples(4096 * 8, 10)).cuda()
    test = random_samples(2048, 11)
    tq = torch.from_numpy(np.ascontiguousarray(test[:, :12])).cuda()

    def test_loss():
        y = nrc.infer(tq).cpu().numpy()
        t = test[:, 12:15]
        return float(np.mean((y - t) ** 2 / (y ** 2 + 0.01)))

    l0 = test_loss()
    for _ in range(30):
        nrc.train(train)
    l1 = test_loss()
    steps, last = nrc.status()
    _say(f"  held-out relative L2: {l0:.4f} -> {l1:.5f} after {steps} Adam steps (last batch loss {last:.5f})")
    assert steps == 240
    assert l1 < 0.05 * l0
    nrc.close()


def _cornell_small(ctx, W, H):
    from pyrtx import scenes as S
    sc, lights = S.cornell_scene(5, width=W, height=H)
    blases, tlas = ctx.upload_scene(sc)
    rec = torch.from_numpy(np.ascontiguousarray(sc.records).view(np.uint8).reshape(-1).copy()).cuda()
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12).copy()).cuda()
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
    for k in range(16):
        ctx.render(tlas, rec, sc.camera, lt, W, H, spp=64, sample_offset=(1 << 20) + 64 * k, accum64=acc)
    ref = (acc / 1024).cpu().numpy()
    return sc, (blases, tlas), rec, lt, ref


def _relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def test_render_nrc_beats_plain_at_equal_samples(ctx):
    """The trained cache is what the NRC renderer gains: at the same number of samples per pixel its image is
    closer to the converged reference than plain path tracing, and its bias (mean radiance) stays small."""
    import pyrtx
    _say("render_nrc vs plain")
    W, H, frames = 320, 180, 48
    sc, (blases, tlas), rec, lt, ref = _cornell_small(ctx, W, H)
    plain = ctx.render(tlas, rec, sc.camera, lt, W, H, spp=frames)[..., :3].cpu().numpy().astype(np.float64)
    nrc = pyrtx.Nrc(ctx, (0, 0, 0), (1, 1, 1), seed=1, batch_size=4096, cache_vertex=1)
    try:
        accn = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
        for fr in range(frames):
            ctx.render_nrc(tlas, rec, sc.camera, lt, nrc, W, H, spp=1, train_paths=16384, sample_offset=fr,
                           accum64=accn)
        ctx.sync()
        img = (accn / frames).cpu().numpy()
        steps, loss = nrc.status()
    finally:
        nrc.close()
    e_plain, e_nrc = _relmse(plain, ref), _relmse(img, ref)
    bias = img.mean() / ref.mean() - 1.0
    _say(f"  relMSE plain {e_plain:.5f}, nrc {e_nrc:.5f}; mean bias {bias:+.4f}; {steps} Adam steps, loss {loss:.4f}")
    assert steps >= frames
    assert np.isfinite(img).all()
    assert e_nrc < 0.7 * e_plain
    assert abs(bias) < 0.05


def test_self_training_converges(ctx):
    """Training paths that end in a cache query of their own (train_suffix > 0): the frame images approach the
    reference as the cache trains (the early, untrained frames are biased)."""
    import pyrtx
    _say("render_nrc self-training")
    W, H = 320, 180
    sc, (blases, tlas), rec, lt, ref = _cornell_small(ctx, W, H)
    nrc = pyrtx.Nrc(ctx, (0, 0, 0), (1, 1, 1), seed=1, batch_size=4096, cache_vertex=1, train_suffix=3)
    errs = []
    try:
        for block in range(4):
            accn = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
            for fr in range(24):
                ctx.render_nrc(tlas, rec, sc.camera, lt, nrc, W, H, spp=1, train_paths=16384,
                               sample_offset=24 * block + fr, accum64=accn)
            errs.append(_relmse((accn / 24).cpu().numpy(), ref))
    finally:
        nrc.close()
    _say(f"  relMSE of 24-frame blocks: {[round(e, 5) for e in errs]}")
    assert all(np.isfinite(errs))
    assert errs[-1] < 0.5 * errs[0]


def test_cache_outliving_its_context(ctx):
    """Destroying a context before its cache leaves the cache destroyable (and every other call an error)."""
    import pyrtx
    c2 = pyrtx.Context(0)
    n2 = pyrtx.Nrc(c2, (0, 0, 0), (1, 1, 1), seed=2, batch_size=64)
    c2.close()
    with pytest.raises(pyrtx.RtxError):
        n2.status()
    n2.close()


def test_render_nrc_at_1080p_with_65536_training_paths(ctx):
    """The radiance-cache workload at its own size: the displaced level-8 icosphere Cornell box at 1920x1080, 65,536 training
    paths per frame, batches of 16,384, paths ending in the cache at the second bounce.  Over 32 frames the
    accumulated image approaches a 256-spp plain reference (error after 32 frames below half that after 8), beats
    plain path tracing at the same samples per pixel, and its mean radiance stays within 5% of the reference's
    (the cache is biased while it trains)."""
    import pyrtx
    from pyrtx import scenes as S
    _say("render_nrc, the radiance-cache workload at 1920x1080")
    W, H, frames = 1920, 1080, 32
    sc, lights = S.cornell_scene(8, width=W, height=H)
    blases, tlas = ctx.upload_scene(sc)
    rec = torch.from_numpy(np.ascontiguousarray(sc.records).view(np.uint8).reshape(-1).copy()).cuda()
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12).copy()).cuda()
    acc = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
    for k in range(4):
        ctx.render(tlas, rec, sc.camera, lt, W, H, spp=64, sample_offset=(1 << 20) + 64 * k, accum64=acc)
    ref = (acc / 256).cpu().numpy()
    plain = ctx.render(tlas, rec, sc.camera, lt, W, H, spp=frames)[..., :3].cpu().numpy().astype(np.float64)
    nrc = pyrtx.Nrc(ctx, (0, 0, 0), (1, 1, 1), seed=1, batch_size=16384, cache_vertex=1)
    try:
        accn = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
        errs = {}
        for fr in range(frames):
            ctx.render_nrc(tlas, rec, sc.camera, lt, nrc, W, H, spp=1, train_paths=65536, sample_offset=fr,
                           accum64=accn)
            if fr + 1 in (8, frames):
                ctx.sync()
                errs[fr + 1] = _relmse((accn / (fr + 1)).cpu().numpy(), ref)
        img = (accn / frames).cpu().numpy()
        steps, loss = nrc.status()
    finally:
        nrc.close()
    e_plain = _relmse(plain, ref)
    bias = img.mean() / ref.mean() - 1.0
    _say(f"  relMSE vs 256-spp reference: nrc after 8 frames {errs[8]:.5f}, after {frames} {errs[frames]:.5f}; "
         f"plain {frames} spp {e_plain:.5f}; mean bias {bias:+.4f}; {steps} Adam steps, loss {loss:.4f}")
    assert np.isfinite(img).all()
    assert steps >= 8 * frames
    # measured: 0.0189 after 8 frames, 0.0057 after 32, plain 32 spp 0.0117; mean bias -2.6%
    assert errs[frames] < 0.5 * errs[8]
    assert errs[frames] < 0.7 * e_plain
    assert abs(bias) < 0.05
