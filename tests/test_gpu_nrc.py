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
    assert rel < 1e-2
    assert abs(loss - oloss) / oloss < 1e-3
    # per layer
    offs = [0, 6208, 10368, 14528, 18688, 22848, 23888]
    for l in range(6):
        a, b = g[offs[l]:offs[l + 1]], og[offs[l]:offs[l + 1]]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-9, l
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
    assert err <= 1e-3 * max(scale, 1.0)
    nrc.close()


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
