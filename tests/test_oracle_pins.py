"""CPU pins for the oracle's neural-radiance-cache MLP, its Adam update and pcg4d.

None of these three has a counterpart in the reference (its arithmetic lives in the Vulkan driver and it has no
tests); they come from the radiance-cache specification this library implements: a frequency-encoded MLP (position 3 x 12 frequencies x
sin / cos, normal, direction, albedo; 5 hidden layers of 64, ReLU, RGB output; BF16 weights and activations with
FP32 accumulation, FP32 biases; relative L2 loss), Adam (Kingma & Ba 2015) and pcg4d (Jarzynski & Olano 2020,
JCGT 9(3)).  Each function is checked against something it shares no code with:

- the MLP forward and backward against an independent torch (CPU, float64) model built from that specification,
  with the BF16 rounding points as explicit torch.bfloat16 casts, in FP32 and in BF16 form;
- the input encoding, the bias placement and the ReLU placement against closed forms, read through selector
  weights (a network that copies one feature to an output);
- Adam against hand-derived known answers (first step -lr g / (|g| + eps), a two-step example, the constant-gradient
  closed form, scale invariance);
- pcg4d against hand-derived known answers of the published code and its avalanche structure, with negative
  controls (mutated variants written here) that each check must reject.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O

P = O.NRC_PARAM_COUNT
IN = [96, 64, 64, 64, 64, 64]
OUT = [64, 64, 64, 64, 64, 16]


def w_off(l):
    return sum(OUT[k] * IN[k] + OUT[k] for k in range(l))


def b_off(l):
    return w_off(l) + OUT[l] * IN[l]


assert b_off(5) + 16 == P


# ---- independent torch model of the cache network ------------------------------------------------------------------
def _bf16(x: torch.Tensor) -> torch.Tensor:
    """Round a float64 tensor as the GPU does: the FP32 value, rounded to nearest-even BF16."""
    return x.to(torch.float32).to(torch.bfloat16).to(torch.float64)


class _RoundGrad(torch.autograd.Function):
    """Identity forward; the incoming gradient is rounded to BF16 (the backward pass keeps layer gradients in
    BF16 MFMA operands)."""

    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return _bf16(g)


class _RoundST(torch.autograd.Function):
    """BF16 rounding forward, identity backward (straight-through to the FP32 master value)."""

    @staticmethod
    def forward(ctx, x):
        return _bf16(x)

    @staticmethod
    def backward(ctx, g):
        return g


def torch_nrc(params: np.ndarray, samples: np.ndarray, lo, hi, bf16: bool):
    """Prediction, gradient of the mean loss with respect to the FP32 parameters, and the mean loss, from the
    specification alone.  Features: position p normalised to [lo, hi] (FP32: (q - lo) * (1 / (hi - lo))); for
    axis a and frequency k < 12, feature 24 a + 2 k = sin(2^k pi p_a) and 24 a + 2 k + 1 = cos(2^k pi p_a); then
    normal (72-74), direction (75-77), albedo (78-80), zeros (81-95).  Layer l: z = W_l a + b_l; hidden a = ReLU(z)
    (rounded to BF16); output rows 0..2 = RGB.  Loss: mean over samples of sum over RGB of (y - t)^2 / (y^2 + 0.01)
    with the denominator held constant."""
    f64 = torch.float64
    rnd = _RoundST.apply if bf16 else (lambda x: x)
    rgrad = _RoundGrad.apply if bf16 else (lambda x: x)
    smp = torch.from_numpy(np.ascontiguousarray(samples, np.float32))
    q = smp[:, :12]
    lo32 = torch.tensor(lo, dtype=torch.float32)
    hi32 = torch.tensor(hi, dtype=torch.float32)
    inv = torch.where(hi32 - lo32 > 0, 1.0 / (hi32 - lo32), torch.zeros(3))
    pn = ((q[:, :3] - lo32) * inv).to(f64)
    n = smp.shape[0]
    feats = torch.zeros(n, 96, dtype=f64)
    for a in range(3):
        for k in range(12):
            x = torch.fmod(pn[:, a] * 2.0 ** k, 2.0) * math.pi
            feats[:, 24 * a + 2 * k] = torch.sin(x)
            feats[:, 24 * a + 2 * k + 1] = torch.cos(x)
    feats[:, 72:81] = q[:, 3:12].to(f64)
    if bf16:
        feats = _bf16(feats)
    p = torch.tensor(np.asarray(params, np.float32).astype(np.float64), requires_grad=True)
    a = feats
    for l in range(6):
        W = p[w_off(l):w_off(l) + OUT[l] * IN[l]].view(OUT[l], IN[l])
        b = p[b_off(l):b_off(l) + OUT[l]]
        z = rgrad(a @ rnd(W).T + b)
        a = rnd(torch.relu(z)) if l < 5 else z
    y = a[:, :3]
    t = smp[:, 12:15].to(f64)
    loss = (((y - t) ** 2) / (y.detach() ** 2 + 0.01)).sum(1).mean()
    loss.backward()
    return y.detach().numpy(), p.grad.numpy(), float(loss.detach())


def _random_case(seed, n=48):
    rng = np.random.default_rng(seed)
    params = np.zeros(P, np.float32)
    for l in range(6):
        bound = math.sqrt(6.0 / IN[l])
        params[w_off(l):b_off(l)] = rng.uniform(-bound, bound, OUT[l] * IN[l])
        params[b_off(l):b_off(l) + OUT[l]] = rng.uniform(-0.1, 0.1, OUT[l])
    lo = np.array([-1.5, 0.0, -3.0], np.float32)
    hi = np.array([2.5, 1.0, 1.0], np.float32)
    smp = np.zeros((n, 16), np.float32)
    smp[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    nrm = rng.normal(size=(n, 3))
    smp[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    smp[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    smp[:, 9:12] = rng.uniform(0, 1, (n, 3))
    smp[:, 12:15] = rng.uniform(0, 3, (n, 3))
    return params, smp, lo, hi


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("seed", [0, 1])
def test_nrc_oracle_matches_independent_torch_model(bf16, seed):
    params, smp, lo, hi = _random_case(seed)
    y, g, loss = O.nrc(params, smp, lo, hi, bf16=bf16)
    ty, tg, tloss = torch_nrc(params, smp, lo, hi, bf16)
    assert np.abs(y - ty).max() <= 1e-9 * max(1.0, np.abs(ty).max())
    assert abs(loss - tloss) <= 1e-9 * tloss
    scale = np.abs(tg).max()
    assert scale > 0
    # every layer's weights and biases separately (a dropped or misplaced term shows in its own block)
    for l in range(6):
        for lo_, hi_ in ((w_off(l), b_off(l)), (b_off(l), b_off(l) + OUT[l])):
            blk = slice(lo_, hi_)
            err = np.abs(g[blk] - tg[blk]).max()
            assert err <= 1e-9 * scale, (l, lo_, err, scale)
    # the parameters that no feature or output reaches get no gradient
    dead = np.zeros(P, bool)
    dead[w_off(0):b_off(0)] = (np.arange(64 * 96) % 96) >= 81
    dead[w_off(5):b_off(5)] = (np.arange(16 * 64) // 64) >= 3
    dead[b_off(5) + 3:b_off(5) + 16] = True
    assert np.all(g[dead] == 0.0)


def test_bf16_form_differs_from_fp32_form():
    """Negative control for the comparison above: the BF16 and FP32 forms disagree far beyond its tolerance, so the
    torch model does see the rounding points."""
    params, smp, lo, hi = _random_case(3)
    y32 = torch_nrc(params, smp, lo, hi, False)[0]
    y16 = O.nrc(params, smp, lo, hi, bf16=True)[0]
    assert np.abs(y32 - y16).max() > 1e-4


def _selector(feature, out_row=0, neg_row=1):
    """Parameters of a network whose output row out_row is ReLU(f) - ReLU(-f) = f for input feature `feature`, and
    whose row neg_row is ReLU(-f): layer 0 puts +f in unit 0 and -f in unit 1, hidden layers copy units 0 and 1."""
    p = np.zeros(P, np.float32)
    p[w_off(0) + 0 * 96 + feature] = 1.0
    p[w_off(0) + 1 * 96 + feature] = -1.0
    for l in range(1, 5):
        p[w_off(l) + 0 * 64 + 0] = 1.0
        p[w_off(l) + 1 * 64 + 1] = 1.0
    p[w_off(5) + out_row * 64 + 0] = 1.0
    p[w_off(5) + out_row * 64 + 1] = -1.0
    p[w_off(5) + neg_row * 64 + 1] = 1.0
    return p


def _bf16_np(x):
    return _bf16(torch.tensor(np.asarray(x, np.float64))).numpy()


def test_nrc_encoding_closed_form():
    """Every one of the 96 input features read through selector weights, at positions whose normalised coordinate
    is 0, 1/4, 1/2 and 1: sin / cos(2^k pi p) take the closed-form values 0, +-1, +-sqrt(2)/2 (k = 0 at p = 1/4)."""
    lo = np.array([-2.0, -2.0, -2.0], np.float32)
    hi = np.array([2.0, 2.0, 2.0], np.float32)
    pos = np.array([-2.0, -1.0, 0.0, 2.0], np.float32)   # normalised 0, 1/4, 1/2, 1
    pn = [0.0, 0.25, 0.5, 1.0]
    smp = np.zeros((12, 16), np.float32)
    # sample 4 a + i: axis a at pos[i], the other axes at 0 (normalised 1/2)
    for a in range(3):
        for i in range(4):
            smp[4 * a + i, 0:3] = 0.0
            smp[4 * a + i, a] = pos[i]
    smp[:, 3:12] = np.arange(1, 10, dtype=np.float32) * np.float32(0.0625) - np.float32(0.3)

    def closed(k, p, cos):
        # sin / cos(2^k pi p) for p in {0, 1/4, 1/2, 1}, by hand
        x = (2 ** k) * p  # multiples of pi
        if x == 0.25:
            return math.sqrt(0.5)
        if x == 0.5:
            return 0.0 if cos else 1.0
        half_turns = int(x)  # x is an integer here
        return (1.0 if half_turns % 2 == 0 else -1.0) if cos else 0.0

    for f in range(96):
        y, _, _ = O.nrc(_selector(f), smp, lo, hi, bf16=True, want_grad=False)
        for s in range(12):
            a_s, i = divmod(s, 4)
            if f < 72:
                a, r = divmod(f, 24)
                k, cos = divmod(r, 2)
                p = pn[i] if a == a_s else 0.5
                want = closed(k, p, cos)
            elif f < 81:
                want = float(smp[s, 3 + f - 72])
            else:
                want = 0.0
            want = float(_bf16_np(want))
            assert y[s, 0] == want, (f, s, y[s, 0], want)
            assert y[s, 1] == max(0.0, -want)


def test_nrc_bias_and_relu_placement():
    """Biases are added before each ReLU and to the output; hidden layers clamp at 0, the output does not."""
    smp = np.zeros((2, 16), np.float32)
    smp[:, 0:3] = 0.5
    lo, hi = np.zeros(3, np.float32), np.ones(3, np.float32)
    p = np.zeros(P, np.float32)
    p[b_off(5) + 0] = -0.5          # output bias, no ReLU after it
    p[b_off(0) + 2] = -1.0          # hidden unit 2 of layer 0: ReLU(-1) = 0
    p[b_off(0) + 3] = 1.25          # hidden unit 3: 1.25, copied by layers 1..4
    for l in range(1, 5):
        p[w_off(l) + 2 * 64 + 2] = 1.0
        p[w_off(l) + 3 * 64 + 3] = 1.0
    p[b_off(4) + 4] = 0.75          # a bias of the last hidden layer
    p[w_off(5) + 1 * 64 + 2] = 1.0  # output 1 = unit 2 (0 after the ReLU)
    p[w_off(5) + 1 * 64 + 3] = 1.0  # + unit 3
    p[w_off(5) + 2 * 64 + 4] = -2.0  # output 2 = -2 x unit 4 of layer 4 (negative, kept)
    y, _, _ = O.nrc(p, smp, lo, hi, bf16=True, want_grad=False)
    assert np.all(y[:, 0] == -0.5)
    assert np.all(y[:, 1] == 1.25)
    assert np.all(y[:, 2] == -1.5)


# ---- Adam ----------------------------------------------------------------------------------------------------------
def _adam(p0, grads, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    p = np.array(p0, np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    out = []
    for s, g in enumerate(grads, start=1):
        O.adam(p, m, v, np.asarray(g, np.float64), lr, b1, b2, eps, s)
        out.append(p.copy())
    return out, m, v


def test_adam_known_answers():
    # step 1: m = (1 - b1) g, v = (1 - b2) g^2; the bias corrections 1/(1 - b1), 1/(1 - b2) undo the factors, so
    # the update is -lr g / (|g| + eps)
    (p1,), m, v = _adam([1.0, -2.0, 0.0], [[0.5, -4.0, 0.0]])
    assert np.allclose(m, [0.05, -0.4, 0.0], rtol=1e-6) and np.allclose(v, [0.0025, 0.16, 0.0], rtol=1e-6)
    assert np.allclose(p1, [1.0 - 1e-3 * 0.5 / (0.5 + 1e-8), -2.0 + 1e-3 * 4.0 / (4.0 + 1e-8), 0.0], rtol=0, atol=1e-7)
    # step 2 with g = -0.25 after g = 0.5 (b1 0.9, b2 0.99, by hand):
    #   m2 = 0.9 * 0.05 + 0.1 * (-0.25) = 0.02                 m^ = 0.02 / (1 - 0.81) = 0.105263158
    #   v2 = 0.99 * 0.0025 + 0.01 * 0.0625 = 0.0031            v^ = 0.0031 / (1 - 0.9801) = 0.155778894
    #   update = -1e-3 * 0.105263158 / (sqrt(0.155778894) + 1e-8) = -2.66699e-4
    (_, p2), m, v = _adam([1.0], [[0.5], [-0.25]])
    # m and v are stored in FP32 between the steps and b1, b2 are FP32 (0.99f = 0.99000001)
    assert abs(float(m[0]) - 0.02) < 2e-6 * 0.02 and abs(float(v[0]) - 0.0031) < 2e-6 * 0.0031
    assert abs(float(p2[0]) - (1.0 - 1e-3 * 0.5 / (0.5 + 1e-8) - 2.66699e-4)) < 2e-7


def test_adam_constant_gradient_and_scale_invariance():
    # a constant gradient g gives m^ = g and v^ = g^2 at every step (bias correction exact), so with eps = 0 each
    # step moves by exactly lr against sign(g); a wrong correction exponent breaks this from step 2 on
    ps, _, _ = _adam([0.0, 0.0], [[3.0, -0.125]] * 6, lr=0.01, eps=0.0)
    for s, p in enumerate(ps, start=1):
        assert np.allclose(p, [-0.01 * s, 0.01 * s], rtol=1e-5)
    # scaling every gradient by 1000 leaves the updates unchanged (eps = 0)
    rng = np.random.default_rng(5)
    g = rng.normal(size=(5, 8))
    a, _, _ = _adam(np.ones(8), g, eps=0.0)
    b, _, _ = _adam(np.ones(8), g * 1000.0, eps=0.0)
    assert np.allclose(a[-1], b[-1], rtol=1e-6)


# ---- pcg4d ---------------------------------------------------------------------------------------------------------
_MUL, _ADD = 1664525, 1013904223


def _lcg_preimage(v):
    """Inputs whose first step (v * 1664525 + 1013904223 mod 2^32, the published constants) gives v."""
    inv = pow(_MUL, -1, 1 << 32)
    return np.array([((int(x) - _ADD) * inv) % (1 << 32) for x in v], np.uint32)


def test_pcg4d_known_answers():
    # after the LCG step, pcg4d runs x += y w; y += z x; z += x y; w += y z; v ^= v >> 16; the same four updates
    # again.  By hand for small post-LCG values:
    #   (0, 0, 0, 0)       -> (0, 0, 0, 0)
    #   (1, 1, 1, 1)       -> round 1: x = 1 + 1 = 2, y = 1 + 1*2 = 3, z = 1 + 2*3 = 7, w = 1 + 3*7 = 22 (shift: no
    #                         change below 2^16); round 2: x = 2 + 3*22 = 68, y = 3 + 7*68 = 479,
    #                         z = 7 + 68*479 = 32579, w = 22 + 479*32579 = 15605363
    #   (2^16, 0, 0, 0)    -> round 1 leaves it; the shift gives x = 2^16 ^ 1 = 65537; round 2 leaves it
    cases = {(0, 0, 0, 0): (0, 0, 0, 0), (1, 1, 1, 1): (68, 479, 32579, 15605363),
             (65536, 0, 0, 0): (65537, 0, 0, 0), (0, 65536, 0, 0): (0, 65537, 0, 0)}
    for post, want in cases.items():
        got = O.pcg4d(_lcg_preimage(post)[None, :])[0]
        assert tuple(int(x) for x in got) == want, (post, got)
    # negative control: a mutated LCG multiplier written here misses the first known answer
    inv_wrong = pow(_MUL + 2, -1, 1 << 32)
    pre = np.array([((0 - _ADD) * inv_wrong) % (1 << 32)] * 4, np.uint32)
    assert tuple(int(x) for x in O.pcg4d(pre[None, :])[0]) != (0, 0, 0, 0)


def _pcg4d_variant(v, rounds=2, shift=16, xorshift=True):
    """The published pcg4d with one mutation (negative controls of the avalanche test)."""
    v = v.astype(np.uint32).copy()
    with np.errstate(over="ignore"):
        v = v * np.uint32(_MUL) + np.uint32(_ADD)
        x, y, z, w = (v[:, i].copy() for i in range(4))
        for r in range(rounds):
            x += y * w
            y += z * x
            z += x * y
            w += y * z
            if r == 0 and rounds > 1 and xorshift:
                s = np.uint32(shift)
                x ^= x >> s
                y ^= y >> s
                z ^= z >> s
                w ^= w >> s
    return np.stack([x, y, z, w], 1)


def _avalanche(f, v):
    """P[i, j]: fraction of inputs for which flipping input bit i flips output bit j (bit 32 c + b = bit b of
    component c)."""
    base = f(v)
    n = v.shape[0]
    out = np.zeros((128, 128))
    sh = np.arange(32, dtype=np.uint32)
    for i in range(128):
        w = v.copy()
        w[:, i // 32] ^= np.uint32(1 << (i % 32))
        d = f(w) ^ base
        out[i] = ((d[:, :, None] >> sh) & 1).reshape(n, 128).mean(0)
    return out


def _avalanche_checks(P_):
    """The avalanche structure of the published pcg4d: additions and products only carry upward and the one
    16-bit xorshift carries 16 bits down, so input bit b can only reach output bits >= b - 16 (never flipped
    outside that cone); inside it every output bit flips often (p > 0.25); and the low 16 input bits of each word
    reach every output bit with p within 0.07 of 1/2 (output bits 0 and 1 aside, which sit on the cone's edge)."""
    ib = np.arange(128)[:, None] % 32
    ob = np.arange(128)[None, :] % 32
    cone = ob >= np.maximum(0, ib - 16)
    low = (ib < 16) & (ob >= 2)
    return (bool(np.all(P_[~cone] == 0.0)), bool(np.all(P_[cone] > 0.25)),
            bool(np.all(np.abs(P_[low & cone] - 0.5) < 0.07)))


def test_pcg4d_avalanche_with_negative_controls():
    rng = np.random.default_rng(1)
    v = rng.integers(0, 1 << 32, (4096, 4), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(_pcg4d_variant(v), O.pcg4d(v))  # the variant function unmutated is the oracle
    assert _avalanche_checks(_avalanche(O.pcg4d, v)) == (True, True, True)
    # each mutation fails at least one check
    for kw in ({"xorshift": False}, {"rounds": 1}, {"shift": 15}, {"shift": 17}):
        res = _avalanche_checks(_avalanche(lambda a: _pcg4d_variant(a, **kw), v))
        assert not all(res), kw
