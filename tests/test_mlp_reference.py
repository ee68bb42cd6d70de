"""CPU: the surrogate network's reference (oracle/mlp_ref.py) pinned by
central finite differences of its own loss, and BF16 rounding by its
definition (nearest even on the float32 bit pattern)."""
import numpy as np

from oracle import mlp_ref as M


def _setup(rng, n=7):
    # weights ~ 1/sqrt(fan-in): outputs O(1), the log-variance clamp (+-20) inactive
    layers = [(rng.normal(0, 1.0 / np.sqrt(k), (k, m)), rng.normal(0, 0.1, m)) for k, m in M.SIZES]
    p = M.pack(layers)
    x = rng.normal(0, 1, (n, 144))
    t = rng.normal(0, 1, (n, 6))
    xm, xs = rng.normal(0, 0.1, 144), rng.uniform(0.5, 2.0, 144)
    tm, ts = rng.normal(0, 0.1, 6), rng.uniform(0.5, 2.0, 6)
    return p, x, t, xm, xs, tm, ts


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p, x, t, xm, xs, tm, ts = _setup(rng)
    _, g = M.loss_grad(p, x, t, xm, xs, tm, ts, emulate_bf16=False)
    idx = rng.choice(len(p), 60, replace=False)
    # one coordinate in each layer's weights and biases too
    o = 0
    for k, n in M.SIZES:
        idx = np.append(idx, [o, o + k * n])
        o += k * n + n
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, x, t, xm, xs, tm, ts, False)[0] - M.loss_grad(pm, x, t, xm, xs, tm, ts, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    r = M.bf16(a)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    np.testing.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0])
    x = np.random.default_rng(1).normal(0, 1, 1000).astype(np.float32)
    assert np.all(np.abs(M.bf16(x) - x) <= np.abs(x) * 2 ** -8 + 1e-45)


def test_forward_hand_computed_network():
    """A sparse network whose outputs follow by hand: the layer wiring (W is
    K x N, row = input), the biases, SiLU z / (1 + e^-z) on the three hidden
    layers and none on the output, and the input standardisation."""
    import math

    def silu(z):
        return z / (1.0 + math.exp(-z))

    layers = [(np.zeros((k, n)), np.zeros(n)) for k, n in M.SIZES]
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = layers
    W1[0, 0], W1[1, 1], b1[2] = 1.0, -0.5, 0.25
    W2[0, 0], W2[1, 0], W2[2, 1] = 2.0, 1.0, 1.0
    W3[0, 0], W3[1, 1], b3[1] = 1.0, 1.0, -3.0
    W4[0, 0], W4[1, 6], b4[3], b4[0] = 1.0, 2.0, 0.5, -4.0
    p = M.pack(layers)
    raw = np.zeros((1, 144))
    raw[0, 0], raw[0, 1] = 4.0, 5.0
    x_mean, x_std = np.ones(144), np.full(144, 2.0)  # standardised: x0 = 1.5, x1 = 2
    y, _ = M.forward(p, raw, x_mean, x_std, emulate_bf16=False)
    h1 = [silu(1.5), silu(-1.0), silu(0.25)]
    h2 = [silu(2 * h1[0] + h1[1]), silu(h1[2])]
    h3 = [silu(h2[0]), silu(h2[1] - 3.0)]
    expect = np.zeros(12)
    expect[0] = h3[0] - 4.0   # negative: no activation on the output layer
    expect[6] = 2.0 * h3[1]
    expect[3] = 0.5
    np.testing.assert_allclose(y[0], expect, rtol=1e-12, atol=1e-15)
    assert y[0, 0] < -2.0
