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
