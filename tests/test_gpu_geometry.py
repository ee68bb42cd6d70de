"""GPU: the periodic-geometry helpers of the kernels against the formulas
they restate, bit for bit, on the inputs where a shortcut could go wrong.

img(a, L) = a - L*round(a/L) and mod(a, L) = a - L*round((a - L/2)/L) with
round half away from zero (cpp/parameters.hpp:74-76): the device forms skip
the division where the rounded quotient is certain, so the test covers the
thresholds of those shortcuts (|a| near 0.4999 L, L/2, 0.5001 L, 1.4999 L,
1.5 L), signed zeros and random values; a mod n and a / b (the tile index)
are checked over their whole ranges."""
import numpy as np
import pytest

import brownian_amd as B

pytestmark = pytest.mark.gpu


def round_away(x):
    """C round(): half away from zero, exact for |x| < 2^52, zero signed as x."""
    t = np.trunc(x)
    f = x - t
    r = t + np.where(np.abs(f) >= 0.5, np.sign(x), 0.0)
    return np.where(r == 0, np.copysign(0.0, x), r)


def img_ref(a, L):
    return a - L * round_away(a / L)


def mod_ref(a, L):
    return a - L * round_away((a - L / 2) / L)


def adversarial(L, rng):
    base = []
    for f in (0.0, 0.4999, 0.5, 0.5001, 1.0, 1.4999, 1.5, 2.0, 2.5):
        x = f * L
        base += [x, np.nextafter(x, np.inf), np.nextafter(x, -np.inf)]
    vals = np.array(base + [1e-300, 5e-324, 0.25, 0.75 * L, L - 1e-12])
    vals = np.concatenate([vals, -vals, [0.0, -0.0]])
    rnd = rng.uniform(-3 * L, 3 * L, 20000)
    near = L * (rng.integers(-5, 6, 5000) * 0.5) + rng.normal(0, 1e-9 * L, 5000)
    return np.concatenate([vals, rnd, near])


@pytest.mark.parametrize("L", [30.0, 128.0, 24.0, 16.0])
def test_img_and_mod_match_the_reference_formulas_bit_for_bit(L):
    rng = np.random.default_rng(int(L))
    s = B.System(L=16, density=2, n_colloid=0, n_sub=2, precision=B.BW_FP32)
    a = adversarial(L, rng)
    ia = np.zeros(len(a), np.int32)
    img, mod, _, _ = s.debug_geometry(a, L, ia, 16, 4)
    np.testing.assert_array_equal(img.view(np.uint64), img_ref(a, L).view(np.uint64))
    np.testing.assert_array_equal(mod.view(np.uint64), mod_ref(a, L).view(np.uint64))


def test_wrap_and_tile_division_over_their_ranges():
    s = B.System(L=16, density=2, n_colloid=0, n_sub=2, precision=B.BW_FP32)
    ia = np.arange(1024, dtype=np.int32)
    for Li in (1, 7, 16, 30, 128, 1023):
        w = np.arange(-5 * Li, 6 * Li, dtype=np.int32)
        _, _, wrap, _ = s.debug_geometry(np.zeros(len(w)), 1.0, w, Li, 1)
        np.testing.assert_array_equal(wrap, np.mod(w, Li))
    for b in range(1, 64):
        _, _, _, div = s.debug_geometry(np.zeros(len(ia)), 1.0, ia, 1024, b)
        np.testing.assert_array_equal(div, ia // b)
