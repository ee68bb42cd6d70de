"""The benchmarked workload at its full size, checked through properties that
do not depend on size: 128^3 cells x 10 fluid particles per cell (20,971,520),
2400 colloids (WCA sigma 4.2, mass 50), 20 substeps per MPCD step, FP32
storage with FP64 cell sums, tile (8,4,4), device draws.

- Blocked streaming (far particles streamed through the whole step by the
  rotation pass) equals per-substep streaming of every particle, bit for bit,
  with the benchmark's activities v0 ~ U[0, 0.05].
- Without activity the total momentum (fluid + colloids, FP64 sums) is
  conserved to 1e-10 of its scale, the thermostat holds the fluid at kT
  (1 %), and no status flag (overflow, drift, margin) is raised."""
import numpy as np
import pytest

import brownian_amd as B

pytestmark = pytest.mark.gpu

PRM = dict(L=128, density=10, n_colloid=2400, n_sub=20, ndt=0.1, kbt=1.0, mass_fl=1.0, mass_colloid=50.0,
           sig_colloid=4.2, eps=1.0, sigma=4.0, theta_deg=90.0, v0=0.0, seed=20240,
           precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, vel_law=B.BW_VEL_MAXWELL,
           orient_law=B.BW_ORIENT_SPHERE, tile=(8, 4, 4))


def started(v0=None, **over):
    s = B.System(**dict(PRM, **over))
    s.startup()
    if v0 is not None:
        s.set_v0(v0)
    return s


def test_full_size_blocked_streaming_equals_per_substep():
    v0 = np.random.default_rng(PRM["seed"]).uniform(0.0, 0.05, PRM["n_colloid"])
    a = started(v0, streaming=B.BW_STREAM_BLOCKED)
    a.step(2)
    assert a.status() == 0
    pa, va, ca = a.download_fluid()
    da = a.download_colloid()
    del a
    b = started(v0, streaming=B.BW_STREAM_PER_SUBSTEP)
    b.step(2)
    assert b.status() == 0
    pb, vb, cb = b.download_fluid()
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)
    db = b.download_colloid()
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)


def test_full_size_momentum_temperature_and_flags():
    s = started()
    assert s.info[0]["nf"] == 128 ** 3 * 10
    a0 = s.audit()
    for _ in range(3):
        s.step(1)
        assert s.status() == 0
        a = s.audit()
        drift = np.abs(a["momentum"] - a0["momentum"]).max() / a["momentum_scale"]
        assert drift < 1e-10, drift
    T = a["ke_fluid"] / (1.5 * s.info[0]["nf"])
    assert abs(T - 1.0) < 0.01, T
