"""Fluid start-up on the GPU (BW_INIT_DEVICE): the laws of
cpp/initialize.cpp:62-83 with counter-based draws, checked by what the laws
fix -- no particle closer than sigma/2 to a colloid centre, uniform density
elsewhere, zero total momentum, per-component velocity variance kT/m with the
kurtosis of the law (9/5 for the reference's uniform law, 3 for
Maxwell-Boltzmann) -- and by the layout: a context stepped from the device
start-up equals one that uploaded the same state from the host, bit for bit."""
import time

import numpy as np
import pytest

import brownian_amd as B

pytestmark = pytest.mark.gpu

BASE = dict(L=32, density=10, n_colloid=40, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
            theta_deg=90.0, v0=0.02, seed=606, draws=B.BW_DRAWS_DEVICE, orient_law=B.BW_ORIENT_SPHERE,
            fluid_init=B.BW_INIT_DEVICE)


def _pdiff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


@pytest.mark.parametrize("vel_law,kurt", [(B.BW_VEL_REFERENCE, 1.8), (B.BW_VEL_MAXWELL, 3.0)])
@pytest.mark.parametrize("precision", [B.BW_FP32, B.BW_FP64])
def test_device_start_up_laws(vel_law, kurt, precision):
    s = B.System(precision=precision, vel_law=vel_law, **BASE)
    s.startup()
    assert s.status() == 0
    pos, vel, _ = s.download_fluid()
    c = s.download_colloid()
    L, rad = BASE["L"], 0.5 * BASE["sigma"]
    n = len(pos)
    assert n == L ** 3 * BASE["density"]
    dmin = np.inf
    for j in range(len(c["pos"])):
        d = np.sqrt((_pdiff(pos, c["pos"][j], L) ** 2).sum(1))
        dmin = min(dmin, d.min())
    assert dmin >= rad * (1 - 1e-7), dmin
    assert np.all((pos >= 0) & (pos < L))
    # density: counts in cells of 4^3 unit cells are Poisson-like around the mean
    # of the free volume (blocks touching a colloid's exclusion excluded)
    blk = (pos // 4).astype(int)
    counts = np.bincount(blk[:, 0] + 8 * (blk[:, 1] + 8 * blk[:, 2]), minlength=512)
    centres = (np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)[:, [2, 1, 0]] * 4 + 2.0)
    near = np.zeros(512, bool)
    for j in range(len(c["pos"])):
        near |= np.sqrt((_pdiff(centres, c["pos"][j], L) ** 2).sum(1)) < rad + 3.5
    free = counts[~near]
    mu = free.mean()
    assert abs(mu - n / (L ** 3 - len(c["pos"]) * 4 / 3 * np.pi * rad ** 3) * 64) < 4 * np.sqrt(mu / len(free)) + 1
    assert 0.7 * mu < free.var() < 1.3 * mu
    # velocities
    scale = np.abs(vel).sum()
    assert np.abs(vel.sum(0)).max() < (1e-12 if precision == B.BW_FP64 else 1e-6) * scale
    var = vel.var(0)
    assert np.all(np.abs(var - 1.0) < 5 * np.sqrt(2.0 / n) * 1.5), var
    k = ((vel - vel.mean(0)) ** 4).mean(0) / var ** 2
    assert np.all(np.abs(k - kurt) < 0.05), k


def test_device_start_up_equals_upload_and_is_reproducible():
    a = B.System(precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL, **BASE)
    a.startup()
    pos, vel, _ = a.download_fluid()
    c = a.download_colloid()
    b = B.System(precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL, **dict(BASE, fluid_init=B.BW_INIT_HOST))
    b.startup()
    b.upload_fluid(pos, vel)
    b.upload_colloid(**c)
    b.neighbour_list_md(); b.compute_force_md()
    a2 = B.System(precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL, **BASE)
    a2.startup()
    for x, y in zip(a.download_fluid(), a2.download_fluid()):
        np.testing.assert_array_equal(x, y)
    a.step(3); b.step(3)
    assert a.status() == 0 and b.status() == 0
    for x, y in zip(a.download_fluid(), b.download_fluid()):
        np.testing.assert_array_equal(x, y)
    ca, cb = a.download_colloid(), b.download_colloid()
    for k in ca:
        np.testing.assert_array_equal(ca[k], cb[k], err_msg=k)
    d = B.System(precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL, **dict(BASE, seed=607))
    d.startup()
    assert not np.array_equal(d.download_fluid()[0], a2.download_fluid()[0])


def test_device_start_up_at_full_size_is_fast():
    """128^3 x 10 fluid with 2400 colloids: the start-up sequence with the
    fluid placed on the GPU."""
    prm = dict(BASE, L=128, n_colloid=2400, precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL, tile=(8, 4, 4))
    t0 = time.time()
    s = B.System(**prm)
    s.startup()
    s.sync()
    dt = time.time() - t0
    assert s.status() == 0
    print(f"128^3 start-up with the device fluid placement: {dt:.2f} s")
    s.step(1)
    assert s.status() == 0
    assert dt < 5.0
