"""The benchmark configurations as parity cases, through the C-ABI.

* The CPU-runnable system (20^3 cells x 10 fluid per cell = 80,000 particles,
  8 colloids of radius 2, WCA sigma 4.2, mass 50, 20 substeps, SRD angle 90
  degrees, per-colloid active speed near 0.01, Maxwell velocities, orientations
  uniform on the sphere): FP64 on the GPU against the oracle over five
  consecutive MPCD steps with the oracle generator's draws injected (1e-10),
  the FP64 audit against the oracle's (cpp/active_brownian.cpp:39-55), and the
  FP32 storage mode per step from the same input state (1e-5).
* The batched ensemble (256 replicas, box edge in {16,...,48}, colloid count
  1..64 capped at 4 % volume fraction, v0 ~ U[0, 0.1], seeds 1000 + r): two
  steps with no status flag, replicas of edge 40 and 48 bit-identical to the
  same systems run alone, and with every v0 set to zero each replica's total
  momentum conserved to 1e-10 of its scale.
* The single-system configuration (64^3 x 10, 300 colloids, v0 = 0.02):
  blocked == per-substep streaming, and 100 steps with the thermostat and
  momentum bookkeeping.
"""
import sys
import os

import numpy as np
import pytest

import brownian_amd as B
import oracle as O
from tests.parity import compare_state, inject_step, oracle_step_draws

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

C0 = dict(L=20, density=10, n_colloid=8, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20, ndt=0.1,
          kbt=1.0, mass_fl=1.0, eps=1.0, theta_deg=90.0, v0=0.01, seed=77777)


def _oracle_from(s, prm, v0):
    o = O.Oracle(**prm)
    o.create_box()
    pos, vel, _ = s.download_fluid()
    o.set_fluid(pos, vel)
    o.set_colloid(**s.download_colloid())
    o.set_rng_state(s.rng_state())
    o.set_v0(v0)
    o.neighbour_list_mpcd()
    o.neighbour_list_md()
    o.compute_force_md()
    return o


def test_c0_fp64_five_steps_vs_oracle():
    v0 = np.random.default_rng(1).uniform(0.005, 0.015, C0["n_colloid"])
    s = B.System(precision=B.BW_FP64, draws=B.BW_DRAWS_INJECTED, vel_law=B.BW_VEL_MAXWELL,
                 orient_law=B.BW_ORIENT_SPHERE, **C0)
    s.startup()
    s.set_v0(v0)
    o = _oracle_from(s, C0, v0)
    # the start-up forces and potential agree
    c, oc = s.download_colloid(), o.get_colloid()
    np.testing.assert_allclose(c["f"], oc["f"], rtol=1e-12, atol=1e-14)
    events = 0
    for step in range(5):
        shift, axes, refl = oracle_step_draws(o, C0["n_sub"])
        events += sum(len(r[0]) for r in refl)
        inject_step(s, shift, axes, refl)
        s.step(1)
        assert s.status() == 0, step
        compare_state(s, o, C0["L"], 1e-10, 1e-10, 1e-10, label=f"C0 step {step}")
        a, oa = s.audit(), o.audit()
        np.testing.assert_allclose(a["ke_fluid"], oa[3], rtol=1e-12)
        np.testing.assert_allclose(a["ke_colloid"], oa[0], rtol=1e-10)
        np.testing.assert_allclose(a["ang_ke_colloid"], oa[1], rtol=1e-10)
        np.testing.assert_allclose(a["potential"], oa[2], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(a["momentum"], oa[4:7], atol=1e-10 * a["momentum_scale"])
    assert events > 50  # the reflections and their injected draws were exercised


def test_c0_fp32_per_step_vs_oracle():
    v0 = np.random.default_rng(2).uniform(0.005, 0.015, C0["n_colloid"])
    s = B.System(precision=B.BW_FP32, draws=B.BW_DRAWS_INJECTED, vel_law=B.BW_VEL_MAXWELL,
                 orient_law=B.BW_ORIENT_SPHERE, **C0)
    s.startup()
    s.set_v0(v0)
    for step in range(3):
        o = _oracle_from(s, C0, v0)
        shift, axes, refl = oracle_step_draws(o, C0["n_sub"])
        inject_step(s, shift, axes, refl)
        s.step(1)
        assert s.status() == 0, step
        compare_state(s, o, C0["L"], 1e-5, 1e-5, 1e-5, label=f"C0 fp32 step {step}")


def _ensemble_params():
    sys.path.insert(0, ROOT)
    from bench import ensemble_params
    common = dict(precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, vel_law=B.BW_VEL_MAXWELL,
                  orient_law=B.BW_ORIENT_SPHERE)
    return [dict(r, **common) for r in ensemble_params(256)]


@pytest.mark.timeout(400)
def test_c3_ensemble_256_replicas():
    reps = _ensemble_params()
    Ls = [r["L"] for r in reps]
    assert set(Ls) == {16, 24, 32, 40, 48}
    e = B.System(replicas=reps)
    e.startup()
    e.step(2)
    assert e.status() == 0
    # one replica of edge 40 and one of edge 48 (with the most colloids) alone
    for L in (40, 48):
        r = max((i for i in range(len(reps)) if reps[i]["L"] == L), key=lambda i: reps[i]["n_colloid"])
        s = B.System(**reps[r])
        s.startup()
        s.step(2)
        assert s.status() == 0
        for a, b in zip(e.download_fluid(r), s.download_fluid()):
            np.testing.assert_array_equal(a, b, err_msg=f"replica {r} (L={L})")
        ce, cs = e.download_colloid(r), s.download_colloid()
        for k in ce:
            np.testing.assert_array_equal(ce[k], cs[k], err_msg=f"replica {r} (L={L}) {k}")
        del s
    # v0 = 0 everywhere: each replica's momentum is conserved.  FP32 storage
    # rounds every stored velocity by up to 2^-24 of itself, so the audit's
    # momentum moves by a random walk of about 2^-24 sqrt(3 N k) over k
    # velocity updates of N particles; that bound (times 4) applies to every
    # replica, and 1e-10 of the momentum scale to the large ones (L >= 40).
    for r, p in enumerate(reps):
        e.set_v0(np.zeros(p["n_colloid"]), replica=r)
    before = [e.audit(r) for r in range(len(reps))]
    e.step(2)
    assert e.status() == 0
    worst_large = 0.0
    for r in range(len(reps)):
        a = e.audit(r)
        drift = np.abs(a["momentum"] - before[r]["momentum"]).max() / a["momentum_scale"]
        nf = e.info[r]["nf"]
        bound = 4 * 2.0 ** -24 * np.sqrt(3 * nf * 2) / a["momentum_scale"]
        assert drift <= max(1e-10, bound), (r, reps[r]["L"], drift, bound)
        if reps[r]["L"] >= 40:
            worst_large = max(worst_large, drift)
    assert worst_large < 1e-10, worst_large


def test_c1_single_system_with_activity():
    """The single-system configuration with its activity (64^3 x 10 =
    2,621,440 fluid particles, 300 colloids of radius 2, v0 = 0.02, FP32
    integer-cell storage, fluid placed on the GPU): one step with blocked
    streaming equals per-substep streaming (velocities and every colloid
    quantity bit for bit, positions to FP32 rounding); over 100 steps no
    status flag, the fluid held at kT by the thermostat (1 %), and the total
    momentum conserved -- run() removes each colloid's kick from the fluid in
    its rear shell (cpp/run.cpp:14-24) -- to the random walk of the FP32
    roundings."""
    from tests.parity import blocked_equals_per_substep
    prm = dict(L=64, density=10, n_colloid=300, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20, ndt=0.1,
               theta_deg=90.0, v0=0.02, seed=6401, precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE,
               vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE, fluid_init=B.BW_INIT_DEVICE)
    a = B.System(**prm)
    b = B.System(**dict(prm, streaming=B.BW_STREAM_PER_SUBSTEP))
    a.startup(); b.startup()
    dp = blocked_equals_per_substep(a, b, 64, pos_tol=2e-6)
    print(f"64^3 blocked vs per-substep: largest position difference {dp:.2e}")
    b.close()
    before = a.audit()
    a.step(100)
    assert a.status() == 0
    after = a.audit()
    nf = a.info[0]["nf"]
    obs_t = after["ke_fluid"] * 2.0 / (3.0 * nf)
    assert abs(obs_t - 1.0) < 1e-2, obs_t
    drift = np.abs(after["momentum"] - before["momentum"]).max() / after["momentum_scale"]
    bound = 4 * 2.0 ** -24 * np.sqrt(3 * nf * 100 * 3) / after["momentum_scale"]
    print(f"64^3, 100 steps: T_fluid {obs_t:.5f}, momentum drift {drift:.2e} (bound {bound:.2e})")
    assert drift <= bound, (drift, bound)


def test_single_system_64cubed_runs_are_bit_reproducible():
    """Two runs of the single-system configuration (64^3 x 10, 300 colloids,
    v0 = 0.02, FP32, device draws, fused graph-captured step) from the same
    seed agree bit for bit after 20 steps: every sum whose order the GPU's
    scheduling could change is taken in fixed point or re-sorted (events by
    stencil position and id, serial lists by id, tile colloid lists by index)."""
    prm = dict(L=64, density=10, n_colloid=300, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20, ndt=0.1,
               theta_deg=90.0, v0=0.02, seed=6402, precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE,
               vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE, fluid_init=B.BW_INIT_DEVICE)
    out = []
    for _ in range(2):
        s = B.System(**prm)
        s.startup()
        s.step(20)
        assert s.status() == 0
        pos, vel, cell = s.download_fluid()
        out.append((pos, vel, cell, s.download_colloid()))
        s.close()
    (pa, va, ca, da), (pb, vb, cb, db) = out
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)
