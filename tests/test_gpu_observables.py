"""GPU: observables (colloid MSD, orientation autocorrelation, temperatures),
the SRD thermostat's per-cell closed forms, and the FP32 error report."""
import numpy as np
import pytest

import brownian_amd as B
from brownian_amd.error_report import report

pytestmark = pytest.mark.gpu


def test_msd_and_orientation_closed_forms():
    """One free colloid (no force, no collisions called): x(t) = x0 + v t and
    ra rotates about z by w dt per update (cpp/update.cpp:34-54)."""
    prm = dict(L=12, density=3, n_colloid=1, sig_colloid=2.5, sigma=2.0, mass_colloid=50.0, n_sub=10,
               precision=B.BW_FP64)
    s = B.System(**prm)
    s.startup()
    z = np.zeros((1, 3))
    v, w = 0.5, 2.0
    s.upload_colloid(pos=np.array([[6.0, 6.0, 6.0]]), vel=np.array([[v, 0.0, 0.0]]),
                     omega=np.array([[0.0, 0.0, w]]), f=z, old_f=z, ra=np.array([[1.0, 0.0, 0.0]]))
    s.observables_reset()
    dt = 0.1 / prm["n_sub"]
    k = 7
    for _ in range(k):
        s.update_activity_direction()
        s.update_pos_md()
    o = s.observables()
    assert o["lag_steps"] == 0
    assert o["msd"] == pytest.approx((k * v * dt) ** 2, rel=1e-12)
    assert o["orient_corr"] == pytest.approx(np.cos(k * w * dt), rel=1e-12)
    assert o["t_colloid"] == pytest.approx(50.0 * v * v / 3, rel=1e-14)
    I = 0.1 * 50.0 * 2.0 ** 2
    assert o["t_rot"] == pytest.approx(I * w * w / 3, rel=1e-14)
    c = s.download_colloid()
    np.testing.assert_allclose(c["pos"][0], [6.0 + k * v * dt, 6.0, 6.0], rtol=0, atol=1e-13)


@pytest.mark.parametrize("precision", [B.BW_FP64, B.BW_FP32])
def test_thermostat_per_cell_closed_form(precision):
    """After rotation_mpcd every shifted cell with N > 1 keeps its momentum and
    has relative kinetic energy 3 (N-1) kT / 2 exactly (cpp/rotation_mpcd.cpp:37-49)."""
    prm = dict(L=10, density=5, n_colloid=2, sig_colloid=2.5, sigma=2.0, mass_colloid=50.0, n_sub=10,
               theta_deg=90.0, precision=precision, vel_law=B.BW_VEL_MAXWELL)
    s = B.System(**prm)
    s.startup()
    s.step(2)
    s.status()
    _, v0, _ = s.download_fluid()
    s.rotation_mpcd()
    assert s.status() == 0
    _, v1, _ = s.download_fluid()
    cells = s.shifted_cells()
    ncell = prm["L"] ** 3
    cnt = np.bincount(cells, minlength=ncell)
    tol = 1e-12 if precision == B.BW_FP64 else 2e-6
    for comp in range(3):
        p0 = np.bincount(cells, weights=v0[:, comp], minlength=ncell)
        p1 = np.bincount(cells, weights=v1[:, comp], minlength=ncell)
        assert (np.abs(p1 - p0) / np.maximum(cnt, 1)).max() < tol
    mean = np.stack([np.bincount(cells, weights=v1[:, k], minlength=ncell) for k in range(3)], 1)
    mean /= np.maximum(cnt, 1)[:, None]
    rel = ((v1 - mean[cells]) ** 2).sum(1)
    e = np.bincount(cells, weights=rel, minlength=ncell)
    m = cnt > 1
    assert np.abs(e[m] / (3.0 * (cnt[m] - 1)) - 1).max() < 10 * tol


@pytest.mark.parametrize("spread", [1e-4, 1.0])
def test_fp32_thermostat_closed_form_cold_and_drifting_cells(spread):
    """FP32 storage with a common drift of (3, -2, 1) and a thermal spread of
    1e-4 (every cell's relative energy far below what the exact cell sums
    resolve, so each takes the exact per-particle pass) or 1 (most cells take
    the one-pass form): each shifted cell with N > 1 keeps its momentum and
    ends with relative kinetic energy 3 (N-1) kT / 2 (cpp/rotation_mpcd.cpp:37-49)."""
    prm = dict(L=10, density=5, n_colloid=0, sig_colloid=2.5, sigma=2.0, mass_colloid=50.0, n_sub=10,
               theta_deg=90.0, precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL)
    s = B.System(**prm)
    s.startup()
    pos, _, _ = s.download_fluid()
    rng = np.random.default_rng(7)
    vel = np.array([3.0, -2.0, 1.0]) + spread * rng.standard_normal(pos.shape)
    s.upload_fluid(pos, vel)
    _, v0, _ = s.download_fluid()
    s.rotation_mpcd()
    assert s.status() == 0
    _, v1, _ = s.download_fluid()
    cells = s.shifted_cells()
    ncell = prm["L"] ** 3
    cnt = np.bincount(cells, minlength=ncell)
    for comp in range(3):
        p0 = np.bincount(cells, weights=v0[:, comp], minlength=ncell)
        p1 = np.bincount(cells, weights=v1[:, comp], minlength=ncell)
        # one FP32 rounding of one particle per cell (the compensated residual)
        assert (np.abs(p1 - p0) / np.maximum(cnt, 1)).max() < 2e-6
    mean = np.stack([np.bincount(cells, weights=v1[:, k], minlength=ncell) for k in range(3)], 1)
    mean /= np.maximum(cnt, 1)[:, None]
    rel = ((v1 - mean[cells]) ** 2).sum(1)
    e = np.bincount(cells, weights=rel, minlength=ncell)
    m = cnt > 1
    # FP32 velocities of magnitude ~3 around a relative speed ~1: ~1e-6 relative
    assert np.abs(e[m] / (3.0 * (cnt[m] - 1)) - 1).max() < 2e-5


def test_fluid_temperature_observable_matches_state():
    prm = dict(L=12, density=10, n_colloid=3, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               theta_deg=90.0, v0=0.02, precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL,
               orient_law=B.BW_ORIENT_SPHERE)
    s = B.System(**prm)
    s.startup()
    s.observables_reset()
    s.step(5)
    o = s.observables()
    assert o["lag_steps"] == 5
    _, v, _ = s.download_fluid()
    N = len(v)
    t = (np.sum(v * v) - np.sum(np.sum(v, 0) ** 2) / N) / (3 * (N - 1))
    assert o["t_fluid"] == pytest.approx(t, rel=1e-10)
    assert 0.9 < o["t_fluid"] < 1.1       # the thermostat holds kT = 1
    assert -1.0 <= o["orient_corr"] <= 1.0 and o["msd"] > 0


def test_fp32_error_report_within_tolerance():
    """FP32 storage vs an FP64 run from the same state: per-step errors far
    below 1e-5 relative, apart from discrete collision decisions within a
    rounding distance of a colloid surface (counted by the report)."""
    r = report(L=16, ncol=6, steps=2)
    w = r["per_step_worst"]
    assert w["fluid_vel_rel_q999999"] < 1e-5
    assert w["fluid_pos_abs_q999999"] < 1e-5 * 16
    assert w["fluid_n_vel_rel_over_1e-5"] <= 4
    for k in ("colloid_pos_rel", "colloid_vel_rel", "colloid_omega_rel", "colloid_ra_rel"):
        assert w[k] < 1e-5, k


def test_autotuner_picks_a_candidate_with_identical_results():
    """The autotuner times every accepted tile shape; whichever it picks, the
    trajectory is bit-identical to the default shape's."""
    from brownian_amd.autotune import autotune, candidates
    prm = dict(L=16, density=10, n_colloid=4, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               theta_deg=90.0, v0=0.03, precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL,
               orient_law=B.BW_ORIENT_SPHERE)
    res = autotune(prm, steps=2, warmup=1)
    assert set(res["ms_per_step"]) == set(candidates(16, 10))
    assert res["best"] in res["ms_per_step"]
    a = B.System(**prm)
    b = B.System(tile=res["best"], **prm)
    a.startup(); b.startup()
    a.step(3); b.step(3)
    assert a.status() == 0 and b.status() == 0
    pa, va, _ = a.download_fluid()
    pb, vb, _ = b.download_fluid()
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)


def test_ensemble_autotuner_per_box_shapes_give_identical_results():
    """The ensemble tuner's per-box-edge refinement returns, for each edge, a
    shape that edge accepts; the ensemble stepped with those shapes (one slot
    capacity and SRD workgroup size for all, the largest tile's) gives the
    common shape's trajectory bit for bit."""
    from brownian_amd.autotune import autotune_ensemble, candidates
    reps = [dict(L=L, density=10, n_colloid=nc, n_sub=20, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0,
                 theta_deg=90.0, v0=0.03, seed=1100 + i, precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL,
                 orient_law=B.BW_ORIENT_SPHERE)
            for i, (L, nc) in enumerate([(16, 3), (24, 8), (16, 2)])]
    res = autotune_ensemble(reps, steps=2, warmup=1, per_box=True, max_cells=128)
    assert set(res["per_box"]) == {16, 24}
    for L, shape in res["per_box"].items():
        assert shape in candidates(L, 10)
    # a mixed assignment, whatever the timing chose
    mixed = {16: (4, 4, 4), 24: (8, 4, 4)}
    a = B.System(replicas=[dict(r, tile=res["best"]) for r in reps])
    b = B.System(replicas=[dict(r, tile=mixed[r["L"]]) for r in reps])
    a.startup(); b.startup()
    a.step(3); b.step(3)
    assert a.status() == 0 and b.status() == 0
    for r in range(len(reps)):
        pa, va, _ = a.download_fluid(r)
        pb, vb, _ = b.download_fluid(r)
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(va, vb)
        ca, cb = a.download_colloid(r), b.download_colloid(r)
        for k in ca:
            np.testing.assert_array_equal(ca[k], cb[k], err_msg=f"replica {r} {k}")


def test_small_ensemble_of_mixed_boxes_equals_single_runs():
    """Replicas of different edge, colloid count and v0 stepped together give
    each replica's single-system result bit for bit."""
    reps = [dict(L=L, density=10, n_colloid=nc, n_sub=20, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0,
                 theta_deg=90.0, v0=v0, seed=1000 + i, precision=B.BW_FP32, vel_law=B.BW_VEL_MAXWELL,
                 orient_law=B.BW_ORIENT_SPHERE)
            for i, (L, nc, v0) in enumerate([(16, 3, 0.05), (24, 10, 0.02), (32, 20, 0.08), (16, 1, 0.0)])]
    e = B.System(replicas=reps)
    e.startup()
    e.step(3)
    assert e.status() == 0
    for r, p in enumerate(reps):
        s = B.System(**p)
        s.startup()
        s.step(3)
        assert s.status() == 0
        pe, ve, _ = e.download_fluid(r)
        ps, vs, _ = s.download_fluid()
        np.testing.assert_array_equal(pe, ps)
        np.testing.assert_array_equal(ve, vs)
        ce, cs = e.download_colloid(r), s.download_colloid()
        for k in ce:
            np.testing.assert_array_equal(ce[k], cs[k], err_msg=f"replica {r} {k}")
