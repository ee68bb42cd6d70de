"""GPU parity: libbrownian (HIP, gfx950) through its C-ABI against the
compiled reference's golden vectors and the CPU oracle, with the reference
generator's draws injected into the GPU run.

Tolerances: FP64 runs agree with the FP64 reference to 1e-10 (only libm vs
device-libm ulps and the integer-cell position representation differ); FP32
runs (integer cell + FP32 offset/velocity) agree per step from the same input
state to 1e-5.  Maps (cell lists, neigh_fl, pair lists, shifted cells) are
compared bit-exactly."""
import ast
import os

import numpy as np
import pytest

import brownian_amd as B
import oracle as O
from tests.parity import compare_state, inject_step, oracle_step_draws, pdiff

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    d = dict(np.load(os.path.join(GOLD, name + ".npz")))
    params = dict(ast.literal_eval(str(d.pop("params_json"))))
    return d, params


def sys_params(p, **over):
    q = {k: v for k, v in p.items() if k in ("L", "density", "n_colloid", "n_sub", "ndt", "kbt", "mass_fl",
                                            "mass_colloid", "sig_colloid", "eps", "sigma", "theta_deg", "v0",
                                            "seed")}
    q.update(precision=B.BW_FP64, draws=B.BW_DRAWS_INJECTED)
    q.update(over)
    return q


def lists_from_cells(cells, ncell):
    order = np.lexsort((np.arange(len(cells)), cells))
    return np.bincount(cells, minlength=ncell), order.astype(np.int32)


def test_startup_and_maps_bit_exact():
    g, p = load("small")
    s = B.System(**sys_params(p))
    s.startup()
    assert s.status() == 0
    pos, vel, cell = s.download_fluid()
    np.testing.assert_array_equal(pos, g["start_pos"])
    np.testing.assert_array_equal(vel, g["start_vel"])
    c = s.download_colloid()
    for k in ("pos", "vel", "omega", "ra"):
        np.testing.assert_array_equal(c[k], g["start_c_" + k])
    np.testing.assert_allclose(c["f"], g["force0"], rtol=1e-13, atol=1e-15)
    counts, lst = s.cell_lists()
    np.testing.assert_array_equal(counts, g["nl_fluid_no"])
    np.testing.assert_array_equal(lst, g["nl_box_part"])
    nn, nl = s.neigh_fl()
    np.testing.assert_array_equal(nn, g["nl_no_neigh"])
    np.testing.assert_array_equal(nl, g["nl_neigh_fl"])
    mc, ml = s.md_list()
    np.testing.assert_array_equal(mc, g["md_count"])
    np.testing.assert_array_equal(ml, g["md_list"])


def test_function_trace_vs_reference():
    """Every entry point of one MPCD step against the reference's own outputs."""
    g, p = load("small")
    s = B.System(**sys_params(p))
    s.startup()
    o = O.Oracle(**p)
    o.startup()
    shift, axes, refl = oracle_step_draws(o, p["n_sub"])
    inject_step(s, shift, axes, refl)
    L = p["L"]

    def check(prefix, tol=1e-11):
        pos, vel, _ = s.download_fluid()
        assert np.abs(pdiff(pos, g[prefix + "pos"], L)).max() <= tol, prefix
        assert np.abs(vel - g[prefix + "vel"]).max() <= tol, prefix
        c = s.download_colloid()
        for k in c:
            ref = g[prefix + "c_" + k]
            d = pdiff(c[k], ref, L) if k == "pos" else c[k] - ref
            assert np.abs(d).max() <= tol, prefix + k

    s.rotation_mpcd()
    assert s.status() == 0
    check("rot_")
    cnt, lst = lists_from_cells(s.shifted_cells(), L ** 3)
    np.testing.assert_array_equal(cnt, g["rot_fluid_no"])
    np.testing.assert_array_equal(lst, g["rot_cell_part"])
    s.run()
    assert s.status() == 0
    check("run_")
    np.testing.assert_array_equal(s.run_counts(), g["run_cnt"])
    seq = ["copy_old_force", "update_pos_md", "neighbour_list_md", "update_pos_mpcd", "neighbour_list_mpcd",
           "fluid_colloid_collision", "update_activity_direction", "compute_force_md", "update_velocity_colloid"]
    for l in range(p["n_sub"]):
        for fn in seq:
            getattr(s, fn)()
            if l == 0:
                assert s.status() == 0, fn
                check(f"s1_{fn}_")
                if fn == "neighbour_list_mpcd":
                    counts, lst = s.cell_lists()
                    np.testing.assert_array_equal(counts, g["s1_nl_fluid_no"])
                    np.testing.assert_array_equal(lst, g["s1_nl_box_part"])
                    nn, nl = s.neigh_fl()
                    np.testing.assert_array_equal(nn, g["s1_nl_no_neigh"])
                    np.testing.assert_array_equal(nl, g["s1_nl_neigh_fl"])
                if fn == "neighbour_list_md":
                    mc, ml = s.md_list()
                    np.testing.assert_array_equal(mc, g["s1_md_count"])
                    np.testing.assert_array_equal(ml, g["s1_md_list"])
    assert s.status() == 0
    check("step1_")
    # two more steps through the fused, graph-replayed path
    for _ in range(2):
        shift, axes, refl = oracle_step_draws(o, p["n_sub"])
        inject_step(s, shift, axes, refl)
        s.step(1)
        assert s.status() == 0
    check("step3_", tol=1e-10)


def test_c0like_fused_vs_reference():
    g, p = load("c0like")
    s = B.System(**sys_params(p))
    s.startup()
    o = O.Oracle(**p)
    o.startup()
    for _ in range(2):
        shift, axes, refl = oracle_step_draws(o, p["n_sub"])
        inject_step(s, shift, axes, refl)
        s.step(1)
        assert s.status() == 0
    compare_state(s, o, p["L"], 1e-10, 1e-10, 1e-10, label="c0like")
    pos, vel, _ = s.download_fluid()
    assert np.abs(pdiff(pos, g["step2_pos"], p["L"])).max() < 1e-10
    assert np.abs(vel - g["step2_vel"]).max() < 1e-10


def test_overlapping_colloids_chain_particles():
    """Fluid inside two colloids: the serial per-colloid replay."""
    g, p = load("overlap")
    # the colloids here spin up far beyond any physical speed within the step,
    # past the displacement bound of blocked streaming: stream every substep
    s = B.System(**sys_params(p, streaming=B.BW_STREAM_PER_SUBSTEP))
    s.startup()
    o = O.Oracle(**p)
    o.startup()
    c = o.get_colloid()
    c["pos"][0] = [5.0, 5.0, 5.0]
    c["pos"][1] = [8.1, 5.4, 4.8]
    o.set_colloid(**c)
    o.neighbour_list_mpcd(); o.neighbour_list_md(); o.compute_force_md()
    s.upload_colloid(pos=c["pos"])
    s.neighbour_list_mpcd(); s.neighbour_list_md(); s.compute_force_md()
    shift, axes, refl = oracle_step_draws(o, p["n_sub"])
    inject_step(s, shift, axes, refl)
    s.step(1)
    assert s.status() == 0  # every GPU event found its injected draw; counts equal
    # This start is violent: ~300 reflections per substep (the event-bucket
    # overflow and serial-replay paths) spin the colloids up to |w| ~ 2e4 and
    # the fluid to |v| ~ 4e4 within the step, so FP64 rounding grows in
    # absolute terms; compare relative to the state's magnitude (1e-9).
    pos, vel, _ = s.download_fluid()
    opos, ovel = o.get_fluid()
    vs = np.abs(ovel).max()
    assert np.abs(vel - ovel).max() <= 1e-9 * vs
    assert np.abs(pdiff(pos, opos, p["L"])).max() <= 1e-9 * p["L"]
    assert np.abs(vel - g["step1_vel"]).max() <= 1e-9 * vs
    c, oc = s.download_colloid(), o.get_colloid()
    for k in ("vel", "omega"):
        assert np.abs(c[k] - oc[k]).max() <= 1e-9 * np.abs(oc[k]).max(), k


@pytest.mark.parametrize("precision", [B.BW_FP32, B.BW_FP64])
def test_fused_step_equals_entry_point_sequence(precision):
    """bw_step == rotation_mpcd, run, n_sub x substep.  FP64: bit-exact over
    two steps.  FP32: the fused step flies far particles through the step in
    one rounding, the entry points stream every substep, so after one step
    velocities and colloids are bit-exact and positions agree to 2e-6 cells."""
    prm = dict(L=12, density=10, n_colloid=4, sig_colloid=2.5, sigma=2.0, mass_colloid=50.0, n_sub=5,
               v0=0.05, theta_deg=90.0, seed=4321, precision=precision, draws=B.BW_DRAWS_DEVICE)
    a = B.System(**prm)
    b = B.System(**prm)
    a.startup(); b.startup()
    nsteps = 2 if precision == B.BW_FP64 else 1
    for _ in range(nsteps):
        a.step_sequence(prm["n_sub"])
    b.step(nsteps)
    assert a.status() == 0 and b.status() == 0
    pa, va, ca = a.download_fluid()
    pb, vb, cb = b.download_fluid()
    if precision == B.BW_FP64:
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(ca, cb)
    else:
        assert np.abs(pdiff(pa, pb, prm["L"])).max() <= 2e-6
    np.testing.assert_array_equal(va, vb)
    da, db = a.download_colloid(), b.download_colloid()
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)


def test_blocked_streaming_fp32_one_step_at_a_time():
    """FP32 storage: particles out of every colloid's reach flown through the
    step by the rotation pass in one rounding == every particle streamed every
    substep: velocities and colloids bit for bit, positions within 2e-6 cells
    (n_sub FP32 roundings), one step from each of several start-ups."""
    from tests.parity import blocked_equals_per_substep
    prm = dict(L=16, density=10, n_colloid=8, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               v0=0.05, theta_deg=90.0, precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE,
               vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE)
    worst = 0.0
    for seed, nc in [(2468, 8), (2469, 8), (2470, 12), (2471, 0)]:
        a = B.System(streaming=B.BW_STREAM_BLOCKED, seed=seed, **dict(prm, n_colloid=nc))
        b = B.System(streaming=B.BW_STREAM_PER_SUBSTEP, seed=seed, **dict(prm, n_colloid=nc))
        a.startup(); b.startup()
        worst = max(worst, blocked_equals_per_substep(a, b, prm["L"], 2e-6))
    assert worst > 0  # the two flights do round differently


@pytest.mark.parametrize("precision", [B.BW_FP64])
def test_blocked_streaming_equals_per_substep(precision):
    """FP64 storage: particles out of every colloid's reach streamed by the
    rotation pass through all substeps == every particle streamed every
    substep (bit-exact)."""
    prm = dict(L=16, density=10, n_colloid=8, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               v0=0.05, theta_deg=90.0, seed=2468, precision=precision, draws=B.BW_DRAWS_DEVICE,
               vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE)
    a = B.System(streaming=B.BW_STREAM_BLOCKED, **prm)
    b = B.System(streaming=B.BW_STREAM_PER_SUBSTEP, **prm)
    a.startup(); b.startup()
    a.step(3); b.step(3)
    assert a.status() == 0 and b.status() == 0
    pa, va, ca = a.download_fluid()
    pb, vb, cb = b.download_fluid()
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(ca, cb)
    da, db = a.download_colloid(), b.download_colloid()
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)
    # no colloids: the rotation pass does all of the streaming
    prm0 = dict(prm, n_colloid=0)
    a0 = B.System(streaming=B.BW_STREAM_BLOCKED, **prm0)
    b0 = B.System(streaming=B.BW_STREAM_PER_SUBSTEP, **prm0)
    a0.startup(); b0.startup()
    a0.step(2); b0.step(2)
    assert a0.status() == 0 and b0.status() == 0
    pa, va, _ = a0.download_fluid()
    pb, vb, _ = b0.download_fluid()
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)


def test_blocked_streaming_margin_violation_is_reported():
    """A colloid driven (from the host) past the displacement bound within a
    step raises BW_ERR_PHYSICS with the margin flag in blocked mode only."""
    prm = dict(L=12, density=5, n_colloid=2, sig_colloid=2.5, sigma=2.0, mass_colloid=50.0, n_sub=10,
               v0=0.0, theta_deg=90.0, seed=1357, precision=B.BW_FP64, draws=B.BW_DRAWS_DEVICE)
    flags = {}
    for mode in (B.BW_STREAM_BLOCKED, B.BW_STREAM_PER_SUBSTEP):
        s = B.System(streaming=mode, **prm)
        s.startup()
        c = s.download_colloid()
        c["vel"][0] = [6.0, 0.0, 0.0]  # 0.06 cells per substep: 0.25 passed within the step
        s.upload_colloid(vel=c["vel"])
        s.step(1)
        flags[mode] = s.status(raise_on_error=False)
    assert flags[B.BW_STREAM_BLOCKED] & 0x200
    assert not flags[B.BW_STREAM_PER_SUBSTEP] & 0x200


def test_fp32_per_step_vs_oracle():
    """FP32 storage: each step from the same input state within 1e-5."""
    prm = dict(L=12, density=10, n_colloid=3, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               v0=0.01, theta_deg=90.0, seed=97531)
    s = B.System(precision=B.BW_FP32, draws=B.BW_DRAWS_INJECTED, **prm)
    s.startup()
    o = O.Oracle(**prm)
    o.create_box()
    o.set_rng_state(s.rng_state())
    for step in range(3):
        pos, vel, _ = s.download_fluid()
        o.set_fluid(pos, vel)
        o.set_colloid(**s.download_colloid())
        o.neighbour_list_mpcd()
        o.neighbour_list_md()
        shift, axes, refl = oracle_step_draws(o, prm["n_sub"])
        inject_step(s, shift, axes, refl)
        s.step(1)
        assert s.status() == 0, step
        compare_state(s, o, prm["L"], 1e-5, 1e-5, 1e-5, label=f"fp32 step {step}")


def test_shifted_binning_matches_oracle_binning():
    """Bit-exact shifted-cell map on the GPU's own positions (FP32)."""
    prm = dict(L=16, density=10, n_colloid=2, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=4,
               theta_deg=90.0, seed=2468)
    s = B.System(precision=B.BW_FP32, draws=B.BW_DRAWS_INJECTED, **prm)
    s.startup()
    o = O.Oracle(**prm)
    pos, vel, _ = s.download_fluid()
    o.set_fluid(pos, vel)
    o.set_colloid(**s.download_colloid())
    o.set_rng_state(s.rng_state())
    o.clear_drawlog()
    o.rotation_mpcd()
    k, keys, vals = o.drawlog()
    shift = vals[k == O.DRAW_SHIFT][0][:3]
    axes = np.full((prm["L"] ** 3, 2), np.nan)
    axes[keys[k == O.DRAW_AXIS, 0]] = vals[k == O.DRAW_AXIS, :2]
    s.inject_rotation(shift, axes)
    s.rotation_mpcd()
    assert s.status() == 0
    cnt, lst = lists_from_cells(s.shifted_cells(), prm["L"] ** 3)
    np.testing.assert_array_equal(lst, o.cell_lists(1))
    _, vel2, _ = s.download_fluid()
    _, ovel = o.get_fluid()
    assert np.abs(vel2 - ovel).max() < 1e-5


def test_tile_shape_invariance():
    prm = dict(L=24, density=10, n_colloid=6, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=10,
               v0=0.05, theta_deg=90.0, seed=1357, precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE)
    ref = None
    for tile in [(8, 8, 4), (6, 6, 6), (12, 4, 4), (4, 4, 4)]:
        s = B.System(tile=tile, **prm)
        assert s.info[0]["tile"] == tile
        s.startup()
        s.step(3)
        assert s.status() == 0
        st = s.download_fluid() + tuple(s.download_colloid().values())
        if ref is None:
            ref = st
        else:
            for a, b in zip(ref, st):
                np.testing.assert_array_equal(a, b, err_msg=str(tile))


def test_ensemble_equals_single_runs():
    common = dict(sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=10, theta_deg=90.0,
                  precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, density=10)
    reps = [dict(L=16, n_colloid=2, v0=0.05, seed=1001, **common),
            dict(L=24, n_colloid=5, v0=0.02, seed=1002, **common),
            dict(L=16, n_colloid=1, v0=0.0, seed=1003, **common)]
    e = B.System(replicas=reps)
    e.startup()
    e.step(3)
    assert e.status() == 0
    for r, prm in enumerate(reps):
        s = B.System(**prm)
        s.startup()
        s.step(3)
        assert s.status() == 0
        for a, b in zip(e.download_fluid(r), s.download_fluid()):
            np.testing.assert_array_equal(a, b)
        ce, cs = e.download_colloid(r), s.download_colloid()
        for k in ce:
            np.testing.assert_array_equal(ce[k], cs[k], err_msg=f"replica {r} {k}")


def test_graph_replay_equals_direct_launches():
    prm = dict(L=16, density=10, n_colloid=3, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=8,
               v0=0.03, theta_deg=90.0, seed=777, precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE)
    a = B.System(**prm); b = B.System(**prm)
    a.startup(); b.startup()
    a.step(4)
    b.profile_steps(4)
    assert a.status() == 0 and b.status() == 0
    for x, y in zip(a.download_fluid(), b.download_fluid()):
        np.testing.assert_array_equal(x, y)


def test_momentum_and_temperature_64cubed():
    """64^3 x 10 fluid, 300 colloids, FP32: total momentum conserved (v0 = 0)
    and the thermostat holds kT."""
    prm = dict(L=64, density=10, n_colloid=300, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               v0=0.0, theta_deg=90.0, seed=24680, precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE,
               vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE)
    s = B.System(**prm)
    s.startup()
    a0 = s.audit()
    for _ in range(3):
        s.step(1)
        assert s.status() == 0
        a = s.audit()
        drift = np.abs(a["momentum"] - a0["momentum"]).max() / a["momentum_scale"]
        assert drift < 1e-10, drift
    nf = s.info[0]["nf"]
    T = a["ke_fluid"] / (1.5 * nf)
    assert abs(T - 1.0) < 0.01, T


def test_injected_draw_mismatch_is_reported():
    prm = dict(L=10, density=5, n_colloid=3, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20,
               v0=0.01, theta_deg=130.0, seed=77777)
    s = B.System(precision=B.BW_FP64, draws=B.BW_DRAWS_INJECTED, **prm)
    s.startup()
    o = O.Oracle(**prm)
    o.startup()
    shift, axes, refl = oracle_step_draws(o, prm["n_sub"])
    # drop the first reflection tuple of the first substep that has any
    l = next(i for i, r in enumerate(refl) if len(r[0]))
    c, pp, v = refl[l]
    refl[l] = (c[1:], pp[1:], v[1:])
    inject_step(s, shift, axes, refl)
    s.step(1)
    with pytest.raises(B.BrownianError) as ei:
        s.status()
    assert ei.value.code == -4
