"""Pins of the oracle functions that the reference's golden vectors leave
open, each by what the mathematics fixes rather than by restating the
oracle's own formula:

* ``orc_audit`` (cpp/active_brownian.cpp:39-55): the kinetic energies and the
  momentum are the definitions 1/2 m sum|v|^2 and sum m v, evaluated here from
  the state with numpy; the potential is the WCA sum whose pair force the
  goldens pin, checked by finite differences (F = -dU/dx) and by its zero at
  the cutoff.
* per-colloid v0 (``orc_set_v0``, used by cpp/run.cpp:10-26 through the
  scalar v0): run() must change colloid j's velocity by exactly v0_j ra_j and
  remove m_c v0_j ra_j from the fluid of its rear shell whenever the shell is
  not empty; a uniform array must equal the scalar path bit for bit.
* the SRD rotation at 90 degrees (cpp/rotation_mpcd.cpp:19-49, goldens exist
  only at the reference's 130): in every cell the map v - u -> v' - u is
  s R with R orthonormal, det R = +1, rotation angle theta, axis the drawn
  one, and s the thermostat factor that makes the relative kinetic energy
  3 (N - 1) kT / 2 (the velocity-rescaling thermostat's target).
"""
import os

import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(O.ORACLE_SO):
        O.build(with_reference=False)


def _random_state(o, seed, vscale=1.0):
    rng = np.random.default_rng(seed)
    L = o.L
    pos = rng.uniform(0, L, (o.nf, 3))
    vel = rng.normal(0, vscale, (o.nf, 3))
    c = o.get_colloid()
    # colloids on a loose grid so that some pairs sit inside the WCA range
    base = np.array([[1.0, 1.0, 1.0], [5.3, 1.2, 1.1], [1.2, 5.5, 0.9], [7.0, 7.0, 7.5]])[: o.nc]
    c["pos"] = base + rng.uniform(-0.05, 0.05, base.shape)
    c["vel"] = rng.normal(0, 0.2, (o.nc, 3))
    c["omega"] = rng.normal(0, 0.1, (o.nc, 3))
    ra = rng.normal(size=(o.nc, 3))
    c["ra"] = ra / np.linalg.norm(ra, axis=1, keepdims=True)
    o.set_fluid(pos, vel)
    o.set_colloid(**c)
    return pos, vel, c


def test_audit_closed_form():
    p = dict(L=12, density=3, n_colloid=4, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, mass_fl=1.3)
    o = O.Oracle(**p)
    pos, vel, c = _random_state(o, 11)
    o.create_box(); o.neighbour_list_md(); o.compute_force_md()
    a = o.audit()
    I = 0.1 * p["mass_colloid"] * p["sigma"] ** 2  # cpp/active_brownian.cpp:7
    np.testing.assert_allclose(a[0], 0.5 * p["mass_colloid"] * (c["vel"] ** 2).sum(), rtol=1e-13)
    np.testing.assert_allclose(a[1], 0.5 * I * (c["omega"] ** 2).sum(), rtol=1e-13)
    np.testing.assert_allclose(a[3], 0.5 * p["mass_fl"] * (vel ** 2).sum(), rtol=1e-12)
    mom = p["mass_colloid"] * c["vel"].sum(0) + p["mass_fl"] * vel.sum(0)
    np.testing.assert_allclose(a[4:7], mom, rtol=1e-10, atol=1e-10 * np.abs(p["mass_fl"] * vel).sum())
    assert a[2] == o.potential()
    # the potential: F = -dU/dx by central differences on colloid 1, which
    # interacts with colloid 0 (distance ~4.3 < r_c = 2^(1/6) 4.2 = 4.71)
    f = o.get_colloid()["f"]
    assert np.abs(f[1]).max() > 1e-3
    h = 1e-6
    for k in range(3):
        cc = o.get_colloid()
        U = []
        for sgn in (+1, -1):
            q = {kk: v.copy() for kk, v in cc.items()}
            q["pos"][1, k] += sgn * h
            o.set_colloid(**q)
            o.neighbour_list_md(); o.compute_force_md()
            U.append(o.potential())
        o.set_colloid(**cc)
        np.testing.assert_allclose(-(U[0] - U[1]) / (2 * h), f[1, k], rtol=1e-5, atol=1e-7)
    # shifted potential: a lone pair at the cutoff contributes (almost) nothing
    o2 = O.Oracle(**dict(p, n_colloid=2))
    rc = 2 ** (1 / 6) * p["sig_colloid"]
    c2 = o2.get_colloid()
    c2["pos"] = np.array([[2.0, 2.0, 2.0], [2.0 + rc * (1 - 1e-9), 2.0, 2.0]])
    o2.set_colloid(**c2)
    o2.create_box(); o2.neighbour_list_md(); o2.compute_force_md()
    assert abs(o2.potential()) < 1e-8


def test_per_colloid_v0_run_momentum_identity():
    p = dict(L=12, density=10, n_colloid=4, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, v0=0.0)
    o = O.Oracle(**p)
    _random_state(o, 5)
    o.create_box(); o.neighbour_list_mpcd(); o.neighbour_list_md(); o.compute_force_md()
    v0 = np.array([0.03, 0.0, 0.011, 0.05])
    o.set_v0(v0)
    _, vel0 = o.get_fluid()
    c0 = o.get_colloid()
    o.run()
    _, vel1 = o.get_fluid()
    c1 = o.get_colloid()
    cnt, _ = o.run_lists()
    m_c, m_f = p["mass_colloid"], 1.0
    # colloid j: v += v0_j ra_j (cpp/run.cpp:14)
    np.testing.assert_allclose(c1["vel"] - c0["vel"], v0[:, None] * c0["ra"], rtol=1e-12, atol=1e-16)
    # fluid: minus the kick momentum of every colloid with a non-empty shell
    expect = -(m_c * v0[:, None] * c0["ra"])[cnt > 0].sum(0)
    got = m_f * (vel1 - vel0).sum(0)
    np.testing.assert_allclose(got, expect, rtol=1e-10, atol=1e-14)
    assert (cnt > 0).sum() >= 2
    # a colloid with v0 = 0 changes nothing: total momentum conserved exactly
    # up to rounding when every active colloid has a shell
    tot0 = m_c * c0["vel"].sum(0) + m_f * vel0.sum(0)
    tot1 = m_c * c1["vel"].sum(0) + m_f * vel1.sum(0)
    if np.all(cnt[v0 > 0] > 0):
        np.testing.assert_allclose(tot1, tot0, atol=1e-11 * np.abs(vel0).sum())
    # a uniform array equals the scalar path bit for bit
    a = O.Oracle(**dict(p, v0=0.02))
    b = O.Oracle(**dict(p, v0=0.0))
    for x in (a, b):
        _random_state(x, 5)
        x.create_box(); x.neighbour_list_mpcd(); x.neighbour_list_md(); x.compute_force_md()
    b.set_v0(np.full(p["n_colloid"], 0.02))
    a.run(); b.run()
    for u, v in zip(a.get_fluid(), b.get_fluid()):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(a.get_colloid()["vel"], b.get_colloid()["vel"])


def rotation_invariants(vel0, vel1, cells, theta_deg, kT=1.0, m=1.0, min_n=4):
    """Per cell with at least min_n particles: the least-squares map M with
    v1 - u = M (v0 - u); returns the largest deviations of M/s from an
    orthonormal matrix, of det, of the rotation angle and of the axis, and of
    the relative kinetic energy from 3 (N - 1) kT / 2."""
    order = np.argsort(cells, kind="stable")
    cs = cells[order]
    bounds = np.flatnonzero(np.diff(cs)) + 1
    groups = np.split(order, bounds)
    worst = dict(orth=0.0, det=0.0, angle=0.0, axis=0.0, energy=0.0, mean=0.0, cells=0)
    ct = np.cos(np.radians(theta_deg))
    for g in groups:
        N = len(g)
        if N < min_n:
            continue
        u0 = vel0[g].mean(0)
        u1 = vel1[g].mean(0)
        worst["mean"] = max(worst["mean"], np.abs(u1 - u0).max())
        X = vel0[g] - u0
        Y = vel1[g] - u0
        Mt, *_ = np.linalg.lstsq(X, Y, rcond=None)
        M = Mt.T
        s = np.sqrt(np.trace(M.T @ M) / 3)
        R = M / s
        worst["orth"] = max(worst["orth"], np.abs(R.T @ R - np.eye(3)).max())
        worst["det"] = max(worst["det"], abs(np.linalg.det(R) - 1))
        worst["angle"] = max(worst["angle"], abs((np.trace(R) - 1) / 2 - ct))
        e = 0.5 * m * (Y ** 2).sum()
        worst["energy"] = max(worst["energy"], abs(e / (1.5 * (N - 1) * kT) - 1))
        worst["cells"] += 1
    return worst


def rotation_axes_check(vel0, vel1, cells, axes, theta_deg):
    """The signed axis of each cell's rotation -- from the antisymmetric part
    of R, (R - R^T)/2 = sin(theta) [r]_x for a right-handed rotation by theta
    about r -- against the axis drawn for the cell (a uniform point on the
    sphere: z = 2 U1 - 1, azimuth 2 pi U2).  Returns the largest deviation."""
    dev = 0.0
    st = np.sin(np.radians(theta_deg))
    for c in np.unique(cells):
        g = np.flatnonzero(cells == c)
        if len(g) < 4 or np.isnan(axes[c, 0]):
            continue
        u0 = vel0[g].mean(0)
        X, Y = vel0[g] - u0, vel1[g] - u0
        M = np.linalg.lstsq(X, Y, rcond=None)[0].T
        R = M / np.sqrt(np.trace(M.T @ M) / 3)
        a = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / (2 * st)
        rho = 2 * axes[c, 0] - 1
        phi = 2 * np.pi * axes[c, 1]
        r = np.array([np.cos(phi) * np.sqrt(1 - rho ** 2), np.sin(phi) * np.sqrt(1 - rho ** 2), rho])
        dev = max(dev, np.abs(a - r).max())
    return dev


def oracle_cells_after_rotation(o):
    counts = o.cell_no()
    lst = o.cell_lists(1)
    cells = np.empty(o.nf, np.int64)
    cells[lst] = np.repeat(np.arange(len(counts)), counts)
    return cells


@pytest.mark.parametrize("theta", [90.0, 130.0, 45.0])
def test_oracle_rotation_invariants(theta):
    p = dict(L=8, density=10, n_colloid=1, theta_deg=theta, seed=3141)
    o = O.Oracle(**p)
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 8, (o.nf, 3))
    vel = rng.normal(0, 1, (o.nf, 3)) + np.array([0.3, -0.2, 0.1])
    o.set_fluid(pos, vel)
    c = o.get_colloid()
    c["pos"][:] = 4.0
    o.set_colloid(**c)
    o.clear_drawlog()
    o.rotation_mpcd()
    _, vel1 = o.get_fluid()
    cells = oracle_cells_after_rotation(o)
    w = rotation_invariants(vel, vel1, cells, theta)
    assert w["cells"] > 400
    assert w["orth"] < 1e-10 and w["det"] < 1e-10, w
    assert w["angle"] < 1e-10, w
    assert w["energy"] < 1e-10, w
    assert w["mean"] < 1e-13, w
    k, keys, vals = o.drawlog()
    axes = np.full((o.cells, 2), np.nan)
    m = k == O.DRAW_AXIS
    axes[keys[m, 0]] = vals[m, :2]
    assert rotation_axes_check(vel, vel1, cells, axes, theta) < 1e-9


def updown_brute_force(pos, vel, cpos, cvel, L, sigma, neigh):
    """cpp/run.cpp:28-50 evaluated directly: per colloid the fluid of its
    neigh_fl list behind it (r . v_c <= 0) within sigma/2 + 0.5 (rear shell) and
    sigma/2 + 0.1 (near surface), the mean velocity relative to the colloid,
    with the reference's running (never reset) up_vel and its up_cnt guard on
    both divisions."""
    out_v, out_u, up = [], [], np.zeros(3)
    for i, lst in enumerate(neigh):
        r = pos[lst] - cpos[i]
        r -= L * np.round(r / L)
        d2 = (r ** 2).sum(1)
        behind = r @ cvel[i] <= 0
        rear = lst[(d2 <= (sigma / 2 + 0.5) ** 2) & behind]
        near = lst[(d2 <= (sigma / 2 + 0.1) ** 2) & behind]
        v = vel[rear].sum(0) if len(rear) else np.zeros(3)
        up = up + (vel[near].sum(0) if len(near) else 0.0)
        if len(near):
            up = up / len(near) - cvel[i]
            v = v / len(rear) - cvel[i]
        out_v.append(v)
        out_u.append(up.copy())
    return np.array(out_v), np.array(out_u)


def test_updown_velocity_against_reference_lists_and_brute_force():
    import ast
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "updown.npz")))
    p = dict(ast.literal_eval(str(g.pop("params_json"))))
    o = O.Oracle(**p)
    o.set_fluid(g["pre_pos"], g["pre_vel"])
    o.set_colloid(**{k: g["pre_c_" + k] for k in ("pos", "vel", "omega", "f", "old_f", "ra")})
    o.create_box()
    o.neighbour_list_mpcd()
    o.updown_velocity()
    # the lists the compiled reference wrote (cnt/nbr and up_cnt/up_nbr), bit for bit
    cnt, nbr = o.run_lists()
    np.testing.assert_array_equal(cnt, g["cnt"])
    np.testing.assert_array_equal(nbr, g["nbr"])
    up_cnt, up_nbr = o.up_lists()
    np.testing.assert_array_equal(up_cnt, g["up_cnt"])
    np.testing.assert_array_equal(up_nbr, g["up_nbr"])
    res = o.updown()
    counts, flat = o.neigh_fl()
    neigh = np.split(flat, np.cumsum(counts)[:-1])
    v, u = updown_brute_force(g["pre_pos"], g["pre_vel"], g["pre_c_pos"], g["pre_c_vel"], p["L"], p["sigma"], neigh)
    np.testing.assert_allclose(res["vel"], v, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(res["up_vel"], u, rtol=1e-13, atol=1e-14)
    assert np.all(res["up_cnt"] > 0)
