"""Helpers for the parity tests: drive the oracle one MPCD step, turn its
draw log into the GPU's injected draws, and compare states."""
import numpy as np

import oracle as O


def oracle_from_gpu(sysm, params, replica=0, rng_state=None):
    """An oracle holding the GPU context's current state (widened to FP64)."""
    o = O.Oracle(**params)
    pos, vel, _ = sysm.download_fluid(replica)
    o.set_fluid(pos, vel)
    c = sysm.download_colloid(replica)
    o.set_colloid(**c)
    if rng_state is not None:
        o.set_rng_state(rng_state)
    return o


def oracle_step_draws(o, n_sub):
    """Run one oracle MPCD step; return (shift raw 3, axes L^3 x 2, [per-substep
    (colloid, particle, vals6)])."""
    cells = o.cells
    o.clear_drawlog()
    o.rotation_mpcd()
    k, keys, vals = o.drawlog()
    shift = vals[k == O.DRAW_SHIFT][0][:3].copy()
    axes = np.full((cells, 2), np.nan)
    ax = k == O.DRAW_AXIS
    axes[keys[ax, 0]] = vals[ax, :2]
    o.run()
    refl = []
    for _ in range(n_sub):
        o.clear_drawlog()
        o.substep()
        k, keys, vals = o.drawlog()
        m = k == O.DRAW_REFLECT
        refl.append((keys[m, 0].copy(), keys[m, 1].copy(), vals[m, :6].copy()))
    return shift, axes, refl


def inject_step(sysm, shift, axes, refl, replica=0):
    sysm.inject_rotation(shift, axes, replica)
    for l, (c, p, v) in enumerate(refl):
        sysm.inject_reflections(l, c, p, v, replica)


def pdiff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def compare_state(sysm, o, L, pos_tol, vel_tol, col_tol, replica=0, label=""):
    pos, vel, _ = sysm.download_fluid(replica)
    opos, ovel = o.get_fluid()
    dp = np.abs(pdiff(pos, opos, L)).max() if len(pos) else 0.0
    dv = np.abs(vel - ovel).max() if len(vel) else 0.0
    out = {"pos": dp, "vel": dv}
    c = sysm.download_colloid(replica)
    oc = o.get_colloid()
    for k in c:
        if k == "pos":
            out["c_" + k] = np.abs(pdiff(c[k], oc[k], L)).max() if len(c[k]) else 0.0
        else:
            out["c_" + k] = np.abs(c[k] - oc[k]).max() if len(c[k]) else 0.0
    assert dp <= pos_tol, f"{label} fluid positions differ by {dp}"
    assert dv <= vel_tol, f"{label} fluid velocities differ by {dv}"
    for k in c:
        assert out["c_" + k] <= col_tol, f"{label} colloid {k} differs by {out['c_' + k]}"
    return out


def blocked_equals_per_substep(a, b, L, pos_tol, replica=0):
    """a (blocked streaming) and b (per-substep streaming) holding the same
    start-up state take one step: velocities and every colloid quantity equal
    bit for bit, positions within pos_tol (FP32 storage flies far particles
    through the step in one rounding, the per-substep kernel in n_sub).  One
    step from a start-up only: a reflection can leave an offset outside
    [0,1), and reloading a state re-normalises such a (cell, offset) pair, so
    a reloaded context would round that particle's next flight differently.
    Returns the largest position difference."""
    a.step(1)
    b.step(1)
    assert a.status() == 0 and b.status() == 0
    pa, va, _ = a.download_fluid(replica)
    pb, vb, _ = b.download_fluid(replica)
    np.testing.assert_array_equal(va, vb)
    dp = np.abs(pdiff(pa, pb, L)).max() if len(pa) else 0.0
    assert dp <= pos_tol, dp
    da, db = a.download_colloid(replica), b.download_colloid(replica)
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)
    return dp
