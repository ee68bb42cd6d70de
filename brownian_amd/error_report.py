"""FP32-storage error report against an all-FP64 GPU run.

Per-step error: before every MPCD step the FP64 context is loaded with the
FP32 context's state (FP32 values convert to FP64 exactly, so both start
from identical numbers); both take one step with the same device draws
(Philox keys do not depend on precision); the difference after the step is
the error FP32 storage adds in one step.  Trajectory drift: a second FP64
context steps from the common start without re-loading.

  python -m brownian_amd.error_report [--L 128 --ncol 2400 --steps 3] > report.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import (BW_DRAWS_DEVICE, BW_FP32, BW_FP64, BW_ORIENT_SPHERE, BW_VEL_MAXWELL, System)


def _pdiff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def _copy_state(src, dst):
    pos, vel, _ = src.download_fluid()
    dst.upload_fluid(pos, vel)
    c = src.download_colloid()
    dst.upload_colloid(**c)


def compare(a, b, L) -> dict:
    """Max absolute and max relative (to the quantity's RMS) errors of b vs a."""
    pa, va, _ = a.download_fluid()
    pb, vb, _ = b.download_fluid()
    ca, cb = a.download_colloid(), b.download_colloid()
    vrms = float(np.sqrt(np.mean(vb ** 2)))
    dp = np.abs(_pdiff(pa, pb, L)).max(axis=1)
    dv = np.abs(va - vb).max(axis=1) / vrms
    out = {"fluid_pos_abs": float(dp.max()), "fluid_vel_abs": float(dv.max() * vrms), "fluid_vel_rms": vrms}
    out["fluid_pos_rel"] = out["fluid_pos_abs"] / L
    out["fluid_vel_rel"] = float(dv.max())
    # a particle within ~1e-7 of a colloid surface may reflect in one precision
    # and not the other: such discrete events are counted, and the bulk error
    # is the 1 - 1e-6 quantile
    out["fluid_vel_rel_q999999"] = float(np.quantile(dv, 1 - 1e-6))
    out["fluid_pos_abs_q999999"] = float(np.quantile(dp, 1 - 1e-6))
    out["fluid_n_vel_rel_over_1e-5"] = int((dv > 1e-5).sum())
    for k in ("pos", "vel", "omega", "f", "ra"):
        if len(cb[k]) == 0:
            continue
        d = _pdiff(ca[k], cb[k], L) if k == "pos" else ca[k] - cb[k]
        scale = L if k == "pos" else float(np.sqrt(np.mean(cb[k] ** 2))) or 1.0
        out[f"colloid_{k}_abs"] = float(np.abs(d).max())
        out[f"colloid_{k}_rel"] = float(np.abs(d).max() / scale)
    return out


def attribute(a, b, L, params, tol=1e-5) -> dict:
    """Per-step errors of FP32 storage b against FP64 a, with discrete
    collision decisions separated from the bulk.

    A fluid particle whose velocity error exceeds tol (relative to the RMS
    velocity) took a different collision decision near a colloid surface (or
    followed one).  Each such particle is attributed to the colloid nearest to
    it.  A colloid's velocity can then differ by at most the momentum those
    particles exchanged differently, (m_f / m_c) sum |dv_i|, and its angular
    velocity by (m_f / I) (R + 1) sum |dv_i| (collisions conserve momentum and
    angular momentum, cpp/fluid_colloid_collision.cpp:34-47); a colloid with no
    such particle must agree to tol.  Returns the bulk quantiles, how far from
    a colloid the differing particles lie, the particles whose position
    differs beyond tol without a differing velocity, and the colloids outside
    their bound."""
    pa, va, _ = a.download_fluid()
    pb, vb, _ = b.download_fluid()
    ca, cb = a.download_colloid(), b.download_colloid()
    vrms = float(np.sqrt(np.mean(va ** 2)))
    dv = np.abs(va - vb).max(axis=1) / vrms
    dp = np.abs(_pdiff(pa, pb, L)).max(axis=1)
    odd = np.flatnonzero(dv > tol)
    nc = len(ca["pos"])
    m_f, m_c = params["mass_fl"], params["mass_colloid"]
    R = 0.5 * params["sigma"]
    I = 0.1 * m_c * params["sigma"] ** 2
    budget = np.zeros(nc)
    surf = 0.0  # largest distance of a differing particle from its nearest colloid's centre
    for i in odd:
        r = _pdiff(pa[i][None, :], ca["pos"], L)
        d2 = (r ** 2).sum(1)
        j = int(np.argmin(d2))
        surf = max(surf, float(np.sqrt(d2[j])))
        budget[j] += np.abs(va[i] - vb[i]).sum()
    pos_odd = np.flatnonzero(dp > tol)
    vsc = float(np.sqrt(np.mean(ca["vel"] ** 2))) or 1.0
    wsc = float(np.sqrt(np.mean(ca["omega"] ** 2))) or 1.0
    dvc = np.abs(ca["vel"] - cb["vel"]).max(axis=1)
    dwc = np.abs(ca["omega"] - cb["omega"]).max(axis=1)
    dpc = np.abs(_pdiff(ca["pos"], cb["pos"], L)).max(axis=1)
    lim_v = tol * vsc + (m_f / m_c) * budget
    lim_w = tol * wsc + (m_f / I) * (R + 1.0) * budget
    clean = budget == 0
    return {
        "fluid_vel_rel_q999999": float(np.quantile(dv, 1 - 1e-6)),
        "fluid_pos_abs_q999999": float(np.quantile(dp, 1 - 1e-6)),
        "fluid_vel_rel_q99999": float(np.quantile(dv, 1 - 1e-5)),
        "fluid_pos_abs_q99999": float(np.quantile(dp, 1 - 1e-5)),
        "fluid_vel_rel_max": float(dv.max()), "fluid_n_decisions": int(len(odd)),
        "decisions_max_dist_from_centre": surf,
        "fluid_pos_off_not_decisions": int(len(np.setdiff1d(pos_odd, odd))),
        "colloids_with_decisions": int((~clean).sum()),
        "clean_colloid_vel_rel": float(dvc[clean].max() / vsc) if clean.any() else 0.0,
        "clean_colloid_omega_rel": float(dwc[clean].max() / wsc) if clean.any() else 0.0,
        "colloid_pos_abs": float(dpc.max()),
        "colloids_over_vel_bound": int((dvc > lim_v).sum()),
        "colloids_over_omega_bound": int((dwc > lim_w).sum()),
    }


def report(L=128, ncol=2400, steps=3, seed=20240, n_sub=20, log=None) -> dict:
    prm = dict(L=L, density=10, n_colloid=ncol, n_sub=n_sub, ndt=0.1, kbt=1.0, mass_fl=1.0, mass_colloid=50.0,
               sig_colloid=4.2, eps=1.0, sigma=4.0, theta_deg=90.0, v0=0.0, seed=seed, draws=BW_DRAWS_DEVICE,
               vel_law=BW_VEL_MAXWELL, orient_law=BW_ORIENT_SPHERE)
    v0 = np.random.default_rng(seed).uniform(0.0, 0.05, ncol)
    ctx = {}
    for name, prec in (("fp32", BW_FP32), ("fp64", BW_FP64), ("fp64_free", BW_FP64)):
        s = System(precision=prec, **prm)
        s.startup()
        s.set_v0(v0)
        ctx[name] = s
    s32, s64, f64 = ctx["fp32"], ctx["fp64"], ctx["fp64_free"]
    _copy_state(s32, f64)  # common start: the FP32 start-up state
    per_step, drift = [], []
    for k in range(steps):
        _copy_state(s32, s64)
        s32.step(1)
        s64.step(1)
        f64.step(1)
        for s in (s32, s64, f64):
            s.status()
        per_step.append(compare(s64, s32, L))
        per_step[-1]["attribution"] = attribute(s64, s32, L, prm)
        drift.append(compare(f64, s32, L))
        if log:
            log(f"step {k + 1}: per-step fluid vel rel {per_step[-1]['fluid_vel_rel']:.2e}, "
                f"pos abs {per_step[-1]['fluid_pos_abs']:.2e}")
    worst = {key: max(e[key] for e in per_step) for key in per_step[0] if key != "attribution"}
    return {"workload": f"{L}^3 x 10, {ncol} colloids, {n_sub} substeps, Maxwell, v0~U[0,0.05]",
            "method": "per step: FP64 context loaded with the FP32 state before each step, same device draws; "
                      "drift: FP64 context from the common start, never reloaded",
            "steps": steps, "per_step": per_step, "per_step_worst": worst, "trajectory_drift": drift}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--ncol", type=int, default=2400)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=20240)
    args = ap.parse_args()
    t0 = time.time()
    res = report(args.L, args.ncol, args.steps, args.seed,
                 log=lambda m: print(f"[error_report {time.time() - t0:6.1f}s] {m}", file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
