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
        drift.append(compare(f64, s32, L))
        if log:
            log(f"step {k + 1}: per-step fluid vel rel {per_step[-1]['fluid_vel_rel']:.2e}, "
                f"pos abs {per_step[-1]['fluid_pos_abs']:.2e}")
    worst = {key: max(e[key] for e in per_step) for key in per_step[0]}
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
