"""Long-run statistics against the reference path (2 sigma).

The CPU-runnable configuration (20^3 cells x 10 fluid per cell, 8 colloids of
radius 2, WCA sigma 4.2, mass 50, 20 substeps, SRD angle 90 degrees,
v0 = 0.01) from the reference's start-up laws, 1,000 MPCD steps:

* the oracle -- the serial FP64 restatement of cpp/, pinned bit-exactly to
  the compiled reference -- ran 64 seeds with the reference generator
  (tests/golden/make_longrun.py, fixture longrun_c0.npz);
* the GPU runs 256 replicas (seeds 1001..1256, the first 64 starting from the
  very states the oracle started from) in one ensemble, FP32 storage, device
  draws, blocked streaming.

Each trajectory is one sample (box averages over its 8 colloids); the means
of the two sets must agree within 2 sigma, sigma the combined standard error,
for the fluid temperature, the colloid mean squared displacement and the
orientation autocorrelation <ra(t).ra(0)> at the end of the run (the
quantities of cpp/active_brownian.cpp:39-55 and cpp/update.cpp:34-46), and
within 3 sigma at the intermediate lags and for the colloid temperatures.
"""
import ast
import os

import numpy as np
import pytest

import brownian_amd as B

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gpu_series(seeds, steps, every, params):
    reps = [dict(params, seed=int(s), precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE) for s in seeds]
    e = B.System(replicas=reps)
    e.startup()
    e.observables_reset()
    keys = ["msd", "orient_corr", "t_fluid", "t_colloid", "t_rot"]

    def record():
        return [[e.observables(r)[k] for k in keys] for r in range(len(reps))]

    rows = [record()]
    for _ in range(steps // every):
        e.step(every)
        assert e.status() == 0
        rows.append(record())
    return np.array(rows).transpose(1, 0, 2)  # replica, time, quantity


def _z(a, b):
    ma, mb = a.mean(), b.mean()
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return (ma - mb) / se, ma, mb, se


@pytest.mark.timeout(600)
def test_long_run_statistics_match_the_oracle():
    g = np.load(os.path.join(GOLD, "longrun_c0.npz"))
    params = dict(ast.literal_eval(str(g["params_json"])))
    ora = g["series"]  # seed, time, quantity
    steps_axis = g["steps"]
    every = int(steps_axis[1] - steps_axis[0])
    steps = int(steps_axis[-1])
    gpu = _gpu_series(range(1001, 1001 + 256), steps, every, params)
    assert gpu.shape[1] == ora.shape[1]
    report = {}
    # end of the run: MSD and orientation autocorrelation; fluid temperature
    # averaged over the run
    checks = [("msd_end", gpu[:, -1, 0], ora[:, -1, 0], 2.0),
              ("orient_end", gpu[:, -1, 1], ora[:, -1, 1], 2.0),
              ("t_fluid", gpu[:, 1:, 2].mean(1), ora[:, 1:, 2].mean(1), 2.0),
              ("t_colloid", gpu[:, 1:, 3].mean(1), ora[:, 1:, 3].mean(1), 3.0),
              ("t_rot", gpu[:, 1:, 4].mean(1), ora[:, 1:, 4].mean(1), 3.0)]
    for lag in (100, 300, 600):
        i = lag // every
        checks.append((f"msd_{lag}", gpu[:, i, 0], ora[:, i, 0], 3.0))
        checks.append((f"orient_{lag}", gpu[:, i, 1], ora[:, i, 1], 3.0))
    bad = []
    for name, a, b, lim in checks:
        z, ma, mb, se = _z(a, b)
        report[name] = (round(float(z), 2), float(ma), float(mb), float(se))
        if abs(z) >= lim:
            bad.append(name)
    print("long-run z-scores (gpu - oracle) / sigma:", report)
    assert not bad, (bad, report)
    # the thermostat holds the fluid at kT = 1 (to its finite-cell correction)
    assert abs(gpu[:, 1:, 2].mean() - 1.0) < 0.01
