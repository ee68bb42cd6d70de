"""Long-run statistics of the CPU oracle (the serial FP64 restatement of the
reference cpp/ step, pinned bit-exactly against the compiled reference by
tests/test_oracle_golden.py) for the CPU-runnable configuration: 20^3 cells x
10 fluid per cell, 8 colloids (radius 2, WCA sigma 4.2, mass 50), 20
substeps, SRD angle 90 degrees, v0 = 0.01, the reference's start-up laws.

For each seed the oracle runs `steps` MPCD steps with its own generator and
records every `every` steps the box averages of the colloid mean squared
displacement (unwrapped), the orientation autocorrelation ra(t).ra(0), the
fluid temperature m (sum|v|^2 - |sum v|^2/N) / (3 (N-1)) and the colloid
translational and rotational temperatures.  Output: longrun_c0.npz.

  python tests/golden/make_longrun.py [--seeds 32] [--steps 1000] [--every 10] [--jobs 8]
"""
import argparse
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

C0 = dict(L=20, density=10, n_colloid=8, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0, n_sub=20, ndt=0.1,
          kbt=1.0, mass_fl=1.0, eps=1.0, theta_deg=90.0, v0=0.01)
SEED0 = 1001


def observables(o, disp, ra0, prm):
    _, vel = o.get_fluid()
    c = o.get_colloid()
    n = len(vel)
    t_f = prm["mass_fl"] * ((vel ** 2).sum() - (vel.sum(0) ** 2).sum() / n) / (3 * (n - 1))
    I = 0.1 * prm["mass_colloid"] * prm["sigma"] ** 2
    return [float(((disp ** 2).sum(1)).mean()), float((c["ra"] * ra0).sum(1).mean()), float(t_f),
            float(prm["mass_colloid"] * (c["vel"] ** 2).sum(1).mean() / 3), float(I * (c["omega"] ** 2).sum(1).mean() / 3)]


def run(args):
    seed, steps, every = args
    import oracle as O
    prm = dict(C0, seed=seed)
    o = O.Oracle(**prm)
    o.startup()
    L = prm["L"]
    c = o.get_colloid()
    pos_prev, ra0 = c["pos"].copy(), c["ra"].copy()
    disp = np.zeros_like(pos_prev)
    rows = [observables(o, disp, ra0, prm)]
    for s in range(1, steps + 1):
        o.step()
        p = o.get_colloid()["pos"]
        d = p - pos_prev
        disp += d - L * np.round(d / L)
        pos_prev = p.copy()
        if s % every == 0:
            rows.append(observables(o, disp, ra0, prm))
    print(f"seed {seed} done", flush=True)
    return np.array(rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    import oracle as O
    O.build(with_reference=False)
    seeds = list(range(SEED0, SEED0 + a.seeds))
    with Pool(a.jobs) as pool:
        series = pool.map(run, [(s, a.steps, a.every) for s in seeds])
    np.savez_compressed(os.path.join(HERE, "longrun_c0.npz"), series=np.stack(series), seeds=np.array(seeds),
                        steps=np.arange(0, a.steps + 1, a.every),
                        columns=np.array(["msd", "orient_corr", "t_fluid", "t_colloid", "t_rot"]),
                        params_json=json.dumps(C0))


if __name__ == "__main__":
    main()
