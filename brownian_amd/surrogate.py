"""Surrogate closure workflow: collect (features, impulse) pairs from the
explicit collision kernel, fit the network on the GPU, attach it to a system.

A pair is one colloid over one MPCD step: the 144 features at the step's
start (27 coarse-cell fluid moments around the colloid, its velocity,
angular velocity and orientation) and the momentum and angular momentum the
fluid collisions gave it during the step (6 values, units of the fluid
particle mass).  Attached, the network's Gaussian prediction of that impulse
replaces the collisions for every colloid whose features lie inside the
training range; the others keep the explicit collisions.
"""
from __future__ import annotations

import numpy as np

from . import MLP


def collect(system, steps, replicas=None, explicit_only=True):
    """Run `steps` fused MPCD steps with moments on and return the pairs
    (X: n x 144 float32, T: n x 6 float64) of every colloid and step."""
    reps = range(system.n_rep) if replicas is None else replicas
    system.set_moments(True)
    system.step(1)  # the moments describe the fluid from here on
    system.status()
    xs, ts = [], []
    for _ in range(steps):
        system.step(1)
        for r in reps:
            f, imp, mode = system.surrogate_data(r)
            keep = mode == 0 if explicit_only else np.ones(len(mode), bool)
            xs.append(f[keep])
            ts.append(imp[keep])
    system.status()
    return np.concatenate(xs), np.concatenate(ts)


def feature_range(X, margin=0.05):
    """Per-feature training range widened by `margin` of its span."""
    lo, hi = X.min(0), X.max(0)
    span = np.maximum(hi - lo, 1e-6)
    return (lo - margin * span).astype(np.float32), (hi + margin * span).astype(np.float32)


def fit(X, T, epochs=10, batch=4096, lr=1e-3, seed=1, val_frac=0.1):
    """Standardise, train on the GPU, report the held-out loss."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(X))
    nv = max(1, int(val_frac * len(X)))
    va, tr = idx[:nv], idx[nv:]
    xm, xsd = X[tr].mean(0), X[tr].std(0)
    xsd = np.where(xsd > 1e-6, xsd, 1.0)
    tm, tsd = T[tr].mean(0), T[tr].std(0)
    tsd = np.where(tsd > 1e-12, tsd, 1.0)
    net = MLP(seed)
    net.set_normalisation(xm, xsd, tm, tsd)
    v0, _ = net.loss_grad(X[va], T[va])
    hist = net.train(X[tr], T[tr], epochs=epochs, batch=batch, lr=lr, seed=seed)
    v1, _ = net.loss_grad(X[va], T[va])
    lo, hi = feature_range(X[tr])
    t_lo, t_hi = feature_range(T[tr], margin=0.0)
    return net, {"train_loss": hist.tolist(), "val_loss_before": v0, "val_loss_after": v1, "lo": lo, "hi": hi,
                 "t_lo": t_lo, "t_hi": t_hi, "n_train": int(len(tr)), "n_val": int(nv)}


def attach(system, net, info):
    system.surrogate_attach(net, info["lo"], info["hi"], info["t_lo"], info["t_hi"])


def c1_replicas(n=8, seed0=5000, **common):
    """Training systems of the single-system configuration: 64^3 cells x 10,
    300 colloids of radius 2 (about 4 % volume fraction), v0 = 0.02, seeds
    seed0 + r (disjoint from the ensemble's evaluation seeds)."""
    return [dict(L=64, density=10, n_colloid=300, n_sub=20, ndt=0.1, kbt=1.0, mass_fl=1.0, mass_colloid=50.0,
                 sig_colloid=4.2, eps=1.0, sigma=4.0, theta_deg=90.0, v0=0.02, seed=seed0 + r, **common)
            for r in range(n)]


def collect_pairs(system_factory, n_pairs):
    """Pairs from a fresh system (system_factory()), stepped until at least
    n_pairs explicit (features, impulse) pairs are held."""
    s = system_factory()
    s.startup()
    nc = sum(i["nc"] for i in s.info)
    steps = max(1, -(-n_pairs // nc))
    X, T = collect(s, steps)
    s.close()
    return X, T


def paired_msd(plain, sur, weights):
    """Pooled mean squared displacement of the two runs and their ratio, with
    the standard error of the ratio from the replicas as independent blocks
    (both runs start from the same state, so replica-to-replica differences
    in box, colloid count and activity cancel in the paired difference)."""
    w = np.asarray(weights, float) / np.sum(weights)
    a, b = np.asarray(plain, float), np.asarray(sur, float)
    ma, mb = float((w * a).sum()), float((w * b).sum())
    n = len(w)
    # delta method for the ratio of two weighted means over the same blocks
    ra = w * (a - ma) / ma
    rb = w * (b - mb) / mb
    se_ratio = (mb / ma) * float(np.sqrt(((rb - ra) ** 2).sum() * n / (n - 1)))
    se_a = float(np.sqrt(((w * (a - ma)) ** 2).sum() * n / (n - 1)))
    se_b = float(np.sqrt(((w * (b - mb)) ** 2).sum() * n / (n - 1)))
    return {"plain": ma, "plain_se": se_a, "surrogate": mb, "surrogate_se": se_b, "ratio": mb / ma,
            "ratio_se": se_ratio}


def compare_ensemble(system_factory, net, info, steps, lags, warmup=3, log=None):
    """Plain and surrogate runs of the same ensemble from one start-up state
    (system_factory() builds it): ms per step, colloid-steps/s, the fraction
    of colloid-steps the network drove, and at each lag (steps) the pooled
    MSD, orientation autocorrelation and their paired ratios; the
    diffusivity is the MSD slope between the first and last lag."""
    out = {}
    series = {}
    for name in ("plain", "surrogate"):
        s = system_factory()
        s.startup()
        if name == "surrogate":
            attach(s, net, info)
        s.step(warmup)
        s.status()
        s.observables_reset()
        nc = np.array([i["nc"] for i in s.info])
        msd, oc, done, ms = [], [], 0, 0.0
        for lag in lags:
            ms += s.time_steps(lag - done)
            done = lag
            s.status()
            obs = [s.observables(r) for r in range(s.n_rep)]
            msd.append([o["msd"] for o in obs])
            oc.append([o["orient_corr"] for o in obs])
        frac = None
        if name == "surrogate":
            frac = float(np.mean(np.concatenate([s.surrogate_data(r)[2] for r in range(s.n_rep)])))
        series[name] = (np.array(msd), np.array(oc))
        w = nc / nc.sum()
        out[name] = {"ms_per_step": ms / done, "colloid_steps_per_s": nc.sum() / (ms / 1e3 / done),
                     "surrogate_fraction": frac,
                     "t_colloid": float((w * [o["t_colloid"] for o in obs]).sum()),
                     "t_rot": float((w * [o["t_rot"] for o in obs]).sum())}
        if log:
            log(f"surrogate comparison: {name} {ms / done:.3f} ms/step")
        s.close()
    w = nc
    out["lags"] = list(lags)
    out["msd"] = [paired_msd(series["plain"][0][i], series["surrogate"][0][i], w) for i in range(len(lags))]
    out["orient_corr"] = [paired_msd(series["plain"][1][i], series["surrogate"][1][i], w) for i in range(len(lags))]
    d_pl = (series["plain"][0][-1] - series["plain"][0][0])
    d_su = (series["surrogate"][0][-1] - series["surrogate"][0][0])
    out["diffusivity"] = paired_msd(d_pl, d_su, w)
    out["speedup_colloid_steps"] = out["surrogate"]["colloid_steps_per_s"] / out["plain"]["colloid_steps_per_s"]
    return out
