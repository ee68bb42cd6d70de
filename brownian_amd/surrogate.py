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
