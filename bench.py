"""Benchmark: MPCD steps of active Brownian colloids on one MI355X.

Workload (the roofline-scale system): periodic 128^3 cells x 10 fluid
particles per cell = 20,971,520 fluid particles, 2,400 colloids of radius 2
(LJ/WCA sigma 4.2, mass 50), per-colloid active speed v0 ~ U[0, 0.05],
SRD angle 90 degrees about random axes, random grid shift, 20 streaming
substeps per MPCD step (ndt = 0.1), FP32 integer-cell + offset storage with
FP64 per-cell sums; Maxwell-Boltzmann velocities at kT = 1, orientations
uniform on the sphere.  Synthetic initial state from a fixed seed.

One line of JSON: fluid particle-steps/s of the fused, graph-replayed step
(inputs resident in HBM); the HBM roofline of the step's dominant kernel from
HIP events on the launch stream (algorithmic bytes per launch / average launch
time); the plain method (every particle streamed every substep) timed beside
it; and the CPU oracle (the serial restatement of the reference cpp/ step,
built from oracle/oracle.c at -O2 and -O1) timed on one host core on a bounded
sample of the same state (rotation + run() + the first substeps, extrapolated
to one MPCD step).

The batched ensemble (256 replicas, boxes 16-48, 1-64 colloids, v0 up to
0.1) is timed with the tile shape its box edges share tuned on the whole
ensemble, then refined per box edge (kept where faster), and compared with
the surrogate closure trained on 2M pairs from
8 single-system-style replicas (64^3, 300 colloids): colloid-steps/s, and
from one start-up state the plain and surrogate runs' pooled MSD and
orientation autocorrelation at 100..1000 steps with paired standard errors.

  python bench.py [--steps K --warmup W --L 128 --ncol 2400]
  python bench.py --pmc-summary FETCH_DIR WRITE_DIR --out FILE   (rocprofv3 --pmc outputs -> bytes per launch)
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12        # MI355X HBM3E, bytes/s
METRIC = "fluid particle-steps/s & %HBM peak; ensemble+surrogate vs plain at matched MSD"
COPY_BYTES = 1 << 30     # calibration copy: reads and writes this many bytes
# Algorithmic HBM bytes per particle and launch (FP32 storage: packed cell 4 B,
# offsets 3x4 B, velocities 3x4 B, id 4 B):
#   srd     read the state 32 + write it 32 (the tile re-sort moves every field)
#   stream  read cell+offsets 16 + velocities 12, write cell+offsets 16 (near particles)
#   depart  read cell+offsets 16, write the departure index 4
ALG_BYTES = {"srd": 64, "stream": 44, "depart": 20}
PMC_FILE = os.path.join(ROOT, "profiles", "srd_kernel_128cubed_pmc.json")


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def workload_params(args):
    return dict(L=args.L, density=10, n_colloid=args.ncol, n_sub=args.nsub, ndt=0.1, kbt=1.0, mass_fl=1.0,
                mass_colloid=50.0, sig_colloid=4.2, eps=1.0, sigma=4.0, theta_deg=90.0, v0=0.0, seed=args.seed)


def cpu_baseline(args, opt):
    """Time the CPU oracle -- the project's serial FP64 C restatement of the
    reference cpp/ step (oracle/oracle.c, bit-exact with the compiled reference
    on the golden vectors), built at -O2 or -O1 (the reference Makefile's
    level, cpp/Makefile:3) -- on one host core, in a child process, on a
    bounded sample of this workload: the same start-up state (recomputed on
    the host), one rotation_mpcd + run() and the first substeps; the step time
    is rotation + run + n_sub x the mean substep."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--cpu-worker", opt, "--L", str(args.L), "--ncol",
           str(args.ncol), "--nsub", str(args.nsub), "--seed", str(args.seed)]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    if p.returncode != 0:
        return dict(error=f"worker exit {p.returncode}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def cpu_worker(args):
    """Child process: no GPU use.  Prints progress on stderr, one JSON line."""
    import oracle as O
    import brownian_amd as B
    opt = args.cpu_worker
    so = {"O2": O.ORACLE_SO, "O1": O.ORACLE_SO_O1, "ref": O.REF_SO}[opt]
    if not os.path.exists(so):
        print(json.dumps(dict(error=f"{so} not built")))
        return
    prm = workload_params(args)
    t0 = time.time()
    st = B.host_initialize(vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE, **prm)
    v0 = np.random.default_rng(args.seed).uniform(0.0, 0.05, args.ncol)
    log(f"cpu worker {opt}: start-up state recomputed on the host ({time.time() - t0:.1f} s)")
    z = np.zeros_like(st["c_pos"])
    if opt == "ref":  # the compiled reference: 130 deg and one v0 (the mean) are fixed in its source
        r = O.Reference(**{**prm, "v0": float(np.mean(v0)), "theta_deg": 130.0})
        r.initialize()
    else:
        r = O.Oracle(so=so, **prm)
        r.set_v0(v0)
    r.set_fluid(st["pos"], st["vel"])
    r.set_colloid(st["c_pos"], st["c_vel"], st["c_omega"], z, z, st["c_ra"])
    del st
    r.create_box()
    r.neighbour_list_mpcd()
    r.neighbour_list_md()
    r.compute_force_md()
    log(f"cpu worker {opt}: state loaded ({time.time() - t0:.1f} s); timing rotation_mpcd + run")
    c0 = time.process_time()
    if opt == "ref":
        r.rotation_mpcd_bigstack()
    else:
        r.rotation_mpcd()
    r.run()
    t_rot = time.process_time() - c0
    log(f"cpu worker {opt}: rotation_mpcd + run {t_rot:.2f} s; timing substeps")
    subs = []
    # one substep as in cpp/active_brownian.cpp:27-37 (no updown_velocity: nn <= 10000)
    while len(subs) < args.nsub and (not subs or sum(subs) < 10.0):
        c0 = time.process_time()
        r.copy_old_force()
        r.update_pos_md()
        r.neighbour_list_md()
        r.update_pos_mpcd()
        r.neighbour_list_mpcd()
        r.fluid_colloid_collision()
        r.update_activity_direction()
        r.compute_force_md()
        r.update_velocity_colloid()
        subs.append(time.process_time() - c0)
        log(f"cpu worker {opt}: substep {len(subs)} {subs[-1]:.2f} s")
    step_s = t_rot + args.nsub * float(np.mean(subs))
    print(json.dumps(dict(step_s=step_s, rot_run_s=t_rot, substep_s=subs)))


# ------------------------------------------------------------- PMC summary
def _pmc_rows(d):
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _per_kernel(rows, counter):
    """Mean over dispatches of a per-dispatch counter (summed over its
    dimensions), keyed by the kernel's base name (k_srd, k_stream, ...)."""
    per = {}
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        name = r.get("Kernel_Name", "")
        base = next((k for k in ("k_srd", "k_stream", "k_depart", "k_collide", "k_colloid_update", "k_copy",
                                 "k_md_candidates", "k_tile_colloids") if k in name), None)
        if base is None:
            continue
        key = (base, r.get("Dispatch_Id") or r.get("Correlation_Id"))
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    out = {}
    for (base, _), v in per.items():
        out.setdefault(base, []).append(v)
    return {k: (float(np.mean(v)), len(v)) for k, v in out.items()}


def pmc_summary(fetch_dir, write_dir, out_file, command):
    fetch = _per_kernel(_pmc_rows(fetch_dir), "FETCH_SIZE")
    write = _per_kernel(_pmc_rows(write_dir), "WRITE_SIZE")
    # FETCH_SIZE / WRITE_SIZE are KiB per dispatch; the plain copy of
    # COPY_BYTES gives the factor from counter to bytes
    f_fac = COPY_BYTES / (fetch["k_copy"][0] * 1024.0)
    w_fac = COPY_BYTES / (write["k_copy"][0] * 1024.0)
    res = {"counters": "FETCH_SIZE, WRITE_SIZE (KiB per dispatch), one rocprofv3 --pmc pass each",
           "copy_kernel_bytes": COPY_BYTES, "fetch_factor": f_fac, "write_factor": w_fac,
           "command": command, "kernels": {}}
    for k in sorted(set(fetch) & set(write)):
        fb = fetch[k][0] * 1024.0 * f_fac
        wb = write[k][0] * 1024.0 * w_fac
        res["kernels"][k] = {"fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                             "bytes_per_launch": fb + wb, "dispatches": fetch[k][1]}
    with open(out_file, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res, indent=1))


# ---------------------------------------------------------------- ensemble
def ensemble_params(n_rep, seed0=1000):
    """Batched ensemble: box edge uniform in {16,24,32,40,48} at 10 particles
    per cell, colloid count uniform in 1..cap(L) with the 4 % volume-fraction
    cap (16->4, 24->16, 32->39, >=40->64), v0 uniform in [0, 0.1], seeds
    seed0 + r."""
    rng = np.random.default_rng(seed0)
    cap = {16: 4, 24: 16, 32: 39}
    reps = []
    for r in range(n_rep):
        L = int(rng.choice([16, 24, 32, 40, 48]))
        nc = int(rng.integers(1, cap.get(L, 64) + 1))
        reps.append(dict(L=L, density=10, n_colloid=nc, n_sub=20, ndt=0.1, kbt=1.0, mass_fl=1.0,
                         mass_colloid=50.0, sig_colloid=4.2, eps=1.0, sigma=4.0, theta_deg=90.0,
                         v0=float(rng.uniform(0.0, 0.1)), seed=seed0 + r))
    return reps


def ensemble_replicas(B, n_rep, tile=None, seed0=1000):
    """The batched configuration's replicas; tile: one shape for all, or a
    dict {box edge: shape}."""
    common = dict(precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, vel_law=B.BW_VEL_MAXWELL,
                  orient_law=B.BW_ORIENT_SPHERE, fluid_init=B.BW_INIT_DEVICE)
    reps = [dict(r, **common) for r in ensemble_params(n_rep, seed0)]
    if isinstance(tile, dict):
        for r in reps:
            r["tile"] = tuple(tile[r["L"]])
    elif tile is not None:
        for r in reps:
            r["tile"] = tuple(tile)
    return reps


def parse_tiles(text):
    """'16:8x4x4,48:6x6x4' -> {16: (8, 4, 4), 48: (6, 6, 4)}; None -> None."""
    if not text:
        return None
    out = {}
    for item in text.split(","):
        L, shape = item.split(":")
        out[int(L)] = tuple(int(x) for x in shape.split("x"))
    return out


def run_ensemble(B, n_rep, steps, warmup, tile=None):
    reps = ensemble_replicas(B, n_rep, tile)
    t0 = time.time()
    s = B.System(replicas=reps)
    s.startup()
    s.sync()
    t_setup = time.time() - t0
    nf = sum(i["nf"] for i in s.info)
    nc = sum(i["nc"] for i in s.info)
    log(f"ensemble: {n_rep} replicas, {nf:,} fluid particles, {nc} colloids, start-up {t_setup:.1f} s")
    s.step(warmup)
    s.status()
    ms = s.time_steps(steps)
    s.status()
    near = sum(s.near_particles(r) for r in range(n_rep))
    prof = s.profile_steps(3)
    s.status()
    step_s = ms / 1e3 / steps
    log(f"ensemble: {steps} timed steps: {ms / steps:.3f} ms/step")
    return {"workload": f"{n_rep} replicas, L~U{{16,24,32,40,48}} x 10/cell, colloids U{{1..cap(L)}} (4 % cap), "
                        f"v0~U[0,0.1], seeds 1000+r, 20 substeps, FP32, one set of launches per step",
            "replicas": n_rep, "n_fluid": nf, "n_colloid": nc, "steps": steps, "ms_per_step": ms / steps,
            "tile": list(s.info[0]["tile"]),
            "tiles_per_box": {str(i["L"]): list(i["tile"]) for i in s.info},
            "fluid_particle_steps_per_s": nf / step_s, "colloid_steps_per_s": nc / step_s,
            "near_fraction": near / nf, "setup_s": t_setup,
            "kernels_ms_per_step": {k: v[0] / 3 for k, v in prof.items()}}


def run_surrogate(B, n_rep, n_pairs, epochs, lags, tile=None):
    """The surrogate closure as specified: trained on n_pairs (features,
    impulse) pairs from 8 single-system-style replicas (64^3 x 10, 300
    colloids, v0 0.02, seeds 5000+r), run inside the batched ensemble (seeds
    1000+r), against the plain ensemble from the same start: colloid-steps/s,
    pooled MSD and orientation autocorrelation at each lag, diffusivity (MSD
    slope) ratio, standard errors from the replicas as paired blocks."""
    from brownian_amd import surrogate as S
    common = dict(precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, vel_law=B.BW_VEL_MAXWELL,
                  orient_law=B.BW_ORIENT_SPHERE, fluid_init=B.BW_INIT_DEVICE)
    t0 = time.time()
    X, T = S.collect_pairs(lambda: B.System(replicas=S.c1_replicas(8, **common)), n_pairs)
    log(f"surrogate: {len(X):,} training pairs collected ({time.time() - t0:.1f} s)")
    t1 = time.time()
    net, info = S.fit(X, T, epochs=epochs, batch=4096, lr=1e-3, seed=1)
    t_train = time.time() - t1
    log(f"surrogate: trained {epochs} epochs in {t_train:.1f} s, held-out NLL "
        f"{info['val_loss_before']:.3f} -> {info['val_loss_after']:.3f}")
    reps = ensemble_replicas(B, n_rep, tile)
    out = S.compare_ensemble(lambda: B.System(replicas=reps), net, info, None, lags, log=log)
    out.update({"training_pairs": int(len(X)), "training_systems": "8 x (64^3 x 10, 300 colloids, v0 0.02), "
                "seeds 5000-5007", "epochs": epochs, "train_s": t_train,
                "val_nll_before": info["val_loss_before"], "val_nll_after": info["val_loss_after"]})
    m = out["msd"][-1]
    log(f"surrogate: MSD ratio at {lags[-1]} steps {m['ratio']:.4f} +- {m['ratio_se']:.4f}, diffusivity ratio "
        f"{out['diffusivity']['ratio']:.4f} +- {out['diffusivity']['ratio_se']:.4f}, colloid-steps/s x"
        f"{out['speedup_colloid_steps']:.3f}")
    return out


# ------------------------------------------------------------------- bench
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--ncol", type=int, default=2400)
    ap.add_argument("--nsub", type=int, default=20)
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--streaming", choices=["blocked", "per-substep"], default="blocked")
    ap.add_argument("--copy-reps", type=int, default=10)
    ap.add_argument("--ensemble", type=int, default=256, help="replicas of the batched run (0 = skip)")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("--ensemble-tiles", metavar="L:BXxBYxBZ,...",
                    help="the ensemble's tile shape per box edge (skips its autotuning), e.g. 16:8x4x4,48:6x6x4")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="time only the batched ensemble (one JSON line of its figures)")
    ap.add_argument("--tile", type=int, nargs=3, metavar=("BX", "BY", "BZ"),
                    help="tile shape (skips the autotuner)")
    ap.add_argument("--surrogate", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("--surrogate-pairs", type=int, default=2_000_000)
    ap.add_argument("--surrogate-epochs", type=int, default=8)
    ap.add_argument("--surrogate-lags", default="100,200,500,1000",
                    help="steps at which the plain and surrogate ensembles are compared")
    ap.add_argument("--plain-steps", type=int, default=10, help="timed steps of per-substep streaming (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-ref", action="store_true", help="also time the compiled reference (oracle/_ref)")
    ap.add_argument("--cpu-worker", choices=["O2", "O1", "ref"], help=argparse.SUPPRESS)
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--out")
    ap.add_argument("--command", default="")
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.out, args.command)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the step runs on one GPU; a multi-GPU figure is not produced "
              "(replicas or box slabs would be split over the GPUs).", file=sys.stderr)
        sys.exit(2)

    import brownian_amd as B
    if args.ensemble_only:
        ens = run_ensemble(B, args.ensemble, args.ensemble_steps, 5, parse_tiles(args.ensemble_tiles))
        print(json.dumps({"ensemble": ens}), flush=True)
        return
    prm = workload_params(args)
    t0 = time.time()
    mode = B.BW_STREAM_BLOCKED if args.streaming == "blocked" else B.BW_STREAM_PER_SUBSTEP
    sprm = dict(precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, vel_law=B.BW_VEL_MAXWELL,
                orient_law=B.BW_ORIENT_SPHERE, streaming=mode, fluid_init=B.BW_INIT_DEVICE, **prm)
    rng = np.random.default_rng(args.seed)
    v0 = rng.uniform(0.0, 0.05, args.ncol)
    tune = None
    if args.tile:
        sprm["tile"] = tuple(args.tile)
    elif args.autotune:
        from brownian_amd.autotune import autotune
        log("autotuning the tile shape (a few steps per candidate)")
        tune = autotune(sprm, steps=3, warmup=1, v0=v0, log=log)
        sprm["tile"] = tune["best"]
        log(f"tile {tune['best']} chosen")
    s = B.System(**sprm)
    s.initialize()
    s.initialize_colloid()
    log("colloids placed")
    s.initialize_fluid()
    log(f"fluid placed and uploaded ({time.time() - t0:.1f} s)")
    s.create_box()
    s.neighbour_list_md()
    s.compute_force_md()
    s.tumble()
    s.set_v0(v0)
    s.sync()
    t_setup = time.time() - t0
    nf = s.info[0]["nf"]

    tile_used = list(s.info[0]["tile"])
    log(f"start-up done in {t_setup:.1f} s; {nf:,} fluid particles, tile {tuple(tile_used)}")
    s.step(args.warmup)
    s.status()
    s.sync()
    log(f"{args.warmup} warm-up steps done")
    ms = s.time_steps(args.steps)           # HIP events around the graph replays, on the launch stream
    s.status()
    log(f"{args.steps} timed steps: {ms / args.steps:.3f} ms/step")
    step_s = ms / 1e3 / args.steps
    value = nf / step_s
    n_near = s.near_particles()

    prof = s.profile_steps(args.profile_steps)  # HIP events around every launch, same stream
    s.status()
    log("per-kernel timing: " + ", ".join(f"{k} {v[0] / args.profile_steps:.3f} ms/step ({v[1]} launches)"
                                          for k, v in prof.items()))
    units = {"srd": nf, "depart": nf, "stream": n_near}
    kern = {}
    for k, (kms, kn) in prof.items():
        e = {"ms_per_step": kms / args.profile_steps, "launches_per_step": kn / args.profile_steps,
             "avg_launch_us": 1e3 * kms / max(kn, 1)}
        if k in ALG_BYTES and kn:
            b = ALG_BYTES[k] * units[k]
            e.update(alg_bytes_per_launch=b, achieved_GBps=b / (kms / 1e3 / kn) / 1e9)
        kern[k] = e
    dom = max(kern, key=lambda k: kern[k]["ms_per_step"])
    rk = "srd"  # the HBM kernel that moves every particle of the blocked step
    traffic = None
    if os.path.exists(PMC_FILE):
        try:
            traffic = json.load(open(PMC_FILE))["kernels"]["k_srd"]["bytes_per_launch"]
        except Exception:
            traffic = None
    copy_ms = s.debug_copy_ms(COPY_BYTES, reps=args.copy_reps)
    copy_bw = 2 * COPY_BYTES / (copy_ms / 1e3)

    # the plain method for comparison: every particle streamed every substep
    plain = None
    if args.plain_steps > 0 and args.streaming == "blocked":
        s2 = B.System(**dict(sprm, streaming=B.BW_STREAM_PER_SUBSTEP))
        del s  # one full-size context at a time
        s2.initialize(); s2.initialize_colloid(); s2.initialize_fluid(); s2.create_box()
        s2.neighbour_list_md(); s2.compute_force_md(); s2.tumble(); s2.set_v0(v0)
        s2.step(2)
        s2.status()
        pms = s2.time_steps(args.plain_steps)
        s2.status()
        pval = nf / (pms / 1e3 / args.plain_steps)
        plain = {"streaming": "per-substep", "steps": args.plain_steps, "ms_per_step": pms / args.plain_steps,
                 "fluid_particle_steps_per_s": pval, "alg_bytes_per_particle_step": 44 * args.nsub + 40,
                 "hbm_frac": (44 * args.nsub + 40) * pval / HBM_PEAK}
        log(f"plain per-substep streaming: {pms / args.plain_steps:.3f} ms/step, "
            f"{plain['hbm_frac']:.3f} of HBM peak at {44 * args.nsub + 40} B/particle-step")
        del s2
    else:
        del s

    cpu = None
    if not args.no_cpu_baseline:
        runs = {}
        for opt in ["O2", "O1"] + (["ref"] if args.cpu_ref else []):
            s_ref = time.time()
            log(f"CPU oracle ({opt}) on one host core: bounded sample (progress follows)")
            runs[opt] = cpu_baseline(args, opt)
            log(f"CPU sample {opt} done in {time.time() - s_ref:.1f} s: {runs[opt]}")
        r = runs["O2"]
        if r and "step_s" in r:
            cpu = {"value": nf / r["step_s"], "unit": "fluid particle-steps/s", "cores": 1, "kind": "port",
                   "sample": f"CPU oracle (oracle/oracle.c, the serial FP64 restatement of the reference cpp/ "
                             f"step, bit-exact with it on the golden vectors) compiled -O2, one host core, process "
                             f"CPU time, on this workload's start-up state (90 deg, per-colloid v0): rotation_mpcd "
                             f"+ run {r['rot_run_s']:.2f} s and {len(r['substep_s'])} substep(s) of mean "
                             f"{np.mean(r['substep_s']):.2f} s; step time = rotation + run + {args.nsub} x substep "
                             f"= {r['step_s']:.1f} s",
                   "host_cores_total": os.cpu_count(),
                   "gpu_over_cpu": value / (nf / r["step_s"])}
            for opt, ro in runs.items():
                if opt != "O2" and ro and "step_s" in ro:
                    cpu[f"value_{opt}"] = nf / ro["step_s"]
                    cpu[f"step_s_{opt}"] = ro["step_s"]
        elif r:
            cpu = {"value": None, "unit": "fluid particle-steps/s", "cores": 1, "kind": "port",
                   "sample": "failed: " + str(r.get("error"))}
    ens = None
    if args.ensemble > 0:
        ens_tile, ens_tune = None, None
        try:
            if args.ensemble_tiles:
                ens_tile = parse_tiles(args.ensemble_tiles)
            elif args.autotune:
                from brownian_amd.autotune import autotune_ensemble
                log("autotuning the ensemble's tile shape")
                ens_tune = autotune_ensemble(ensemble_replicas(B, args.ensemble), per_box=True,
                                             even=args.surrogate, log=log)
                ens_tile = ens_tune["best"]
                if ens_tune["per_box_ms"] < ens_tune["ms_per_step"][ens_tile]:
                    ens_tile = ens_tune["per_box"]  # a shape per box edge
            ens = run_ensemble(B, args.ensemble, args.ensemble_steps, 5, ens_tile)
            if ens_tune:
                ens["autotune"] = {"best": list(ens_tune["best"]),
                                   "ms_per_step": {str(k): v for k, v in ens_tune["ms_per_step"].items()},
                                   "per_box": {str(k): list(v) for k, v in ens_tune["per_box"].items()},
                                   "per_box_ms": ens_tune["per_box_ms"],
                                   "per_box_tried": ens_tune["per_box_tried"]}
        except Exception as e:  # reported, the single-system figure stands
            ens = {"error": repr(e)}
        if args.surrogate:
            try:
                lags = [int(x) for x in args.surrogate_lags.split(",")]
                sur = run_surrogate(B, args.ensemble, args.surrogate_pairs, args.surrogate_epochs, lags, ens_tile)
                # the plain ensemble's pooled statistics over the run (same start, replicas as blocks)
                ens["pooled_plain"] = {"lags": sur["lags"],
                                       "msd": [[m["plain"], m["plain_se"]] for m in sur["msd"]],
                                       "orient_corr": [[m["plain"], m["plain_se"]] for m in sur["orient_corr"]]}
                ens["surrogate"] = sur
            except Exception as e:
                ens["surrogate"] = {"error": repr(e)}
    ach = kern[rk]["achieved_GBps"] * 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "fluid particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{args.L}^3 cells x 10 fluid/cell = {nf:,} fluid particles, {args.ncol} colloids "
                               f"R=2 (WCA sigma 4.2, m 50), v0~U[0,0.05], SRD 90 deg, {args.nsub} substeps/step, "
                               f"FP32 int-cell storage, FP64 cell sums, {args.streaming} streaming",
                   "L": args.L, "n_fluid": nf, "n_colloid": args.ncol, "n_sub": args.nsub,
                   "tile": tile_used, "streaming": args.streaming, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_srd (gather, sort, rotate, stream far particles)",
                     "kernel_avg_us": kern[rk]["avg_launch_us"], "alg_bytes_per_launch": kern[rk]["alg_bytes_per_launch"]},
        "dominant_kernel_by_time": dom,
        "kernels": kern,
        "near_fraction": n_near / nf,
        "step_bytes_per_particle": {"plain_per_substep": 44 * args.nsub + 40, "fused_declared": 84,
                                    "fused_hbm_frac": 84 * value / HBM_PEAK},
        "plain_per_substep": plain,
        "autotune": None if tune is None else {"best": list(tune["best"]),
                                                "ms_per_step": {str(k): v for k, v in tune["ms_per_step"].items()}},
        "ensemble": ens,
        "copy_kernel_GBps": copy_bw / 1e9,
        "cpu_baseline": cpu,
        "setup_s": t_setup,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
