"""Benchmark: MPCD steps of active Brownian colloids on one MI355X.

Workload (the roofline-scale system): periodic 128^3 cells x 10 fluid
particles per cell = 20,971,520 fluid particles, 2,400 colloids of radius 2
(LJ/WCA sigma 4.2, mass 50), per-colloid active speed v0 ~ U[0, 0.05],
SRD angle 90 degrees about random axes, random grid shift, 20 streaming
substeps per MPCD step (ndt = 0.1), FP32 integer-cell + offset storage with
FP64 per-cell sums; Maxwell-Boltzmann velocities at kT = 1, orientations
uniform on the sphere.  Synthetic initial state from a fixed seed.

One line of JSON: fluid particle-steps/s of the fused, graph-replayed step
(inputs resident in HBM), the HBM roofline of the dominant kernel (streaming)
from HIP events, and the reference cpp/ program (compiled from its sources
into oracle/_ref) timed on one host core on a bounded sample of the same state
(rotation + run() + the first substeps, extrapolated to one MPCD step).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12        # MI355X HBM3E, bytes/s
STREAM_BYTES = 44        # per particle per substep: R pos 16 + R vel 12 + W pos 16
SRD_BYTES = 40           # per particle per step:    R pos 16 + R vel 12 + W vel 12
METRIC = "fluid particle-steps/s & %HBM peak; ensemble+surrogate vs plain at matched MSD"
PMC_FILE = os.path.join(ROOT, "profiles", "stream_kernel_c2_pmc.json")


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def workload_params(args):
    return dict(L=args.L, density=10, n_colloid=args.ncol, n_sub=args.nsub, ndt=0.1, kbt=1.0, mass_fl=1.0,
                mass_colloid=50.0, sig_colloid=4.2, eps=1.0, sigma=4.0, theta_deg=90.0, v0=0.0, seed=args.seed)


def cpu_reference_baseline(args):
    """Time the reference cpp/ (compiled from its sources into oracle/_ref) on
    one host core, in a child process, on a bounded sample of this workload:
    the same start-up state (recomputed on the host), one rotation_mpcd + run()
    and the first substep(s); the step time is rotation + run + n_sub x the mean
    substep."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--cpu-worker", "--L", str(args.L), "--ncol",
           str(args.ncol), "--nsub", str(args.nsub), "--seed", str(args.seed)]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    if p.returncode != 0:
        return dict(error=f"worker exit {p.returncode}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def cpu_worker(args):
    """Child process: no GPU use.  Prints progress on stderr, one JSON line."""
    import oracle as O
    import brownian_amd as B
    if not os.path.exists(O.REF_SO):
        print(json.dumps(dict(error="oracle/_ref not built")))
        return
    prm = workload_params(args)
    t0 = time.time()
    st = B.host_initialize(vel_law=B.BW_VEL_MAXWELL, orient_law=B.BW_ORIENT_SPHERE, **prm)
    v0 = np.random.default_rng(args.seed).uniform(0.0, 0.05, args.ncol)
    log(f"cpu worker: start-up state recomputed on the host ({time.time() - t0:.1f} s)")
    r = O.Reference(**{**prm, "v0": float(np.mean(v0)), "theta_deg": 130.0})
    r.initialize()
    r.set_fluid(st["pos"], st["vel"])
    z = np.zeros_like(st["c_pos"])
    r.set_colloid(st["c_pos"], st["c_vel"], st["c_omega"], z, z, st["c_ra"])
    del st
    r.create_box()
    r.neighbour_list_mpcd()
    r.neighbour_list_md()
    r.compute_force_md()
    log(f"cpu worker: reference loaded ({time.time() - t0:.1f} s); timing rotation_mpcd + run")
    c0 = time.process_time()
    r.rotation_mpcd_bigstack()
    r.run()
    t_rot = time.process_time() - c0
    log(f"cpu worker: rotation_mpcd + run {t_rot:.2f} s; timing substeps")
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
        log(f"cpu worker: substep {len(subs)} {subs[-1]:.2f} s")
    step_s = t_rot + args.nsub * float(np.mean(subs))
    print(json.dumps(dict(step_s=step_s, rot_run_s=t_rot, substep_s=subs)))


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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; a multi-GPU figure is not produced "
              "(the slab decomposition is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import brownian_amd as B
    prm = workload_params(args)
    t0 = time.time()
    s = B.System(precision=B.BW_FP32, draws=B.BW_DRAWS_DEVICE, vel_law=B.BW_VEL_MAXWELL,
                 orient_law=B.BW_ORIENT_SPHERE, **prm)
    rng = np.random.default_rng(args.seed)
    v0 = rng.uniform(0.0, 0.05, args.ncol)
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

    log(f"start-up done in {t_setup:.1f} s; {nf:,} fluid particles, tile {s.info[0]['tile']}")
    s.step(args.warmup)
    s.status()
    s.sync()
    log(f"{args.warmup} warm-up steps done")
    ms = s.time_steps(args.steps)           # HIP events around the graph replays
    s.status()
    log(f"{args.steps} timed steps: {ms / args.steps:.3f} ms/step")
    step_s = ms / 1e3 / args.steps
    value = nf / step_s
    b_alg = STREAM_BYTES * args.nsub + SRD_BYTES
    step_frac = nf * b_alg / step_s / HBM_PEAK

    prof = s.profile_steps(args.profile_steps)  # HIP events around every launch
    s.status()
    log("per-kernel timing: " + ", ".join(f"{k} {v[0] / args.profile_steps:.3f} ms/step ({v[1]} launches)"
                                          for k, v in prof.items()))
    st_ms, st_n = prof["stream"]
    stream_avg_s = st_ms / 1e3 / max(st_n, 1)
    achieved = nf * STREAM_BYTES / stream_avg_s
    traffic = None
    if os.path.exists(PMC_FILE):
        try:
            traffic = json.load(open(PMC_FILE)).get("bytes_per_launch")
        except Exception:
            traffic = None
    copy_bytes = 1 << 30
    copy_ms = s.debug_copy_ms(copy_bytes, reps=10)
    copy_bw = 2 * copy_bytes / (copy_ms / 1e3)

    cpu = None
    if not args.no_cpu_baseline:
        s_ref = time.time()
        log("reference cpp/ on one host core: bounded sample (progress follows)")
        r = cpu_reference_baseline(args)
        log(f"reference sample done in {time.time() - s_ref:.1f} s: {r}")
        if r and "step_s" in r:
            cpu = {"value": nf / r["step_s"], "unit": "fluid particle-steps/s", "cores": 1, "kind": "reference",
                   "sample": f"reference cpp/ compiled -O2 (oracle/_ref), one host core, process CPU time, on this "
                             f"workload's start-up state: rotation_mpcd + run {r['rot_run_s']:.2f} s and "
                             f"{len(r['substep_s'])} substep(s) of mean {np.mean(r['substep_s']):.2f} s, step time = "
                             f"rotation + run + {args.nsub} x substep = {r['step_s']:.1f} s; the reference fixes the "
                             f"SRD angle at 130 deg and one v0 (the mean)",
                   "gpu_over_cpu": value / (nf / r["step_s"])}
        elif r:
            cpu = {"value": None, "unit": "fluid particle-steps/s", "cores": 1, "kind": "reference",
                   "sample": "failed: " + str(r.get("error"))}
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
                               f"FP32 int-cell storage, FP64 cell sums",
                   "L": args.L, "n_fluid": nf, "n_colloid": args.ncol, "n_sub": args.nsub,
                   "tile": list(s.info[0]["tile"]), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": "k_stream (streaming + detection)",
                     "kernel_avg_us": stream_avg_s * 1e6, "bytes_alg_per_launch": nf * STREAM_BYTES},
        "step_hbm_frac": step_frac,
        "step_bytes_alg_per_particle": b_alg,
        "kernel_ms_per_step": {k: v[0] / args.profile_steps for k, v in prof.items()},
        "copy_kernel_GBps": copy_bw / 1e9,
        "cpu_baseline": cpu,
        "setup_s": t_setup,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
