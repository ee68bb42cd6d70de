"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs the reference cpp/ sources of shashankgandham/brownian, compiled in place
by ``make -C oracle ref`` into oracle/_ref/libbrownian_ref.so, each case in a
fresh process (the reference keeps its generator state in function statics).
The fixtures are data only: inputs, parameters and the reference's outputs.

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

# Cases.  "small": L=7 so that the 7^3 stencil has no duplicate cells; colloid
# sigma reduced so that two colloids fit (placement needs 1.3*sig_colloid).
SMALL = dict(L=7, density=3, n_colloid=2, sig_colloid=3.0, v0=0.05, n_sub=10)
# "c0like": the shape of the 20^3 workload scaled down (decoupled diameter 4,
# LJ sigma 4.2, colloid mass 50, 20 substeps); 130 degrees as hard-coded.
C0LIKE = dict(L=10, density=5, n_colloid=3, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0,
              n_sub=20, v0=0.01)
# "overlap": two colloids moved to overlap after start-up, so that fluid lies
# inside both and the sequential per-colloid semantics matter.
OVERLAP = dict(L=10, density=5, n_colloid=2, sig_colloid=4.2, sigma=4.0, mass_colloid=50.0,
               n_sub=10, v0=0.02)


def _state(r):
    pos, vel = r.get_fluid()
    c = r.get_colloid()
    out = {"pos": pos, "vel": vel}
    out.update({"c_" + k: v for k, v in c.items()})
    return out


def case_rng(q):
    r = O.Reference(L=7, density=1, n_colloid=1)
    r.initialize()  # allocates the generator table (iv)
    a = np.array([r.ran() for _ in range(20000)])
    more = np.array([r.ran() for _ in range(180000)])
    q.put({"draws_77777": a, "sum_200k": np.array([a.sum() + more.sum()]), "last_200k": more[-1:]})


def case_rng_neg(q):
    r = O.Reference(L=7, density=1, n_colloid=1, seed=-12345)
    r.initialize()
    a = np.array([r.ran() for _ in range(20000)])
    q.put({"draws_m12345": a})


def case_small(q):
    r = O.Reference(**SMALL)
    out = {}
    r.initialize()
    r.initialize_colloid()
    out.update({"init_colloid_" + k: v for k, v in r.get_colloid().items()})
    r.initialize_fluid()
    out.update({"init_" + k: v for k, v in _state(r).items()})
    r.create_box()
    out["box_neigh"] = r.box_neigh().astype(np.int16)
    r.neighbour_list_mpcd()
    out["nl_fluid_no"] = r.fluid_no()
    out["nl_box_part"] = r.cell_lists(0)
    out["nl_no_neigh"], out["nl_neigh_fl"] = r.neigh_fl()
    r.neighbour_list_md()
    out["md_count"], out["md_list"] = r.md_list()
    r.compute_force_md()
    out["force0"] = r.get_colloid()["f"]
    out["potential0"] = np.array([r.potential()])
    r.tumble()
    out.update({"start_" + k: v for k, v in _state(r).items()})
    # function trace of one MPCD step (cpp/active_brownian.cpp:25-38)
    r.rotation_mpcd_bigstack()
    out["rot_fluid_no"] = r.fluid_no()
    out["rot_cell_part"] = r.cell_lists(1)
    out.update({"rot_" + k: v for k, v in _state(r).items()})
    r.run()
    out["run_cnt"], out["run_nbr"] = r.run_lists()
    out.update({"run_" + k: v for k, v in _state(r).items()})
    seq = ["copy_old_force", "update_pos_md", "neighbour_list_md", "update_pos_mpcd",
           "neighbour_list_mpcd", "fluid_colloid_collision", "update_activity_direction",
           "compute_force_md", "update_velocity_colloid"]
    for l in range(1, SMALL["n_sub"] + 1):
        for fn in seq:
            getattr(r, fn)()
            if l == 1:
                out.update({f"s1_{fn}_" + k: v for k, v in _state(r).items()})
                if fn == "neighbour_list_mpcd":
                    out["s1_nl_fluid_no"] = r.fluid_no()
                    out["s1_nl_box_part"] = r.cell_lists(0)
                    out["s1_nl_no_neigh"], out["s1_nl_neigh_fl"] = r.neigh_fl()
                if fn == "neighbour_list_md":
                    out["s1_md_count"], out["s1_md_list"] = r.md_list()
        out[f"pot_sub{l}"] = np.array([r.potential()])
    out.update({"step1_" + k: v for k, v in _state(r).items()})
    r.steps(2)
    out.update({"step3_" + k: v for k, v in _state(r).items()})
    out["step3_potential"] = np.array([r.potential()])
    q.put(out)


def case_c0like(q):
    r = O.Reference(**C0LIKE)
    r.startup()
    out = {"start_" + k: v for k, v in _state(r).items()}
    r.steps(2)
    out.update({"step2_" + k: v for k, v in _state(r).items()})
    out["step2_potential"] = np.array([r.potential()])
    q.put(out)


def case_overlap(q):
    r = O.Reference(**OVERLAP)
    r.startup()
    c = r.get_colloid()
    c["pos"][0] = [5.0, 5.0, 5.0]
    c["pos"][1] = [8.1, 5.4, 4.8]
    r.set_colloid(**c)
    r.neighbour_list_mpcd()
    r.neighbour_list_md()
    r.compute_force_md()
    out = {"start_" + k: v for k, v in _state(r).items()}
    r.steps(1)
    out.update({"step1_" + k: v for k, v in _state(r).items()})
    out["step1_potential"] = np.array([r.potential()])
    q.put(out)


def case_updown(q):
    # cpp/run.cpp:28-50 where the reference calls it (cpp/active_brownian.cpp:32):
    # after neighbour_list_mpcd of a substep, here the first substep of step 2
    r = O.Reference(**C0LIKE)
    r.startup()
    r.steps(1)
    r.rotation_mpcd_bigstack()
    r.run()
    r.copy_old_force(); r.update_pos_md(); r.neighbour_list_md(); r.update_pos_mpcd(); r.neighbour_list_mpcd()
    out = {"pre_" + k: v for k, v in _state(r).items()}
    r.updown_velocity()
    out["cnt"], out["nbr"] = r.run_lists()
    out["up_cnt"], out["up_nbr"] = r.up_lists()
    q.put(out)


def _run(fn):
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p = ctx.Process(target=fn, args=(q,))
    p.start()
    out = q.get()
    p.join()
    if p.exitcode != 0:
        raise RuntimeError(f"{fn.__name__} failed with {p.exitcode}")
    return out


def main():
    if not os.path.exists(O.REF_SO):
        O.build(with_reference=True)
    if "--only" in sys.argv:
        which = sys.argv[sys.argv.index("--only") + 1]
        fn, params = {"updown": (case_updown, C0LIKE)}[which]
        out = _run(fn)
        out["params_json"] = np.array(repr(sorted(O.reference_defaults(**params).items())))
        np.savez_compressed(os.path.join(HERE, f"{which}.npz"), **out)
        print(which, "ok", len(out), "arrays")
        return
    rng = _run(case_rng)
    rng.update(_run(case_rng_neg))
    np.savez_compressed(os.path.join(HERE, "rng.npz"), **rng)
    for name, fn, params in [("small", case_small, SMALL), ("c0like", case_c0like, C0LIKE),
                             ("overlap", case_overlap, OVERLAP)]:
        out = _run(fn)
        p = O.reference_defaults(**params)
        out["params_json"] = np.array(repr(sorted(p.items())))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, "ok", len(out), "arrays")


if __name__ == "__main__":
    main()
