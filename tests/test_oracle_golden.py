"""Pin the CPU oracle (oracle/oracle.c) against golden vectors written by the
compiled reference cpp/ (tests/golden/make_golden.py).  Every comparison is
bit-exact: the oracle keeps the reference's operation order and libm calls."""
import ast
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    d = dict(np.load(os.path.join(GOLD, name + ".npz")))
    params = dict(ast.literal_eval(str(d.pop("params_json")))) if "params_json" in d else None
    return d, params


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(O.ORACLE_SO):
        O.build(with_reference=False)


def state_eq(o, g, prefix):
    pos, vel = o.get_fluid()
    np.testing.assert_array_equal(pos, g[prefix + "pos"], err_msg=prefix + "pos")
    np.testing.assert_array_equal(vel, g[prefix + "vel"], err_msg=prefix + "vel")
    c = o.get_colloid()
    for k, v in c.items():
        np.testing.assert_array_equal(v, g[prefix + "c_" + k], err_msg=prefix + "c_" + k)


def test_generator_stream():
    g, _ = load("rng")
    o = O.Oracle(L=7, density=1, n_colloid=1)
    a = np.array([o.ran() for _ in range(20000)])
    np.testing.assert_array_equal(a, g["draws_77777"])
    more = np.array([o.ran() for _ in range(180000)])
    assert a.sum() + more.sum() == g["sum_200k"][0]
    assert more[-1] == g["last_200k"][0]
    o2 = O.Oracle(L=7, density=1, n_colloid=1, seed=-12345)
    np.testing.assert_array_equal(np.array([o2.ran() for _ in range(20000)]), g["draws_m12345"])


def test_generator_range_and_moments():
    # closed-form checks independent of the golden stream: U(0,1) moments
    o = O.Oracle(L=7, density=1, n_colloid=1, seed=4242)
    a = np.array([o.ran() for _ in range(200000)])
    assert a.min() > 0 and a.max() < 1
    assert abs(a.mean() - 0.5) < 4 * (1 / np.sqrt(12 * len(a)))
    assert abs(a.var() - 1 / 12) < 0.002


def test_small_function_trace():
    g, p = load("small")
    o = O.Oracle(**p)
    o.initialize_colloid()
    c = o.get_colloid()
    for k in ("pos", "vel", "omega"):
        np.testing.assert_array_equal(c[k], g["init_colloid_" + k])
    o.initialize_fluid()
    state_eq(o, g, "init_")
    o.create_box()
    np.testing.assert_array_equal(o.box_neigh(), g["box_neigh"].astype(np.int32))
    o.neighbour_list_mpcd()
    np.testing.assert_array_equal(o.fluid_no(), g["nl_fluid_no"])
    np.testing.assert_array_equal(o.cell_lists(0), g["nl_box_part"])
    cnt, lst = o.neigh_fl()
    np.testing.assert_array_equal(cnt, g["nl_no_neigh"])
    np.testing.assert_array_equal(lst, g["nl_neigh_fl"])
    o.neighbour_list_md()
    cnt, lst = o.md_list()
    np.testing.assert_array_equal(cnt, g["md_count"])
    np.testing.assert_array_equal(lst, g["md_list"])
    o.compute_force_md()
    np.testing.assert_array_equal(o.get_colloid()["f"], g["force0"])
    assert o.potential() == g["potential0"][0]
    o.tumble()
    state_eq(o, g, "start_")
    o.rotation_mpcd()
    # the reference's fluid_no is shared: after rotation_mpcd it holds shifted counts
    np.testing.assert_array_equal(o.cell_lists(1), g["rot_cell_part"])
    state_eq(o, g, "rot_")
    o.run()
    cnt, lst = o.run_lists()
    np.testing.assert_array_equal(cnt, g["run_cnt"])
    np.testing.assert_array_equal(lst, g["run_nbr"])
    state_eq(o, g, "run_")
    seq = ["copy_old_force", "update_pos_md", "neighbour_list_md", "update_pos_mpcd",
           "neighbour_list_mpcd", "fluid_colloid_collision", "update_activity_direction",
           "compute_force_md", "update_velocity_colloid"]
    for l in range(1, p["n_sub"] + 1):
        for fn in seq:
            getattr(o, fn)()
            if l == 1:
                state_eq(o, g, f"s1_{fn}_")
                if fn == "neighbour_list_mpcd":
                    np.testing.assert_array_equal(o.fluid_no(), g["s1_nl_fluid_no"])
                    np.testing.assert_array_equal(o.cell_lists(0), g["s1_nl_box_part"])
                    cnt, lst = o.neigh_fl()
                    np.testing.assert_array_equal(cnt, g["s1_nl_no_neigh"])
                    np.testing.assert_array_equal(lst, g["s1_nl_neigh_fl"])
                if fn == "neighbour_list_md":
                    cnt, lst = o.md_list()
                    np.testing.assert_array_equal(cnt, g["s1_md_count"])
                    np.testing.assert_array_equal(lst, g["s1_md_list"])
        assert o.potential() == g[f"pot_sub{l}"][0]
    state_eq(o, g, "step1_")
    o.step(); o.step()
    state_eq(o, g, "step3_")
    assert o.potential() == g["step3_potential"][0]


def test_c0like_trajectory():
    g, p = load("c0like")
    o = O.Oracle(**p)
    o.startup()
    state_eq(o, g, "start_")
    o.step(); o.step()
    state_eq(o, g, "step2_")
    assert o.potential() == g["step2_potential"][0]
    kinds, keys, vals = o.drawlog()
    assert (kinds == O.DRAW_REFLECT).sum() > 20  # collisions were exercised


def test_overlapping_colloids_sequential_semantics():
    g, p = load("overlap")
    o = O.Oracle(**p)
    o.startup()
    c = o.get_colloid()
    c["pos"][0] = [5.0, 5.0, 5.0]
    c["pos"][1] = [8.1, 5.4, 4.8]
    o.set_colloid(**c)
    o.neighbour_list_mpcd(); o.neighbour_list_md(); o.compute_force_md()
    state_eq(o, g, "start_")
    o.step()
    state_eq(o, g, "step1_")
    assert o.potential() == g["step1_potential"][0]


def _run_twin(*args):
    """Run the Fortran twin with a 256 MiB stack: its binning and rotation keep
    cell lists of 100 x 27,000 entries in local arrays."""
    import resource
    import subprocess

    def big_stack():
        hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
        want = 256 << 20
        if hard != resource.RLIM_INFINITY:
            want = min(want, hard)
        resource.setrlimit(resource.RLIMIT_STACK, (want, hard))
    subprocess.run([O.TWIN_EXE, *map(str, args)], check=True, preexec_fn=big_stack)


def test_fortran_twin_generator_and_stencil(tmp_path):
    """The reference's Fortran twin as a second oracle (fortran/ran1.f90 and
    fortran/create_box.f90, compiled from the reference's sources with the
    flags of fortran/Makefile:3 by oracle/Makefile's `twin` target): its
    generator gives the C++ reference's stream bit for bit (the golden draws
    from seeds 77777 and -12345, cpp/parameters.hpp:16-42), and its colloid
    stencil of the 30^3 box is the oracle's create_box list, entry for entry
    (cpp/neighbour_list.cpp:3-25: the 342 cells of the 7^3 cube but the
    colloid's own, in the same order)."""
    if not os.path.exists(O.TWIN_EXE):
        pytest.skip("Fortran twin not built (needs the reference's fortran/ sources and flang)")
    g, _ = load("rng")
    for seed, key in ((77777, "draws_77777"), (-12345, "draws_m12345")):
        out = tmp_path / f"ran1_{seed}.bin"
        _run_twin("ran1", seed, len(g[key]), out)
        np.testing.assert_array_equal(np.fromfile(out, dtype="<f8"), g[key])
    out = tmp_path / "box.bin"
    _run_twin("create_box", out)
    raw = np.fromfile(out, dtype="<i4")
    nbox = int(raw[0])
    o = O.Oracle(L=30, density=1, n_colloid=1)
    o.create_box()
    assert nbox == 342
    np.testing.assert_array_equal(raw[1:].reshape(30 ** 3, nbox) - 1, o.box_neigh())


def test_fortran_twin_streaming_binning_and_rotation(tmp_path):
    """The Fortran twin's streaming with the periodic wrap
    (fortran/update_pos_mpcd.f90), its per-colloid fluid lists
    (fortran/neighbour_list_mpcd.f90) and its SRD rotation at 130 degrees with
    the thermostat (fortran/rotation_mpcd.f90, shift and axes drawn by ran1
    from seed 77777) equal the oracle's cpp restatement bit for bit on the
    twin's compile-time system (30^3 cells, 5 particles per cell, 10
    colloids), from one random state with particles crossing the faces."""
    if not os.path.exists(O.TWIN_EXE):
        pytest.skip("Fortran twin not built (needs the reference's fortran/ sources and flang)")
    rng = np.random.default_rng(5)
    L, dens, nc = 30, 5, 10
    nf = L ** 3 * dens
    pos = rng.uniform(0, L, (nf, 3))
    vel = rng.uniform(-1.7, 1.7, (nf, 3))
    pc = rng.uniform(0, L, (nc, 3))
    pos[:200, 0] = rng.uniform(0, 0.01, 200)  # these leave through x = 0
    vel[:200, 0] = -1.5
    pos[200:400, 1] = rng.uniform(L - 0.01, L, 200)  # and these through y = L
    vel[200:400, 1] = 1.5
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        pos.astype("<f8").tofile(fh)
        vel.astype("<f8").tofile(fh)
        pc.astype("<f8").tofile(fh)
    _run_twin("stream_bin", inp, out)
    raw = open(out, "rb").read()
    f_pos = np.frombuffer(raw[: nf * 24], "<f8").reshape(nf, 3)
    f_cnt = np.frombuffer(raw[nf * 24: nf * 24 + nc * 4], "<i4")
    f_lst = np.frombuffer(raw[nf * 24 + nc * 4:], "<i4").reshape(nc, 10000)
    o = O.Oracle(**O.reference_defaults(density=dens, n_colloid=nc))
    o.set_fluid(pos, vel)
    z = np.zeros((nc, 3))
    o.set_colloid(pc, z, z, z, z, np.tile([1.0, 0.0, 0.0], (nc, 1)))
    o.create_box()
    o.update_pos_mpcd()
    o.neighbour_list_mpcd()
    o_pos, _ = o.get_fluid()
    np.testing.assert_array_equal(o_pos, f_pos)
    cnt, lst = o.neigh_fl()
    np.testing.assert_array_equal(cnt, f_cnt)
    off = 0
    for j in range(nc):
        np.testing.assert_array_equal(lst[off: off + cnt[j]], f_lst[j, : f_cnt[j]] - 1)
        off += cnt[j]
    # one SRD rotation of a fresh state
    pos = rng.uniform(0, L, (nf, 3))
    vel = rng.uniform(-1.7, 1.7, (nf, 3))
    with open(inp, "wb") as fh:
        pos.astype("<f8").tofile(fh)
        vel.astype("<f8").tofile(fh)
    _run_twin("rotate", inp, out)
    f_vel = np.fromfile(out, "<f8").reshape(nf, 3)
    o.set_fluid(pos, vel)
    o.rng_reset(77777)
    o.rotation_mpcd()
    _, o_vel = o.get_fluid()
    np.testing.assert_array_equal(o_vel, f_vel)


def test_fortran_twin_colloid_forces(tmp_path):
    """The Fortran twin's colloid pair list and WCA forces
    (fortran/neighbour_list_md.f90, fortran/compute_force_md.f90) against the
    oracle's cpp restatement (cpp/neighbour_list.cpp:27-39, cpp/update.cpp:3-32)
    for five close pairs, one across a face, one deep in the repulsive core:
    the forces agree to 1e-13 of each colloid's force (plus one unit) and the
    potential to 1e-13 relative.  They are not bit-identical: the twin forms
    sigma^12 / r_cutoff^13 and (sigma / r)^12 by repeated multiplication
    (Fortran integer powers) where cpp/update.cpp:4-5,24 calls pow, which
    moves the cutoff shift fc and the potential by an ulp."""
    if not os.path.exists(O.TWIN_EXE):
        pytest.skip("Fortran twin not built (needs the reference's fortran/ sources and flang)")
    rng = np.random.default_rng(11)
    nc = 10
    pc = np.zeros((nc, 3))
    for k in range(5):
        a = rng.uniform(0, 30, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pc[2 * k] = a
        pc[2 * k + 1] = (a + d * rng.uniform(4.3, 5.5)) % 30
    pc[0] = [0.5, 10.0, 10.0]
    pc[1] = [29.0, 10.2, 10.1]  # 1.5 away through x = 0
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    pc.astype("<f8").tofile(inp)
    _run_twin("force", inp, out)
    raw = np.fromfile(out, "<f8")
    f_f, f_pot = raw[: 3 * nc].reshape(nc, 3), raw[3 * nc]
    o = O.Oracle(**O.reference_defaults(density=1, n_colloid=nc))
    z = np.zeros((nc, 3))
    o.set_colloid(pc, z, z, z, z, np.tile([1.0, 0.0, 0.0], (nc, 1)))
    o.neighbour_list_md()
    o.compute_force_md()
    o_f = o.get_colloid()["f"]
    mag = np.linalg.norm(o_f, axis=1)
    assert (mag > 0).sum() >= 8  # the pairs interact
    assert (np.abs(o_f - f_f).max(axis=1) <= 1e-13 * (mag + 1.0)).all()
    assert abs(o.potential() - f_pot) <= 1e-13 * abs(f_pot)


def test_fortran_twin_stochastic_reflections(tmp_path):
    """The Fortran twin's colloid-fluid collisions
    (fortran/fluid_colloid_collision.f90 with stochastic_reflection.f90 and
    the polar Box-Muller of gauss.f90, draws by ran1 from seed 77777) equal
    the oracle's cpp restatement (cpp/fluid_colloid_collision.cpp:7-49) bit
    for bit: every reflected particle's position and velocity and every
    colloid's velocity and angular velocity after the impulses, on the twin's
    system with 10 colloids dropped into a random fluid (over a thousand
    reflections, overlapping colloids included)."""
    if not os.path.exists(O.TWIN_EXE):
        pytest.skip("Fortran twin not built (needs the reference's fortran/ sources and flang)")
    rng = np.random.default_rng(13)
    L, dens, nc = 30, 5, 10
    nf = L ** 3 * dens
    pos = rng.uniform(0, L, (nf, 3))
    vel = rng.normal(0, 1.0, (nf, 3))
    pc = rng.uniform(0, L, (nc, 3))
    vc = rng.normal(0, 0.05, (nc, 3))
    wc = rng.normal(0, 0.05, (nc, 3))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        for a in (pos, vel, pc, vc, wc):
            a.astype("<f8").tofile(fh)
    _run_twin("collide", inp, out)
    raw = np.fromfile(out, "<f8")
    f_pos, f_vel = raw[: 3 * nf].reshape(nf, 3), raw[3 * nf: 6 * nf].reshape(nf, 3)
    f_vc, f_wc = raw[6 * nf: 6 * nf + 3 * nc].reshape(nc, 3), raw[6 * nf + 3 * nc:].reshape(nc, 3)
    o = O.Oracle(**O.reference_defaults(density=dens, n_colloid=nc))
    o.set_fluid(pos, vel)
    z = np.zeros((nc, 3))
    o.set_colloid(pc, vc, wc, z, z, np.tile([1.0, 0.0, 0.0], (nc, 1)))
    o.rng_reset(77777)
    o.create_box()
    o.neighbour_list_mpcd()
    o.fluid_colloid_collision()
    o_pos, o_vel = o.get_fluid()
    c = o.get_colloid()
    assert (np.abs(o_vel - vel).max(axis=1) > 0).sum() > 1000  # reflections happened
    np.testing.assert_array_equal(o_pos, f_pos)
    np.testing.assert_array_equal(o_vel, f_vel)
    np.testing.assert_array_equal(c["vel"], f_vc)
    np.testing.assert_array_equal(c["omega"], f_wc)


def test_fortran_twin_colloid_updates(tmp_path):
    """The Fortran twin's colloid side of a step -- run() with its rear-shell
    payment (fortran/run.f90), update_activity_direction, update_pos_md with
    the periodic wrap and update_velocity_colloid -- against the oracle's cpp
    restatement (cpp/run.cpp:10-26, cpp/update.cpp:34-67) at the twin's
    v0 = 0.04: the kicked fluid, colloid positions (one crossing two faces),
    velocities and orientations agree bit for bit."""
    if not os.path.exists(O.TWIN_EXE):
        pytest.skip("Fortran twin not built (needs the reference's fortran/ sources and flang)")
    rng = np.random.default_rng(17)
    L, dens, nc = 30, 5, 10
    nf = L ** 3 * dens
    pos = rng.uniform(0, L, (nf, 3))
    vel = rng.normal(0, 1.0, (nf, 3))
    pc = rng.uniform(0, L, (nc, 3))
    vc = rng.normal(0, 0.05, (nc, 3))
    wc = rng.normal(0, 0.5, (nc, 3))
    ra = rng.normal(size=(nc, 3))
    ra /= np.linalg.norm(ra, axis=1, keepdims=True)
    fc = rng.normal(0, 5, (nc, 3))
    of = rng.normal(0, 5, (nc, 3))
    pc[0] = [29.9995, 0.0004, 15.0]  # leaves through x = L and y = 0
    vc[0] = [0.2, -0.2, 0.0]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        for a in (pos, vel, pc, vc, wc, ra, fc, of):
            a.astype("<f8").tofile(fh)
    _run_twin("colloid", inp, out)
    raw = np.fromfile(out, "<f8")
    k = 3 * nf
    f_vel = raw[:k].reshape(nf, 3)
    f_pc, f_vc, f_ra = (raw[k + 3 * nc * i: k + 3 * nc * (i + 1)].reshape(nc, 3) for i in range(3))
    o = O.Oracle(**O.reference_defaults(density=dens, n_colloid=nc, v0=0.04))
    o.set_fluid(pos, vel)
    o.set_colloid(pc, vc, wc, fc, of, ra)
    o.create_box()
    o.neighbour_list_mpcd()
    o.run()
    o.update_activity_direction()
    o.update_pos_md()
    o.update_velocity_colloid()
    _, o_vel = o.get_fluid()
    c = o.get_colloid()
    assert (np.abs(o_vel - vel).max(axis=1) > 0).sum() > 1000  # shells paid the kicks
    assert c["pos"][0][0] < 1.0 and c["pos"][0][1] > 29.0  # wrapped
    np.testing.assert_array_equal(o_vel, f_vel)
    np.testing.assert_array_equal(c["pos"], f_pc)
    np.testing.assert_array_equal(c["vel"], f_vc)
    np.testing.assert_array_equal(c["ra"], f_ra)


def test_fortran_twin_start_up_and_two_steps(tmp_path):
    """End to end: the Fortran twin's own program order
    (fortran/active_brownian.f90:7-30 -- start-up from seed 77777, initial
    lists and forces, tumble, then per step rotation, run and 10 substeps) on
    its compile-time system (30^3 cells x 5, 10 colloids, 130 degrees,
    v0 = 0.04) equals the oracle's start-up and steps bit for bit, after the
    start-up and after two steps: every fluid position and velocity, every
    colloid position, velocity, angular velocity and orientation."""
    if not os.path.exists(O.TWIN_EXE):
        pytest.skip("Fortran twin not built (needs the reference's fortran/ sources and flang)")
    L, dens, nc = 30, 5, 10
    nf = L ** 3 * dens
    o = O.Oracle(**O.reference_defaults(density=dens, n_colloid=nc, v0=0.04))
    o.startup()
    for k, mode in ((0, ("startup",)), (2, ("steps", 2))):
        for _ in range(k):
            o.step()
        out = tmp_path / f"state_{k}.bin"
        _run_twin(*mode, out)
        raw = np.fromfile(out, "<f8")
        sizes = [nf, nf, nc, nc, nc, nc]
        parts, off = [], 0
        for n in sizes:
            parts.append(raw[off: off + 3 * n].reshape(n, 3))
            off += 3 * n
        o_pos, o_vel = o.get_fluid()
        c = o.get_colloid()
        for name, a, b in zip(("pos_fl", "vel_fl", "pos", "vel", "omega", "ra"),
                              (o_pos, o_vel, c["pos"], c["vel"], c["omega"], c["ra"]), parts):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} after {k} steps")
