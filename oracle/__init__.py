"""CPU oracle for the MPCD step (test infrastructure only).

Python bindings (ctypes) for

* ``lib/liboracle.so`` -- the project's serial FP64 restatement of the
  reference ``cpp/`` algorithm (``oracle.c``), and
* ``_ref/libbrownian_ref.so`` -- the reference ``cpp/`` sources themselves,
  compiled in place by ``oracle/Makefile`` (present only where that build ran).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, as the checker.  The product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "lib", "liboracle.so")
ORACLE_SO_O1 = os.path.join(HERE, "lib", "liboracle_O1.so")  # -O1, the reference Makefile's level
REF_SO = os.path.join(HERE, "_ref", "libbrownian_ref.so")
REF_SO_O1 = os.path.join(HERE, "_ref", "libbrownian_ref_O1.so")

DRAW_SHIFT, DRAW_AXIS, DRAW_REFLECT, DRAW_TUMBLE = 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


FLANG = "/opt/rocm/llvm/bin/flang"
TWIN_EXE = os.path.join(HERE, "_ref", "fortran_twin")


def build(with_reference: bool | None = None) -> None:
    """Compile the oracle (and the reference when its sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_reference is None:
        with_reference = os.path.isdir("/root/reference/cpp")
    if with_reference:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
        # the Fortran twin, where its sources and a Fortran compiler are present
        if os.path.isdir("/root/reference/fortran") and os.path.exists(FLANG):
            subprocess.run(["make", "-s", "-C", HERE, "twin", f"FC={FLANG}"], check=True)


class Params(C.Structure):
    _fields_ = [
        ("L", C.c_int), ("density", C.c_int), ("n_colloid", C.c_int), ("n_sub", C.c_int),
        ("ndt", C.c_double), ("kbt", C.c_double), ("mass_fl", C.c_double),
        ("mass_colloid", C.c_double), ("sig_colloid", C.c_double), ("eps", C.c_double),
        ("sigma", C.c_double), ("theta_deg", C.c_double), ("v0", C.c_double), ("seed", C.c_int),
    ]


def reference_defaults(**kw) -> dict:
    """Defaults of cpp/active_brownian.cpp:3-7 (sigma = 0.8*sig_colloid)."""
    d = dict(L=30, density=10, n_colloid=10, n_sub=10, ndt=0.1, kbt=1.0, mass_fl=1.0,
             mass_colloid=654.1, sig_colloid=5.0, eps=1.0, sigma=None, theta_deg=130.0,
             v0=0.0, seed=77777)
    d.update(kw)
    if d["sigma"] is None:
        d["sigma"] = 0.80 * d["sig_colloid"]
    return d


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


_libs = {}


def _load(so: str = ORACLE_SO):
    if so not in _libs:
        if not os.path.exists(so):
            raise RuntimeError(f"oracle library missing: {so} (run oracle.build())")
        lib = C.CDLL(so)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.POINTER(Params)]
        lib.orc_ran.restype = C.c_double
        lib.orc_potential.restype = C.c_double
        for name in ["orc_destroy", "orc_initialize_colloid", "orc_initialize_fluid", "orc_create_box",
                     "orc_neighbour_list_mpcd", "orc_neighbour_list_md", "orc_compute_force_md",
                     "orc_tumble", "orc_rotation_mpcd", "orc_run", "orc_update_pos_md",
                     "orc_update_pos_mpcd", "orc_fluid_colloid_collision",
                     "orc_update_activity_direction", "orc_update_velocity_colloid",
                     "orc_copy_old_force", "orc_startup", "orc_step", "orc_substep",
                     "orc_drawlog_clear", "orc_ran", "orc_potential", "orc_n_fluid", "orc_n_colloid",
                     "orc_nbox", "orc_drawlog_count", "orc_updown_velocity"]:
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.orc_rng_reset.argtypes = [C.c_void_p, C.c_int]
        _libs[so] = lib
    return _libs[so]


class Oracle:
    """One oracle instance: the reference's global state, 0-based."""

    def __init__(self, so: str = ORACLE_SO, **params):
        p = reference_defaults(**params)
        self.params = p
        self.lib = _load(so)
        self._p = Params(**p)
        self.h = self.lib.orc_create(C.byref(self._p))
        self.L = p["L"]
        self.nf = self.lib.orc_n_fluid(self.h)
        self.nc = self.lib.orc_n_colloid(self.h)
        self.cells = self.L ** 3

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def __getattr__(self, name):
        # entry points: o.rotation_mpcd(), o.step(), ...
        fn = getattr(self.__dict__.get("lib") or _load(), "orc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda: fn(self.h)

    # generator
    def ran(self) -> float:
        return self.lib.orc_ran(self.h)

    def rng_reset(self, seed: int):
        self.lib.orc_rng_reset(self.h, seed)

    def rng_state(self) -> np.ndarray:
        o = np.zeros(36, np.int32)
        self.lib.orc_rng_get(C.c_void_p(self.h), _ptr(o, _ip))
        return o

    def set_rng_state(self, st):
        st = np.ascontiguousarray(st, np.int32)
        self.lib.orc_rng_set(C.c_void_p(self.h), _ptr(st, _ip))

    def set_v0(self, v0):
        if v0 is None:
            self.lib.orc_set_v0(C.c_void_p(self.h), None)
        else:
            self._v0 = np.ascontiguousarray(v0, np.float64)
            self.lib.orc_set_v0(C.c_void_p(self.h), _ptr(self._v0))

    # state
    def get_fluid(self):
        pos = np.zeros((self.nf, 3)); vel = np.zeros((self.nf, 3))
        self.lib.orc_get_fluid(C.c_void_p(self.h), _ptr(pos), _ptr(vel))
        return pos, vel

    def set_fluid(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float64); vel = np.ascontiguousarray(vel, np.float64)
        self.lib.orc_set_fluid(C.c_void_p(self.h), _ptr(pos), _ptr(vel))

    def get_colloid(self) -> dict:
        arrs = [np.zeros((self.nc, 3)) for _ in range(6)]
        self.lib.orc_get_colloid(C.c_void_p(self.h), *[_ptr(a) for a in arrs])
        return dict(zip(["pos", "vel", "omega", "f", "old_f", "ra"], arrs))

    def set_colloid(self, pos, vel, omega, f, old_f, ra):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (pos, vel, omega, f, old_f, ra)]
        self.lib.orc_set_colloid(C.c_void_p(self.h), *[_ptr(a) for a in arrs])

    def potential(self) -> float:
        return self.lib.orc_potential(self.h)

    def audit(self) -> np.ndarray:
        o = np.zeros(7)
        self.lib.orc_audit(C.c_void_p(self.h), _ptr(o))
        return o

    # maps
    def box_neigh(self) -> np.ndarray:
        nb = self.lib.orc_nbox(self.h)
        o = np.zeros((self.cells, nb), np.int32)
        self.lib.orc_get_box_neigh(C.c_void_p(self.h), _ptr(o, _ip))
        return o

    def fluid_no(self) -> np.ndarray:
        o = np.zeros(self.cells, np.int32)
        self.lib.orc_get_fluid_no(C.c_void_p(self.h), _ptr(o, _ip))
        return o

    def cell_no(self) -> np.ndarray:
        """Per-cell counts of the last shifted binning (rotation_mpcd)."""
        o = np.zeros(self.cells, np.int32)
        self.lib.orc_get_cell_no(C.c_void_p(self.h), _ptr(o, _ip))
        return o

    def cell_lists(self, which: int) -> np.ndarray:
        o = np.zeros(self.nf, np.int32)
        self.lib.orc_get_cell_lists(C.c_void_p(self.h), which, _ptr(o, _ip))
        return o

    def _lists(self, fn, cap):
        counts = np.zeros(max(self.nc, 1), np.int32)
        out = np.zeros(max(cap, 1), np.int32)
        w = fn(C.c_void_p(self.h), _ptr(counts, _ip), _ptr(out, _ip))
        return counts[: self.nc], out[:w]

    def neigh_fl(self):
        return self._lists(self.lib.orc_get_neigh_fl, self.nc * 343 * max(1, self.params["density"]) * 4)

    def md_list(self):
        return self._lists(self.lib.orc_get_md_list, self.nc * self.nc)

    def run_lists(self):
        return self._lists(self.lib.orc_get_run_lists, self.nc * 343 * max(1, self.params["density"]) * 4)

    def up_lists(self):
        return self._lists(self.lib.orc_get_up_lists, self.nc * 343 * max(1, self.params["density"]) * 4)

    def updown(self) -> dict:
        """Results of the last updown_velocity: shell counts and, per colloid,
        the reference's vel and up_vel after that colloid."""
        cnt = np.zeros(max(self.nc, 1), np.int32)
        up = np.zeros(max(self.nc, 1), np.int32)
        vel = np.zeros((max(self.nc, 1), 3))
        upv = np.zeros((max(self.nc, 1), 3))
        self.lib.orc_get_updown(C.c_void_p(self.h), _ptr(cnt, _ip), _ptr(up, _ip), _ptr(vel), _ptr(upv))
        n = self.nc
        return dict(cnt=cnt[:n], up_cnt=up[:n], vel=vel[:n], up_vel=upv[:n])

    # draw log
    def drawlog(self):
        n = self.lib.orc_drawlog_count(self.h)
        kinds = np.zeros(max(n, 1), np.int32)
        keys = np.zeros((max(n, 1), 3), np.int32)
        vals = np.zeros((max(n, 1), 8))
        self.lib.orc_drawlog_get(C.c_void_p(self.h), _ptr(kinds, _ip), _ptr(keys, _ip), _ptr(vals))
        return kinds[:n], keys[:n], vals[:n]

    def clear_drawlog(self):
        self.lib.orc_drawlog_clear(self.h)


class Reference:
    """The compiled reference cpp/ (oracle/_ref).  One instance per process:
    the reference keeps its generator state in function statics."""

    def __init__(self, so: str = REF_SO, **params):
        if not os.path.exists(so):
            raise RuntimeError(f"reference build missing: {so} (run `make -C oracle ref`)")
        p = reference_defaults(**params)
        self.params = p
        lib = C.CDLL(so)
        lib.ref_ran.restype = C.c_double
        lib.ref_potential.restype = C.c_double
        lib.ref_setup.argtypes = [C.c_int] * 4 + [C.c_double] * 8 + [C.c_int]
        self.lib = lib
        lib.ref_setup(p["L"], p["density"], p["n_colloid"], p["n_sub"], p["ndt"], p["kbt"], p["mass_fl"],
                      p["mass_colloid"], p["sig_colloid"], p["eps"], p["sigma"], p["v0"], p["seed"])
        self.L = p["L"]
        self.nf = lib.ref_no_of_fluid()
        self.nc = lib.ref_no_of_colloid()
        self.cells = self.L ** 3

    def __getattr__(self, name):
        fn = getattr(self.lib, "ref_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return fn

    def get_fluid(self):
        pos = np.zeros((self.nf, 3)); vel = np.zeros((self.nf, 3))
        self.lib.ref_get_fluid(_ptr(pos), _ptr(vel))
        return pos, vel

    def set_fluid(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float64); vel = np.ascontiguousarray(vel, np.float64)
        self.lib.ref_set_fluid(_ptr(pos), _ptr(vel))

    def get_colloid(self) -> dict:
        arrs = [np.zeros((self.nc, 3)) for _ in range(6)]
        self.lib.ref_get_colloid(*[_ptr(a) for a in arrs])
        return dict(zip(["pos", "vel", "omega", "f", "old_f", "ra"], arrs))

    def set_colloid(self, pos, vel, omega, f, old_f, ra):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (pos, vel, omega, f, old_f, ra)]
        self.lib.ref_set_colloid(*[_ptr(a) for a in arrs])

    def box_neigh(self):
        nb = self.lib.ref_nbox()
        o = np.zeros((self.cells, nb), np.int32)
        self.lib.ref_get_box_neigh(_ptr(o, _ip))
        return o - 1

    def fluid_no(self):
        o = np.zeros(self.cells, np.int32)
        self.lib.ref_get_fluid_no(_ptr(o, _ip))
        return o

    def cell_lists(self, which: int):
        o = np.zeros(self.nf, np.int32)
        w = self.lib.ref_get_cell_lists(which, _ptr(o, _ip))
        return o[:w] - 1

    def _lists(self, fn, cap):
        counts = np.zeros(max(self.nc, 1), np.int32)
        out = np.zeros(max(cap, 1), np.int32)
        w = fn(_ptr(counts, _ip), _ptr(out, _ip))
        return counts[: self.nc], out[:w] - 1

    def neigh_fl(self):
        return self._lists(self.lib.ref_get_neigh_fl, self.nc * 10002)

    def md_list(self):
        return self._lists(self.lib.ref_get_md_list, self.nc * 201)

    def run_lists(self):
        return self._lists(self.lib.ref_get_run_lists, self.nc * 7001)

    def up_lists(self):
        return self._lists(self.lib.ref_get_up_lists, self.nc * 7001)
