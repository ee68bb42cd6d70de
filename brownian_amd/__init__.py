"""brownian_amd -- MI355X (gfx950) MPCD step for active Brownian colloids.

Python view of the C-ABI in include/brownian.h (libbrownian.so, built in-tree
by ``brownian_amd._build``).  The compute is the HIP library; this module only
marshals arrays.  There is no fallback: if the library is missing or no GPU is
present, constructing a System raises.

Entry points keep the names of the reference's functions
(cpp/parameters.hpp:12-14): ``s.rotation_mpcd()``, ``s.run()``,
``s.fluid_colloid_collision()``, ...
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libbrownian.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "brownian.h")

BW_FP32, BW_FP64 = 0, 1
BW_VEL_REFERENCE, BW_VEL_MAXWELL = 0, 1
BW_ORIENT_REFERENCE, BW_ORIENT_SPHERE = 0, 1
BW_DRAWS_DEVICE, BW_DRAWS_INJECTED = 0, 1
BW_STREAM_BLOCKED, BW_STREAM_PER_SUBSTEP = 0, 1
BW_INIT_HOST, BW_INIT_DEVICE = 0, 1
KERNEL_FAMILIES = ["stream", "collide", "depart", "srd", "colloid", "other"]

ERRORS = {-1: "BW_ERR_ARG", -2: "BW_ERR_HIP", -3: "BW_ERR_CAPACITY", -4: "BW_ERR_DRAWS",
          -5: "BW_ERR_STATE", -6: "BW_ERR_PHYSICS"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class BrownianError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("L", C.c_int), ("density", C.c_int), ("n_colloid", C.c_int), ("n_sub", C.c_int),
        ("ndt", C.c_double), ("kbt", C.c_double), ("mass_fl", C.c_double), ("mass_colloid", C.c_double),
        ("sig_colloid", C.c_double), ("eps", C.c_double), ("sigma", C.c_double), ("theta_deg", C.c_double),
        ("v0", C.c_double), ("seed", C.c_int), ("precision", C.c_int), ("vel_law", C.c_int),
        ("orient_law", C.c_int), ("draws", C.c_int), ("tile", C.c_int * 3), ("streaming", C.c_int),
        ("fluid_init", C.c_int), ("reserved", C.c_int * 6),
    ]


_lib = None


def lib():
    """Load libbrownian.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libbrownian.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.bw_last_error.restype = C.c_char_p
        L.bw_ctx_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.bw_ctx_create_ensemble.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]
        L.bw_default_params.argtypes = [C.POINTER(Params)]
        L.bw_host_initialize.argtypes = [C.POINTER(Params)] + [_dp] * 6 + [_ip]
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Entry points declared BW_API in include/brownian.h."""
    text = open(HEADER).read()
    return re.findall(r"BW_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)


def default_params(**kw) -> Params:
    p = Params()
    lib().bw_default_params(C.byref(p))
    sigma_given = "sigma" in kw
    for k, v in kw.items():
        if k == "tile":
            for i in range(3):
                p.tile[i] = int(v[i])
        else:
            setattr(p, k, v)
    if "sig_colloid" in kw and not sigma_given:
        p.sigma = 0.8 * p.sig_colloid
    return p


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def host_initialize(**kw):
    """Reference start-up state computed on the host (no GPU needed)."""
    p = default_params(**kw)
    nf = p.L ** 3 * p.density
    nc = p.n_colloid
    out = {k: np.zeros((n, 3)) for k, n in
           [("pos", nf), ("vel", nf), ("c_pos", nc), ("c_vel", nc), ("c_omega", nc), ("c_ra", nc)]}
    rng = np.zeros(36, np.int32)
    rc = lib().bw_host_initialize(C.byref(p), _ptr(out["pos"]), _ptr(out["vel"]), _ptr(out["c_pos"]),
                                  _ptr(out["c_vel"]), _ptr(out["c_omega"]), _ptr(out["c_ra"]), _ptr(rng, _ip))
    if rc:
        raise BrownianError(rc, "bw_host_initialize")
    out["rng"] = rng
    return out


class System:
    """One context: a single system, or an ensemble when given a list of
    parameter dicts (all sharing the physical constants)."""

    ENTRY = ["initialize", "initialize_colloid", "initialize_fluid", "create_box", "neighbour_list_mpcd",
             "neighbour_list_md", "compute_force_md", "tumble", "rotation_mpcd", "run", "update_pos_md",
             "update_pos_mpcd", "fluid_colloid_collision", "update_activity_direction",
             "update_velocity_colloid", "bw_copy_old_force", "bw_startup", "updown_velocity"]

    def __init__(self, replicas=None, **params):
        self.lib = lib()
        if replicas is None:
            replicas = [params]
        arr = (Params * len(replicas))(*[default_params(**r) for r in replicas])
        self._params = arr
        h = C.c_void_p()
        rc = self.lib.bw_ctx_create_ensemble(arr, len(replicas), C.byref(h))
        if rc:
            msg = self.lib.bw_last_error(h).decode() if h.value else "create failed"
            if h.value:  # a partly built context holds device memory and a stream
                self.lib.bw_ctx_destroy(h)
            raise BrownianError(rc, msg)
        self.h = h
        self.n_rep = len(replicas)
        self.info = []
        for r in range(self.n_rep):
            o = np.zeros(7, np.int32)
            self._chk(self.lib.bw_replica_info(self.h, r, _ptr(o, _ip)))
            self.info.append(dict(nf=int(o[0]), nc=int(o[1]), L=int(o[2]), tile=tuple(int(x) for x in o[3:6]),
                                  cap=int(o[6])))

    def close(self):
        if getattr(self, "h", None):
            self.lib.bw_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise BrownianError(rc, self.lib.bw_last_error(self.h).decode())
        return rc

    def __getattr__(self, name):
        if name in System.ENTRY:
            fn = getattr(self.lib, name)
            return lambda: self._chk(fn(self.h))
        raise AttributeError(name)

    def startup(self):
        return self._chk(self.lib.bw_startup(self.h))

    def copy_old_force(self):
        return self._chk(self.lib.bw_copy_old_force(self.h))

    def substep(self):
        """cpp/active_brownian.cpp:28-37 via the individual entry points."""
        self.copy_old_force()
        self.update_pos_md()
        self.neighbour_list_md()
        self.update_pos_mpcd()
        self.neighbour_list_mpcd()
        self.fluid_colloid_collision()
        self.update_activity_direction()
        self.compute_force_md()
        self.update_velocity_colloid()

    def step_sequence(self, n_sub):
        """One MPCD step (cpp/active_brownian.cpp:25-38) as separate calls."""
        self.rotation_mpcd()
        self.run()
        for _ in range(n_sub):
            self.substep()

    def step(self, n=1):
        return self._chk(self.lib.bw_step(self.h, int(n)))

    def sync(self):
        return self._chk(self.lib.bw_sync(self.h))

    def status(self, raise_on_error=True):
        f = C.c_uint(0)
        rc = self.lib.bw_status(self.h, C.byref(f))
        if rc and raise_on_error:
            raise BrownianError(rc, self.lib.bw_last_error(self.h).decode())
        return f.value

    def time_steps(self, n):
        ms = C.c_double(0)
        self._chk(self.lib.bw_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value

    def profile_steps(self, n):
        ms = np.zeros(6)
        cnt = np.zeros(6, np.int32)
        self._chk(self.lib.bw_profile_steps(self.h, int(n), _ptr(ms), _ptr(cnt, _ip)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_FAMILIES)}

    # ---- state
    def set_v0(self, v0, replica=0):
        a = None if v0 is None else np.ascontiguousarray(v0, np.float64)
        self._chk(self.lib.bw_set_v0(self.h, replica, _ptr(a)))

    def upload_fluid(self, pos, vel, replica=0):
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        self._chk(self.lib.bw_upload_fluid(self.h, replica, _ptr(pos), _ptr(vel)))

    def download_fluid(self, replica=0):
        nf = self.info[replica]["nf"]
        pos = np.zeros((nf, 3)); vel = np.zeros((nf, 3)); cell = np.zeros(nf, np.int32)
        self._chk(self.lib.bw_download_fluid(self.h, replica, _ptr(pos), _ptr(vel), _ptr(cell, _ip)))
        return pos, vel, cell

    def upload_colloid(self, replica=0, **kw):
        arrs = [None if kw.get(k) is None else np.ascontiguousarray(kw[k], np.float64)
                for k in ("pos", "vel", "omega", "f", "old_f", "ra")]
        self._chk(self.lib.bw_upload_colloid(self.h, replica, *[_ptr(a) for a in arrs]))

    def download_colloid(self, replica=0) -> dict:
        nc = self.info[replica]["nc"]
        arrs = [np.zeros((nc, 3)) for _ in range(6)]
        self._chk(self.lib.bw_download_colloid(self.h, replica, *[_ptr(a) for a in arrs]))
        return dict(zip(["pos", "vel", "omega", "f", "old_f", "ra"], arrs))

    def download_displacement(self, replica=0):
        d = np.zeros((self.info[replica]["nc"], 3))
        self._chk(self.lib.bw_download_displacement(self.h, replica, _ptr(d)))
        return d

    def rng_state(self, replica=0):
        o = np.zeros(36, np.int32)
        self._chk(self.lib.bw_rng_state(self.h, replica, _ptr(o, _ip)))
        return o

    # ---- draws
    def inject_rotation(self, shift3, axes, replica=0):
        s = np.ascontiguousarray(shift3, np.float64)
        a = np.ascontiguousarray(axes, np.float64)
        self._chk(self.lib.bw_inject_rotation(self.h, replica, _ptr(s), _ptr(a)))

    def inject_reflections(self, substep, colloid, particle, vals6, replica=0):
        c = np.ascontiguousarray(colloid, np.int32)
        p = np.ascontiguousarray(particle, np.int32)
        v = np.ascontiguousarray(vals6, np.float64).reshape(-1)
        self._chk(self.lib.bw_inject_reflections(self.h, replica, int(substep), len(c), _ptr(c, _ip),
                                                 _ptr(p, _ip), _ptr(v)))

    def clear_injected(self):
        self._chk(self.lib.bw_clear_injected(self.h))

    # ---- maps
    def cell_lists(self, replica=0):
        L = self.info[replica]["L"]
        counts = np.zeros(L ** 3, np.int32)
        lst = np.zeros(self.info[replica]["nf"], np.int32)
        self._chk(self.lib.bw_get_cell_lists(self.h, replica, _ptr(counts, _ip), _ptr(lst, _ip)))
        return counts, lst

    def neigh_fl(self, replica=0):
        nc = self.info[replica]["nc"]
        counts = np.zeros(max(nc, 1), np.int32)
        tot = self._chk(self.lib.bw_get_neigh_fl(self.h, replica, _ptr(counts, _ip), None, 0))
        lst = np.zeros(max(tot, 1), np.int32)
        self._chk(self.lib.bw_get_neigh_fl(self.h, replica, _ptr(counts, _ip), _ptr(lst, _ip), len(lst)))
        return counts[:nc], lst[:tot]

    def md_list(self, replica=0):
        nc = self.info[replica]["nc"]
        counts = np.zeros(max(nc, 1), np.int32)
        lst = np.zeros(max(nc * nc, 1), np.int32)
        w = self._chk(self.lib.bw_get_md_list(self.h, replica, _ptr(counts, _ip), _ptr(lst, _ip), len(lst)))
        return counts[:nc], lst[:w]

    def shifted_cells(self, replica=0):
        o = np.zeros(self.info[replica]["nf"], np.int32)
        self._chk(self.lib.bw_get_shifted_cells(self.h, replica, _ptr(o, _ip)))
        return o

    def updown(self, replica=0) -> dict:
        """Results of the last updown_velocity (cpp/run.cpp:28-50)."""
        n = self.info[replica]["nc"]
        cnt, up = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        vel, upv = np.zeros((max(n, 1), 3)), np.zeros((max(n, 1), 3))
        self._chk(self.lib.bw_get_updown(self.h, replica, _ptr(cnt, _ip), _ptr(up, _ip), _ptr(vel), _ptr(upv)))
        return dict(cnt=cnt[:n], up_cnt=up[:n], vel=vel[:n], up_vel=upv[:n])

    def run_counts(self, replica=0):
        o = np.zeros(max(self.info[replica]["nc"], 1), np.int32)
        self._chk(self.lib.bw_get_run_counts(self.h, replica, _ptr(o, _ip)))
        return o[: self.info[replica]["nc"]]

    def audit(self, replica=0) -> dict:
        o = np.zeros(11)
        self._chk(self.lib.bw_audit(self.h, replica, _ptr(o)))
        return dict(ke_colloid=o[0], ang_ke_colloid=o[1], potential=o[2], ke_fluid=o[3], momentum=o[4:7],
                    fluid_momentum=o[7:10], momentum_scale=o[10])

    def observables_reset(self):
        self._chk(self.lib.bw_observables_reset(self.h))

    def observables(self, replica=0) -> dict:
        o = np.zeros(6)
        self._chk(self.lib.bw_observables(self.h, replica, _ptr(o)))
        return dict(lag_steps=int(o[0]), msd=o[1], orient_corr=o[2], t_fluid=o[3], t_colloid=o[4],
                    t_rot=o[5])

    def set_moments(self, on=True):
        self._chk(self.lib.bw_set_moments(self.h, int(bool(on))))

    def surrogate_attach(self, mlp, lo=None, hi=None, t_lo=None, t_hi=None):
        """Drive colloids by the network where every feature lies in [lo, hi]
        (impulses clamped to [t_lo, t_hi]); mlp=None detaches."""
        if mlp is None:
            self._chk(self.lib.bw_surrogate_attach(self.h, None, None, None, None, None))
            self._mlp = None
            return
        a = [np.ascontiguousarray(v, np.float32) for v in (lo, hi, t_lo, t_hi)]
        self._chk(self.lib.bw_surrogate_attach(self.h, mlp.h, *[v.ctypes.data_as(_fp) for v in a]))
        self._mlp = mlp  # keeps the network alive while attached

    def surrogate_data(self, replica=0):
        """features at the last step's start, its collision impulse, colloid modes"""
        nc = self.info[replica]["nc"]
        f = np.zeros((nc, 144), np.float32)
        imp = np.zeros((nc, 6))
        mode = np.zeros(nc, np.int32)
        self._chk(self.lib.bw_get_surrogate_data(self.h, replica, f.ctypes.data_as(_fp), _ptr(imp), _ptr(mode, _ip)))
        return f, imp, mode

    def surrogate_shells(self, replica=0):
        """after a fused step: shell fluid count (n_colloid) and the 15 sums
        its payment was formed from (n_colloid x 15)"""
        nc = self.info[replica]["nc"]
        cnt = np.zeros(nc, np.int32)
        sums = np.zeros((nc, 15))
        self._chk(self.lib.bw_get_surrogate_shells(self.h, replica, _ptr(cnt, _ip), _ptr(sums)))
        return cnt, sums

    def near_particles(self, replica=0) -> int:
        n = C.c_longlong(0)
        self._chk(self.lib.bw_near_particles(self.h, replica, C.byref(n)))
        return n.value

    def debug_geometry(self, a, L, ia, Li, b):
        """Device img/mod of a (float64), ia mod Li and ia / b (int32)."""
        a = np.ascontiguousarray(a, np.float64)
        ia = np.ascontiguousarray(ia, np.int32)
        n = len(a)
        assert len(ia) == n
        img, mod = np.zeros(n), np.zeros(n)
        wrap, div = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self._chk(self.lib.bw_debug_geometry(self.h, n, _ptr(a), C.c_double(L), _ptr(ia, _ip), Li, b, _ptr(img),
                                             _ptr(mod), _ptr(wrap, _ip), _ptr(div, _ip)))
        return img, mod, wrap, div

    def debug_philox(self, ctr, key):
        """Device Philox4x32-10: ctr (n x 4), key (n x 2) uint32 -> n x 4."""
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
        key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
        assert len(ctr) == len(key)
        out = np.zeros_like(ctr)
        up = C.POINTER(C.c_uint)
        self._chk(self.lib.bw_debug_philox(self.h, len(ctr), _ptr(ctr, up), _ptr(key, up), _ptr(out, up)))
        return out

    def debug_reflections(self, n, step, substep, rf, rs):
        """n device-drawn reflection tuples and the velocities they give at
        particle position rf with surface vector rs."""
        rf = np.ascontiguousarray(rf, np.float64)
        rs = np.ascontiguousarray(rs, np.float64)
        vals, vel = np.zeros((n, 6)), np.zeros((n, 3))
        self._chk(self.lib.bw_debug_reflections(self.h, n, C.c_longlong(step), substep, _ptr(rf), _ptr(rs),
                                                _ptr(vals), _ptr(vel)))
        return vals, vel

    def debug_copy_ms(self, nbytes, reps=10):
        ms = C.c_double(0)
        self._chk(self.lib.bw_debug_copy(self.h, C.c_longlong(nbytes), reps, C.byref(ms)))
        return ms.value


class MLP:
    """The surrogate network (144 -> 3 x 256 SiLU -> 12) on the GPU."""

    N_IN, N_OUT, N_T = 144, 12, 6

    def __init__(self, seed=1):
        L = lib()
        L.bw_mlp_last_error.restype = C.c_char_p
        L.bw_mlp_n_params.restype = C.c_longlong
        L.bw_mlp_train.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, _dp]
        self.lib = L
        h = C.c_void_p()
        rc = L.bw_mlp_create(int(seed), C.byref(h))
        if rc:
            msg = L.bw_mlp_last_error(h).decode() if h.value else ""
            if h.value:
                L.bw_mlp_destroy(h)
            raise BrownianError(rc, "bw_mlp_create: " + msg)
        self.h = h
        self.n_params = int(L.bw_mlp_n_params())

    def close(self):
        if getattr(self, "h", None):
            self.lib.bw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise BrownianError(rc, self.lib.bw_mlp_last_error(self.h).decode())
        return rc

    @staticmethod
    def _f32(a, cols):
        a = np.ascontiguousarray(a, np.float32)
        assert a.ndim == 2 and a.shape[1] == cols
        return a

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._chk(self.lib.bw_mlp_get_params(self.h, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._chk(self.lib.bw_mlp_set_params(self.h, p.ctypes.data_as(_fp)))

    def set_normalisation(self, x_mean, x_std, t_mean, t_std):
        a = [np.ascontiguousarray(v, np.float32) for v in (x_mean, x_std, t_mean, t_std)]
        self._chk(self.lib.bw_mlp_set_normalisation(self.h, *[v.ctypes.data_as(_fp) for v in a]))
        self.norm = a

    def forward(self, x) -> np.ndarray:
        x = self._f32(x, self.N_IN)
        y = np.zeros((len(x), self.N_OUT), np.float32)
        self._chk(self.lib.bw_mlp_forward(self.h, x.ctypes.data_as(_fp), len(x), y.ctypes.data_as(_fp)))
        return y

    def loss_grad(self, x, t):
        x = self._f32(x, self.N_IN)
        t = self._f32(t, self.N_T)
        loss = C.c_double(0)
        g = np.zeros(self.n_params, np.float32)
        self._chk(self.lib.bw_mlp_loss_grad(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(x),
                                            C.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def train(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=7) -> np.ndarray:
        x = self._f32(x, self.N_IN)
        t = self._f32(t, self.N_T)
        out = np.zeros(epochs)
        self._chk(self.lib.bw_mlp_train(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(x), int(epochs),
                                        int(batch), float(lr), int(seed), out.ctypes.data_as(_dp)))
        return out
