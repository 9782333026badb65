"""CPU ORACLE — Python handle on the compiled oracle (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  It loads oracle/_build/liboracle_<system>.so
(built from oracle/oracle_capi.cpp + oracle/zamr_oracle.hpp, a restatement of
the reference's Source/make_rhs.H and the single-level Source/ZoomyAmr.cpp
stage/step logic) and drives it with the same case dictionary the product's
Python layer accepts.

Case keys (deck names of the reference, build.py:55-120 / ZoomyAmr.cpp:29-83):
  nx, ny, prob_lo, prob_hi, max_grid_size, spatial_order, implicit_source,
  well_balanced, clamp_positivity, positivity ("none"|"mood"), face_based_rhs,
  cfl, dtmin, dtmax, time_end, max_step, bc (x_lo, x_hi, y_lo, y_hi tag names),
  params ({name: value} overrides), cfl_dimension, cfl_degree_factor.
"""
import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
BUILD = HERE / "_build"

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

DEFAULTS = dict(
    prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=64, spatial_order=1,
    implicit_source=True, implicit_global=False, well_balanced=False, clamp_positivity=True,
    positivity="none", face_based_rhs=1, cfl=0.5, dtmin=1.0e-7, dtmax=1.0e-2,
    time_end=1.0, max_step=-1, bc=("", "", "", ""), params=None,
    cfl_dimension=None, cfl_degree_factor=None)

_libs = {}


def load(system):
    if system in _libs:
        return _libs[system]
    path = BUILD / f"liboracle_{system}.so"
    if not path.exists():
        raise FileNotFoundError(f"oracle library {path} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(path))
    for name in ("orc_nq", "orc_na", "orc_np", "orc_dim", "orc_ntags"):
        getattr(lib, name).restype = ctypes.c_int
    lib.orc_tag_name.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.orc_param_name.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.orc_default_params.argtypes = [_D]
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [_I, _D, _D]
    for name in ("orc_destroy", "orc_init_solution"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.orc_nboxes.argtypes = [ctypes.c_void_p]
    for name in ("orc_set_state", "orc_get_state", "orc_set_aux", "orc_get_aux"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, _D]
    for name in ("orc_max_inv", "orc_compute_dt", "orc_total_mass"):
        getattr(lib, name).restype = ctypes.c_double
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.orc_update_state.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.orc_advance.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    lib.orc_mood_redos.restype = ctypes.c_long
    lib.orc_mood_redos.argtypes = [ctypes.c_void_p]
    lib.orc_source_field.restype = None
    lib.orc_source_field.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_double]
    lib.orc_gmres_dense.restype = ctypes.c_long
    lib.orc_gmres_dense.argtypes = [ctypes.c_int, _D, _D, ctypes.c_double, _D]
    lib.orc_jfnk_iters.restype = None
    lib.orc_jfnk_iters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    lib.orc_set_tape.restype = None
    lib.orc_set_tape.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_evolve.restype = ctypes.c_int
    lib.orc_evolve.argtypes = [ctypes.c_void_p, _D, _D, _D, ctypes.c_int, ctypes.c_int]
    lib.orc_ge_solve.restype = ctypes.c_int
    lib.orc_ge_solve.argtypes = [ctypes.c_int, _D, _D, _D]
    _U8 = ctypes.POINTER(ctypes.c_ubyte)
    V = ctypes.c_void_p
    lib.orc_amr_create.restype = V
    lib.orc_amr_create.argtypes = [_I, _D, _D, _I, _D]
    for name in ("orc_amr_destroy", "orc_amr_init_solution", "orc_amr_init_levels"):
        getattr(lib, name).restype = None
        getattr(lib, name).argtypes = [V]
    lib.orc_amr_average_down.restype = None
    lib.orc_amr_average_down.argtypes = [V, ctypes.c_int]
    lib.orc_amr_finest.restype = ctypes.c_int
    lib.orc_amr_finest.argtypes = [V]
    lib.orc_amr_boxes.restype = ctypes.c_int
    lib.orc_amr_boxes.argtypes = [V, ctypes.c_int, _I, ctypes.c_int]
    lib.orc_amr_get.restype = None
    lib.orc_amr_get.argtypes = [V, ctypes.c_int, ctypes.c_int, _D]
    lib.orc_amr_set.restype = None
    lib.orc_amr_set.argtypes = [V, ctypes.c_int, ctypes.c_int, _D]
    lib.orc_amr_time_step.restype = None
    lib.orc_amr_time_step.argtypes = [V, ctypes.c_double, ctypes.c_int, _D]
    lib.orc_amr_total_mass.restype = ctypes.c_double
    lib.orc_amr_total_mass.argtypes = [V]
    lib.orc_amr_regrid.restype = None
    lib.orc_amr_regrid.argtypes = [V, ctypes.c_int, ctypes.c_double]
    lib.orc_amr_counts.restype = ctypes.c_long
    lib.orc_amr_counts.argtypes = [V, ctypes.c_int]
    lib.orc_amr_error_est.restype = None
    lib.orc_amr_error_est.argtypes = [V, ctypes.c_int, _U8]
    lib.orc_amr_make_grids.restype = ctypes.c_int
    lib.orc_amr_make_grids.argtypes = [V, _U8, _I, ctypes.c_int]
    lib.orc_amr_interp.restype = None
    lib.orc_amr_interp.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
    lib.orc_cluster.restype = ctypes.c_int
    lib.orc_cluster.argtypes = [_I, ctypes.c_int, ctypes.c_double, _I, ctypes.c_int]
    lib.orc_amr_evolve.restype = ctypes.c_int
    lib.orc_amr_evolve.argtypes = [V, _D, _D, _D, _I, ctypes.c_int, ctypes.c_int]
    _libs[system] = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_D)


class Oracle:
    """One single-level CPU run of `system` ("swe1d", "swe2d", "sme2")."""

    def __init__(self, system, case, threads=1, amr=None):
        self.lib = lib = load(system)
        c = dict(DEFAULTS)
        c.update(case)
        self.case = c
        self.nq, self.na, self.np_ = lib.orc_nq(), lib.orc_na(), lib.orc_np()
        self.dim = lib.orc_dim()
        self.tags = []
        buf = ctypes.create_string_buffer(64)
        for k in range(lib.orc_ntags()):
            lib.orc_tag_name(k, buf, 64)
            self.tags.append(buf.value.decode())
        p = np.zeros(self.np_)
        lib.orc_default_params(_dp(p))
        self.param_names = []
        for k in range(self.np_):
            lib.orc_param_name(k, buf, 64)
            self.param_names.append(buf.value.decode())
        for name, v in (c["params"] or {}).items():
            p[self.param_index(name)] = v
        self.params = p
        bcs = []
        for name in c["bc"]:
            if not name:
                bcs.append(0)
            elif name in self.tags:
                bcs.append(self.tags.index(name))
            else:
                raise ValueError(f"boundary tag {name!r} is not declared by this model: {self.tags}")
        order = int(c["spatial_order"])
        cfl_dim = c["cfl_dimension"] if c["cfl_dimension"] is not None else (2 if c["ny"] > 1 or self.dim == 2 else 1)
        cfl_deg = c["cfl_degree_factor"] if c["cfl_degree_factor"] is not None else 2 * (order - 1) + 1
        iset = np.array([c["nx"], c["ny"], c["max_grid_size"], order, int(bool(c["implicit_source"])),
                         int(bool(c["well_balanced"])), int(bool(c["clamp_positivity"])),
                         int(c["positivity"] == "mood"), int(c["face_based_rhs"]), int(c["max_step"]),
                         *bcs, cfl_dim, cfl_deg, threads, int(bool(c["implicit_global"]))], dtype=np.int32)
        dset = np.array([*c["prob_lo"], *c["prob_hi"], c["cfl"], c["dtmin"], c["dtmax"], c["time_end"]],
                        dtype=np.float64)
        self.nx, self.ny = int(c["nx"]), int(c["ny"])
        self.amr = amr
        if amr is None:
            self.h = lib.orc_create(iset.ctypes.data_as(_I), _dp(dset), _dp(p))
        else:
            a = dict(AMR_DEFAULTS)
            a.update(amr)
            self.amr = a
            strip = a["strip"] if a["strip"] is not None else 64 - 2 * (2 if order == 2 else 1)
            ai = np.array([a["max_level"], a["ref_ratio"], a["blocking_factor"], a["n_error_buf"],
                           a["fine_max_grid_size"], a["n_proper"], strip], dtype=np.int32)
            ad = np.array([a["grid_eff"], a["threshold"], a["b_max"]], dtype=np.float64)
            self.h = lib.orc_amr_create(iset.ctypes.data_as(_I), _dp(dset), _dp(p), ai.ctypes.data_as(_I), _dp(ad))
            if not self.h:
                raise ValueError("amr.max_level must be 0..3 and amr.ref_ratio 2 or 4")

    def param_index(self, name):
        return self.param_names.index(name)

    def close(self):
        if getattr(self, "h", None):
            (self.lib.orc_destroy if self.amr is None else self.lib.orc_amr_destroy)(self.h)
            self.h = None

    __del__ = close

    @property
    def nboxes(self):
        return self.lib.orc_nboxes(self.h)

    def init_solution(self):
        self.lib.orc_init_solution(self.h)

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(self.nq, self.ny, self.nx)
        self.lib.orc_set_state(self.h, _dp(Q))

    def get_state(self):
        Q = np.empty((self.nq, self.ny, self.nx))
        self.lib.orc_get_state(self.h, _dp(Q))
        return Q

    def set_aux(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64).reshape(self.na, self.ny, self.nx)
        self.lib.orc_set_aux(self.h, _dp(A))

    def get_aux(self):
        A = np.empty((self.na, self.ny, self.nx))
        self.lib.orc_get_aux(self.h, _dp(A))
        return A

    def max_inv(self):
        return self.lib.orc_max_inv(self.h)

    def compute_dt(self):
        return self.lib.orc_compute_dt(self.h)

    def update_state(self, time=0.0):
        self.lib.orc_update_state(self.h, time)

    def stage(self, order, time, dt):
        self.lib.orc_stage(self.h, order, time, dt)

    def advance(self, time, dt):
        self.lib.orc_advance(self.h, time, dt)

    def total_mass(self):
        return self.lib.orc_total_mass(self.h)

    def mood_redos(self):
        return self.lib.orc_mood_redos(self.h)

    def source_field(self, Q, time=0.0):
        """Model::source of every cell with aux = update_aux_variables(q, 0)."""
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(self.nq, self.ny, self.nx)
        S = np.empty_like(Q)
        self.lib.orc_source_field(self.h, _dp(Q), _dp(S), time)
        return S

    def jfnk_iters(self):
        """(Newton, GMRES) iterations of the last global source solve."""
        a, b = ctypes.c_long(), ctypes.c_long()
        self.lib.orc_jfnk_iters(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def set_tape(self, mode):
        """Limiter tape: 1 records every minmod choice of the following stages,
        2 replays the recorded choices from the start, 0 turns it off."""
        self.lib.orc_set_tape(self.h, mode)

    def evolve(self, step_limit=-1, cap=100000):
        t = np.zeros(1)
        dts = np.zeros(cap)
        ms = np.zeros(cap)
        n = self.lib.orc_evolve(self.h, _dp(t), _dp(dts), _dp(ms), cap, step_limit)
        return n, float(t[0]), dts[:min(n, cap)].copy(), ms[:min(n, cap)].copy()


AMR_DEFAULTS = dict(max_level=1, ref_ratio=2, blocking_factor=2, n_error_buf=1, fine_max_grid_size=64,
                    grid_eff=0.7, threshold=0.01, b_max=1.0e30, n_proper=1, strip=None)


class AmrOracle(Oracle):
    """A run of the AMR TimeStep over levels 0 .. max_level (at most 3); amr
    keys: max_level, ref_ratio (2 or 4), blocking_factor, n_error_buf,
    n_proper, fine_max_grid_size, strip (x cut of fine boxes into whole
    stage strips; default the device kernel's strip width), grid_eff,
    threshold (tagging.threshold), b_max (tagging.b_max)."""

    def __init__(self, system, case, threads=1, amr=None):
        super().__init__(system, case, threads=threads, amr=dict(amr or {}))

    def init_solution(self):
        self.lib.orc_amr_init_solution(self.h)

    def init_levels(self):
        self.lib.orc_amr_init_levels(self.h)

    def finest_level(self):
        return self.lib.orc_amr_finest(self.h)

    def level_shape(self, lev):
        r = self.amr["ref_ratio"] ** lev
        return self.ny * r, self.nx * r

    def boxes(self, lev):
        out = np.zeros(4 * 4096, dtype=np.int32)
        n = self.lib.orc_amr_boxes(self.h, lev, out.ctypes.data_as(_I), 4096)
        return [tuple(int(v) for v in out[4 * k:4 * k + 4]) for k in range(min(n, 4096))]

    def _get(self, lev, field, fill):
        nc = self.nq if field == 0 else self.na
        a = np.full((nc, *self.level_shape(lev)), fill)
        self.lib.orc_amr_get(self.h, lev, field, _dp(a))
        return a

    def get_state(self, lev=0, fill=np.nan):
        return self._get(lev, 0, fill)

    def get_aux(self, lev=0, fill=np.nan):
        return self._get(lev, 1, fill)

    def set_state(self, Q, lev=0):
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(self.nq, *self.level_shape(lev))
        self.lib.orc_amr_set(self.h, lev, 0, _dp(Q))

    def time_step(self, time, iteration):
        dt = np.zeros(1)
        self.lib.orc_amr_time_step(self.h, time, iteration, _dp(dt))
        return float(dt[0])

    def total_mass(self):
        return self.lib.orc_amr_total_mass(self.h)

    def regrid(self, time=0.0, lbase=0):
        self.lib.orc_amr_regrid(self.h, lbase, time)

    def average_down(self, lev=0):
        self.lib.orc_amr_average_down(self.h, lev)

    def counts(self):
        """(regrids, remakes, level-1 MOOD redos)."""
        return tuple(self.lib.orc_amr_counts(self.h, k) for k in range(3))

    def mood_redos(self, lev):
        return self.lib.orc_amr_counts(self.h, lev + 1)

    def error_est(self, lev=0):
        t = np.zeros(self.level_shape(lev), dtype=np.uint8)
        self.lib.orc_amr_error_est(self.h, lev, t.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return t

    def make_grids(self, tags):
        t = np.ascontiguousarray(tags, dtype=np.uint8)
        out = np.zeros(4 * 65536, dtype=np.int32)
        n = self.lib.orc_amr_make_grids(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                        out.ctypes.data_as(_I), 65536)
        return [tuple(int(v) for v in out[4 * k:4 * k + 4]) for k in range(n)]

    def interp(self, i, j, lev=1):
        v = np.zeros(self.nq)
        self.lib.orc_amr_interp(self.h, lev, i, j, _dp(v))
        return v

    def evolve(self, step_limit=-1, cap=100000):
        t = np.zeros(1)
        dts = np.zeros(cap)
        ms = np.zeros(cap)
        nl = np.zeros(cap, dtype=np.int32)
        n = self.lib.orc_amr_evolve(self.h, _dp(t), _dp(dts), _dp(ms), nl.ctypes.data_as(_I), cap, step_limit)
        k = min(n, cap)
        return n, float(t[0]), dts[:k].copy(), ms[:k].copy(), nl[:k].copy()


def cluster(points, eff=0.7, system="swe2d"):
    """Berger-Rigoutsos boxes (lo0, lo1, hi0, hi1) of integer points (x, y)."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 2))
    out = np.zeros(4 * (len(pts) + 1), dtype=np.int32)
    n = load(system).orc_cluster(pts.ctypes.data_as(_I), len(pts), eff, out.ctypes.data_as(_I), len(pts) + 1)
    return [tuple(int(v) for v in out[4 * k:4 * k + 4]) for k in range(n)]


def gmres_dense(A, b, rtol):
    """The global source solve's GMRES on a dense system; (x, iterations)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    its = load("swe2d").orc_gmres_dense(len(b), _dp(A), _dp(b), rtol, _dp(x))
    return x, its


def ge_solve(A, rhs):
    """make_rhs.H's elimination on a dense n x n system (n in 3, 4, 8)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros(len(rhs))
    lib = load("swe2d")
    ok = lib.orc_ge_solve(len(rhs), _dp(A), _dp(rhs), _dp(x))
    return ok == 1, x
