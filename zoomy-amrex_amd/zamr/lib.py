"""ctypes binding of include/zamr.h and a ZoomyAmr-shaped Python handle.

The product path is the compiled HIP library only: if the variant's .so is
missing this module raises; there is no CPU fallback.
"""
import ctypes
import pathlib

import numpy as np

from . import compile as _compile

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p

ERRORS = {1: "EINVAL", 2: "ETAG", 3: "ERASTER", 4: "EHIP", 5: "EUNSUP", 6: "EIO"}


class ZamrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"zamr error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class ZamrBox(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_int * 2), ("hi", ctypes.c_int * 2), ("dx", ctypes.c_double * 2),
                ("prob_lo", ctypes.c_double * 2), ("group", ctypes.c_int), ("bc", ctypes.c_char_p * 4)]


_cache = {}


def load_library(path):
    path = str(pathlib.Path(path).resolve())
    if path in _cache:
        return _cache[path]
    if not pathlib.Path(path).exists():
        raise FileNotFoundError(f"zamr library {path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    sig = {
        "zamr_model_info": [_I, _I, _I, _I, _I],
        "zamr_tag_name": [ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
        "zamr_param_name": [ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
        "zamr_default_param": [ctypes.c_int, _D],
        "zamr_march_constants": [_I, _I],
        "zamr_create": [ctypes.c_char_p, ctypes.POINTER(_VP)],
        "zamr_create_boxlist": [ctypes.POINTER(ZamrBox), ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(_VP)],
        "zamr_destroy": [_VP],
        "zamr_init_data": [_VP],
        "zamr_level_shape": [_VP, ctypes.c_int, _I, _I, _I],
        "zamr_set_state": [_VP, ctypes.c_int, ctypes.c_int, _D, ctypes.c_size_t],
        "zamr_get_state": [_VP, ctypes.c_int, ctypes.c_int, _D, ctypes.c_size_t],
        "zamr_compute_dt": [_VP, ctypes.c_int, _D],
        "zamr_compute_dt_noalloc": [_VP, ctypes.c_int, _VP],
        "zamr_update_state": [_VP, ctypes.c_int, ctypes.c_double],
        "zamr_fill_patch": [_VP, ctypes.c_int, ctypes.c_double],
        "zamr_stage": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double],
        "zamr_advance": [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double],
        "zamr_time_step": [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_int, _D],
        "zamr_run_steps": [_VP, ctypes.c_int, _I, _D],
        "zamr_evolve": [_VP],
        "zamr_total_mass": [_VP, _D],
        "zamr_get_time": [_VP, ctypes.c_int, _D, _I, _D],
        "zamr_reset_clock": [_VP],
        "zamr_write_plotfile": [_VP, ctypes.c_char_p, ctypes.c_int, ctypes.c_double],
        "zamr_device_ptr": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP),
                            ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), _I,
                            ctypes.POINTER(_VP)],
        "zamr_set_stream": [_VP, _VP],
        "zamr_set_tuning": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "zamr_set_graph": [_VP, ctypes.c_int],
        "zamr_autotune": [_VP, ctypes.c_int, _VP, _VP, _VP],
        "zamr_set_launch_shape": [_VP, _VP, _VP],
        "zamr_set_profiling": [_VP, ctypes.c_int],
        "zamr_kernel_times": [_VP, _D, _I],
        "zamr_dt_history": [_VP, ctypes.c_int, _D, _D, _I],
        "zamr_mood_stats": [_VP, ctypes.c_int, _I, ctypes.POINTER(ctypes.c_longlong)],
        "zamr_finest_level": [_VP, _I],
        "zamr_jfnk_stats": [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                            ctypes.POINTER(ctypes.c_longlong)],
        "zamr_level_boxes": [_VP, ctypes.c_int, _I, ctypes.c_int, _I],
        "zamr_init_levels": [_VP],
        "zamr_regrid": [_VP, ctypes.c_double],
        "zamr_error_est": [_VP, ctypes.POINTER(ctypes.c_ubyte), ctypes.c_size_t],
        "zamr_amr_stats": [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), _D],
        "zamr_adj_setup": [_VP, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _I, _I],
        "zamr_adj_forward": [_VP, _D],
        "zamr_adj_backward": [_VP, _D],
        "zamr_adj_gradient": [_VP, _D, ctypes.c_size_t, _D, ctypes.c_size_t],
        "zamr_adj_freeze_limiter": [_VP, ctypes.c_int],
        "zamr_adj_device_io": [_VP, ctypes.c_int],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.zamr_last_error.argtypes = [_VP]
    lib.zamr_last_error.restype = ctypes.c_char_p
    _cache[path] = lib
    return lib


def library_for(system, dim, order, contract=True):
    return load_library(_compile.lib_path(system, dim, order, contract))


def _dp(a):
    return a.ctypes.data_as(_D)


class ZoomyAmr:
    """One context of a system library, mirroring the reference driver class
    (Source/ZoomyAmr.H:14-166): InitData, ComputeDt, Advance, TimeStep,
    Evolve, ComputeTotalMass, WritePlotFile, plus state I/O."""

    def __init__(self, inputs_text, system="swe2d", dim=None, order=None, contract=True, lib=None, boxes=None):
        if lib is None:
            if order is None:
                order = _deck_int(inputs_text, "solver.spatial_order", 1)
            if dim is None:
                dim = 1 if system.endswith("1d") else 2
            lib = library_for(system, dim, order, contract)
        self.lib = lib
        self.ctx = _VP()
        if boxes is None:
            rc = lib.zamr_create(inputs_text.encode(), ctypes.byref(self.ctx))
        else:
            arr = (ZamrBox * len(boxes))(*boxes)
            self._boxes_keepalive = (arr, boxes)
            rc = lib.zamr_create_boxlist(arr, len(boxes), inputs_text.encode(), ctypes.byref(self.ctx))
        if rc != 0:
            msg = lib.zamr_last_error(self.ctx).decode() if self.ctx else "create failed"
            self.close()
            raise ZamrError(rc, msg)
        nq, na, np_, d, nt = (ctypes.c_int() for _ in range(5))
        lib.zamr_model_info(*(ctypes.byref(v) for v in (nq, na, np_, d, nt)))
        self.nq, self.na, self.nparams, self.dim = nq.value, na.value, np_.value, d.value
        buf = ctypes.create_string_buffer(128)
        self.tags = []
        for k in range(nt.value):
            lib.zamr_tag_name(k, buf, 128)
            self.tags.append(buf.value.decode())
        self.param_names = []
        for k in range(self.nparams):
            lib.zamr_param_name(k, buf, 128)
            self.param_names.append(buf.value.decode())

    def _check(self, rc):
        if rc != 0:
            raise ZamrError(rc, self.lib.zamr_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.zamr_destroy(self.ctx)
        self.ctx = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, lev=0):
        nx, ny, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.zamr_level_shape(self.ctx, lev, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nb)))
        return nx.value, ny.value, nb.value

    def init_data(self):
        self._check(self.lib.zamr_init_data(self.ctx))

    def set_state(self, arr, field=0, lev=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self.lib.zamr_set_state(self.ctx, lev, field, _dp(a), a.size))

    def get_state(self, field=0, lev=0, fill=np.nan):
        """Dense (ncomp, ny, nx) copy of level `lev`; on a refined level the
        cells no box covers hold `fill`."""
        nx, ny, _ = self.shape(lev)
        nc = self.nq if field == 0 else self.na
        a = np.full((nc, ny, nx), fill) if lev > 0 else np.empty((nc, ny, nx))
        self._check(self.lib.zamr_get_state(self.ctx, lev, field, _dp(a), a.size))
        return a

    def jfnk_stats(self):
        """(global source solves, Newton iterations, GMRES iterations) so far."""
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.zamr_jfnk_stats(self.ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    # ---- AMR hierarchy ----
    def finest_level(self):
        f = ctypes.c_int()
        self._check(self.lib.zamr_finest_level(self.ctx, ctypes.byref(f)))
        return f.value

    def level_boxes(self, lev):
        """Boxes of level `lev` as (lo0, lo1, hi0, hi1) tuples."""
        n = ctypes.c_int()
        self._check(self.lib.zamr_level_boxes(self.ctx, lev, None, 0, ctypes.byref(n)))
        out = np.zeros(4 * max(n.value, 1), dtype=np.int32)
        self._check(self.lib.zamr_level_boxes(self.ctx, lev, out.ctypes.data_as(_I), n.value, ctypes.byref(n)))
        return [tuple(int(v) for v in out[4 * k:4 * k + 4]) for k in range(n.value)]

    def init_levels(self):
        self._check(self.lib.zamr_init_levels(self.ctx))

    def regrid(self, time=0.0):
        self._check(self.lib.zamr_regrid(self.ctx, time))

    def error_est(self):
        nx, ny, _ = self.shape(0)
        t = np.zeros((ny, nx), dtype=np.uint8)
        self._check(self.lib.zamr_error_est(self.ctx, t.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), t.size))
        return t

    def amr_stats(self, with_time=False):
        """(regrids, levels remade) since creation, and with with_time the host
        milliseconds spent regridding."""
        r, m, t = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        self._check(self.lib.zamr_amr_stats(self.ctx, ctypes.byref(r), ctypes.byref(m), ctypes.byref(t)))
        return (r.value, m.value, t.value) if with_time else (r.value, m.value)

    def compute_dt(self, lev=0):
        dt = ctypes.c_double()
        self._check(self.lib.zamr_compute_dt(self.ctx, lev, ctypes.byref(dt)))
        return dt.value

    def compute_dt_noalloc(self, dt_device_ptr, lev=0):
        """zamr_compute_dt_noalloc: the step size written to the device double
        at dt_device_ptr (an integer address) on the context's stream."""
        self._check(self.lib.zamr_compute_dt_noalloc(self.ctx, lev, ctypes.c_void_p(dt_device_ptr)))

    def update_state(self, time=0.0, lev=0):
        self._check(self.lib.zamr_update_state(self.ctx, lev, time))

    def stage(self, order, time, dt, lev=0):
        self._check(self.lib.zamr_stage(self.ctx, lev, order, time, dt))

    def advance(self, time, dt, lev=0):
        self._check(self.lib.zamr_advance(self.ctx, lev, time, dt))

    def time_step(self, time, iteration=0, lev=0):
        dt = ctypes.c_double()
        self._check(self.lib.zamr_time_step(self.ctx, lev, time, iteration, ctypes.byref(dt)))
        return dt.value

    def run_steps(self, n):
        done, t = ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.zamr_run_steps(self.ctx, n, ctypes.byref(done), ctypes.byref(t)))
        return done.value, t.value

    def evolve(self):
        self._check(self.lib.zamr_evolve(self.ctx))

    def total_mass(self):
        m = ctypes.c_double()
        self._check(self.lib.zamr_total_mass(self.ctx, ctypes.byref(m)))
        return m.value

    def get_time(self, group=0):
        t, s, dt = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.zamr_get_time(self.ctx, group, ctypes.byref(t), ctypes.byref(s), ctypes.byref(dt)))
        return t.value, s.value, dt.value

    def reset_clock(self):
        self._check(self.lib.zamr_reset_clock(self.ctx))

    def write_plotfile(self, directory, step, time):
        self._check(self.lib.zamr_write_plotfile(self.ctx, str(directory).encode(), step, time))

    def device_ptr(self, field=0, box=0, lev=0):
        p, rs, cs, ng, st = _VP(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int(), _VP()
        self._check(self.lib.zamr_device_ptr(self.ctx, lev, box, field, ctypes.byref(p), ctypes.byref(rs),
                                             ctypes.byref(cs), ctypes.byref(ng), ctypes.byref(st)))
        return p.value, rs.value, cs.value, ng.value, st.value

    def set_stream(self, stream_ptr):
        self._check(self.lib.zamr_set_stream(self.ctx, _VP(stream_ptr)))

    def set_tuning(self, rows=0, per_box_launch=0, xcd_remap=1):
        self._check(self.lib.zamr_set_tuning(self.ctx, rows, per_box_launch, xcd_remap))

    def set_graph(self, on=True):
        self._check(self.lib.zamr_set_graph(self.ctx, 1 if on else 0))

    def autotune(self, reps=10):
        """Measure the stage kernel's launch shape (include/zamr.h
        zamr_autotune); returns (rows per size class, grid blocks,
        (ms before, ms after))."""
        rows = np.zeros(8, dtype=np.int32)
        g = np.zeros(2, dtype=np.int32)
        ms = np.zeros(2)
        self._check(self.lib.zamr_autotune(self.ctx, reps, rows.ctypes.data, g.ctypes.data, _dp(ms)))
        return rows.tolist(), g.tolist(), (float(ms[0]), float(ms[1]))

    def set_launch_shape(self, class_rows, grid_blocks=(0, 0)):
        rows = np.ascontiguousarray(class_rows, dtype=np.int32)
        g = np.ascontiguousarray(grid_blocks, dtype=np.int32)
        if rows.shape != (8,) or g.shape != (2,):
            raise ValueError("class_rows has 8 entries, grid_blocks 2")
        self._check(self.lib.zamr_set_launch_shape(self.ctx, rows.ctypes.data, g.ctypes.data))

    def set_profiling(self, on=True):
        self._check(self.lib.zamr_set_profiling(self.ctx, 1 if on else 0))

    def kernel_times(self):
        out = np.zeros(3)
        n = ctypes.c_int()
        self._check(self.lib.zamr_kernel_times(self.ctx, _dp(out), ctypes.byref(n)))
        return out, n.value

    def mood_stats(self, group=0):
        """(troubled cells of the last order-2 candidate, steps redone at order 1)."""
        t, r = ctypes.c_int(), ctypes.c_longlong()
        self._check(self.lib.zamr_mood_stats(self.ctx, group, ctypes.byref(t), ctypes.byref(r)))
        return t.value, r.value

    def dt_history(self, cap=1 << 16):
        """(dts, times) of the steps taken on the device clock since the last reset."""
        d, t, n = np.zeros(cap), np.zeros(cap), ctypes.c_int()
        self._check(self.lib.zamr_dt_history(self.ctx, cap, _dp(d), _dp(t), ctypes.byref(n)))
        return d[:n.value].copy(), t[:n.value].copy()

    # ---- adjoint (systems with an Adjoint.H; see include/zamr.h) ----
    def adj_setup(self, dts, gauges_i, gauges_j, seg_steps=20):
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        gi = np.ascontiguousarray(gauges_i, dtype=np.int32)
        gj = np.ascontiguousarray(gauges_j, dtype=np.int32)
        self._adj = (len(dts), len(gi))
        self._check(self.lib.zamr_adj_setup(self.ctx, len(dts), seg_steps, _dp(dts), len(gi),
                                            gi.ctypes.data_as(_I), gj.ctypes.data_as(_I)))

    def _adj_io(self, device):
        """Put the adjoint calls in device or host I/O mode (zamr_adj_device_io)."""
        if getattr(self, "_adj_device", False) != device:
            self._check(self.lib.zamr_adj_device_io(self.ctx, 1 if device else 0))
            self._adj_device = device

    def adj_forward(self):
        """Frozen-dt forward from the current state; returns the gauge depths
        after every step, shape (nsteps, ngauges)."""
        self._adj_io(False)
        ns, ng = self._adj
        gh = np.zeros((ns, ng))
        self._check(self.lib.zamr_adj_forward(self.ctx, _dp(gh)))
        return gh

    def adj_backward(self, gbar):
        ns, ng = self._adj
        gbar = np.ascontiguousarray(gbar, dtype=np.float64).reshape(ns, ng)
        self._adj_io(False)
        self._check(self.lib.zamr_adj_backward(self.ctx, _dp(gbar)))

    def adj_freeze_limiter(self, on=True):
        self._check(self.lib.zamr_adj_freeze_limiter(self.ctx, 1 if on else 0))

    def adj_forward_device(self, gauge_h_ptr):
        """Frozen-dt forward; gauge depths (nsteps, ngauges) to the device
        address gauge_h_ptr.  The adj_*_device calls queue on the context's
        stream and return without waiting."""
        self._adj_io(True)
        self._check(self.lib.zamr_adj_forward(self.ctx, ctypes.cast(_VP(gauge_h_ptr), _D)))

    def adj_backward_device(self, gbar_ptr):
        self._adj_io(True)
        self._check(self.lib.zamr_adj_backward(self.ctx, ctypes.cast(_VP(gbar_ptr), _D)))

    def adj_gradient_device(self, qbar_ptr, auxbar_ptr):
        """The gradients into dense device arrays (n_q, ny, nx) and (ny, nx)."""
        self._adj_io(True)
        nx, ny, _ = self.shape()
        self._check(self.lib.zamr_adj_gradient(self.ctx, ctypes.cast(_VP(qbar_ptr), _D), self.nq * nx * ny,
                                               ctypes.cast(_VP(auxbar_ptr), _D), nx * ny))

    def adj_gradient(self):
        """(adjoint of the initial state (n_q, ny, nx), adjoint of the friction aux row (ny, nx))."""
        self._adj_io(False)
        nx, ny, _ = self.shape()
        qb = np.zeros((self.nq, ny, nx))
        fb = np.zeros((ny, nx))
        self._check(self.lib.zamr_adj_gradient(self.ctx, _dp(qb), qb.size, _dp(fb), fb.size))
        return qb, fb


def solver_deck(keys):
    """solver.* / params.* deck lines from a dictionary of solver keys."""
    lines = []
    for k in ("time_end", "max_step", "cfl", "dtmin", "dtmax", "spatial_order"):
        if k in keys:
            lines.append(f"solver.{k} = {keys[k]!r}")
    for k in ("implicit_source", "well_balanced", "clamp_positivity"):
        if k in keys:
            lines.append(f"solver.{k} = {'true' if keys[k] else 'false'}")
    if "positivity" in keys:
        lines.append(f"solver.positivity = {keys['positivity']}")
    for k, v in (keys.get("params") or {}).items():
        lines.append(f"params.{k} = {v!r}")
    return "\n".join(lines) + "\n"


def from_boxes(boxes, states, keys, system="swe2d", contract=True, lib=None):
    """A batched context over independent boxes (see zamr_create_boxlist):
    boxes are dicts with lo, hi, dx, prob_lo, group and optional bc (four tag
    names); states the (n_q, ny, nx) initial state of each box."""
    zb = []
    keep = []
    for b in boxes:
        z = ZamrBox()
        z.lo[:] = list(b["lo"])
        z.hi[:] = list(b["hi"])
        z.dx[:] = list(b["dx"])
        z.prob_lo[:] = list(b["prob_lo"])
        z.group = int(b["group"])
        names = [s.encode() for s in b.get("bc", ("", "", "", ""))]
        keep.append(names)
        for k in range(4):
            z.bc[k] = names[k]
        zb.append(z)
    order = int(keys.get("spatial_order", 2))
    ctx = ZoomyAmr(solver_deck(keys), system=system, dim=2, order=order, contract=contract, lib=lib, boxes=zb)
    ctx._bc_keepalive = keep
    ctx.init_data()
    for k, s in enumerate(states):
        ctx.set_state(s, lev=k)
    return ctx


def _deck_int(text, key, default):
    for line in text.splitlines():
        line = line.split("#", 1)[0]
        if "=" in line and line.split("=", 1)[0].strip() == key:
            return int(float(line.split("=", 1)[1].split()[0]))
    return default
