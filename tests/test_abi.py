"""The C-ABI libraries load and export every entry point include/zamr.h
declares; model introspection and deck validation (no device work).  CPU."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

from zamr import compile as zc
from zamr.deck import deck_from_case, write_inputs
from zamr.lib import ZamrError, ZoomyAmr, load_library

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "zamr.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(zamr_\w+)\s*\(", text, re.M)))


ALL_LIBS = [zc.lib_path(s, d, o) for s, d, o in zc.VARIANTS] + [zc.lib_path("swe2d", 2, 2, contract=False)]


def test_header_declares_the_api():
    names = declared()
    for n in ("zamr_create", "zamr_advance", "zamr_compute_dt", "zamr_evolve", "zamr_run_steps",
              "zamr_device_ptr", "zamr_write_plotfile", "zamr_create_boxlist"):
        assert n in names


@pytest.mark.parametrize("path", ALL_LIBS, ids=lambda p: p.name)
def test_library_exports_every_symbol(path):
    lib = load_library(path)
    for name in declared():
        assert hasattr(lib, name), f"{path.name} lacks {name}"


@pytest.mark.parametrize("system,dim,order,nq,na", [("swe1d", 1, 1, 3, 4), ("swe2d", 2, 1, 4, 7),
                                                     ("swe2d", 2, 2, 4, 7)])
def test_model_info_and_march_constants(system, dim, order, nq, na):
    lib = load_library(zc.lib_path(system, dim, order))
    v = [ctypes.c_int() for _ in range(5)]
    assert lib.zamr_model_info(*(ctypes.byref(x) for x in v)) == 0
    assert (v[0].value, v[1].value, v[3].value) == (nq, na, dim)
    d, k = ctypes.c_int(), ctypes.c_int()
    lib.zamr_march_constants(ctypes.byref(d), ctypes.byref(k))
    assert (d.value, k.value) == (dim, 2 * (order - 1) + 1)
    buf = ctypes.create_string_buffer(64)
    names = []
    for i in range(9):
        assert lib.zamr_param_name(i, buf, 64) == 0
        names.append(buf.value.decode())
    assert names[0] == "g" and names[5] == "wet_dry_eps"  # p(0) and p(5) of the reference driver
    g = ctypes.c_double()
    lib.zamr_default_param(0, ctypes.byref(g))
    assert g.value == 9.81


@pytest.mark.parametrize("bad,code", [
    ("bc.x_lo = Nowhere\n", 2),
    ("amr.max_level = 4\n", 5),
    ("amr.max_level = 1\namr.ref_ratio = 3\n", 5),
    ("solver.spatial_order = 3\n", 1),
    ("solver.cfl = abc\n", 1),
])
def test_deck_errors_before_any_device_work(bad, code):
    base = deck_from_case(dict(nx=16, ny=16, spatial_order=2))
    with pytest.raises(ZamrError) as e:
        ZoomyAmr(base + bad, system="swe2d", dim=2, order=2)
    assert e.value.code == code


def test_write_inputs_keys(tmp_path):
    """Same keys and values as the reference writer (build.py:97-120) for the
    drop-in call run_case makes (run_case.py:235-288)."""
    text = write_inputs(tmp_path / "inputs", 100, 2, tend=0.1, order=2, plot_dt=0.02, cfl=0.45,
                        geom={"nx": 100, "ny": 100, "prob_lo": (0.0, 0.0), "prob_hi": (1.0, 1.0)},
                        bc_sides={"x_lo": "West", "x_hi": "East", "y_lo": "South", "y_hi": "North"},
                        clamp_positivity=False, max_step=7, friction=0.03)
    kv = {}
    for line in text.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    assert kv["amr.n_cell"] == "100 100" and kv["amr.max_grid_size"] == "64"
    assert kv["bc.x_lo"] == "West" and kv["bc.y_hi"] == "North"
    assert kv["solver.spatial_order"] == "2" and kv["solver.max_step"] == "7"
    assert kv["solver.dtmax"] == "0.1" and kv["solver.dtmin"] == "1.e-7"
    assert kv["solver.clamp_positivity"] == "false" and kv["solver.positivity"] == "none"
    assert kv["params.n_m"] == "0.03"
    assert (tmp_path / "inputs").read_text() == text


def test_probe_library_exports():
    """The bandwidth-calibration library bench.py times the stage kernels'
    access pattern with is built beside the system libraries."""
    path = zc.LIBDIR / "libzamr_probe.so"
    assert path.exists(), "run __graft_entry__.build()"
    lib = ctypes.CDLL(str(path))
    for name in ("zamr_probe_pattern", "zamr_probe_copy"):
        assert hasattr(lib, name)
