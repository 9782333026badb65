"""The process seam end to end, as the reference's own tests drive it
(tests/conftest.py:73-111 march -> run_case.py:346-426): the initial
condition sampled on the cell centres into rasters, the inputs deck written by
write_inputs, the drop-in executable `main_<variant>.ex inputs` run as a
subprocess, the last plotfile read back (deliverable.py:33-129), and the
result held to the reference's fixture under its np.allclose rule
(tests/refs.py:33) and to the asserts of the reference test itself."""
import numpy as np
import pytest

from helpers import fixture, stoker_ic
from zamr.run import read_plotfile, run_case

pytestmark = pytest.mark.gpu


def gaussian_pulse_2d(X):  # reference tests/test_swe_2d.py: h = 1 + 0.1 exp(-(x^2 + y^2) / 0.05)
    x, y = X[0], X[1]
    h = 1.0 + 0.1 * np.exp(-(x ** 2 + y ** 2) / 0.05)
    z = 0.0 * h
    return np.array([z, h, z, z])


def test_swe_2d_pulse_through_the_executable(tmp_path):
    print("[dropin] swe_2d pulse via main_swe2d_d2o1.ex", flush=True)
    settings = {"dimension": 2, "domain": [-1.0, 1.0, -1.0, 1.0], "n_cells": [32, 32], "spatial_order": 1,
                "output_snapshots": 1, "max_level": 0, "time_end": 0.1, "cfl": 0.9}
    plt = run_case("swe2d", settings, tmp_path, ic=gaussian_pulse_2d)
    arr, names, t = read_plotfile(plt)
    assert names[:4] == ["var0", "var1", "var2", "var3"] and names[4] == "aux0"
    assert abs(t - 0.1) < 1e-12
    Q, A = arr[:4], arr[4:]
    # the reference test's asserts (tests/test_swe_2d.py)
    assert np.isfinite(Q).all()
    assert Q[1].min() > 0.0
    assert np.abs(Q[2]).max() > 0.0 and np.abs(Q[3]).max() > 0.0
    Qr, Ar = fixture("swe_2d")
    assert np.allclose(Q, Qr) and np.allclose(A, Ar)


def test_stoker_1d_through_the_executable(tmp_path):
    print("[dropin] stoker_wet via main_swe1d_d1o1.ex", flush=True)
    settings = {"dimension": 1, "domain": [0.0, 10.0], "n_cells": 100, "spatial_order": 1,
                "output_snapshots": 1, "max_level": 0, "time_end": 1.0, "cfl": 0.9}
    plt = run_case("swe1d", settings, tmp_path, ic=lambda X: stoker_ic(X[0]))
    arr, names, t = read_plotfile(plt)
    Q, A = arr[:3], arr[3:]
    assert abs(t - 1.0) < 1e-12
    # the reference test's asserts (tests/test_swe_stoker_wet.py:36-41)
    assert np.isfinite(Q).all() and Q[1].min() > 0.0 and np.abs(Q[2]).max() > 0.0
    Qr, Ar = fixture("stoker_wet")
    assert np.allclose(Q, Qr.reshape(Q.shape)) and np.allclose(A, Ar.reshape(A.shape))
