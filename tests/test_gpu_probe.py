"""The bandwidth-calibration library (csrc/zamr_probe.hip) on the device:
the stage kernels' access pattern and a plain copy run and report a time;
bad arguments are rejected before any device work."""
import ctypes

import pytest

from zamr import compile as zc

pytestmark = pytest.mark.gpu


def lib():
    import torch
    assert torch.cuda.is_available()
    L = ctypes.CDLL(str(zc.LIBDIR / "libzamr_probe.so"))
    L.zamr_probe_pattern.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
    L.zamr_probe_copy.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    return L


@pytest.mark.parametrize("nq,heun", [(4, 0), (4, 1), (8, 1)])
def test_pattern_runs(nq, heun):
    print("probe pattern", nq, heun, flush=True)
    ms = ctypes.c_double(-1.0)
    assert lib().zamr_probe_pattern(300, nq, heun, 3, ctypes.byref(ms)) == 0
    assert 0.0 < ms.value < 1e3


def test_copy_runs_and_bad_arguments_are_rejected():
    print("probe copy", flush=True)
    L = lib()
    ms = ctypes.c_double(-1.0)
    assert L.zamr_probe_copy(1 << 20, 3, ctypes.byref(ms)) == 0 and 0.0 < ms.value < 1e3
    assert L.zamr_probe_copy(1000, 3, ctypes.byref(ms)) == 1       # not a multiple of 32
    assert L.zamr_probe_pattern(300, 5, 0, 3, ctypes.byref(ms)) == 1  # 4 or 8 rows only
    assert L.zamr_probe_pattern(32, 4, 0, 3, ctypes.byref(ms)) == 1   # box too small
