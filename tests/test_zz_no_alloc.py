"""zamr_compute_dt_noalloc (the allocation-free ComputeDt) against
zamr_compute_dt through the C-ABI: equal step sizes, bit for bit, on two calls
in a row, the second with a smaller input; and a box list of several
time-step groups."""
import ctypes
import sys

import numpy as np
import pytest

from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr, from_boxes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_hip = None


def hip():
    """The HIP runtime the library runs on (the test allocates the output)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
    return _hip


def device_dt(z, lev=0):
    h = hip()
    ptr = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(ptr), 8) == 0
    try:
        init = ctypes.c_double(-1.0)
        assert h.hipMemcpy(ptr, ctypes.byref(init), 8, 1) == 0  # host to device
        z.compute_dt_noalloc(ptr.value, lev=lev)
        assert h.hipDeviceSynchronize() == 0
        out = ctypes.c_double()
        assert h.hipMemcpy(ctypes.byref(out), ptr, 8, 2) == 0  # device to host
        return out.value
    finally:
        h.hipFree(ptr)


@pytest.mark.parametrize("contract", [True, False])
def test_noalloc_dt_equals_compute_dt(contract):
    for n in (384, 96):  # the second call on a smaller input
        say("noalloc dt", n, contract)
        c, Q0 = cases.swe_bumps(n=n, seed=3, steps=2, max_grid_size=128)
        z = ZoomyAmr(deck_from_case(c), system="swe2d", dim=2, order=2, contract=contract)
        z.init_data()
        z.set_state(Q0)
        got = device_dt(z)
        want = z.compute_dt()
        assert got == want and got > 0.0
        # after steps (stored aux stale on the device): still the same law
        z.run_steps(2)
        got = device_dt(z)
        assert got == z.compute_dt()
        z.close()


def test_noalloc_dt_box_list_groups():
    boxes, states, keys = cases.box_list(seed=4, n_amr=14, n_members=3, steps=2, member_range=(64, 96))
    z = from_boxes(boxes, states, keys)
    got = device_dt(z)
    assert got == z.compute_dt()
