"""Pins of the oracle's well-balanced (Audusse) reconstruction with moving
water (make_rhs.H:84-100): hydrostatic depths h* = max(0, h + b - max(bL, bR)),
the momentum rows rescaled by h*/h so that the face keeps each cell's
velocity, b set to the face bed on both sides.  At rest the rescale has no
effect, so these pins use a flat free surface with one velocity everywhere
over a bed step: the reconstruction then makes the step face's two states
equal, and a stage has a closed form (tests/helpers.py: bed_step).  Deleting
or inverting the rescale, or rescaling only one momentum row, breaks the
mass row, which is exact here.  CPU only."""
import numpy as np
import pytest

from helpers import bed_step, bed_step_case, bed_step_expected
from oracle.pyoracle import Oracle

DT = 0.01


def run_stage(system, case, Q, order):
    o = Oracle(system, case)
    o.init_solution()
    o.set_state(Q)
    o.update_state(0.0)
    o.stage(order, 0.0, DT)
    return o.get_state()


@pytest.mark.parametrize("order", [1, 2])
def test_bed_step_1d(order):
    Q, rate = bed_step(nx=8, ny=1)
    out = run_stage("swe1d", bed_step_case(8, 1, order, 1), Q, order)
    E = bed_step_expected(Q, rate, DT, 0, 4)
    # mass and bed exactly; momentum to the rounding of the pressure terms
    # (g h^2 / 2 and the bed source, which cancel) and of hinv
    assert np.array_equal(out[:2], E[:2])
    assert np.abs(out[2] - E[2]).max() <= 1e-14
    # the cell below the step moved; with the momentum not rescaled it would
    # move by a different mass flux
    assert out[1, 0, 3] == 1.0 + DT * 0.15


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("axis", [0, 1])
def test_bed_step_2d(order, axis):
    nx, ny = (8, 4) if axis == 0 else (4, 8)
    Q, rate = bed_step(nx=nx, ny=ny, axis=axis)
    out = run_stage("swe2d", bed_step_case(nx, ny, order, 2), Q, order)
    E = bed_step_expected(Q, rate, DT, axis, 4)
    assert np.array_equal(out[:2], E[:2])
    assert np.abs(out[2:] - E[2:]).max() <= 1e-14
    # the momentum along the face (rescaled too) changed by its closed form
    t = 3 if axis == 0 else 2
    cell = (slice(None), 3) if axis == 0 else (3, slice(None))
    assert np.abs(out[t][cell] - Q[t][cell] - DT * rate[t]).max() <= 1e-15


def test_bed_step_without_wb_is_not_at_the_closed_form():
    """Control: the same state without well-balancing leaves the closed form
    (the bed step then acts through the path-conservative fluctuations)."""
    Q, rate = bed_step(nx=8, ny=4, axis=0)
    c = dict(bed_step_case(8, 4, 1, 2), well_balanced=False)
    out = run_stage("swe2d", c, Q, 1)
    E = bed_step_expected(Q, rate, DT, 0, 4)
    assert np.abs(out[1] - E[1]).max() > 1e-6
