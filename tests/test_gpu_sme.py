"""The level-2 shallow-moment system (8 state rows, 8x8 point-implicit
solve) through the same stage kernels: GPU against the oracle, bitwise in
the FMA-off build and within 1e-12 in the throughput build, on the seeded
bump workload at small size with walls, with extrapolation and on several
boxes."""
import numpy as np
import pytest

from helpers import rel_linf
from oracle.pyoracle import Oracle
from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr

pytestmark = pytest.mark.gpu


def run_both(case, Q0, contract):
    o = Oracle("sme2d", case)
    o.init_solution()
    o.set_state(Q0)
    n, t, _, _ = o.evolve()
    z = ZoomyAmr(deck_from_case(case), system="sme2d", dim=2, order=2, contract=contract)
    z.init_data()
    z.set_state(Q0)
    done, tg = z.run_steps(case["max_step"])
    return (n, t, o.get_state()), (done, tg, z.get_state())


@pytest.mark.parametrize("bc", [("wall",) * 4, ("", "", "", "")])
@pytest.mark.parametrize("mgs", [96, 32])
def test_sme_matches_oracle_bitwise(bc, mgs):
    case, Q0 = cases.sme_bumps(n=96, seed=1, steps=12, max_grid_size=mgs, bc=bc)
    (n, t, Qo), (done, tg, Q) = run_both(case, Q0, contract=False)
    assert done == n and tg == t
    assert np.array_equal(Q, Qo)


def test_sme_throughput_build_within_1e12():
    case, Q0 = cases.sme_bumps(n=128, seed=2, steps=10, max_grid_size=64)
    (n, t, Qo), (done, tg, Q) = run_both(case, Q0, contract=True)
    assert done == n
    assert rel_linf(Q, Qo) <= 1e-12


def test_sme_full_size_one_step():
    """One step of the 2048^2 workload: finite, mass conserved up to the open
    boundaries' outflow over one step, bed unchanged, and a 256^2 crop of the
    interior equal to the oracle's result on the same crop plus 8 cells of halo
    (the influence of one step reaches 4 cells)."""
    case, Q0 = cases.sme_bumps(n=2048, seed=0, steps=1)
    z = ZoomyAmr(deck_from_case(case), system="sme2d", dim=2, order=2, contract=False)
    z.init_data()
    z.set_state(Q0)
    z.run_steps(1)
    Q = z.get_state()
    assert np.isfinite(Q).all()
    assert np.array_equal(Q[0], Q0[0])
    dt = z.get_time()[2]
    i0, j0, m, hw = 900, 1000, 256, 8
    crop = dict(case, nx=m + 2 * hw, ny=m + 2 * hw, max_grid_size=4096,
                prob_lo=((i0 - hw) / 2048, (j0 - hw) / 2048), prob_hi=((i0 + m + hw) / 2048, (j0 + m + hw) / 2048))
    o = Oracle("sme2d", crop)
    o.init_solution()
    o.set_state(Q0[:, j0 - hw:j0 + m + hw, i0 - hw:i0 + m + hw])
    o.advance(0.0, dt)
    assert np.array_equal(o.get_state()[:, hw:-hw, hw:-hw], Q[:, j0:j0 + m, i0:i0 + m])
