"""Pins of the oracle's point-implicit source solve and its explicit fallback
(make_rhs.H:429-509) driven through a whole stage, on the test system
tests/systems/swe2d_fallback (S_m = -kappa m / h for both momenta) with
kappa = -1/dt:
  * h = 1: the momentum rows of I - dt J are zero, the pivot search finds 0
    and the stage applies the explicit q + dt S = 2 m (exact);
  * h = 4: the solve succeeds, (1 - 1/4) dm = dt S = m/4, so m -> 4m/3
    (the explicit update would give 5m/4);
  * dry (h = 0, m = 0): S and J are NaN, the explicit value too, so the cell
    keeps q_old.
Cells near the band edges also see flux; only band interiors are checked.
CPU only."""
import numpy as np
import pytest

from helpers import FALLBACK_DT, fallback_case, fallback_state
from oracle.pyoracle import Oracle


@pytest.mark.parametrize("order", [1, 2])
def test_fallback_branches(order):
    Q = fallback_state()
    o = Oracle("swe2d_fallback", fallback_case(order))
    o.init_solution()
    o.set_state(Q)
    o.update_state(0.0)
    o.stage(order, 0.0, FALLBACK_DT)
    out = o.get_state()
    inner = lambda a, b: (slice(None), slice(1, -1), slice(a + 3, b - 3))  # noqa: E731
    for sl, factor in ((inner(0, 20), 2.0), (inner(20, 40), 4.0 / 3.0), (inner(40, 60), 1.0)):
        assert np.array_equal(out[:2][sl[0], sl[1], sl[2]], Q[:2][sl[0], sl[1], sl[2]])
        m, m0 = out[2:][sl[0], sl[1], sl[2]], Q[2:][sl[0], sl[1], sl[2]]
        assert np.abs(m - factor * m0).max() <= 1e-15 * max(np.abs(m0).max(), 1.0)
    # the singular band is exactly the explicit update
    assert np.array_equal(out[2, 1:-1, 3:17], 2.0 * Q[2, 1:-1, 3:17])
    assert np.isfinite(out).all()
