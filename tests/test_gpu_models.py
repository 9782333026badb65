"""Model features beyond the inviscid systems, HIP path against the oracle
through the C-ABI: the diffusive flux (make_rhs.H:166-229, on the viscous
system swe2d_visc) and the point-implicit solve's explicit fallback
(make_rhs.H:469-508, on the test system tests/systems/swe2d_fallback).
Bar: bitwise for the FMA-off build, 1e-12 relative for the FMA build, and
the closed forms the oracle is pinned by."""
import sys

import numpy as np
import pytest

from helpers import FALLBACK_DT, fallback_case, fallback_state, rel_linf, smooth_state
from oracle.pyoracle import Oracle
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu(system, case, Q0, contract=True):
    z = ZoomyAmr(deck_from_case(case), system=system, dim=2, order=case["spatial_order"], contract=contract)
    z.init_data()
    z.set_state(Q0)
    return z


def orc(system, case, Q0):
    o = Oracle(system, case)
    o.init_solution()
    o.set_state(Q0)
    return o


def visc_case(order, wb, mgs, nu=2e-3, impl=True, n=70, ny=54, bc=("wall", "", "", "wall")):
    return dict(nx=n, ny=ny, prob_lo=(0.0, 0.0), prob_hi=(1.0, ny / n), max_grid_size=mgs, spatial_order=order,
                implicit_source=impl, well_balanced=wb, clamp_positivity=True, cfl=0.45, dtmin=1e-7, dtmax=1.0,
                time_end=1e30, max_step=-1, bc=bc, params={"nu": nu, "n_m": 0.03 if impl else 0.0})


@pytest.mark.parametrize("order,wb,mgs", [(o, wb, m) for o in (1, 2) for wb in (False, True) for m in (300, 24)])
def test_viscous_stage_parity(order, wb, mgs):
    say("viscous stage", order, wb, mgs)
    n, ny = 70, 54
    Q0 = smooth_state(n, ny, seed=3 + order + 2 * wb)
    Q0[1, 10:14, 20:30] = 0.0
    Q0[2:, 10:14, 20:30] = 0.0
    c = visc_case(order, wb, mgs)
    dt = 2.5e-3
    o = orc("swe2d_visc", c, Q0)
    o.update_state(0.0)
    o.stage(order, 0.0, dt)
    ref = o.get_state()
    zb = gpu("swe2d_visc", c, Q0, contract=False)
    zb.stage(order, 0.0, dt)
    assert np.array_equal(zb.get_state(), ref), "FMA-off build must match the oracle bitwise"
    z = gpu("swe2d_visc", c, Q0)
    z.stage(order, 0.0, dt)
    assert rel_linf(z.get_state(), ref) <= 1e-12
    # the diffusion acted: the inviscid stage differs
    o0 = orc("swe2d_visc", dict(c, params={"nu": 0.0, "n_m": c["params"]["n_m"]}), Q0)
    o0.update_state(0.0)
    o0.stage(order, 0.0, dt)
    assert np.abs(o0.get_state() - ref).max() > 1e-8


@pytest.mark.parametrize("contract", [False, True])
def test_viscous_run_matches_oracle(contract):
    """30 well-balanced order-2 steps over four boxes with walls and moving
    water, graph-captured on the device: the oracle's Evolve, and mass kept."""
    say("viscous run", contract)
    n, ny = 96, 80
    Q0 = smooth_state(n, ny, seed=5)
    c = dict(visc_case(2, True, 48, n=n, ny=ny, bc=("wall",) * 4), max_step=30)
    o = orc("swe2d_visc", c, Q0)
    no, to, _, _ = o.evolve()
    z = gpu("swe2d_visc", c, Q0, contract=contract)
    done, t = z.run_steps(30)
    assert done == no == 30
    Q = z.get_state()
    if contract:
        assert rel_linf(Q, o.get_state()) <= 1e-12
        assert abs(Q[1].sum() - Q0[1].sum()) <= 1e-12 * Q0[1].sum()
    else:
        assert t == to and np.array_equal(Q, o.get_state())


def test_viscous_mode_decay_closed_form():
    """g = 0, no momentum: the depth's Neumann cosine mode decays by the
    discrete heat equation's SSP-RK2 factor (the closed form the oracle is
    pinned by), on the device clock."""
    nx, ny, lx, ly, nu = 48, 32, 1.5, 0.5, 0.01
    dx, dy = lx / nx, ly / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="xy")
    kx, ky = np.pi * 3 / lx, np.pi * 2 / ly
    mode = np.cos(kx * X) * np.cos(ky * Y)
    lam = -4 / dx ** 2 * np.sin(kx * dx / 2) ** 2 - 4 / dy ** 2 * np.sin(ky * dy / 2) ** 2
    dt = 0.2 * min(dx, dy) ** 2 / nu
    zf = nu * dt * lam
    G = 1 + zf + 0.5 * zf * zf
    c = dict(nx=nx, ny=ny, prob_lo=(0.0, 0.0), prob_hi=(lx, ly), max_grid_size=32, spatial_order=2,
             implicit_source=False, well_balanced=False, clamp_positivity=False, cfl=0.45, dtmin=1e-12, dtmax=dt,
             time_end=1e30, max_step=-1, bc=("West", "East", "South", "North"), params={"g": 0.0, "nu": nu})
    h = 1.0 + 0.1 * mode
    z0 = np.zeros_like(h)
    Q0 = np.array([z0, h, z0, z0])
    for contract in (False, True):
        z = gpu("swe2d_visc", c, Q0, contract=contract)
        assert z.run_steps(20)[0] == 20
        Q = z.get_state()
        assert np.abs(Q[1] - (1.0 + 0.1 * G ** 20 * mode)).max() <= 1e-13
        assert not Q[2:].any()


@pytest.mark.parametrize("mgs", [64, 24])
def test_implicit_fallback_through_the_stage(mgs):
    """The point-implicit solve's three outcomes in one stage (solved,
    singular -> explicit q + dt S, non-finite -> q_old), bitwise against the
    oracle in the FMA-off build and within 1e-12 in the FMA build; the
    singular band is exactly the explicit update."""
    say("fallback", mgs)
    Q0 = fallback_state()
    c = fallback_case(2, mgs=mgs)
    o = orc("swe2d_fallback", c, Q0)
    o.update_state(0.0)
    o.stage(2, 0.0, FALLBACK_DT)
    ref = o.get_state()
    zb = gpu("swe2d_fallback", c, Q0, contract=False)
    zb.stage(2, 0.0, FALLBACK_DT)
    assert np.array_equal(zb.get_state(), ref)
    z = gpu("swe2d_fallback", c, Q0)
    z.stage(2, 0.0, FALLBACK_DT)
    Q = z.get_state()
    assert np.isfinite(Q).all() and rel_linf(Q, ref) <= 1e-12
    assert np.array_equal(Q[2, 1:-1, 3:17], 2.0 * Q0[2, 1:-1, 3:17])
