"""solver.implicit_global = true (ZoomyAmr.cpp:647-651, Source/imex_solver.H:
102-160): after each stage's flux update the source is solved globally by
Newton + GMRES with a finite-difference Jacobian-vector product.  Through the
C-ABI, against the oracle's restatement and against closed forms.

Both sides stop Newton at |F| < 1e-9 |F_0| (2-norm over the level) and GMRES
at 1e-4 relative; their reductions sum in different orders, so the two solves
agree to that tolerance, not bit for bit: the bar is 1e-10 of the state's
scale, far below the O(dt^2) gap to the point-implicit step."""
import numpy as np
import pytest

from helpers import manning_backward_euler, manning_restart_case, rel_linf
from oracle.pyoracle import Oracle
from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr

pytestmark = pytest.mark.gpu


def _case(n=48, nm=0.08, steps=6, order=2, implicit_global=True, mgs=24):
    case, Q0 = cases.swe_bumps(n=n, seed=3, nbumps=4, steps=steps, manning=nm)
    case.update(implicit_global=implicit_global, spatial_order=order, max_grid_size=mgs, well_balanced=False,
                clamp_positivity=False)
    Q0 = Q0.copy()
    x = np.linspace(0, 1, n)
    Q0[2] = 0.3 * Q0[1] * np.sin(2 * np.pi * x)[None, :]
    Q0[3] = -0.2 * Q0[1] * np.cos(2 * np.pi * x)[:, None]
    return case, Q0


def _gpu(case, Q0, contract):
    z = ZoomyAmr(deck_from_case(case), system="swe2d", dim=2, order=case["spatial_order"], contract=contract)
    z.init_data()
    z.set_state(Q0)
    return z


@pytest.mark.parametrize("contract,order", [(False, 2), (True, 2), (False, 1)])
def test_global_source_steps_match_oracle(contract, order):
    case, Q0 = _case(order=order)
    print(f"[jfnk] steps vs oracle contract={contract} order={order}", flush=True)
    o = Oracle("swe2d", case)
    o.set_state(Q0)
    n, t, dts, _ = o.evolve(step_limit=6)
    z = _gpu(case, Q0, contract)
    done, tg = z.run_steps(6)
    assert done == n == 6
    hd, _ = z.dt_history()
    assert np.allclose(hd[:6], dts, rtol=1e-10, atol=0)
    Q = z.get_state()
    assert rel_linf(Q, o.get_state()) <= 1e-10
    # not the point-implicit run
    p = Oracle("swe2d", dict(case, implicit_global=False))
    p.set_state(Q0)
    p.evolve(step_limit=6)
    assert rel_linf(Q, p.get_state()) > 1e3 * max(rel_linf(Q, o.get_state()), 1e-16)


def test_stage_is_the_closed_form_backward_euler():
    """One stage through zamr_stage: the momentum is the exact backward-Euler
    solution of the Manning source per cell (|m| (1 + dt c |m|) = |m_exp|)."""
    g, nm, eps, dt = 9.81, 0.08, 1e-4, 2e-3
    case, Q0 = _case(nm=nm)
    print("[jfnk] closed-form stage", flush=True)
    f = _gpu(dict(case, params={"n_m": 0.0}), Q0, False)
    f.stage(2, 0.0, dt)
    Qexp = f.get_state()
    z = _gpu(case, Q0, False)
    z.stage(2, 0.0, dt)
    Q = z.get_state()
    h = Qexp[1]
    hinv = np.sqrt(2.0) * h / np.sqrt(h ** 4 + np.maximum(h ** 4, eps ** 4))
    c = g * nm * nm * hinv * np.maximum(h, eps) ** (-4.0 / 3.0)
    s = np.hypot(Qexp[2], Qexp[3])
    sc = np.where(s > 0, 2.0 / (1.0 + np.sqrt(1.0 + 4.0 * dt * c * s)), 1.0)
    assert np.array_equal(Q[:2], Qexp[:2])
    err = max(np.abs(Q[2] - Qexp[2] * sc).max(), np.abs(Q[3] - Qexp[3] * sc).max())
    corr = np.sqrt(np.sum((Qexp[2] - Qexp[2] * sc) ** 2) + np.sum((Qexp[3] - Qexp[3] * sc) ** 2))
    assert err <= 1e-9 * corr, (err, corr)


def test_stage_with_gmres_restarts_is_the_closed_form_backward_euler():
    """The closed form on a state whose depths fall from 1 m to 1 mm under
    strong friction: GMRES takes more than 30 steps per Newton iteration on
    average, so the device's restart path (the cycle's update with partial
    sums, the restart residual, a new cycle from it, the Q update after a
    restarted cycle, and rounds the host queues after reading the state
    back) is exercised, and the result is still the exact backward-Euler
    step to Newton's tolerance."""
    nm, dt = 3.0, 2e-3
    case, Q0 = manning_restart_case(nm=nm)
    for contract in (False, True):
        print(f"[jfnk] closed-form stage with restarts contract={contract}", flush=True)
        f = _gpu(dict(case, params={"n_m": 0.0}), Q0, contract)
        f.stage(2, 0.0, dt)
        Qexp = f.get_state()
        z = _gpu(case, Q0, contract)
        z.stage(2, 0.0, dt)
        Q = z.get_state()
        solves, newton, gm = z.jfnk_stats()
        assert solves == 1 and newton < 12 and gm > 30 * newton, (solves, newton, gm)
        be = manning_backward_euler(Qexp, nm, dt)
        assert np.array_equal(Q[:2], Qexp[:2])
        err = np.abs(Q[2:] - be[2:]).max()
        assert err <= 1e-9 * np.linalg.norm(Qexp[2:] - be[2:]), (contract, err)


def test_no_source_equals_point_implicit_bitwise():
    case, Q0 = _case(nm=0.0)
    print("[jfnk] no source", flush=True)
    a = _gpu(case, Q0, False)
    a.run_steps(6)
    b = _gpu(dict(case, implicit_global=False), Q0, False)
    b.run_steps(6)
    assert np.array_equal(a.get_state(), b.get_state())


def test_global_source_with_mood_matches_oracle():
    """MOOD with the global solve: every stage of the order-2 candidate and of
    the order-1 redo solves the source globally (ZoomyAmr.cpp:640-651 inside
    do_stage, :699-720); a collapsing wet column at Courant 2.5 with Manning
    friction redoes steps.  Within 1e-10 of the oracle."""
    n = 40
    c = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=20, spatial_order=2,
             implicit_source=True, implicit_global=True, well_balanced=False, clamp_positivity=False,
             positivity="mood", cfl=2.5, dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=5, bc=("wall",) * 4,
             params={"n_m": 0.05})
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.04, 1.0, 0.0)
    z0 = np.zeros_like(h)
    Q0 = np.stack([z0, h, 0.2 * h, z0])
    print("[jfnk] MOOD + global source", flush=True)
    o = Oracle("swe2d", c)
    o.set_state(Q0)
    n_o, _, dts, _ = o.evolve(step_limit=5)
    assert o.mood_redos() > 0
    z = _gpu(c, Q0, False)
    done, _ = z.run_steps(5)
    assert done == n_o == 5
    assert z.mood_stats()[1] == o.mood_redos()
    assert np.allclose(z.dt_history()[0][:5], dts, rtol=1e-10, atol=0)
    assert rel_linf(z.get_state(), o.get_state()) <= 1e-10


@pytest.mark.parametrize("levels", [1, 2])
def test_global_source_on_amr_levels_matches_oracle(levels):
    """The global solve on every level of an AMR hierarchy (each level's Advance
    solves its own valid cells, ZoomyAmr.cpp:640-645): within 1e-10 of the
    oracle's hierarchy, the grids identical."""
    from oracle.pyoracle import AmrOracle
    n = 32
    case, Q0 = _case(n=n, steps=4, mgs=16)
    case = dict(case, bc=("wall",) * 4)
    amr = dict(max_level=levels, ref_ratio=2, blocking_factor=2, n_error_buf=1, grid_eff=0.7, threshold=0.02,
               b_max=0.25)
    print(f"[jfnk] AMR {levels + 1} levels + global source", flush=True)
    o = AmrOracle("swe2d", case, amr=dict(amr, fine_max_grid_size=16))
    o.init_solution()
    o.set_state(Q0)
    o.init_levels()
    z = ZoomyAmr(deck_from_case(case, amr=amr), system="swe2d", dim=2, order=2, contract=False)
    z.init_data()
    z.set_state(Q0)
    z.init_levels()
    assert o.finest_level() >= 1
    t = 0.0
    for k in range(4):
        t += o.time_step(t, k)
    assert z.run_steps(4)[0] == 4
    assert z.finest_level() == o.finest_level()
    for lev in range(o.finest_level() + 1):
        a, b = z.get_state(lev=lev), o.get_state(lev)
        assert np.array_equal(np.isnan(a), np.isnan(b)), f"level {lev}"
        m = ~np.isnan(b)
        assert rel_linf(a[m], b[m]) <= 1e-10, f"level {lev}"
