"""Pins of the CPU oracle (oracle/zamr_oracle.hpp) before anything is compared
against it: the reference's own golden fixtures, closed forms, invariants and
brute force.  CPU only."""
import numpy as np
import pytest

from helpers import FIXTURES_1D, case_1d, case_2d_pulse, fixture, smooth_state
from oracle.pyoracle import Oracle, ge_solve


def run_fixture_1d(name, max_grid_size=64):
    ic, nx, tend, wb, wall = FIXTURES_1D[name]
    c = case_1d(nx, tend, wb, wall)
    c["max_grid_size"] = max_grid_size
    o = Oracle("swe1d", c)
    o.init_solution()
    x = (np.arange(nx) + 0.5) * (10.0 / nx)
    o.set_state(ic(x)[:, None, :])
    n, t, dts, ms = o.evolve()
    return o, n, t


@pytest.mark.parametrize("name", sorted(FIXTURES_1D))
def test_reference_fixture_1d(name):
    """tests/refs/<name>.npz of the reference, under its own rule np.allclose
    (refs.py:33) — and, far tighter, to 1e-13 absolute."""
    o, n, t = run_fixture_1d(name)
    Qr, Ar = fixture(name)
    Q, A = o.get_state(), o.get_aux()
    assert np.allclose(Q, Qr.reshape(Q.shape)) and np.allclose(A, Ar.reshape(A.shape))
    assert np.abs(Q - Qr.reshape(Q.shape)).max() < 1e-13
    assert np.abs(A - Ar.reshape(A.shape)).max() < 1e-10 * np.abs(Ar).max()


def test_reference_fixture_2d_pulse():
    c, Q0 = case_2d_pulse()
    o = Oracle("swe2d", c)
    o.init_solution()
    o.set_state(Q0)
    o.evolve()
    Qr, Ar = fixture("swe_2d")
    Q, A = o.get_state(), o.get_aux()
    assert np.allclose(Q, Qr) and np.allclose(A, Ar)
    assert np.abs(Q - Qr).max() < 1e-13
    # the reference's run is bitwise transpose-symmetric; so is the oracle's
    assert np.array_equal(Q[1], Q[1].T) and np.array_equal(Q[2], Q[3].T)


def test_two_rank_analogue_is_bitwise():
    """tests/test_parallel_two_ranks.py: decomposition must not change a bit.
    Here 128 cells as 1 box, 2 boxes and 4 boxes."""
    outs = [run_fixture_1d("parallel_2rank", mgs)[0] for mgs in (128, 64, 32)]
    for o in outs[1:]:
        assert np.array_equal(o.get_state(), outs[0].get_state())
        assert np.array_equal(o.get_aux(), outs[0].get_aux())


def test_ritter_mass_and_positivity():
    """Ritter dry dam break: mass is exactly 0.025 (fixture) and h >= 0 with no floor."""
    o, _, _ = run_fixture_1d("ritter_dry")
    Q = o.get_state()
    assert Q[1].min() >= 0.0
    assert abs(Q[1].sum() * 0.1 - 0.025) < 1e-15


def _bump_case(n, order, wb, bc=("wall", "wall", "wall", "wall"), impl=False, manning=0.0, mgs=64):
    return dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=mgs, spatial_order=order,
                implicit_source=impl, well_balanced=wb, clamp_positivity=False, cfl=0.45, dtmin=1e-7, dtmax=1.0,
                time_end=1e30, max_step=-1, bc=bc, params={"n_m": manning})


def test_lake_at_rest_order2_2d():
    """Well-balanced order 2 over bumps: eta stays flat and u stays zero to
    round-off (the reference's gate, tests/test_wb_lake_at_rest.py:54-59)."""
    n = 48
    Q = smooth_state(n, n, bed=True, amp=0.0)
    Q[2:] = 0.0
    o = Oracle("swe2d", _bump_case(n, 2, True, impl=True, manning=0.03))
    o.init_solution()
    o.set_state(Q)
    o.evolve(step_limit=20)
    R = o.get_state()
    eta = R[0] + R[1]
    assert np.abs(eta - 1.0).max() < 1e-13
    assert np.abs(R[2:]).max() < 1e-13


def test_constant_state_is_steady():
    """A constant state has zero flux divergence: one stage leaves it unchanged."""
    n = 20
    Q = np.zeros((4, n, n))
    Q[1] = 1.3
    Q[2] = 0.2
    Q[3] = -0.1
    o = Oracle("swe2d", _bump_case(n, 2, False, bc=("", "", "", "")))
    o.init_solution()
    o.set_state(Q)
    o.update_state(0.0)
    o.stage(2, 0.0, 1e-3)
    assert np.abs(o.get_state() - Q).max() < 1e-15


def test_wall_box_conserves_mass():
    """Reflective walls + no clamp: total water is conserved to round-off at
    order 2 (ZoomyAmr.H:110-113 comment; mirror ghosts ZoomyAmr.cpp:390-398)."""
    n = 40
    Q = smooth_state(n, n, bed=True, amp=0.1)
    o = Oracle("swe2d", _bump_case(n, 2, True))
    o.init_solution()
    o.set_state(Q)
    m0 = o.total_mass()
    o.evolve(step_limit=30)
    assert abs(o.total_mass() - m0) / m0 < 1e-13


def test_decomposition_invariance_order2():
    n = 40
    Q = smooth_state(n, n, bed=True, amp=0.1)
    res = []
    for mgs in (40, 16, 7):
        o = Oracle("swe2d", _bump_case(n, 2, True, bc=("", "wall", "", ""), impl=True, manning=0.03, mgs=mgs))
        o.init_solution()
        o.set_state(Q)
        o.evolve(step_limit=5)
        res.append(o.get_state())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])


def test_dt_closed_form_and_dry_domain():
    """Uniform state: dt = cfl / (d (2k+1) (|u| + sqrt(g h)) / dx); after an
    UpdateState the aux hinv is live.  A dry domain steps at dtmax."""
    n = 16
    c = _bump_case(n, 2, False, bc=("", "", "", ""))
    o = Oracle("swe2d", c)
    o.init_solution()
    Q = np.zeros((4, n, n))
    Q[1] = 2.0
    Q[2] = 2.0 * 0.5
    o.set_state(Q)
    o.update_state(0.0)
    dx = 1.0 / n
    lam = 0.5 + np.sqrt(9.81 * 2.0)
    assert o.compute_dt() == pytest.approx(0.45 / (2 * 3 * lam / dx), rel=1e-15)
    Q[1] = 0.0
    Q[2] = 0.0
    o.set_state(Q)
    o.update_state(0.0)
    assert o.compute_dt() == 1.0


def test_first_dt_sees_init_aux():
    """The first ComputeDt reads init_solution's aux, whose hinv row is 0
    (ZoomyAmr.cpp:154-155,174-175): momentum does not enter the first speed."""
    n = 16
    o = Oracle("swe2d", _bump_case(n, 2, False, bc=("", "", "", "")))
    o.init_solution()
    Q = np.zeros((4, n, n))
    Q[1] = 1.0
    Q[2] = 5.0
    o.set_state(Q)
    lam = np.sqrt(9.81)
    assert o.compute_dt() == pytest.approx(0.45 / (6 * lam * n), rel=1e-15)


@pytest.mark.parametrize("n", [3, 4, 8])
def test_ge_matches_linalg_solve(n):
    rng = np.random.default_rng(n)
    for _ in range(200):
        A = rng.normal(size=(n, n)) + n * np.eye(n)
        b = rng.normal(size=n)
        ok, x = ge_solve(A, b)
        assert ok
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    # pivoting actually happens: a zero leading pivot still solves
    A = np.eye(n)[::-1].copy()
    b = np.arange(1.0, n + 1)
    ok, x = ge_solve(A, b)
    assert ok and np.allclose(A @ x, b)
    # singular system -> the fallback path
    ok, _ = ge_solve(np.zeros((n, n)), b)
    assert not ok


def test_implicit_friction_is_one_backward_euler_step():
    """With the Jacobian exact for a momentum-only change, one point-implicit
    step solves (1 + dt k(1 + u^2/|u|^2)) dq = -dt k hu for a uniform 1-D flow:
    checked against the closed-form linear solve."""
    n = 8
    c = _bump_case(n, 2, False, bc=("", "", "", ""), impl=True, manning=0.05)
    o = Oracle("swe2d", c)
    o.init_solution()
    Q = np.zeros((4, n, n))
    Q[1] = 0.5
    Q[2] = 0.5 * 1.2
    o.set_state(Q)
    o.update_state(0.0)
    dt = 0.01
    o.stage(2, 0.0, dt)
    R = o.get_state()
    h, hu = 0.5, 0.6
    hinv = 1.0 / h
    u = hu * hinv
    k = 9.81 * 0.05 ** 2 * abs(u) / h ** (4.0 / 3.0)
    dk_dhu = 9.81 * 0.05 ** 2 / h ** (4.0 / 3.0) * hinv * np.sign(u)
    J = -k - hu * dk_dhu
    dq = dt * (-k * hu) / (1 - dt * J)
    assert np.allclose(R[2], hu + dq, rtol=1e-12)
    assert np.array_equal(R[1], Q[1])


def test_order2_converges():
    """A smooth travelling wave in 1-D (periodic-free window, extrapolation):
    the order-2 error against a 4x finer run falls at rate >= 1.9
    (conftest.py:38 ORDER_FLOOR of the reference)."""
    def run(nx):
        L = 1.0
        x = (np.arange(nx) + 0.5) * (L / nx)
        c = dict(nx=nx, ny=1, prob_lo=(0.0, 0.0), prob_hi=(L, L / nx), max_grid_size=1024, spatial_order=2,
                 implicit_source=False, well_balanced=False, clamp_positivity=False, cfl=0.45, dtmin=1e-12,
                 dtmax=1e-3, time_end=0.02, max_step=-1, bc=("West", "East", "", ""), cfl_dimension=1,
                 cfl_degree_factor=3)
        o = Oracle("swe1d", c)
        o.init_solution()
        h = 1.0 + 0.01 * np.exp(-((x - 0.5) ** 2) / 0.005)
        o.set_state(np.array([0 * x, h, 0.3 * h])[:, None, :])
        o.evolve()
        return o.get_state()[1, 0]
    ref = run(1600)
    errs = []
    sizes = (100, 200, 400)
    for nx in sizes:
        r = ref.reshape(nx, -1).mean(axis=1)
        errs.append(np.abs(run(nx) - r).mean())
    rate = -np.polyfit(np.log(sizes), np.log(errs), 1)[0]
    assert rate >= 1.9, (errs, rate)
