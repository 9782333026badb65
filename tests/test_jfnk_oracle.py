"""Pins of the oracle's global implicit-source solve (solver.implicit_global,
Source/imex_solver.H:102-160: Newton on F(Q) = Q - Qexp - dt S(Q), each step
by GMRES with a finite-difference Jacobian-vector product), none re-typing the
solver:
  * its GMRES(30) against numpy.linalg.solve on dense systems, with and
    without restarts;
  * with no source the solve leaves the flux update alone (the run equals the
    point-implicit run bit for bit);
  * the Manning source of swe2d, -g n^2 |m| hinv max(h, eps)^{-4/3} m,
    changes only the momentum m, along its own direction, so one
    backward-Euler step has a closed form per cell: |m| solves
    |m| (1 + dt c |m|) = |m_exp| with c = g n^2 hinv max(h, eps)^{-4/3}
    (also on a state that makes GMRES restart);
  * on the shallow-moment system the returned state satisfies the
    backward-Euler equation with the source ComputeSourceField forms.
"""
import numpy as np
import pytest

from helpers import manning_backward_euler, manning_restart_case
from oracle.pyoracle import Oracle, gmres_dense
from zamr import cases


def test_gmres_dense_against_numpy():
    rng = np.random.default_rng(0)
    n = 80
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, its = gmres_dense(A, b, 1e-10)
    assert its < 30  # no restart needed
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-8)


def test_gmres_dense_with_restarts():
    rng = np.random.default_rng(1)
    n = 200
    A = np.diag(np.linspace(0.1, 10.0, n)) + 0.05 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, its = gmres_dense(A, b, 1e-9)
    assert 30 < its < 2000  # restarted, converged
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b) * 1.01
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-6)


def _case(n=32, nm=0.05, order=2, implicit_global=True, steps=3):
    case, Q0 = cases.swe_bumps(n=n, seed=3, nbumps=4, steps=steps, manning=nm)
    case.update(implicit_global=implicit_global, spatial_order=order, max_grid_size=16, well_balanced=False,
                clamp_positivity=False)
    return case, Q0


def _moving(Q0):
    Q = Q0.copy()
    x = np.linspace(0, 1, Q.shape[2])
    Q[2] = 0.3 * Q[1] * np.sin(2 * np.pi * x)[None, :]
    Q[3] = -0.2 * Q[1] * np.cos(2 * np.pi * x)[:, None]
    return Q


def test_no_source_equals_point_implicit_bitwise():
    case, Q0 = _case(nm=0.0)
    Q0 = _moving(Q0)
    a = Oracle("swe2d", case)
    a.set_state(Q0)
    a.evolve(step_limit=3)
    b = Oracle("swe2d", dict(case, implicit_global=False))
    b.set_state(Q0)
    b.evolve(step_limit=3)
    assert np.array_equal(a.get_state(), b.get_state())


def test_manning_stage_is_the_closed_form_backward_euler():
    g, nm, eps = 9.81, 0.08, 1e-4
    case, Q0 = _case(nm=nm)
    Q0 = _moving(Q0)
    dt = 2e-3
    # flux-only update: the same stage with no friction (implicit_source keeps S out of the RHS)
    f = Oracle("swe2d", dict(case, params={"n_m": 0.0}))
    f.set_state(Q0)
    f.stage(2, 0.0, dt)
    Qexp = f.get_state()
    o = Oracle("swe2d", case)
    o.set_state(Q0)
    o.stage(2, 0.0, dt)
    Q = o.get_state()
    newton, gm = o.jfnk_iters()
    assert newton >= 2 and gm >= newton
    h = Qexp[1]
    hinv = np.sqrt(2.0) * h / np.sqrt(h ** 4 + np.maximum(h ** 4, eps ** 4))
    c = g * nm * nm * hinv * np.maximum(h, eps) ** (-4.0 / 3.0)
    s_exp = np.hypot(Qexp[2], Qexp[3])
    scale = np.where(s_exp > 0, 2.0 / (1.0 + np.sqrt(1.0 + 4.0 * dt * c * s_exp)), 1.0)
    be = Qexp.copy()
    be[2] *= scale
    be[3] *= scale
    assert np.array_equal(Q[:2], Qexp[:2])
    # Newton stops at |F| < 1e-9 |F_0| in the 2-norm over the level, and
    # F_0 = -dt S(Qexp) is the correction to first order
    err = np.abs(Q[2:] - be[2:]).max()
    assert err <= 1e-9 * np.linalg.norm(Qexp[2:] - be[2:]), err
    # and the solve is not the one-step linearisation
    p = Oracle("swe2d", dict(case, implicit_global=False))
    p.set_state(Q0)
    p.stage(2, 0.0, dt)
    assert np.abs(p.get_state()[2:] - be[2:]).max() > 1e3 * max(err, 1e-16)


def test_manning_stage_with_gmres_restarts_is_the_closed_form():
    """The same closed form on a state whose depths fall from 1 m to 1 mm
    under strong friction: the per-cell rates spread over decades, GMRES(30)
    takes more than 30 steps per Newton iteration on average, so restarts
    (restart residual, a new cycle from it, the Q update after a restarted
    cycle) are on the path."""
    nm, dt = 3.0, 2e-3
    case, Q0 = manning_restart_case(nm=nm)
    f = Oracle("swe2d", dict(case, params={"n_m": 0.0}))
    f.set_state(Q0)
    f.stage(2, 0.0, dt)
    Qexp = f.get_state()
    o = Oracle("swe2d", case)
    o.set_state(Q0)
    o.stage(2, 0.0, dt)
    Q = o.get_state()
    newton, gm = o.jfnk_iters()
    assert newton < 12 and gm > 30 * newton  # converged, and some cycle restarted
    be = manning_backward_euler(Qexp, nm, dt)
    assert np.array_equal(Q[:2], Qexp[:2])
    err = np.abs(Q[2:] - be[2:]).max()
    assert err <= 1e-9 * np.linalg.norm(Qexp[2:] - be[2:]), err


def test_sme_stage_solves_backward_euler():
    """Level-2 shallow moments with slip + Newtonian friction: the returned
    state satisfies Q = Qexp + dt S(Q) (S with the aux of Q itself, as
    ComputeSourceField forms it) to Newton's tolerance, and differs from the
    point-implicit step, whose source reads the stage-start aux."""
    case, Q0 = cases.sme_bumps(n=32, seed=1, steps=2)
    case.update(max_grid_size=16)
    dt = 1e-3
    f = Oracle("sme2d", dict(case, implicit_global=True, params=dict(case["params"], nu=0.0)))
    f.set_state(Q0)
    f.stage(2, 0.0, dt)
    Qexp = f.get_state()
    assert np.abs(f.source_field(Qexp)).max() == 0.0  # nu = 0: no friction at all
    o = Oracle("sme2d", dict(case, implicit_global=True))
    o.set_state(Q0)
    o.stage(2, 0.0, dt)
    Q = o.get_state()
    F = Q - Qexp - dt * o.source_field(Q)
    F0 = dt * o.source_field(Qexp)
    assert np.linalg.norm(F) <= 1e-9 * np.linalg.norm(F0)
    p = Oracle("sme2d", dict(case, implicit_global=False))
    p.set_state(Q0)
    p.stage(2, 0.0, dt)
    assert np.abs(p.get_state() - Q).max() > 1e3 * np.linalg.norm(F)
