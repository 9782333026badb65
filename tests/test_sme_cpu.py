"""CPU checks of the level-2 shallow-moment system (systems/sme2d) and of the
oracle running it.  The reference generates its moment models from
zoomy_core, which is absent, so there is no reference fixture: the system is
checked against its own definition (flux Jacobian + non-conservative matrix =
quasilinear matrix, the speed bound against the true spectrum, the path
integral against a fine quadrature) and against what the mathematics fixes
(zero moments stay zero without friction, mass conservation in a walled box,
the friction step = backward Euler, a uniform state is steady)."""
import ctypes
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle.pyoracle import Oracle

SYS = ROOT / "zoomy-amrex_amd" / "systems" / "sme2d"
COMPAT = ROOT / "zoomy-amrex_amd" / "compat"
_D = ctypes.POINTER(ctypes.c_double)
P = np.array([9.81, 0.0, 0.0, 1.0, 1.0, 1e-4, 0.01, 0.0, 0.1])


def dp(a):
    return a.ctypes.data_as(_D)


@pytest.fixture(scope="module")
def sz(tmp_path_factory):
    out = tmp_path_factory.mktemp("sz") / "libsz.so"
    subprocess.run([shutil.which("g++") or "g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    f"-I{COMPAT}", f"-I{SYS}", str(ROOT / "tests" / "sme_harness.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    lib.sz_speed.restype = ctypes.c_double
    lib.sz_speed.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D]
    return lib


def rand_state(rng, moments=True):
    h = rng.uniform(0.3, 2.0)
    q = np.zeros(8)
    q[0] = rng.uniform(-0.2, 0.2)
    q[1] = h
    q[2], q[5] = rng.uniform(-1, 1, 2) * h
    if moments:
        q[[3, 4, 6, 7]] = rng.uniform(-0.4, 0.4, 4) * h
    return q


def mat(sz, fn, q, d):
    out = np.zeros(64)
    getattr(sz, fn)(dp(q), d, dp(P), dp(out))
    return out.reshape(8, 8)


def test_quasilinear_is_flux_jacobian_plus_ncp(sz):
    rng = np.random.default_rng(0)
    for _ in range(30):
        q = rand_state(rng)
        for d in (0, 1):
            J = np.zeros((8, 8))
            for k in range(8):
                e = 1e-6 * max(1.0, abs(q[k]))
                qp, qm = q.copy(), q.copy()
                qp[k] += e
                qm[k] -= e
                fp, fm = np.zeros(8), np.zeros(8)
                sz.sz_flux(dp(qp), d, dp(P), dp(fp))
                sz.sz_flux(dp(qm), d, dp(P), dp(fm))
                J[:, k] = (fp - fm) / (2 * e)
            A = mat(sz, "sz_quasi", q, d)
            B = mat(sz, "sz_ncp", q, d)
            assert np.allclose(A, J + B, rtol=1e-6, atol=1e-6), (d, np.abs(A - J - B).max())


def test_speed_bounds_the_spectrum(sz):
    """The Rusanov/CFL speed is an upper bound of the spectral radius of the
    quasilinear matrix along n, and reduces to a shallow-water speed when the
    moments vanish."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rand_state(rng)
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(th), np.sin(th)]) if rng.uniform() < 0.5 else np.eye(2)[rng.integers(2)]
        A = n[0] * mat(sz, "sz_quasi", q, 0) + n[1] * mat(sz, "sz_quasi", q, 1)
        rho = np.abs(np.linalg.eigvals(A[1:, 1:])).max()
        s = sz.sz_speed(dp(q), n[0], n[1], dp(P))
        assert s >= rho * (1 - 1e-12), (s, rho)
    q = rand_state(rng, moments=False)
    un = q[2] / q[1]
    c = np.sqrt(9.81 * q[1])
    s = sz.sz_speed(dp(q), 1.0, 0.0, dp(P))
    assert abs(un) + c <= s <= c + 2 * abs(un) + 1e-12


def test_fluctuations_are_the_path_integral(sz):
    """D+ = D- = half the straight-path integral of B(q(s)) (qR - qL); the
    3-point rule against a 40-point one, and exact on the bed coupling."""
    rng = np.random.default_rng(2)
    xg, wg = np.polynomial.legendre.leggauss(40)
    s40, w40 = 0.5 * (xg + 1), 0.5 * wg
    for _ in range(20):
        qL, qR = rand_state(rng), rand_state(rng)
        qR[1] = qL[1] * rng.uniform(0.9, 1.1)  # a smooth-flow jump
        for d in (0, 1):
            out = np.zeros(16)
            sz.sz_fluct(dp(qL), dp(qR), d, dp(P), dp(out))
            assert np.array_equal(out[:8], out[8:])
            I = sum(w * mat(sz, "sz_ncp", qL + s * (qR - qL), d) @ (qR - qL) for s, w in zip(s40, w40))
            assert np.allclose(2 * out[:8], I, rtol=1e-4, atol=1e-8)
    qL = np.array([0.1, 1.0, 0, 0, 0, 0, 0, 0.0])
    qR = np.array([0.3, 0.7, 0, 0, 0, 0, 0, 0.0])
    out = np.zeros(16)
    sz.sz_fluct(dp(qL), dp(qR), 0, dp(P), dp(out))
    assert abs(out[2] - 0.25 * 9.81 * (1.0 + 0.7) * 0.2) <= 1e-15


def case(n=32, bc=("wall",) * 4, nu=0.0, lam=0.0, order=2, steps=10):
    return dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=16, spatial_order=order,
                implicit_source=True, well_balanced=False, clamp_positivity=False, cfl=0.45, dtmin=1e-9,
                dtmax=1.0, time_end=1e30, max_step=steps, bc=bc, params={"nu": nu, "lambda_s": lam})


def bumpy(n, rng, moments):
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    b = 0.3 * np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2) / 0.02)
    h = 1.0 - b + 0.05 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    Q = np.zeros((8, n, n))
    Q[0], Q[1] = b, h
    Q[2] = 0.1 * h * np.cos(2 * np.pi * Y)
    Q[5] = -0.1 * h * np.sin(2 * np.pi * X)
    if moments:
        for k in (3, 4, 6, 7):
            Q[k] = rng.uniform(-0.05, 0.05, (n, n)) * h
    return Q


def test_zero_moments_stay_zero_and_mass_is_conserved():
    o = Oracle("sme2d", case())
    o.init_solution()
    Q0 = bumpy(32, None, moments=False)
    o.set_state(Q0)
    o.evolve()
    Q = o.get_state()
    assert not Q[[3, 4, 6, 7]].any()
    assert abs(Q[1].sum() - Q0[1].sum()) <= 1e-12 * Q0[1].sum()
    assert np.array_equal(Q[0], Q0[0])


def test_moments_and_friction_conserve_mass_and_stay_finite():
    o = Oracle("sme2d", case(nu=0.01, lam=0.1))
    o.init_solution()
    Q0 = bumpy(32, np.random.default_rng(3), moments=True)
    o.set_state(Q0)
    o.evolve()
    Q = o.get_state()
    assert np.isfinite(Q).all()
    assert abs(Q[1].sum() - Q0[1].sum()) <= 1e-12 * Q0[1].sum()


def test_friction_step_is_backward_euler(sz):
    """A uniform state in a walled box feels no flux; one stage is then the
    point-implicit friction step, exactly (I - dt J)^-1 Q since the friction
    is linear in Q at fixed hinv."""
    q = np.array([0.0, 0.8, 0.3, 0.05, -0.02, -0.1, 0.03, 0.01])
    o = Oracle("sme2d", case(n=16, nu=0.01, lam=0.1, order=1))
    o.init_solution()
    Q0 = np.broadcast_to(q[:, None, None], (8, 16, 16)).copy()
    o.set_state(Q0)
    dt = 0.05
    o.stage(1, 0.0, dt)
    W = o.get_state()[:, 5, 7]
    S, J = np.zeros(8), np.zeros(64)
    p = P.copy()
    sz.sz_source(dp(q), dp(p), dp(S), dp(J))
    J = J.reshape(8, 8)
    assert np.allclose(J @ q, S, rtol=1e-14, atol=1e-15)
    expect = np.linalg.solve(np.eye(8) - dt * J, q)
    assert np.allclose(W, expect, rtol=1e-13, atol=1e-15)
    assert np.abs(W[[2, 3, 4, 5, 6, 7]]).max() < np.abs(q[[2, 3, 4, 5, 6, 7]]).max()


def test_uniform_state_is_steady_without_friction():
    q = np.array([0.0, 0.8, 0.3, 0.05, -0.02, -0.1, 0.03, 0.01])
    o = Oracle("sme2d", case(n=16, bc=("", "", "", ""), order=2, steps=5))
    o.init_solution()
    Q0 = np.broadcast_to(q[:, None, None], (8, 16, 16)).copy()
    o.set_state(Q0)
    o.evolve()
    assert np.allclose(o.get_state(), Q0, rtol=0, atol=1e-15)
