"""CPU checks of the hand-written reverse functions of the per-cell-Manning
system (systems/swe2d_nf/Adjoint.H) against central finite differences of
the forward model functions they reverse (Numerics.H flux / fluctuations, the
point-implicit source solved by Gaussian elimination), on random wet states.

The forward functions are the ones the HIP kernels and the oracle run; the
reverse functions are compiled from the same headers by tests/adjoint_harness.cpp."""
import ctypes
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT as REPO

SYS = REPO / "zoomy-amrex_amd" / "systems" / "swe2d_nf"
COMPAT = REPO / "zoomy-amrex_amd" / "compat"
ORACLE = REPO / "oracle"

P = np.array([9.81, 0.0, 0.0, 1.0, 1.0, 1e-4, 0.0, 0.0, 0.0])
_D = ctypes.POINTER(ctypes.c_double)


def dp(a):
    return a.ctypes.data_as(_D)


@pytest.fixture(scope="module")
def hz(tmp_path_factory):
    out = tmp_path_factory.mktemp("hz") / "libhz.so"
    cxx = shutil.which("g++") or "g++"
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{COMPAT}", f"-I{SYS}",
                    f"-I{ORACLE}", str(REPO / "tests" / "adjoint_harness.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    lib.hz_hinv.restype = ctypes.c_double
    lib.hz_dhinv.restype = ctypes.c_double
    lib.hz_hinv.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.hz_dhinv.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.hz_implicit.restype = ctypes.c_int
    for f in ("hz_flux", "hz_face_reverse", "hz_implicit", "hz_manning", "hz_source_reverse", "hz_bc",
              "hz_bc_reverse", "hz_wb_face", "hz_wb_face_reverse"):
        if f != "hz_implicit":
            getattr(lib, f).restype = None
    lib.hz_wb_face_reverse.argtypes = [_D, _D, ctypes.c_int, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double, _D,
                                       _D]
    return lib


def flux(hz, qL, qR, d):
    out = np.zeros(12)
    hz.hz_flux(dp(qL), dp(qR), d, dp(P), dp(out))
    return out


def rand_face(rng, dry=False):
    h = rng.uniform(0.2, 2.0, 2)
    if dry:
        h[0] = rng.uniform(1e-6, 5e-5)  # below wet_dry_eps: the desingularised branch
    b = rng.uniform(-0.3, 0.3, 2)
    m = rng.uniform(-1.0, 1.0, (2, 2)) * h[:, None]
    qL = np.array([b[0], h[0], m[0, 0], m[0, 1]])
    qR = np.array([b[1], h[1], m[1, 0], m[1, 1]])
    return qL, qR


@pytest.mark.parametrize("d", [0, 1])
@pytest.mark.parametrize("dry", [False, True])
def test_face_reverse_matches_finite_differences(hz, d, dry):
    rng = np.random.default_rng(11 + d + 2 * dry)
    for _ in range(40):
        qL, qR = rand_face(rng, dry)
        w = rng.normal(size=12)  # output adjoint: flux, D+, D-
        w[0] = 0.0  # b carries no flux, so its flux adjoint never enters
        qLb, qRb = np.zeros(4), np.zeros(4)
        hz.hz_face_reverse(dp(qL), dp(qR), d, dp(P), dp(w[:4].copy()), dp(w[4:8].copy()), dp(w[8:].copy()),
                           dp(qLb), dp(qRb))
        for side, q, qb in ((0, qL, qLb), (1, qR, qRb)):
            for n in range(4):
                e = 1e-7 * max(1.0, abs(q[n]))
                qp, qm = q.copy(), q.copy()
                qp[n] += e
                qm[n] -= e
                fp = flux(hz, qp, qR, d) if side == 0 else flux(hz, qL, qp, d)
                fm = flux(hz, qm, qR, d) if side == 0 else flux(hz, qL, qm, d)
                fd = np.dot(w, (fp - fm) / (2 * e))
                assert abs(fd - qb[n]) <= 1e-6 * max(1.0, abs(fd)), (side, n, fd, qb[n])


def wb_face(hz, qL, qR, d):
    out = np.zeros(14)
    hz.hz_wb_face(dp(qL), dp(qR), d, dp(P), dp(out))
    return out


@pytest.mark.parametrize("d", [0, 1])
@pytest.mark.parametrize("kind", ["wet", "dry_side", "dry_cell"])
def test_wb_face_reverse_matches_finite_differences(hz, d, kind):
    """The hydrostatic-reconstruction face reverse against central
    differences of the oracle's well-balanced face (flux, fluctuations and the
    reconstructed depths that feed the balancing bed source), weighted by a
    random output adjoint; `dry_side`: the lower side's water lies below the
    face bed (h* = 0 there), `dry_cell`: one side has h = 0."""
    rng = np.random.default_rng(31 + d + 3 * ["wet", "dry_side", "dry_cell"].index(kind))
    checked = 0
    for _ in range(60):
        qL, qR = rand_face(rng)
        if kind == "dry_side":
            qL[0], qR[0] = 0.0, rng.uniform(0.5, 1.0)
            qL[1] = rng.uniform(0.05, 0.3)  # hL + bL - bf < 0
        elif kind == "dry_cell":
            qL[1] = 0.0
            qL[2:] = 0.0
        w = rng.normal(size=14)  # flux, D+, D-, then the h* weights kL, kR
        w[0] = 0.0
        qLb, qRb = np.zeros(4), np.zeros(4)
        hz.hz_wb_face_reverse(dp(qL), dp(qR), d, dp(P), dp(w[:4].copy()), dp(w[4:8].copy()), dp(w[8:12].copy()),
                              w[12], w[13], dp(qLb), dp(qRb))
        base = wb_face(hz, qL, qR, d)
        if kind == "dry_side":
            assert base[12] == 0.0
        for side, q, qb in ((0, qL, qLb), (1, qR, qRb)):
            for n in range(4):
                if kind == "dry_cell" and side == 0 and n > 0:
                    continue  # h = 0 sits on the wet/dry kink: one-sided there
                e = 1e-7 * max(1.0, abs(q[n]))
                qp, qm = q.copy(), q.copy()
                qp[n] += e
                qm[n] -= e
                fp = wb_face(hz, qp, qR, d) if side == 0 else wb_face(hz, qL, qp, d)
                fm = wb_face(hz, qm, qR, d) if side == 0 else wb_face(hz, qL, qm, d)

                def J(f):
                    return np.dot(w[:12], f[:12]) + 0.5 * w[12] * f[12] ** 2 + 0.5 * w[13] * f[13] ** 2

                fd = (J(fp) - J(fm)) / (2 * e)
                assert abs(fd - qb[n]) <= 1e-6 * max(1.0, abs(fd)), (kind, side, n, fd, qb[n])
                checked += 1
    assert checked > 0


def test_hinv_derivative(hz):
    for eps in (1e-4, 1e-2):
        for h in (3e-5, 9e-5, 2e-4, 0.007, 0.013, 0.5, 3.0):  # off the kink at h = eps
            e = 1e-6 * h
            fd = (hz.hz_hinv(h + e, eps) - hz.hz_hinv(h - e, eps)) / (2 * e)
            an = hz.hz_dhinv(h, eps)
            assert abs(fd - an) <= 1e-6 * abs(an), (h, eps, fd, an)
    # wet branch: hinv = 1/h, so the derivative is -1/h^2
    assert abs(hz.hz_dhinv(2.0, 1e-4) + 0.25) < 1e-15


def implicit(hz, V, U, nn, dt):
    W = np.zeros(4)
    ok = hz.hz_implicit(dp(V), dp(U), ctypes.c_double(nn), ctypes.c_double(dt), dp(P), dp(W))
    assert ok == 1
    return W


def test_manning_closed_form_equals_elimination(hz):
    """The closed form W = m (1 + tau)/(1 + 2 tau) the adjoint differentiates
    is the model's point-implicit step (I - dt J) dq = dt S, solved by the
    oracle's pivoted elimination, to rounding."""
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(500):
        U = np.array([0.1, rng.uniform(0.05, 2.0), *rng.uniform(-2, 2, 2)])
        V = U + rng.normal(scale=0.1, size=4) * np.array([0, 0.1, 1, 1])
        V[1] = abs(V[1]) + 1e-3
        nn = rng.uniform(0.0, 0.2)
        dt = 10 ** rng.uniform(-4, 0)
        We = implicit(hz, V, U, nn, dt)
        Wc = np.zeros(4)
        hz.hz_manning(dp(V), dp(U), ctypes.c_double(nn), ctypes.c_double(dt), dp(P), dp(Wc))
        worst = max(worst, np.max(np.abs(We - Wc)) / max(1.0, np.max(np.abs(We))))
    assert worst <= 1e-13, worst


def test_source_reverse_matches_finite_differences(hz):
    rng = np.random.default_rng(5)
    for _ in range(60):
        U = np.array([0.2, rng.uniform(0.1, 2.0), *rng.uniform(-1, 1, 2)])
        V = U + rng.normal(scale=0.05, size=4)
        V[1] = abs(V[1]) + 0.05
        nn = rng.uniform(0.01, 0.1)
        dt = 10 ** rng.uniform(-3, -1)
        Wb = rng.normal(size=4)
        Vb, Ub, nb = np.zeros(4), np.zeros(4), ctypes.c_double()
        hz.hz_source_reverse(dp(U), dp(V), ctypes.c_double(nn), ctypes.c_double(dt), dp(P), dp(Wb), dp(Vb),
                             dp(Ub), ctypes.byref(nb))

        def fwd(V, U, nn):
            return np.dot(Wb, implicit(hz, V, U, nn, dt))

        for n in range(4):
            e = 1e-6 * max(1.0, abs(V[n]))
            vp, vm = V.copy(), V.copy()
            vp[n] += e
            vm[n] -= e
            fd = (fwd(vp, U, nn) - fwd(vm, U, nn)) / (2 * e)
            assert abs(fd - Vb[n]) <= 1e-6 * max(1.0, abs(fd)), ("V", n, fd, Vb[n])
            up, um = U.copy(), U.copy()
            up[n] += e
            um[n] -= e
            fd = (fwd(V, up, nn) - fwd(V, um, nn)) / (2 * e)
            assert abs(fd - Ub[n]) <= 1e-6 * max(1.0, abs(fd)), ("U", n, fd, Ub[n])
        e = 1e-6 * nn
        fd = (fwd(V, U, nn + e) - fwd(V, U, nn - e)) / (2 * e)
        assert abs(fd - nb.value) <= 1e-6 * max(1.0, abs(fd)), ("n", fd, nb.value)


def test_source_reverse_at_rest(hz):
    """Zero momentum: the Manning step is the identity to first order (the
    source is |m| m, whose derivative vanishes at 0)."""
    U = np.array([0.0, 1.0, 0.0, 0.0])
    Wb = np.array([1.0, 2.0, 3.0, 4.0])
    Vb, Ub, nb = np.zeros(4), np.zeros(4), ctypes.c_double()
    hz.hz_source_reverse(dp(U), dp(U), ctypes.c_double(0.03), ctypes.c_double(0.01), dp(P), dp(Wb), dp(Vb), dp(Ub),
                         ctypes.byref(nb))
    assert np.array_equal(Vb, Wb) and not Ub.any() and nb.value == 0.0


@pytest.mark.parametrize("tag", [0, 4])
@pytest.mark.parametrize("d", [0, 1])
def test_bc_reverse_is_transpose(hz, tag, d):
    rng = np.random.default_rng(tag + d)
    for side in (-1.0, 1.0):
        J = np.zeros((4, 4))
        for k in range(4):
            e = np.zeros(4)
            e[k] = 1.0
            out = np.zeros(4)
            hz.hz_bc(tag, d, ctypes.c_double(side), dp(e), dp(out))
            J[:, k] = out
        v = rng.normal(size=4)
        out = np.zeros(4)
        hz.hz_bc_reverse(tag, d, dp(v), dp(out))
        assert np.allclose(out, J.T @ v, rtol=0, atol=1e-15)


def test_oracle_limiter_tape_replays_its_run():
    """The oracle's limiter tape: a replayed run from the recorded initial
    state is the recorded run bit for bit, and a perturbed replayed run is a
    different (smooth-in-the-perturbation) map than the free one."""
    from oracle.pyoracle import Oracle
    from zamr import cases
    c = cases.swe_adjoint(n=32, seed=2, steps=6, ngauges=4, nfd=2)
    Q0 = cases.initial_state(c["b0"], c["eta0"])
    o = Oracle("swe2d_nf", c["case"])
    o.init_solution()

    def run(Q):
        o.set_state(Q)
        t = 0.0
        for _ in range(6):
            o.advance(t, 1e-3)
            t += 1e-3
        return o.get_state()

    o.set_tape(1)
    A = run(Q0)
    o.set_tape(2)
    B = run(Q0)
    assert np.array_equal(A, B)
    Qp = Q0.copy()
    Qp[0, 16, 16] += 1e-3
    o.set_tape(0)
    free = run(Qp)
    o.set_tape(2)
    frozen = run(Qp)
    assert np.isfinite(frozen).all() and not np.array_equal(free, A)
    o.close()
