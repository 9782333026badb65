"""Shared test helpers: reference fixture cases, comparisons, deck builders."""
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "refs"


def rel_linf(a, b):
    """Relative L-infinity distance of two states: max|a - b| / max|b| over
    every row and cell.  (A per-row scale is ill-posed for momentum rows that
    sit near zero while their fluxes are O(g h^2); the whole-state scale is the
    state's own magnitude.)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_linf_rows(a, b):
    """Per-row relative L-infinity distance (row scale floored at 1e-300)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return [float(np.abs(a[r] - b[r]).max() / max(np.abs(b[r]).max(), 1e-300)) for r in range(b.shape[0])]


def stoker_ic(x):  # reference tests/cases.py:195-204
    h = np.where(x < 5.0, 0.005, 0.001)
    return np.array([0 * x, h, 0 * x])


def ritter_ic(x):  # reference tests/cases.py:207-210
    h = np.where(x < 5.0, 0.005, 0.0)
    return np.array([0 * x, h, 0 * x])


def lake_ic(x):  # reference tests/cases.py:213-221
    b = 0.2 * np.exp(-((x - 5.0) ** 2))
    return np.array([b, 1.0 - b, 0 * b])


# name -> (ic, nx, t_end, well_balanced, wall)   (reference tests/test_*.py)
FIXTURES_1D = {
    "stoker_wet": (stoker_ic, 100, 1.0, False, False),
    "ritter_dry": (ritter_ic, 100, 1.0, False, False),
    "wb_lake": (lake_ic, 100, 1.0, True, True),
    "parity_numpy_n100": (stoker_ic, 100, 0.5, False, False),
    "parity_numpy_n200": (stoker_ic, 200, 0.5, False, False),
    "parallel_2rank": (stoker_ic, 128, 0.5, False, False),
}


def case_1d(nx, t_end, wb, wall):
    """The deck run_case writes for a 1-D test (run_case.py:75-84, 235-288):
    nx x 1 cells on [0,10] x [0,dx], order 1, CFL 0.9 with d = 1, dtmax = t_end,
    clamp off, explicit source.  The reference's wall model names its sides
    West/East with a wall definition; this system's reflective tag is 'wall'."""
    L = 10.0
    dx = L / nx
    side = ("wall", "wall") if wall else ("West", "East")
    return dict(nx=nx, ny=1, prob_lo=(0.0, 0.0), prob_hi=(L, dx), max_grid_size=64, spatial_order=1,
                implicit_source=False, well_balanced=wb, clamp_positivity=False, cfl=0.9, dtmin=1e-7,
                dtmax=t_end, time_end=t_end, max_step=-1, bc=(side[0], side[1], "", ""),
                cfl_dimension=1, cfl_degree_factor=1)


def case_2d_pulse():
    """reference tests/test_swe_2d.py: 32^2 on [-1,1]^2, Gaussian pulse, t=0.1, CFL 0.9, order 1."""
    n = 32
    c = dict(nx=n, ny=n, prob_lo=(-1.0, -1.0), prob_hi=(1.0, 1.0), max_grid_size=64, spatial_order=1,
             implicit_source=False, well_balanced=False, clamp_positivity=False, cfl=0.9, dtmin=1e-7,
             dtmax=0.1, time_end=0.1, max_step=-1, bc=("West", "East", "South", "North"),
             cfl_dimension=2, cfl_degree_factor=1)
    dx = 2.0 / n
    xc = -1 + (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(xc, xc, indexing="xy")
    h = 1.0 + 0.1 * np.exp(-(X ** 2 + Y ** 2) / 0.05)
    z = 0 * h
    return c, np.array([z, h, z, z])


def fixture(name):
    d = np.load(GOLDEN / f"{name}.npz")
    return d["Q"], d["Qaux"]


def smooth_state(nx, ny, seed=1, bed=True, amp=0.05):
    """A smooth wet state with momentum and (optionally) a bed, for stage parity."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="xy")
    b = 0.3 * np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2) / 0.02) if bed else 0 * X
    eta = 1.0 + amp * np.sin(2 * np.pi * X + rng.uniform(0, 6)) * np.cos(2 * np.pi * Y + rng.uniform(0, 6))
    h = np.maximum(eta - b, 0.0)
    hu = h * 0.3 * np.sin(2 * np.pi * Y)
    hv = h * 0.2 * np.cos(2 * np.pi * X)
    return np.array([b, h, hu, hv])


def bed_step(nx=8, ny=4, axis=0, u=0.3, v=-0.2, h_lo=1.0, b_hi=0.5, k=4):
    """Moving water over a bed step: b = 0 below index k along `axis` and
    b_hi from k on, a flat free surface eta = h_lo and one velocity (u, v)
    everywhere (a 1-D state when ny == 1: rows b, h, hu).

    The hydrostatic reconstruction maps the two sides of the step face to
    the same state (b*, h*, h* u, h* v) with h* = h_lo - b_hi, because the
    momentum is rescaled by h*/h; every other face separates equal states.
    So a Rusanov stage has a closed form: the flux through the step face is
    the physical flux of that common state, with no dissipation, and only
    the cell below the step changes:
        d(h)/dt   = (h_lo - h*) u_n / dx
        d(hu_t)/dt = (h_lo - h*) u_n u_t / dx   for the momentum along the face,
        d(hu_n)/dt = (h_lo - h*) u_n u_n / dx   (the pressure terms cancel
                     against the well-balanced bed source).
    Returns (Q, expected d/dt of the cell below the step per row)."""
    hs = h_lo - b_hi
    if ny == 1:
        b = np.where(np.arange(nx) < k, 0.0, b_hi)
        h = h_lo - b
        Q = np.array([b, h, h * u])[:, None, :]
        rate = np.array([0.0, (h_lo - hs) * u, (h_lo - hs) * u * u])
        return Q, rate
    idx = np.arange(nx if axis == 0 else ny)
    line = np.where(idx < k, 0.0, b_hi)
    b = np.tile(line, (ny, 1)) if axis == 0 else np.tile(line[:, None], (1, nx))
    h = h_lo - b
    Q = np.array([b, h, h * u, h * v])
    un = u if axis == 0 else v
    rate = np.array([0.0, (h_lo - hs) * un, (h_lo - hs) * un * u, (h_lo - hs) * un * v])
    return Q, rate


def bed_step_case(nx, ny, order, dim):
    side = ("West", "East", "South", "North")
    return dict(nx=nx, ny=ny, prob_lo=(0.0, 0.0), prob_hi=(float(nx), float(max(ny, 1))), max_grid_size=64,
                spatial_order=order, implicit_source=False, well_balanced=True, clamp_positivity=False, cfl=0.45,
                dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=-1, bc=side if dim == 2 else side[:2] + ("", ""))


def bed_step_expected(Q, rate, dt, axis, k):
    """The state after one stage of length dt: only the cells just below the
    step (index k-1 along `axis`) change, by dt * rate."""
    E = Q.copy()
    if axis == 0:
        E[:, :, k - 1] = Q[:, :, k - 1] + dt * rate[:, None]
    else:
        E[:, k - 1, :] = Q[:, k - 1, :] + dt * rate[:, None]
    return E


def fallback_state(nx=60, ny=8, u=0.3, v=-0.2):
    """Three bands along x for the implicit-source fallback test system
    (tests/systems/swe2d_fallback, S_m = -kappa m / h): h = 1 (cols < nx/3),
    h = 4 (next third), dry (last third), one velocity in the wet bands."""
    h = np.zeros((ny, nx))
    h[:, : nx // 3] = 1.0
    h[:, nx // 3: 2 * nx // 3] = 4.0
    return np.array([0 * h, h, h * u, h * v])


FALLBACK_DT = 1.0 / 256
FALLBACK_KAPPA = -256.0  # -1/dt: I - dt J has zero momentum rows where h = 1


def fallback_case(order, mgs=64, nx=60, ny=8):
    return dict(nx=nx, ny=ny, prob_lo=(0.0, 0.0), prob_hi=(1.0, ny / nx), max_grid_size=mgs, spatial_order=order,
                implicit_source=True, well_balanced=False, clamp_positivity=False, cfl=0.45, dtmin=1e-12, dtmax=1.0,
                time_end=1e30, max_step=-1, bc=("wall", "wall", "wall", "wall"), params={"kappa": FALLBACK_KAPPA})


def manning_restart_case(n=48, nm=3.0, decades=3.0, u=0.3):
    """A global-source stage whose GMRES needs restarts: a flat bed, depths
    falling smoothly from 1 m to 10^-decades m across the domain, a swirl of
    speed u and strong Manning friction.  The per-cell backward-Euler rates
    dt c |m| then span several decades, which GMRES(30) does not resolve in
    one cycle, while Newton still converges (depths stay well above the
    wet-dry threshold)."""
    from zamr import cases
    case, Q0 = cases.swe_bumps(n=n, seed=3, nbumps=4, steps=6, manning=nm)
    case.update(implicit_global=True, spatial_order=2, max_grid_size=24, well_balanced=False,
                clamp_positivity=False)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)  # [row j, column i]
    h = 10.0 ** (-decades * 0.5 * (X + Y))
    Q = np.zeros_like(Q0)
    Q[1] = h
    Q[2] = u * h * np.cos(2 * np.pi * Y)
    Q[3] = -u * h * np.sin(2 * np.pi * X)
    return case, Q


def manning_backward_euler(Qexp, nm, dt, g=9.81, eps=1e-4):
    """The backward-Euler step of swe2d's Manning source from Qexp, per cell:
    the momentum m keeps its direction and |m| (1 + dt c |m|) = |m_exp| with
    c = g n^2 hinv max(h, eps)^(-4/3); h and b do not change."""
    h = Qexp[1]
    hinv = np.sqrt(2.0) * h / np.sqrt(h ** 4 + np.maximum(h ** 4, eps ** 4))
    c = g * nm * nm * hinv * np.maximum(h, eps) ** (-4.0 / 3.0)
    s = np.hypot(Qexp[2], Qexp[3])
    scale = np.where(s > 0, 2.0 / (1.0 + np.sqrt(1.0 + 4.0 * dt * c * s)), 1.0)
    be = Qexp.copy()
    be[2] *= scale
    be[3] *= scale
    return be
