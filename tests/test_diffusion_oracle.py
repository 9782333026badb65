"""Pins of the oracle's add_diffusive_flux (make_rhs.H:166-229) on the viscous
shallow-water system (systems/swe2d_visc: diffusive_flux = nu grad q on the
rows h, hu, hv).  With g = 0 and no momentum the hyperbolic part vanishes
exactly and the depth obeys the discrete heat equation of the 3-point
central stencil, whose Neumann cosine modes decay by closed-form factors
(the mirror ghosts of extrapolation boundaries make them exact eigenvectors
up to the boundary).  CPU only."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle

NX, NY, LX, LY = 24, 16, 1.5, 0.5
NU = 0.01


def case(order, nsteps, dt, bc=("West", "East", "South", "North"), g=0.0, nu=NU):
    return dict(nx=NX, ny=NY, prob_lo=(0.0, 0.0), prob_hi=(LX, LY), max_grid_size=64, spatial_order=order,
                implicit_source=False, well_balanced=False, clamp_positivity=False, cfl=0.45, dtmin=1e-12,
                dtmax=dt, time_end=1e30, max_step=nsteps, bc=bc, params={"g": g, "nu": nu})


def centers():
    dx, dy = LX / NX, LY / NY
    x = (np.arange(NX) + 0.5) * dx
    y = (np.arange(NY) + 0.5) * dy
    return np.meshgrid(x, y, indexing="xy"), dx, dy


def run(c, h):
    o = Oracle("swe2d_visc", c)
    o.init_solution()
    z = np.zeros_like(h)
    o.set_state(np.array([z, h, z, z]))
    n, t, _, _ = o.evolve()
    return n, o.get_state()


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("mx,my", [(2, 3), (1, 0), (0, 2)])
def test_cosine_mode_decays_by_the_discrete_factor(order, mx, my):
    (X, Y), dx, dy = centers()
    kx, ky = np.pi * mx / LX, np.pi * my / LY
    mode = np.cos(kx * X) * np.cos(ky * Y)
    lam = -4.0 / dx ** 2 * np.sin(kx * dx / 2) ** 2 - 4.0 / dy ** 2 * np.sin(ky * dy / 2) ** 2
    dt = 0.2 * min(dx, dy) ** 2 / NU
    z = NU * dt * lam
    # one forward-Euler stage per step at order 1; Heun's two stages at order 2
    G = 1.0 + z if order == 1 else 1.0 + z + 0.5 * z * z
    steps = 6
    n, Q = run(case(order, steps, dt), 1.0 + 0.1 * mode)
    assert n == steps
    assert np.abs(Q[1] - (1.0 + 0.1 * G ** steps * mode)).max() <= 1e-14
    assert not Q[2:].any() and not Q[0].any()
    if mx or my:
        assert abs(G) < 1.0 - 1e-4  # the mode did decay


def test_linear_profile_is_steady_inside():
    """A linear depth has zero discrete Laplacian: one stage leaves every cell
    away from the boundary unchanged (the mirror ghosts bend the profile
    only at the boundary cells)."""
    (X, Y), dx, dy = centers()
    h = 1.0 + 0.3 * X - 0.2 * Y
    n, Q = run(case(1, 1, 0.1 * dx ** 2 / NU), h)
    d = Q[1] - h
    assert np.abs(d[1:-1, 1:-1]).max() <= 1e-15
    assert np.abs(d[:, 0]).min() > 1e-6 and np.abs(d[0, 1:-1]).min() > 1e-6


def test_mass_conserved_with_walls_and_moving_water():
    """g = 9.81, walls, moving water, order 2: the diffusive mass fluxes
    telescope and vanish at the walls (mirror ghosts), so mass is kept to
    round-off; the viscous run differs from the inviscid one."""
    (X, Y), dx, dy = centers()
    h = 1.0 + 0.05 * np.sin(2 * np.pi * X / LX) * np.cos(2 * np.pi * Y / LY)
    hu = 0.1 * h * np.cos(2 * np.pi * Y / LY)
    hv = -0.05 * h * np.sin(2 * np.pi * X / LX)
    Q0 = np.array([0 * h, h, hu, hv])
    outs = []
    for nu in (0.0, 2e-3):
        c = case(2, 20, 1.0, bc=("wall",) * 4, g=9.81, nu=nu)
        o = Oracle("swe2d_visc", c)
        o.init_solution()
        o.set_state(Q0)
        o.evolve()
        outs.append(o.get_state())
    assert abs(outs[1][1].sum() - h.sum()) <= 1e-13 * h.sum()
    assert np.abs(outs[1] - outs[0]).max() > 1e-6


def test_zero_viscosity_equals_the_inviscid_system():
    """nu = 0: the viscous system's run is the swe2d run bit for bit."""
    (X, Y), dx, dy = centers()
    h = 1.0 + 0.05 * np.sin(2 * np.pi * X / LX)
    Q0 = np.array([0 * h, h, 0.1 * h, 0 * h])
    c = case(2, 10, 1.0, g=9.81, nu=0.0)
    res = []
    for system in ("swe2d_visc", "swe2d"):
        cc = dict(c, params={"g": 9.81}) if system == "swe2d" else c
        o = Oracle(system, cc)
        o.init_solution()
        o.set_state(Q0)
        o.evolve()
        res.append(o.get_state())
    assert np.array_equal(res[0], res[1])
