"""Pins of the oracle's AMR hierarchy beyond two levels and at ref_ratio 4
(oracle/zamr_oracle_amr.hpp: nsubsteps, the InitData level loop and the
recursive TimeStep of ZoomyAmr.cpp:96-106, 137-155, 733-762, with AMReX's
MakeNewGrids proper nesting), checked against what mathematics fixes:
  * cell_cons_interp at ratio 4 conserves (16 children average to the
    parent) and reproduces linear data; average_down undoes it;
  * with every cell tagged, the finest level of a 2- or 3-level hierarchy
    at ratio 2 or 4 is a single-level run at r^l times the resolution,
    r^l of its own CFL steps per coarse step, bit for bit;
  * every level nests in the one below it: the cells within n_proper of a
    coarsened finer box are covered, after every coarse step of a moving
    front that creates, moves and clears levels;
  * a third level appears where the front is and nowhere else.
"""
import numpy as np
import pytest

from oracle.pyoracle import AmrOracle, Oracle
from zamr import cases


def column_case(n=32, steps=6, cfl=0.4, order=2, **kw):
    c = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=16, spatial_order=order,
             implicit_source=False, well_balanced=False, clamp_positivity=False, positivity="none", cfl=cfl,
             dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=steps, bc=("wall",) * 4, params={})
    c.update(kw)
    return c


def column_state(n, wet_outside=0.0, radius=0.2):
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < radius ** 2, 1.0, wet_outside)
    z = np.zeros_like(h)
    return np.stack([z, h, z, z])


def cover(boxes, shape):
    m = np.zeros(shape, dtype=int)
    for (x0, y0, x1, y1) in boxes:
        m[y0:y1 + 1, x0:x1 + 1] += 1
    return m


def check_nesting(o, r, n_proper=1):
    for lev in range(1, o.finest_level() + 1):
        shape = o.level_shape(lev)
        c = cover(o.boxes(lev), shape)
        assert c.max() == 1, f"level {lev} boxes overlap"
        assert all(x0 >= 0 and y0 >= 0 and x1 < shape[1] and y1 < shape[0] for (x0, y0, x1, y1) in o.boxes(lev))
        below = cover(o.boxes(lev - 1), o.level_shape(lev - 1)) > 0
        ny, nx = below.shape
        for (x0, y0, x1, y1) in o.boxes(lev):
            a0, b0 = max(x0 // r - n_proper, 0), max(y0 // r - n_proper, 0)
            a1, b1 = min(x1 // r + n_proper, nx - 1), min(y1 // r + n_proper, ny - 1)
            assert below[b0:b1 + 1, a0:a1 + 1].all(), f"level {lev} box {(x0, y0, x1, y1)} not nested"


def test_interp_ratio4_conserves_and_is_exact_on_linear_data():
    n, r = 12, 4
    o = AmrOracle("swe2d", column_case(n=n), amr=dict(ref_ratio=4))
    j, i = np.mgrid[0:n, 0:n].astype(float)
    coef = [(0.1, 0.01, -0.02), (1.0, 0.03, 0.05), (0.2, -0.01, 0.0), (0.3, 0.0, 0.02)]
    Q = np.stack([c0 + cx * i + cy * j for (c0, cx, cy) in coef])
    o.set_state(Q)
    for jc in range(1, n - 1):
        for ic in range(1, n - 1):
            kids = np.array([o.interp(r * ic + a, r * jc + b) for b in range(r) for a in range(r)])
            assert np.allclose(kids.mean(axis=0), Q[:, jc, ic], rtol=0, atol=2e-15)
            for b in range(r):
                for a in range(r):
                    xf, yf = ic + (a + 0.5) / r - 0.5, jc + (b + 0.5) / r - 0.5
                    exact = [c0 + cx * xf + cy * yf for (c0, cx, cy) in coef]
                    assert np.allclose(kids[r * b + a], exact, rtol=0, atol=4e-15)


@pytest.mark.parametrize("r", [2, 4])
def test_average_down_undoes_interp(r):
    n = 16
    o = AmrOracle("swe2d", column_case(n=n), amr=dict(ref_ratio=r))
    j, i = np.mgrid[0:n, 0:n].astype(float)
    Q = np.stack([0 * i, 1.0 + 0.01 * i + 0.02 * j, 0.001 * i, 0.002 * j])
    o.set_state(Q)
    o.init_levels()  # all wet: level 1 covers the domain
    assert (cover(o.boxes(1), o.level_shape(1)) == 1).all()
    o.average_down(0)
    assert np.allclose(o.get_state(0), Q, rtol=0, atol=1e-15)


@pytest.mark.parametrize("r,levels", [(4, 1), (2, 2), (4, 2)])
def test_all_tagged_finest_level_is_the_single_level_run(r, levels):
    """Every cell wet, so every cell of every level tagged: the finest level
    covers the domain, has no coarse-fine ghosts, and takes r^levels of its
    own CFL steps per coarse step -- a single-level run at r^levels times the
    resolution from the interpolated start, bit for bit."""
    n, steps = 8, 3
    c = column_case(n=n, steps=steps, cfl=0.45)
    Q0 = column_state(n, wet_outside=0.5)
    o = AmrOracle("swe2d", c, amr=dict(max_level=levels, ref_ratio=r))
    o.set_state(Q0)
    o.init_levels()
    assert o.finest_level() == levels
    for lev in range(1, levels + 1):
        assert (cover(o.boxes(lev), o.level_shape(lev)) == 1).all()
    Qf = o.get_state(levels)
    t = 0.0
    for k in range(steps):
        t += o.time_step(t, k)
    assert o.finest_level() == levels
    nf = n * r ** levels
    f = Oracle("swe2d", dict(c, nx=nf, ny=nf, max_grid_size=64, max_step=steps * r ** levels))
    f.set_state(Qf)
    f.evolve()
    assert np.array_equal(o.get_state(levels), f.get_state())


@pytest.mark.parametrize("r", [2, 4])
def test_three_levels_nest_through_a_moving_front(r):
    """A collapsing water column over a dry bed: level 1 follows the wet
    region, level 2 its front; every step the hierarchy nests properly, the
    third level exists, and lies where level 1's tags are."""
    n = 24
    c = column_case(n=n, steps=10, cfl=0.4)
    Q0 = column_state(n, radius=0.18)
    o = AmrOracle("swe2d", c, amr=dict(max_level=2, ref_ratio=r, threshold=0.05))
    o.set_state(Q0)
    o.init_levels()
    check_nesting(o, r)
    assert o.finest_level() == 2
    t, seen = 0.0, set()
    for k in range(8):
        t += o.time_step(t, k)
        check_nesting(o, r)
        seen.add(o.finest_level())
        assert np.isfinite(o.get_state(0)).all()
        tags1 = o.error_est(1)
        if o.finest_level() == 2:
            c2 = cover(o.boxes(2), o.level_shape(2)) > 0
            # every level-2 cell lies under a level-1 cell within n_error_buf + 1 of a level-1 tag
            ny, nx = tags1.shape
            near = np.zeros_like(tags1, dtype=bool)
            ty, tx = np.nonzero(tags1)
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    near[np.clip(ty + dy, 0, ny - 1), np.clip(tx + dx, 0, nx - 1)] = True
            assert near[np.nonzero(c2)[0] // r, np.nonzero(c2)[1] // r].mean() > 0.5
    assert 2 in seen
    regrids, remakes, _ = o.counts()
    assert regrids > 8 and remakes > 0


@pytest.mark.parametrize("r", [2, 4])
def test_regrid_from_level1_erodes_by_n_proper(r):
    """regrid(1) (a level-1 substep): with every level-1 cell tagged, level 2
    covers exactly the level-1 region shrunk by n_proper cells (inside the
    domain), the proper-nesting domain of MakeNewGrids."""
    n = 16
    o = AmrOracle("swe2d", column_case(n=n), amr=dict(max_level=2, ref_ratio=r))
    o.set_state(column_state(n, radius=0.2))
    o.init_levels()
    Q1 = o.get_state(1)
    cov1 = ~np.isnan(Q1[1])
    Q1[1] = np.where(cov1, 1.0, 0.0)
    o.set_state(np.nan_to_num(Q1), lev=1)
    o.regrid(0.0, lbase=1)
    ny, nx = cov1.shape
    pad = np.pad(cov1, 1, constant_values=True)  # beyond the domain does not erode
    eroded = np.ones_like(cov1)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            eroded &= pad[1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    assert eroded.any() and not eroded.all()
    c2 = cover(o.boxes(2), o.level_shape(2))
    assert c2.max() == 1
    assert np.array_equal(c2[::r, ::r] > 0, eroded)
    assert np.array_equal(c2 > 0, np.kron(eroded, np.ones((r, r), dtype=bool)))


def test_ratio_and_depth_limits():
    with pytest.raises(Exception):
        AmrOracle("swe2d", column_case(n=8), amr=dict(ref_ratio=3)).init_levels()
