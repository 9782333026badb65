"""AMR beyond two levels and at ref_ratio 4 (amr.max_level = 2, amr.ref_ratio
= 2 or 4; ZoomyAmr.cpp:96-106 nsubsteps, :137-155 InitData's level loop,
:733-762 the recursive subcycled TimeStep with regrid(lev) per substep and
average_down, AMReX's MakeNewGrids proper nesting) through the C-ABI,
against the oracle's hierarchy.

The bitwise build must agree with the oracle bit for bit on every level:
the set of covered cells per level, every state, the step sizes and the
composite mass, on a collapsing water column over a dry bed whose front the
levels follow (regrids from level 0 every coarse step and from level 1
every level-1 substep, levels made, remade and cleared)."""
import numpy as np
import pytest

from helpers import rel_linf
from oracle.pyoracle import AmrOracle
from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr

pytestmark = pytest.mark.gpu


def column(n=32, steps=8, cfl=0.4, mgs=32, positivity="none", wet_outside=0.0, radius2=0.03):
    c = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=mgs, spatial_order=2,
             implicit_source=False, well_balanced=False, clamp_positivity=False, positivity=positivity, cfl=cfl,
             dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=steps, bc=("wall",) * 4, params={})
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.45) ** 2 + (Y - 0.55) ** 2 < radius2, 1.0, wet_outside)
    z = np.zeros_like(h)
    return c, np.stack([z, h, z, z])


def coverage(boxes, shape):
    m = np.zeros(shape, dtype=np.int32)
    for (x0, y0, x1, y1) in boxes:
        m[y0:y1 + 1, x0:x1 + 1] += 1
    return m


def pair(c, Q0, amr, contract=False):
    o = AmrOracle("swe2d", c, amr=dict(amr, fine_max_grid_size=c["max_grid_size"]))
    o.init_solution()
    o.set_state(Q0)
    o.init_levels()
    z = ZoomyAmr(deck_from_case(c, amr=amr), system="swe2d", dim=2, order=2, contract=contract)
    z.init_data()
    z.set_state(Q0)
    z.init_levels()
    return o, z


def assert_same(o, z, exact=True):
    assert z.finest_level() == o.finest_level()
    for lev in range(o.finest_level() + 1):
        shape = o.level_shape(lev)
        assert np.array_equal(coverage(z.level_boxes(lev), shape), coverage(o.boxes(lev), shape)), f"level {lev}"
        a, b = z.get_state(lev=lev), o.get_state(lev)
        if exact:
            assert np.array_equal(a, b, equal_nan=True), f"level {lev}"
        else:
            m = ~np.isnan(b)
            assert rel_linf(a[m], b[m]) <= 1e-12, f"level {lev}"


def amr_of(levels, r, **kw):
    return dict(max_level=levels, ref_ratio=r, blocking_factor=2, n_error_buf=1, grid_eff=0.7, threshold=0.05,
                b_max=1.0e30, **kw)


@pytest.mark.parametrize("levels,r", [(1, 4), (2, 2), (2, 4)])
def test_hierarchy_steps_match_oracle_bitwise(levels, r):
    steps = 8
    c, Q0 = column(steps=steps)
    print(f"[amr-levels] {levels + 1} levels ratio {r}, {steps} coarse steps", flush=True)
    o, z = pair(c, Q0, amr_of(levels, r))
    assert o.finest_level() == levels
    assert_same(o, z)
    zr0, zm0 = z.amr_stats()  # init_data already ran InitData once on the flat state
    or0, om0, _ = o.counts()
    dts, seen = [], set()
    for k in range(steps):
        dts.append(o.time_step(sum(dts), k))
        seen.add(o.finest_level())
    done, t = z.run_steps(steps)
    assert done == steps
    hd, ht = z.dt_history()
    assert np.array_equal(hd[:steps], np.array(dts))
    assert_same(o, z)
    assert abs(z.total_mass() - o.total_mass()) <= 1e-13 * o.total_mass()  # summation orders differ
    regrids, remakes = z.amr_stats()
    oreg, orem, _ = o.counts()
    assert regrids - zr0 == oreg - or0 and remakes - zm0 == orem - om0 >= 1


def test_three_levels_step_by_step_bitwise():
    """Coarse step by coarse step through zamr_time_step: the hierarchy and
    every state agree after each step."""
    c, Q0 = column(n=24, steps=6, radius2=0.04)
    o, z = pair(c, Q0, amr_of(2, 2))
    t = 0.0
    for k in range(6):
        dt_o = o.time_step(t, k)
        dt_z = z.time_step(t, k)
        assert dt_z == dt_o, k
        assert_same(o, z)
        t += dt_o


def test_three_levels_mood_bitwise():
    c, Q0 = column(n=24, steps=6, cfl=2.5, positivity="mood", radius2=0.04)
    print("[amr-levels] MOOD on three levels", flush=True)
    o, z = pair(c, Q0, amr_of(2, 2))
    t = 0.0
    for k in range(6):
        t += o.time_step(t, k)
    done, _ = z.run_steps(6)
    assert done == 6
    assert_same(o, z)
    assert max(o.mood_redos(1), o.mood_redos(2)) > 0


def test_all_tagged_finest_is_the_single_level_run():
    """Every cell wet and tagged on every level: level 2 covers the domain and
    takes four of its own CFL steps per coarse step, so it is the single-level
    run at 4x the resolution from the interpolated start, bit for bit."""
    n, steps = 16, 3
    c, Q0 = column(n=n, steps=steps, cfl=0.45, wet_outside=0.5)
    z = ZoomyAmr(deck_from_case(c, amr=amr_of(2, 2)), system="swe2d", dim=2, order=2, contract=False)
    z.init_data()
    z.set_state(Q0)
    z.init_levels()
    assert z.finest_level() == 2
    Qf = z.get_state(lev=2)
    assert not np.isnan(Qf).any()
    done, _ = z.run_steps(steps)
    assert done == steps and z.finest_level() == 2
    f = ZoomyAmr(deck_from_case(dict(c, nx=4 * n, ny=4 * n, max_step=4 * steps)), system="swe2d", dim=2, order=2,
                 contract=False)
    f.init_data()
    f.set_state(Qf)
    assert f.run_steps(4 * steps)[0] == 4 * steps
    assert np.array_equal(z.get_state(lev=2), f.get_state())


@pytest.mark.parametrize("r", [2, 4])
def test_three_levels_throughput_build_within_1e12(r):
    """FMA build, wet everywhere (a hump on a depth-1 lake over a bed sloping
    up in x; tagging.b_max keeps the upper half coarse, so waves cross both
    coarse-fine boundaries every step): within 1e-12 of the oracle.  Over a
    dry bed the desingularised 1/h amplifies ulp-level differences at the
    front; dry-front parity is held by the bitwise build above."""
    n = 32
    c, _ = column(n=n, steps=4)
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = 1.0 + 0.2 * np.exp(-((X - 0.3) ** 2 + (Y - 0.5) ** 2) / 0.01)
    Q0 = np.stack([0.4 * X, h, 0 * h, 0 * h])
    o, z = pair(c, Q0, amr_of(2, r, **{}) | dict(b_max=0.2), contract=True)
    assert o.finest_level() == 2
    t = 0.0
    for k in range(4):
        t += o.time_step(t, k)
    assert z.run_steps(4)[0] == 4
    assert_same(o, z, exact=False)


def test_deeper_and_other_ratios_are_rejected():
    c, _ = column()
    for bad in (amr_of(4, 2), amr_of(2, 3)):
        with pytest.raises(Exception):
            z = ZoomyAmr(deck_from_case(c, amr=bad), system="swe2d", dim=2, order=2)
            z.init_data()
