"""Two-level AMR advance (amr.max_level = 1; ZoomyAmr.cpp:137-155 InitData's
level loop, :225-253 ErrorEst, :147/:736 regrid, :258-288 FillPatch /
FillCoarsePatch with cell_cons_interp, :733-762 the subcycled TimeStep with
average_down, :517-536 the composite mass) through the C-ABI, against the
oracle's AMR restatement.

The bitwise build (FMA off) must agree with the oracle bit for bit on both
levels, the level-1 grids (compared as the set of covered cells: the split
into boxes does not enter the results), the tags and every step size, on a
collapsing water column over a dry bed whose front the refinement follows
(regrids every coarse step, the fine level is remade as the front moves).
The FMA build agrees to 1e-12."""
import os
import tempfile

import numpy as np
import pytest

from helpers import rel_linf
from oracle.pyoracle import AmrOracle
from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr
from zamr.run import read_levels

pytestmark = pytest.mark.gpu

AMR = dict(max_level=1, ref_ratio=2, blocking_factor=2, n_error_buf=1, grid_eff=0.7, threshold=0.02, b_max=1.0e30)


def column(n=64, steps=10, cfl=0.4, mgs=32, order=2, positivity="none", bc=("wall",) * 4, wet_outside=0.0,
           radius2=0.04, **kw):
    c = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=mgs, spatial_order=order,
             implicit_source=False, well_balanced=False, clamp_positivity=False, positivity=positivity, cfl=cfl,
             dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=steps, bc=tuple(bc), params={})
    c.update(kw)
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.4) ** 2 + (Y - 0.55) ** 2 < radius2, 1.0, wet_outside)
    z = np.zeros_like(h)
    return c, np.stack([z, h, z, z])


def coverage(boxes, shape):
    m = np.zeros(shape, dtype=np.int32)
    for (x0, y0, x1, y1) in boxes:
        m[y0:y1 + 1, x0:x1 + 1] += 1
    return m


def pair(c, Q0, amr=AMR, contract=False):
    o = AmrOracle("swe2d", c, amr=dict(amr, fine_max_grid_size=c["max_grid_size"]))
    o.init_solution()
    o.set_state(Q0)
    o.init_levels()
    z = ZoomyAmr(deck_from_case(c, amr=amr), system="swe2d", dim=2, order=c["spatial_order"], contract=contract)
    z.init_data()
    z.set_state(Q0)
    z.init_levels()
    return o, z


def assert_same_hierarchy(o, z, exact=True):
    assert z.finest_level() == o.finest_level()
    n = o.nx
    assert np.array_equal(coverage(z.level_boxes(0), (n, n)), np.ones((n, n)))
    if o.finest_level() == 1:
        assert np.array_equal(coverage(z.level_boxes(1), (2 * n, 2 * n)), coverage(o.boxes(1), (2 * n, 2 * n)))
    for lev in range(o.finest_level() + 1):
        a, b = z.get_state(lev=lev), o.get_state(lev)
        if exact:
            assert np.array_equal(a, b, equal_nan=True), f"level {lev}"
        else:
            assert np.array_equal(np.isnan(a), np.isnan(b))
            m = ~np.isnan(b)
            err = rel_linf(a[m], b[m])
            assert err <= 1e-12, f"level {lev}: {err:.3e}"


def test_init_levels_grids_and_fill_match_oracle():
    c, Q0 = column()
    print(f"[amr] init_levels column 64^2", flush=True)
    o, z = pair(c, Q0)
    assert o.finest_level() == 1 and len(o.boxes(1)) >= 1
    assert_same_hierarchy(o, z)


def test_subcycled_steps_with_regrids_match_oracle_bitwise():
    c, Q0 = column(steps=12)
    print(f"[amr] 12 coarse steps column 64^2 bitwise", flush=True)
    o, z = pair(c, Q0)
    dts = []
    for k in range(12):
        dts.append(o.time_step(sum(dts), k))
    done, t = z.run_steps(12)
    assert done == 12
    hd, ht = z.dt_history()
    assert np.array_equal(hd[:12], np.array(dts))
    assert t == ht[11]
    assert_same_hierarchy(o, z)
    regrids, remakes = z.amr_stats()
    oreg, orem, _ = o.counts()
    assert remakes >= 1 and orem >= 1  # the front moved the fine level
    assert abs(z.total_mass() - o.total_mass()) <= 1e-13 * o.total_mass()


def test_clustering_grows_its_buffers(monkeypatch):
    """The device clustering started at one region, the minimum count
    space and one cluster: every regrid overflows, grows its buffers and
    repeats until they fit; grids and states stay the oracle's bit for bit."""
    monkeypatch.setenv("ZAMR_BR_CAP", "1,1,1")
    c, Q0 = column(steps=6)
    print(f"[amr] 6 coarse steps column 64^2, clustering from tiny buffers", flush=True)
    o, z = pair(c, Q0)
    monkeypatch.delenv("ZAMR_BR_CAP")
    dts = []
    for k in range(6):
        dts.append(o.time_step(sum(dts), k))
    done, _ = z.run_steps(6)
    assert done == 6
    assert_same_hierarchy(o, z)


def test_regrid_of_a_dry_map_clears_the_fine_level():
    """A regrid of a map without tags (every cell dry): the device clustering
    returns no cluster and level 1 is cleared, as in the oracle; a regrid of
    the wet column makes it again (from the retired level's buffers), with
    the oracle's grids and fill."""
    c, Q0 = column()
    print(f"[amr] regrid: dry map clears level 1, wet map remakes it", flush=True)
    o, z = pair(c, Q0)
    assert z.finest_level() == o.finest_level() == 1
    dry = np.zeros_like(Q0)
    o.set_state(dry)
    z.set_state(dry)
    o.regrid(0.0)
    z.regrid(0.0)
    assert z.finest_level() == o.finest_level() == 0
    o.set_state(Q0)
    z.set_state(Q0)
    o.regrid(0.0)
    z.regrid(0.0)
    assert z.finest_level() == 1
    assert_same_hierarchy(o, z)


def test_error_est_reads_the_last_fillpatch_ghosts():
    """ErrorEst after steps: neighbours across box edges read the ghost cells
    the last stage's FillPatch left (many level-0 boxes)."""
    c, Q0 = column(steps=5, mgs=16)
    print(f"[amr] error_est after 5 steps, 16^2 level-0 boxes", flush=True)
    o, z = pair(c, Q0)
    t = 0.0
    for k in range(5):
        t += o.time_step(t, k)
    z.run_steps(5)
    assert np.array_equal(z.error_est(), o.error_est())
    assert_same_hierarchy(o, z)


def test_many_level0_boxes_match_oracle_bitwise():
    """Level 0 as 16^2 boxes.  ErrorEst reads a box's ghost cells as the last
    stage's FillPatch left them (ZoomyAmr.cpp:236-246 reads Q_arr(i+-1, j)
    without a fill), so the grids depend on the level-0 decomposition, as in
    the reference; the oracle makes the same reads."""
    c, Q0 = column(steps=8, mgs=16)
    print(f"[amr] 16^2 level-0 boxes, 8 steps", flush=True)
    o, z = pair(c, Q0)
    t = 0.0
    for k in range(8):
        t += o.time_step(t, k)
    z.run_steps(8)
    assert_same_hierarchy(o, z)


def test_mood_on_both_levels_bitwise():
    c, Q0 = column(steps=10, cfl=2.5, positivity="mood")
    print(f"[amr] MOOD on both levels", flush=True)
    o, z = pair(c, Q0)
    t = 0.0
    for k in range(10):
        t += o.time_step(t, k)
    z.run_steps(10)
    assert_same_hierarchy(o, z)
    assert z.mood_stats(1)[1] == o.counts()[2]
    assert z.mood_stats(1)[1] > 0


def test_throughput_build_within_1e12():
    """FMA build, wet everywhere: a hump on a depth-1 lake over a bed sloping
    up in x, with tagging.b_max keeping the upper half coarse, so waves cross
    the coarse-fine boundary every step.  Within 1e-12 of the oracle.  (Over
    a dry bed the desingularised 1/h of the model amplifies ulp-level
    differences by up to 1/wet_dry_eps^2 at the front; dry-front parity is
    held by the bitwise build above.)"""
    n = 64
    c, _ = column(n=n, steps=10)
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    b = 0.4 * X
    h = 1.0 + 0.2 * np.exp(-((X - 0.3) ** 2 + (Y - 0.5) ** 2) / 0.01)
    Q0 = np.stack([b, h, 0 * h, 0 * h])
    amr = dict(AMR, b_max=0.2)
    print(f"[amr] FMA build, wet hump, b_max", flush=True)
    o, z = pair(c, Q0, amr=amr, contract=True)
    assert 0 < coverage(o.boxes(1), (2 * n, 2 * n)).mean() < 1
    t = 0.0
    for k in range(10):
        t += o.time_step(t, k)
    z.run_steps(10)
    assert_same_hierarchy(o, z, exact=False)


def test_all_wet_fine_level_is_the_fine_single_level_run():
    """Every cell tagged: level 1 covers the domain and is a single-level run
    at twice the resolution (two CFL steps per coarse step), bit for bit, and
    the composite mass inside walls is conserved to round-off."""
    c, Q0 = column(n=32, steps=6, wet_outside=0.5)
    print(f"[amr] all-wet vs fine single level", flush=True)
    _, z = pair(c, Q0)
    assert z.level_boxes(1) and coverage(z.level_boxes(1), (64, 64)).min() == 1
    Q1 = z.get_state(lev=1)
    m0 = z.total_mass()
    z.run_steps(6)
    f = ZoomyAmr(deck_from_case(dict(c, nx=64, ny=64, max_step=12)), system="swe2d", dim=2, order=2, contract=False)
    f.init_data()
    f.set_state(Q1)
    f.run_steps(12)
    assert np.array_equal(z.get_state(lev=1), f.get_state())
    assert abs(z.total_mass() - m0) <= 1e-13 * m0


def test_evolve_lines_and_multilevel_plotfile(capfd):
    c, Q0 = column(steps=4)
    o, z = pair(c, Q0)
    print(f"[amr] evolve + plotfile", flush=True)
    capfd.readouterr()
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            z.evolve()
            plts = sorted(p for p in os.listdir(d) if p.startswith("plt_"))
            levels, H = read_levels(os.path.join(d, plts[-1]))
        finally:
            os.chdir(cwd)
    out = capfd.readouterr().out
    steps = [l for l in out.splitlines() if l.startswith("Step ")]
    assert len(steps) == 4 and all(" levels: 2 " in l for l in steps)
    assert H["finest_level"] == 1 and H["refs"] == [2] and len(levels) == 2
    assert levels[1]["ref"] == 2
    fine = z.get_state(lev=1)
    for (x0, y0, x1, y1) in levels[1]["boxes"]:
        ilo, jlo = levels[1]["ext"][:2]
        got = levels[1]["arr"][:4, y0 - jlo:y1 - jlo + 1, x0 - ilo:x1 - ilo + 1]
        assert np.array_equal(got, fine[:, y0:y1 + 1, x0:x1 + 1])
