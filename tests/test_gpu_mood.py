"""A-posteriori MOOD positivity (solver.positivity = mood, ZoomyAmr.cpp:699-720):
an order-2 step whose candidate has a troubled cell (h < 0, NaN or |h| >= 1e300)
is redone at order 1 from Q^n.  Checked against the oracle's Advance on a
collapsing water column over a dry bed at Courant number 0.5, where about
half of the steps are redone."""
import numpy as np
import pytest

from helpers import rel_linf
from oracle.pyoracle import Oracle
from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZoomyAmr, from_boxes

pytestmark = pytest.mark.gpu


def column(n=64, cfl=3.0, steps=40, mgs=64, positivity="mood", bc=("", "", "", "")):
    c = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=mgs, spatial_order=2,
             implicit_source=False, well_balanced=False, clamp_positivity=False, positivity=positivity, cfl=cfl,
             dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=steps, bc=tuple(bc), params={})
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.04, 1.0, 0.0)
    z = np.zeros_like(h)
    return c, np.stack([z, h, z, z])


def oracle_run(c, Q0):
    o = Oracle("swe2d", c)
    o.init_solution()
    o.set_state(Q0)
    n, t, _, _ = o.evolve()
    return n, t, o.mood_redos(), o.get_state()


def gpu_run(c, Q0, contract):
    z = ZoomyAmr(deck_from_case(c), system="swe2d", dim=2, order=2, contract=contract)
    z.init_data()
    z.set_state(Q0)
    done, t = z.run_steps(c["max_step"])
    return z, done, t


def test_mood_redo_matches_oracle_bitwise():
    c, Q0 = column()
    n, t, redos, Qo = oracle_run(c, Q0)
    assert redos > 5
    z, done, tg = gpu_run(c, Q0, contract=False)
    assert done == n and tg == t
    assert z.mood_stats()[1] == redos
    assert np.array_equal(z.get_state(), Qo)


def test_mood_multibox_throughput_build():
    """Four boxes, reflective walls, FMA build: same number of redone steps,
    state within 1e-12, mass conserved to round-off (a redo is a whole order-1
    step, so every face flux stays shared by its two cells)."""
    c, Q0 = column(mgs=32, steps=30, bc=("wall",) * 4)
    n, t, redos, Qo = oracle_run(c, Q0)
    z, done, tg = gpu_run(c, Q0, contract=True)
    assert done == n
    assert z.mood_stats()[1] == redos > 0
    Q = z.get_state()
    assert rel_linf(Q, Qo) <= 1e-12
    assert abs(Q[1].sum() - Q0[1].sum()) <= 1e-12 * Q0[1].sum()


def test_mood_off_differs_and_goes_negative():
    """Without MOOD (and without the clamp) the same run leaves negative depths."""
    c, Q0 = column(positivity="none", steps=20)
    z, done, _ = gpu_run(c, Q0, contract=False)
    assert z.get_state()[1].min() < 0.0


def test_mood_step_lines(capfd):
    """Evolve prints the reference's MOOD line before each redone step's Step line."""
    c, Q0 = column(steps=12)
    n, t, redos, Qo = oracle_run(c, Q0)
    z = ZoomyAmr(deck_from_case(dict(c, time_end=1e30)) + "output.plot_dt_interval = 1e30\n", system="swe2d", dim=2,
                 order=2, contract=False)
    z.init_data()
    z.set_state(Q0)
    import os, tempfile
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            z.evolve()
        finally:
            os.chdir(cwd)
    out = capfd.readouterr().out
    lines = [l for l in out.splitlines() if "[MOOD] lev 0 troubled cells:" in l]
    assert len(lines) == redos and all(l.endswith("-> whole-step order-1 redo") for l in lines)
    assert np.array_equal(z.get_state(), Qo)


def test_mood_per_group_in_box_lists():
    """Box lists: each time-step group decides for itself; a troubled member
    redoes its step, a quiet one keeps its order-2 candidate."""
    c, Q0 = column(n=64, steps=24)
    quiet = np.stack([np.zeros((64, 64)), np.full((64, 64), 1.0), np.zeros((64, 64)), np.zeros((64, 64))])
    boxes = [dict(lo=(0, 0), hi=(63, 63), dx=(1 / 64, 1 / 64), prob_lo=(0.0, 0.0), group=0),
             dict(lo=(0, 0), hi=(63, 63), dx=(1 / 64, 1 / 64), prob_lo=(0.0, 0.0), group=1)]
    keys = {k: c[k] for k in ("spatial_order", "implicit_source", "well_balanced", "clamp_positivity", "cfl",
                              "dtmin", "dtmax", "time_end", "max_step")}
    keys["positivity"] = "mood"
    z = from_boxes(boxes, [Q0, quiet], keys, contract=False)
    z.run_steps(24)
    o = Oracle("swe2d", dict(c, max_grid_size=4096))
    o.init_solution()
    o.set_state(Q0)
    o.evolve()
    assert z.mood_stats(0)[1] == o.mood_redos() > 0
    assert z.mood_stats(1)[1] == 0
    assert np.array_equal(z.get_state(lev=0), o.get_state())
    assert np.array_equal(z.get_state(lev=1), quiet)


def test_mood_run_steps_twice_matches_oracle():
    """Two run_steps calls whose first ends with eager steps (20 = one captured
    chunk of 15 + 5 eager): the buffer rotation then starts the second call
    from another buffer than the captured chunk did, and the chunk is captured
    again from there."""
    c, Q0 = column(steps=35)
    n, t, redos, Qo = oracle_run(c, Q0)
    assert n == 35 and redos > 0
    z = ZoomyAmr(deck_from_case(c), system="swe2d", dim=2, order=2, contract=False)
    z.init_data()
    z.set_state(Q0)
    assert z.run_steps(20)[0] == 20
    done, tg = z.run_steps(15)
    assert done == 15 and tg == t
    assert z.mood_stats()[1] == redos
    assert np.array_equal(z.get_state(), Qo)


def test_stage_between_graph_runs_matches_oracle():
    """run_steps (a captured chunk of 16), a single stage (which moves the
    current buffer), run_steps again: the second run starts from the stage's
    output, as the oracle's sequence of the same calls does."""
    c, Q0 = column(steps=-1, positivity="none")
    c = dict(c, cfl=0.45, clamp_positivity=True)
    dt = 1.0e-3
    o = Oracle("swe2d", c)
    o.init_solution()
    o.set_state(Q0)
    o.evolve(step_limit=16)
    o.stage(2, 0.0, dt)
    o.evolve(step_limit=16)
    z = ZoomyAmr(deck_from_case(c), system="swe2d", dim=2, order=2, contract=False)
    z.init_data()
    z.set_state(Q0)
    assert z.run_steps(16)[0] == 16
    z.stage(2, 0.0, dt)
    assert z.run_steps(16)[0] == 16
    assert np.array_equal(z.get_state(), o.get_state())


def test_mood_box_list_groups_finish_apart():
    """MOOD box list with a finite time_end: the two members take different
    dt and finish on different steps; a finished member's state is carried
    through the buffer rotation, so each equals its own oracle run."""
    c, Q0 = column(n=64, steps=10 ** 6)
    c = dict(c, time_end=0.02)
    quiet = np.stack([np.zeros((64, 64)), np.full((64, 64), 1.0), np.zeros((64, 64)), np.zeros((64, 64))])
    quiet[1, :, :32] = 1.5
    boxes = [dict(lo=(0, 0), hi=(63, 63), dx=(1 / 64, 1 / 64), prob_lo=(0.0, 0.0), group=0),
             dict(lo=(0, 0), hi=(63, 63), dx=(1 / 64, 1 / 64), prob_lo=(0.0, 0.0), group=1)]
    keys = {k: c[k] for k in ("spatial_order", "implicit_source", "well_balanced", "clamp_positivity", "cfl",
                              "dtmin", "dtmax", "time_end", "max_step")}
    keys["positivity"] = "mood"
    z = from_boxes(boxes, [Q0, quiet], keys, contract=False)
    z.run_steps(10 ** 5)  # returns once no group is active
    steps = []
    for k, Qk in enumerate((Q0, quiet)):
        o = Oracle("swe2d", dict(c, max_grid_size=4096))
        o.init_solution()
        o.set_state(Qk)
        n, t, _, _ = o.evolve()
        steps.append(n)
        assert z.get_time(k)[1] == n and z.get_time(k)[0] == t
        assert np.array_equal(z.get_state(lev=k), o.get_state()), f"member {k}"
    assert steps[0] != steps[1]
