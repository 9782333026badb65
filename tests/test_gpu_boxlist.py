"""Batched box lists: many independent boxes (AMR patches of three levels and
ensemble members) in one launch per RK stage, with per-group dt on the device.
Checked against the per-box launch path (bitwise) and against the oracle run
box by box."""
import numpy as np
import pytest

from helpers import rel_linf
from oracle.pyoracle import Oracle
from zamr import cases
from zamr.lib import from_boxes

pytestmark = pytest.mark.gpu


def small_list(seed=4, steps=8):
    boxes, states, keys = cases.box_list(seed=seed, n_amr=14, n_members=3, steps=steps, member_range=(64, 96))
    return boxes, states, keys


def run(ctx):
    total = 0
    while True:
        n, _ = ctx.run_steps(64)
        total += n
        if n < 64:
            return total


def oracle_box(b, state, keys):
    nx, ny = b["hi"][0] - b["lo"][0] + 1, b["hi"][1] - b["lo"][1] + 1
    c = dict(nx=nx, ny=ny, prob_lo=b["prob_lo"], prob_hi=(b["prob_lo"][0] + nx * b["dx"][0],
                                                          b["prob_lo"][1] + ny * b["dx"][1]),
             max_grid_size=4096, bc=("", "", "", ""), **{k: v for k, v in keys.items() if k != "params"})
    o = Oracle("swe2d", c)
    o.init_solution()
    o.set_state(state)
    return o


def test_batched_equals_per_box_launches():
    boxes, states, keys = small_list()
    a = from_boxes(boxes, states, keys)
    run(a)
    b = from_boxes(boxes, states, keys)
    b.set_tuning(per_box_launch=1)
    run(b)
    for k in range(len(boxes)):
        assert np.array_equal(a.get_state(lev=k), b.get_state(lev=k))


def test_each_box_its_own_group_matches_oracle():
    boxes, states, keys = small_list(seed=6, steps=5)
    own = [dict(b, group=k) for k, b in enumerate(boxes)]
    zb = from_boxes(own, states, keys, contract=False)
    z = from_boxes(own, states, keys)
    run(zb)
    run(z)
    for k, b in enumerate(own):
        o = oracle_box(b, states[k], keys)
        n, t, _, _ = o.evolve()
        ref = o.get_state()
        assert np.array_equal(zb.get_state(lev=k), ref), f"box {k}"
        assert rel_linf(z.get_state(lev=k), ref) <= 1e-11, f"box {k}"
        assert zb.get_time(k)[1] == n and zb.get_time(k)[0] == t


def test_shared_group_dt_is_the_segmented_max():
    """Boxes of one AMR level share dt: the first step's dt is the smallest of
    the per-box dts (the max of the per-box CFL speeds)."""
    boxes, states, keys = small_list(seed=8, steps=1)
    z = from_boxes(boxes, states, keys, contract=False)
    run(z)
    for g in range(3):
        members = [k for k, b in enumerate(boxes) if b["group"] == g]
        if not members:
            continue
        dts = [oracle_box(boxes[k], states[k], keys).compute_dt() for k in members]
        assert z.get_time(g)[2] == min(dts)
    for m in range(3):
        g = 3 + m
        k = [i for i, b in enumerate(boxes) if b["group"] == g][0]
        assert z.get_time(g)[2] == oracle_box(boxes[k], states[k], keys).compute_dt()


def test_full_box_list_launch_modes_bitwise():
    """The whole 664-box list (600 AMR patches, 64 ensemble members, 26.9M
    cells) for 16 steps, one captured chunk: one launch per box per pass,
    one batched launch per pass eager, and the same in a hipGraph all give the
    same bits in every box."""
    boxes, states, keys = cases.box_list(seed=0, steps=16)
    assert len(boxes) == 664
    per_box = from_boxes(boxes, states, keys)
    per_box.set_tuning(per_box_launch=1)
    assert per_box.run_steps(16)[0] == 16
    ref = [per_box.get_state(lev=k) for k in range(len(boxes))]
    del per_box
    for graph in (False, True):
        z = from_boxes(boxes, states, keys)
        z.set_graph(graph)
        assert z.run_steps(16)[0] == 16
        for k in range(len(boxes)):
            assert np.array_equal(z.get_state(lev=k), ref[k]), f"box {k}, graph {graph}"
        del z


@pytest.mark.parametrize("contract", [True, False])
def test_tuned_launch_shape_changes_no_result(contract):
    """The measured tuner picks rows per item per size class and a grid; the
    result of a run is the same bits whatever it picks, and whatever fixed
    rows per item are set."""
    boxes, states, keys = small_list(seed=9, steps=6)
    base = from_boxes(boxes, states, keys, contract=contract)
    run(base)
    ref = [base.get_state(lev=k) for k in range(len(boxes))]
    del base
    z = from_boxes(boxes, states, keys, contract=contract)
    before = [z.get_state(lev=k) for k in range(len(boxes))]
    rows, grid, ms = z.autotune(reps=3)
    assert len(rows) == 8 and all(r in (0, 4, 8, 16, 32, 64) for r in rows)
    assert len(grid) == 2 and min(grid) >= 0 and ms[0] > 0 and ms[1] > 0
    for k in range(len(boxes)):
        assert np.array_equal(z.get_state(lev=k), before[k]), f"the tuner changed box {k}"
    run(z)
    for k in range(len(boxes)):
        assert np.array_equal(z.get_state(lev=k), ref[k]), f"tuned, box {k}"
    for fixed in (4, 64):
        y = from_boxes(boxes, states, keys, contract=contract)
        y.set_tuning(rows=fixed)
        run(y)
        for k in range(len(boxes)):
            assert np.array_equal(y.get_state(lev=k), ref[k]), f"rows {fixed}, box {k}"
        del y
