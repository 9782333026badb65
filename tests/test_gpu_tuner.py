"""The measured launch-shape tuner (zamr_autotune): rows per work item per box
size class and the grid of each stage kernel are picked by timing, and no
result may depend on the pick.  Runs after tuning are compared bit for bit
with untuned runs of the same deck (two-buffer order 2, three-buffer MOOD,
order 1), in the FMA-off and the throughput build."""
import numpy as np
import pytest

from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZamrError, ZoomyAmr

pytestmark = pytest.mark.gpu


def bumps(n, order=2, steps=12):
    case, Q0 = cases.swe_bumps(n=n, seed=3, steps=steps)
    return dict(case, spatial_order=order), Q0


def run(case, Q0, contract, tune, order=2, shape=None):
    z = ZoomyAmr(deck_from_case(case), system="swe2d", dim=2, order=order, contract=contract)
    z.init_data()
    z.set_state(Q0)
    info = None
    if tune:
        before = z.get_state()
        info = z.autotune(reps=2)
        assert np.array_equal(z.get_state(), before), "the tuner changed the state"
    if shape is not None:
        z.set_launch_shape(*shape)
    done, t = z.run_steps(case["max_step"])
    assert done == case["max_step"]
    return z.get_state(), t, info


@pytest.mark.parametrize("contract", [True, False])
@pytest.mark.parametrize("order", [1, 2])
def test_tuned_single_box_equals_untuned(contract, order):
    case, Q0 = bumps(384, order=order, steps=20)
    ref, t_ref, _ = run(case, Q0, contract, False, order=order)
    got, t_got, info = run(case, Q0, contract, True, order=order)
    rows, grid, ms = info
    assert len(rows) == 8 and len(grid) == 2 and ms[0] > 0 and ms[1] > 0
    assert np.array_equal(got, ref) and t_got == t_ref


@pytest.mark.parametrize("contract", [True, False])
def test_set_launch_shapes_equal_untuned(contract):
    """Shapes set by hand, over the 4-box decomposition: shallow items with a
    persistent grid of 64 blocks, deep items with 8 blocks."""
    case, Q0 = bumps(256, steps=16)
    case = dict(case, max_grid_size=128)
    ref, _, _ = run(case, Q0, contract, False)
    for rows, grid in (([4] * 8, [64, 64]), ([64] * 8, [8, 16]), ([0, 0, 8, 0, 0, 0, 0, 0], [0, 24])):
        got, _, _ = run(case, Q0, contract, False, shape=(rows, grid))
        assert np.array_equal(got, ref), (rows, grid)


def test_tuned_mood_three_buffers_equals_untuned():
    c = dict(nx=96, ny=96, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=48, spatial_order=2,
             implicit_source=False, well_balanced=False, clamp_positivity=False, positivity="mood", cfl=3.0,
             dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=24, bc=("", "", "", ""), params={})
    x, y = cases.cell_centers(96, 96)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.04, 1.0, 0.0)
    z = np.zeros_like(h)
    Q0 = np.stack([z, h, z, z])
    ref, t_ref, _ = run(c, Q0, False, False)
    got, t_got, _ = run(c, Q0, False, True)
    assert np.array_equal(got, ref) and t_got == t_ref


def test_tuner_rejects_amr_and_bad_shapes():
    case, Q0 = bumps(64, steps=2)
    z = ZoomyAmr(deck_from_case(case, amr=dict(max_level=1, threshold=0.02)), system="swe2d", dim=2, order=2)
    z.init_data()
    with pytest.raises(ZamrError):
        z.autotune()
    y = ZoomyAmr(deck_from_case(case), system="swe2d", dim=2, order=2)
    y.init_data()
    with pytest.raises(ZamrError):
        y.set_launch_shape([-1] * 8, [0, 0])
    with pytest.raises(ValueError):
        y.set_launch_shape([4] * 7, [0, 0])
